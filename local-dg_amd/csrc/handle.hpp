// handle.hpp -- the solver handle shared by the translation units of
// libldg.so (ldg.hip: solver; closure.hip: learned closure).  Not part of
// the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "basis.hpp"
#include "ldg.h"
#include "ldg_kernels.hpp"

struct ldg_closure;

namespace ldg_host {

using namespace ldg;

// Candidate tilings (elements per workgroup) per kernel and order.
constexpr int kMaxCand = 2;
// steps of kinetic-energy history held on the device (per instance); a call
// recording more steps copies the history out in chunks of this many
constexpr int kKeHistSteps = 1024;

struct Inst {
  int p, nx, ny, nz;
  double Re, M0, V0, phase[3];
  double dx[3];
  long K;
  int group, slot;        // group and position inside it
};

// Per (group, candidate E): instance table and block -> instance map.
struct Tiling {
  int E = 1, nblocks = 0;
  InstDesc *d_inst = nullptr;
  int *d_blk = nullptr;   // nullptr when the group has one instance
  std::vector<InstDesc> h_inst;
};

// Learned-closure state of one group (inference buffers, device).
struct GroupClosure {
  const ldg_closure *net = nullptr;
  double bound = 0.0;                 // gate: correction RMS in units of the training spread
  int dissipative = 0;                // energy gate: +1 off when a correction raises the element's KE, -1 when it lowers it
  bool per_step = false;              // correction evaluated in stage 1 of each RK4 step and applied in all four
  float *X = nullptr, *H1 = nullptr, *H2 = nullptr, *Y = nullptr;
  double *corr = nullptr;             // per element, state layout
  unsigned char *gate = nullptr;      // per element: 1 = correction applied
  long *gated_count = nullptr;        // elements switched off in the last stage
};

struct Group {
  int n = 0;              // nodes per direction (p + 1)
  std::vector<int> members;   // handle-level instance indices
  long K = 0;
  size_t nU = 0;
  double *u = nullptr, *acc = nullptr, *k = nullptr;
  void *ua = nullptr, *ub = nullptr;  // stage inputs: FP64 state, or FP32 fluctuations in the mixed path
  void *q = nullptr;
  void *ff = nullptr;     // face fluxes written by the gradient kernel, lifted by the rhs kernel
  double *ke_part = nullptr, *diag_part = nullptr, *d_scale = nullptr;
  void *ops = nullptr;
  void *sops = nullptr;   // the rhs kernel's operator images, one per member
  Tiling rhs[kMaxCand], grad[kMaxCand];
  int crhs = 0, cgrad = 0;   // chosen candidates
  Basis1D basis;
  GroupClosure cl;
};

}  // namespace ldg_host

struct ldg_handle {
  ldg_config cfg{};
  bool mixed = false;
  std::vector<ldg_host::Inst> inst;
  std::vector<ldg_host::Group> groups;
  hipStream_t stream = nullptr;     // the handle's stream: every call is ordered on it
  // ensembles of several orders: one stream per group (default priority),
  // forked from and joined back to `stream` in every RK4 step, so that the
  // groups' launches overlap (graph capture records them as parallel branches)
  std::vector<hipStream_t> gstreams;
  hipEvent_t ev_fork = nullptr;
  std::vector<hipEvent_t> ev_join;
  double *ke_hist = nullptr, *scal = nullptr;
  int *d_step = nullptr;
  size_t ke_hist_cap = 0;          // kKeHistSteps x instances, allocated at create
  std::vector<double> dt_cur;       // dt currently written into the instance tables
  bool use_graph = false;
  struct GraphSlot {
    hipGraphExec_t exec = nullptr;
    bool ke = false;
    bool first = false;            // first step reads the FP64 state (first replay of a call)
    const double *u = nullptr;     // groups[0].u when captured
  } graphs[4];
  int graph_next = 0;              // slot replaced next when all are taken
  std::string err;
};

namespace ldg_host {

// Records the message on the handle (or as the last create error) and
// returns code.
int fail(ldg_handle *h, int code, const std::string &msg);
int fail(ldg_handle *h, int code, const char *msg);   // records without a heap copy when h is null

#define HIPCHECK(h, call)                                                                         \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      return ::ldg_host::fail((h), LDG_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Input of a stage's kernels: the conserved state (FP64), or in the mixed
// path the FP32 fluctuation about the background that the previous stage's
// epilogue wrote.
struct StageIn {
  const void *p;
  bool fluct;
};
inline StageIn state_in(const double *u) { return StageIn{u, false}; }

StageArgs empty_stage();
// wq: write the nodal gradient Q (1), not (0), or only when a closure is
// attached (-1)
int launch_grad(ldg_handle *h, Group &g, StageIn U, int c = -1, int wq = -1);
int launch_rhs(ldg_handle *h, Group &g, StageIn U, const StageArgs &st, int c = -1);
void drop_graph(ldg_handle *h);

// Offset (in doubles) and size of instance i inside its group's arrays.
size_t inst_offset(const ldg_handle *h, int i);
size_t inst_values(const ldg_handle *h, int i);

// closure.hip: before a stage's rhs launch, evaluates the group's closure on
// the stage input U (and the gradient just computed into g.q) and points st
// at the correction; no-op without a closure.
int closure_stage(ldg_handle *h, Group &g, StageIn U, StageArgs &st, int stage = 0);
void closure_release(Group &g);

}  // namespace ldg_host
