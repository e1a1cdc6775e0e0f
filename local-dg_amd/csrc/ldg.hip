// ldg.hip -- C-ABI implementation of libldg.so (see include/ldg.h).
//
// A handle owns one or more solver instances.  Instances of the same order
// form a group whose arrays are stored back to back, so one launch of each
// kernel advances every instance of the group (a plain solver is a group of
// one instance).  Per group, on the handle's HIP stream:
//   u      u^n (the reference's Solver::u)
//   acc    u^n + dt sum_s b_s k_s, becomes u^{n+1} after stage 4 (pointer swap)
//   ua/ub  stage inputs U_s (double-buffered: stage s+1 is written while
//          neighbours of stage s are still being read)
//   q      LDG auxiliary gradient (15 fields per element; FP64, or FP32 in
//          the mixed mode)
// Steps can be captured into a HIP graph: two steps per graph, so that the
// u/acc swap returns the arrays to their starting roles.  An odd number of
// eager steps leaves the roles swapped, so graphs are kept per arrangement of
// the arrays and per kind of first step (FP64 state or FP32 carried input).
#include "ldg.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "handle.hpp"

using namespace ldg;

namespace ldg_host {

// message of the last failed call without a handle (create, allocation-free
// right-hand side); a fixed buffer, so that recording it allocates nothing
char g_create_error[512] = "";

void set_global_error(const char *msg) { std::snprintf(g_create_error, sizeof(g_create_error), "%s", msg); }

int fail(ldg_handle *h, int code, const std::string &msg) {
  if (h) h->err = msg; else set_global_error(msg.c_str());
  return code;
}

int fail(ldg_handle *h, int code, const char *msg) {
  if (h) h->err = msg; else set_global_error(msg);
  return code;
}

// Candidate elements per workgroup of each kernel, per order (the
// autotuner picks among them; index 0 is the default).
template <int N> struct Cand;
template <> struct Cand<2> { static constexpr int rhs[kMaxCand] = {16, 8}, grad[kMaxCand] = {16, 8}; };
template <> struct Cand<3> { static constexpr int rhs[kMaxCand] = {8, 4}, grad[kMaxCand] = {8, 4}; };
template <> struct Cand<4> { static constexpr int rhs[kMaxCand] = {4, 2}, grad[kMaxCand] = {4, 2}; };
template <> struct Cand<5> { static constexpr int rhs[kMaxCand] = {2, 1}, grad[kMaxCand] = {2, 4}; };
template <> struct Cand<6> { static constexpr int rhs[kMaxCand] = {1, 1}, grad[kMaxCand] = {1, 2}; };

int cand_e(int n, bool rhs, int c) {
  switch (n) {
    case 2: return rhs ? Cand<2>::rhs[c] : Cand<2>::grad[c];
    case 3: return rhs ? Cand<3>::rhs[c] : Cand<3>::grad[c];
    case 4: return rhs ? Cand<4>::rhs[c] : Cand<4>::grad[c];
    case 5: return rhs ? Cand<5>::rhs[c] : Cand<5>::grad[c];
    case 6: return rhs ? Cand<6>::rhs[c] : Cand<6>::grad[c];
  }
  return 1;
}

// largest change made by symmetrize (A is replaced by its symmetric part)
template <int N>
double symmetrize_checked(double (&A)[N][N], int S) {
  double B[N][N], dev = 0.0;
  for (int a = 0; a < N; ++a)
    for (int b = 0; b < N; ++b) B[a][b] = A[a][b];
  symmetrize<N>(A, S);
  for (int a = 0; a < N; ++a)
    for (int b = 0; b < N; ++b) dev = std::max(dev, std::fabs(A[a][b] - B[a][b]));
  return dev;
}

template <int N>
Ops<N, double> make_ops(const Basis1D &B, double &asym) {
  Ops<N, double> o;
  for (int a = 0; a < N; ++a) {
    for (int b = 0; b < N; ++b) {
      o.I[a][b] = B.at(B.I, a, b);
      o.D[a][b] = B.at(B.D, a, b);
      o.Dn[a][b] = B.at(B.Dn, a, b);
      o.Minv[a][b] = B.at(B.M1inv, a, b);
    }
    o.L0[a] = B.at(B.M1inv, a, 0);
    o.w[a] = B.w[a];
  }
  for (int q = 0; q < N; ++q) {
    o.IL0[q] = 0.0;
    for (int j = 0; j < N; ++j) o.IL0[q] += o.I[q][j] * o.L0[j];
  }
  // sweep operators in even/odd split form, from their (anti-)centro-symmetric
  // parts (the node sets are symmetric; this removes construction round-off)
  double I[N][N], It[N][N], Dt[N][N], Mi[N][N], D[N][N];
  for (int a = 0; a < N; ++a)
    for (int b = 0; b < N; ++b) {
      I[a][b] = o.I[a][b];
      It[a][b] = o.I[b][a];
      Dt[a][b] = o.D[b][a];
      Mi[a][b] = o.Minv[a][b];
      D[a][b] = o.D[a][b];
    }
  asym = std::max(std::max(symmetrize_checked<N>(I, 1), symmetrize_checked<N>(It, 1)),
                  std::max(symmetrize_checked<N>(Dt, -1), symmetrize_checked<N>(Mi, 1)));
  asym = std::max(asym, symmetrize_checked<N>(D, -1));
  make_symmat<N, double>(I, o.sI);
  make_symmat<N, double>(It, o.sIt);
  make_symmat<N, double>(Dt, o.sDt);
  make_symmat<N, double>(Mi, o.sMi);
  make_symmat<N, double>(D, o.sD);
  return o;
}

PhysArgs phys_args(const ldg_config &c, const Inst &in) {
  PhysArgs p;
  p.gamma = c.gamma;
  p.gm1 = c.gamma - 1.0;
  p.mu = 1.0 / in.Re;
  p.kappa = p.mu * c.gamma / c.Pr;
  const double c0 = in.V0 / in.M0;
  const double p0 = c.rho0 * c0 * c0 / c.gamma;   // solver.cpp:40-41
  p.rho0 = c.rho0;
  p.E0 = p0 / p.gm1;
  p.H0 = p.E0 + p0;
  return p;
}

double ke_scale(const ldg_config &c, const Inst &in) {
  const double J = (in.dx[0] / 2) * (in.dx[1] / 2) * (in.dx[2] / 2);
  return J / (c.rho0 * std::pow(2 * M_PI * c.L, 3.0));
}

StageArgs empty_stage() {
  StageArgs s;
  std::memset(&s, 0, sizeof(s));
  return s;
}

// ---------------------------------------------------------------- launches

template <int N, int E>
void launch_grad_e(ldg_handle *h, Group &g, const Tiling &t, StageIn U, bool wq) {
  LaunchArgs la{t.d_inst, t.d_blk, t.nblocks, (t.nblocks + 7) / 8};
  const Ops<N, double> *op = (const Ops<N, double> *)g.ops;
  const dim3 grid(8 * la.per_xcd), block(NT);
  if (h->mixed && U.fluct) {
    if (wq)
      hipLaunchKernelGGL((ldg_grad_kernel<N, E, float, float, true>), grid, block, 0, h->stream, (const float *)U.p,
                         (float *)g.q, (float *)g.ff, la, op);
    else
      hipLaunchKernelGGL((ldg_grad_kernel<N, E, float, float, false>), grid, block, 0, h->stream, (const float *)U.p,
                         (float *)g.q, (float *)g.ff, la, op);
  } else if (h->mixed) {
    if (wq)
      hipLaunchKernelGGL((ldg_grad_kernel<N, E, float, double, true>), grid, block, 0, h->stream, (const double *)U.p,
                         (float *)g.q, (float *)g.ff, la, op);
    else
      hipLaunchKernelGGL((ldg_grad_kernel<N, E, float, double, false>), grid, block, 0, h->stream,
                         (const double *)U.p, (float *)g.q, (float *)g.ff, la, op);
  } else {
    if (wq)
      hipLaunchKernelGGL((ldg_grad_kernel<N, E, double, double, true>), grid, block, 0, h->stream,
                         (const double *)U.p, (double *)g.q, (double *)g.ff, la, op);
    else
      hipLaunchKernelGGL((ldg_grad_kernel<N, E, double, double, false>), grid, block, 0, h->stream,
                         (const double *)U.p, (double *)g.q, (double *)g.ff, la, op);
  }
}

template <int N, int E>
void launch_rhs_e(ldg_handle *h, Group &g, const Tiling &t, StageIn U, const StageArgs &st) {
  LaunchArgs la{t.d_inst, t.d_blk, t.nblocks, (t.nblocks + 7) / 8};
  const unsigned char *img = (const unsigned char *)g.sops;
  const dim3 grid(8 * la.per_xcd), block(NT);
  if (h->mixed && U.fluct)
    hipLaunchKernelGGL((ldg_rhs_kernel<N, E, float, float>), grid, block, 0, h->stream, (const float *)U.p,
                       (const float *)g.ff, la, img, st);
  else if (h->mixed)
    hipLaunchKernelGGL((ldg_rhs_kernel<N, E, float, double>), grid, block, 0, h->stream, (const double *)U.p,
                       (const float *)g.ff, la, img, st);
  else
    hipLaunchKernelGGL((ldg_rhs_kernel<N, E, double, double>), grid, block, 0, h->stream, (const double *)U.p,
                       (const double *)g.ff, la, img, st);
}

template <int N>
void launch_grad_n(ldg_handle *h, Group &g, int c, StageIn U, bool wq) {
  if (c == 0) launch_grad_e<N, Cand<N>::grad[0]>(h, g, g.grad[0], U, wq);
  else launch_grad_e<N, Cand<N>::grad[1]>(h, g, g.grad[1], U, wq);
}

template <int N>
void launch_rhs_n(ldg_handle *h, Group &g, int c, StageIn U, const StageArgs &st) {
  if (c == 0) launch_rhs_e<N, Cand<N>::rhs[0]>(h, g, g.rhs[0], U, st);
  else launch_rhs_e<N, Cand<N>::rhs[1]>(h, g, g.rhs[1], U, st);
}

int launch_grad(ldg_handle *h, Group &g, StageIn U, int c, int wq) {
  if (U.fluct && !h->mixed) return fail(h, LDG_EINVAL, "FP32 stage input outside the mixed path");
  if (c < 0) c = g.cgrad;
  // the nodal gradient is written when asked for, and whenever a closure
  // (whose features read it) is attached
  const bool w = wq > 0 || (wq < 0 && g.cl.net != nullptr);
  switch (g.n) {
    case 2: launch_grad_n<2>(h, g, c, U, w); break;
    case 3: launch_grad_n<3>(h, g, c, U, w); break;
    case 4: launch_grad_n<4>(h, g, c, U, w); break;
    case 5: launch_grad_n<5>(h, g, c, U, w); break;
    case 6: launch_grad_n<6>(h, g, c, U, w); break;
    default: return fail(h, LDG_EINVAL, "unsupported order");
  }
  HIPCHECK(h, hipGetLastError());
  return LDG_OK;
}

int launch_rhs(ldg_handle *h, Group &g, StageIn U, const StageArgs &st, int c) {
  if (U.fluct && !h->mixed) return fail(h, LDG_EINVAL, "FP32 stage input outside the mixed path");
  if (c < 0) c = g.crhs;
  switch (g.n) {
    case 2: launch_rhs_n<2>(h, g, c, U, st); break;
    case 3: launch_rhs_n<3>(h, g, c, U, st); break;
    case 4: launch_rhs_n<4>(h, g, c, U, st); break;
    case 5: launch_rhs_n<5>(h, g, c, U, st); break;
    case 6: launch_rhs_n<6>(h, g, c, U, st); break;
    default: return fail(h, LDG_EINVAL, "unsupported order");
  }
  HIPCHECK(h, hipGetLastError());
  return LDG_OK;
}

// ----------------------------------------------------------------- set-up

template <int N>
int upload_ops_t(ldg_handle *h, Group &g) {
  double asym = 0.0;
  const Ops<N, double> o = make_ops<N>(g.basis, asym);
  if (!(asym < 1e-9)) return fail(h, LDG_EINVAL, "1D operators are not centro-symmetric (node sets must be symmetric)");
  HIPCHECK(h, hipMalloc(&g.ops, sizeof(o)));
  HIPCHECK(h, hipMemcpy(g.ops, &o, sizeof(o), hipMemcpyHostToDevice));
  // the rhs kernel's operator images, one per member (slot order)
  std::vector<unsigned char> img((size_t)SOpsImage<N>::BYTES * g.members.size(), 0);
  for (size_t j = 0; j < g.members.size(); ++j) {
    const Inst &in = h->inst[g.members[j]];
    const double a[3] = {2.0 / in.dx[0], 2.0 / in.dx[1], 2.0 / in.dx[2]};
    unsigned char *b = img.data() + j * SOpsImage<N>::BYTES;
    fill_sops<N, float>(*reinterpret_cast<SOps<N, float> *>(b + SOpsImage<N>::template offset<float>()), o, a);
    fill_sops<N, double>(*reinterpret_cast<SOps<N, double> *>(b + SOpsImage<N>::template offset<double>()), o, a);
  }
  HIPCHECK(h, hipMalloc(&g.sops, img.size()));
  HIPCHECK(h, hipMemcpy(g.sops, img.data(), img.size(), hipMemcpyHostToDevice));
  return LDG_OK;
}

int upload_ops(ldg_handle *h, Group &g) {
  switch (g.n) {
    case 2: return upload_ops_t<2>(h, g);
    case 3: return upload_ops_t<3>(h, g);
    case 4: return upload_ops_t<4>(h, g);
    case 5: return upload_ops_t<5>(h, g);
    case 6: return upload_ops_t<6>(h, g);
  }
  return fail(h, LDG_EINVAL, "unsupported order");
}

// Instance table entry of instance `in` for a tiling of E elements per
// workgroup, with its first element and first logical workgroup.
int inst_desc(const ldg_config &cfg, const Inst &in, int E, long elem_off, int block_off, InstDesc &d) {
  std::memset(&d, 0, sizeof(d));
  d.nx = in.nx; d.ny = in.ny; d.nz = in.nz;
  d.nelem = in.K;
  d.elem_off = elem_off;
  // blocks of E elements along x-rows, visited in bricks of about 8 x 8 x 4
  // elements (whole bricks; blocks outside the mesh are padding)
  d.nbx = (in.nx + E - 1) / E;
  d.tb[0] = std::min(std::max(1, 8 / E), d.nbx);
  d.tb[1] = std::min(8, in.ny);
  d.tb[2] = std::min(4, in.nz);
  d.nbrick[0] = (d.nbx + d.tb[0] - 1) / d.tb[0];
  d.nbrick[1] = (in.ny + d.tb[1] - 1) / d.tb[1];
  d.nbrick[2] = (in.nz + d.tb[2] - 1) / d.tb[2];
  const long nb = (long)d.nbrick[0] * d.nbrick[1] * d.nbrick[2] * d.tb[0] * d.tb[1] * d.tb[2];
  if (nb > (1L << 30)) return LDG_EINVAL;
  d.nblocks = (int)nb;
  d.block_off = block_off;
  for (int l = 0; l < 3; ++l) d.ialpha[l] = 2.0 / in.dx[l];
  d.dt = 0.0;
  d.ph = phys_args(cfg, in);
  d.div_bs = make_fastdiv(d.tb[0] * d.tb[1] * d.tb[2]);
  d.div_nb0 = make_fastdiv(d.nbrick[0]);
  d.div_nb1 = make_fastdiv(d.nbrick[1]);
  d.div_tb0 = make_fastdiv(d.tb[0]);
  d.div_tb1 = make_fastdiv(d.tb[1]);
  return LDG_OK;
}

int build_tiling(ldg_handle *h, Group &g, Tiling &t, int E) {
  t.E = E;
  t.h_inst.clear();
  std::vector<int> blk;
  long off = 0;
  int boff = 0;
  for (int idx : g.members) {
    const Inst &in = h->inst[idx];
    InstDesc d;
    if (inst_desc(h->cfg, in, E, off, boff, d) != LDG_OK)
      return fail(h, LDG_EINVAL, "mesh too large for the block numbering");
    for (int b = 0; b < d.nblocks; ++b) blk.push_back(in.slot);
    off += in.K;
    boff += d.nblocks;
    t.h_inst.push_back(d);
  }
  t.nblocks = boff;
  HIPCHECK(h, hipMalloc((void **)&t.d_inst, sizeof(InstDesc) * t.h_inst.size()));
  HIPCHECK(h, hipMemcpy(t.d_inst, t.h_inst.data(), sizeof(InstDesc) * t.h_inst.size(), hipMemcpyHostToDevice));
  if (g.members.size() > 1) {
    HIPCHECK(h, hipMalloc((void **)&t.d_blk, sizeof(int) * blk.size()));
    HIPCHECK(h, hipMemcpy(t.d_blk, blk.data(), sizeof(int) * blk.size(), hipMemcpyHostToDevice));
  }
  return LDG_OK;
}

void free_tiling(Tiling &t) {
  if (t.d_inst) (void)hipFree(t.d_inst);
  if (t.d_blk) (void)hipFree(t.d_blk);
  t.d_inst = nullptr;
  t.d_blk = nullptr;
}

int validate_instance(const ldg_config &c, const ldg_instance &in) {
  if (in.p < 1 || in.p > 5) return fail(nullptr, LDG_EINVAL, "order p must be in 1..5");
  if (in.nx < 3 || in.ny < 3 || in.nz < 3)
    return fail(nullptr, LDG_EINVAL, "need at least 3 elements per direction (mesh.cpp:102-107)");
  if (!(in.Re > 0) || !(in.M0 > 0) || !(in.V0 > 0)) return fail(nullptr, LDG_EINVAL, "Re, M0 and V0 must be positive");
  for (int l = 0; l < 3; ++l)
    if (!(c.xmax[l] > c.xmin[l])) return fail(nullptr, LDG_EINVAL, "empty domain");
  if (!(c.Pr > 0) || !(c.gamma > 1) || !(c.rho0 > 0) || !(c.L > 0) || !(c.R > 0))
    return fail(nullptr, LDG_EINVAL, "physics parameters must be positive (gamma > 1)");
  if (c.precision != LDG_FP64 && c.precision != LDG_MIXED) return fail(nullptr, LDG_EINVAL, "unknown precision mode");
  return LDG_OK;
}

int create_impl(const ldg_config *cfg, const ldg_instance *insts, int n, ldg_handle **out) {
  if (!cfg || !insts || !out || n < 1) return fail(nullptr, LDG_EINVAL, "null argument or empty ensemble");
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    int rc = validate_instance(*cfg, insts[i]);
    if (rc) return rc;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, LDG_EDEVICE, "no HIP device");

  ldg_handle *h = new ldg_handle();
  h->cfg = *cfg;
  h->mixed = cfg->precision == LDG_MIXED;
  auto bad = [&](int code) { set_global_error(h->err.c_str()); ldg_destroy(h); return code; };
  for (int i = 0; i < n; ++i) {
    Inst in;
    in.p = insts[i].p; in.nx = insts[i].nx; in.ny = insts[i].ny; in.nz = insts[i].nz;
    in.Re = insts[i].Re; in.M0 = insts[i].M0; in.V0 = insts[i].V0;
    const int ns[3] = {in.nx, in.ny, in.nz};
    for (int l = 0; l < 3; ++l) {
      in.phase[l] = insts[i].phase[l];
      in.dx[l] = (cfg->xmax[l] - cfg->xmin[l]) / ns[l];
    }
    in.K = (long)in.nx * in.ny * in.nz;
    in.group = -1;
    for (size_t gi = 0; gi < h->groups.size(); ++gi)
      if (h->groups[gi].n == in.p + 1) in.group = (int)gi;
    if (in.group < 0) {
      h->groups.emplace_back();
      h->groups.back().n = in.p + 1;
      in.group = (int)h->groups.size() - 1;
    }
    Group &g = h->groups[in.group];
    in.slot = (int)g.members.size();
    g.members.push_back(i);
    g.K += in.K;
    h->inst.push_back(in);
  }
  h->dt_cur.assign(n, 0.0);
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    h->err = "hipStreamCreate failed";
    return bad(LDG_EDEVICE);
  }
  if (h->groups.size() > 1) {
    h->gstreams.assign(h->groups.size(), nullptr);
    h->ev_join.assign(h->groups.size(), nullptr);
    bool ok = hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) == hipSuccess;
    for (size_t gi = 0; ok && gi < h->groups.size(); ++gi)
      ok = hipStreamCreateWithFlags(&h->gstreams[gi], hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&h->ev_join[gi], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
      h->err = "hipStreamCreate failed (group streams)";
      return bad(LDG_EDEVICE);
    }
  }
  for (Group &g : h->groups) {
    g.nU = (size_t)g.n * g.n * g.n * NSTATE * g.K;
    g.basis = make_basis(g.n - 1);
    const size_t bytes = g.nU * sizeof(double);
    double **arrs[3] = {&g.u, &g.acc, &g.k};
    for (double **p : arrs)
      if (hipMalloc((void **)p, bytes) != hipSuccess) { h->err = "device allocation failed (state)"; return bad(LDG_ENOMEM); }
    const size_t sbytes = g.nU * (h->mixed ? sizeof(float) : sizeof(double));
    if (hipMalloc(&g.ua, sbytes) != hipSuccess || hipMalloc(&g.ub, sbytes) != hipSuccess) {
      h->err = "device allocation failed (stage inputs)";
      return bad(LDG_ENOMEM);
    }
    if (hipMalloc(&g.q, g.nU * 3 * (h->mixed ? sizeof(float) : sizeof(double))) != hipSuccess) {
      h->err = "device allocation failed (gradient)";
      return bad(LDG_ENOMEM);
    }
    // face fluxes: per element 3 directions x 2 sides x 5 states x n^2 face nodes
    const int wb = h->mixed ? (int)sizeof(float) : (int)sizeof(double);
    if (hipMalloc(&g.ff, (size_t)g.K * 6 * ff_chunk(g.n, wb) * wb) != hipSuccess) {
      h->err = "device allocation failed (face fluxes)";
      return bad(LDG_ENOMEM);
    }
    int maxblk = 0;
    for (int c = 0; c < kMaxCand; ++c) {
      if (build_tiling(h, g, g.rhs[c], cand_e(g.n, true, c)) || build_tiling(h, g, g.grad[c], cand_e(g.n, false, c)))
        return bad(LDG_EDEVICE);
      maxblk = std::max(maxblk, g.rhs[c].nblocks);
    }
    std::vector<double> sc;
    for (int idx : g.members) sc.push_back(ke_scale(h->cfg, h->inst[idx]));
    if (hipMalloc((void **)&g.ke_part, sizeof(double) * (maxblk + 1)) != hipSuccess ||
        hipMalloc((void **)&g.diag_part, sizeof(double) * 2 * (maxblk + 1)) != hipSuccess ||
        hipMalloc((void **)&g.d_scale, sizeof(double) * sc.size()) != hipSuccess) {
      h->err = "device allocation failed (partials)";
      return bad(LDG_ENOMEM);
    }
    if (hipMemcpy(g.d_scale, sc.data(), sizeof(double) * sc.size(), hipMemcpyHostToDevice) != hipSuccess ||
        upload_ops(h, g) != LDG_OK)
      return bad(LDG_EDEVICE);
    if (hipMemsetAsync(g.u, 0, bytes, h->stream) != hipSuccess) { h->err = "device initialisation failed"; return bad(LDG_EDEVICE); }
  }
  // the device KE history holds kKeHistSteps steps; longer calls copy it
  // out chunk by chunk, so a step call never allocates
  h->ke_hist_cap = (size_t)kKeHistSteps * n;
  if (hipMalloc((void **)&h->scal, sizeof(double) * 3 * n) != hipSuccess ||
      hipMalloc((void **)&h->d_step, sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&h->ke_hist, sizeof(double) * h->ke_hist_cap) != hipSuccess) {
    h->err = "device allocation failed (scalars)";
    return bad(LDG_ENOMEM);
  }
  if (hipStreamSynchronize(h->stream) != hipSuccess) { h->err = "device initialisation failed"; return bad(LDG_EDEVICE); }
  *out = h;
  return LDG_OK;
}

// ------------------------------------------------------------- time steps

// Write per-instance dt into every instance table (when it changed).
int set_dt(ldg_handle *h, const double *dt) {
  bool changed = false;
  for (size_t i = 0; i < h->inst.size(); ++i) changed |= (h->dt_cur[i] != dt[i]);
  if (!changed) return LDG_OK;
  for (size_t i = 0; i < h->inst.size(); ++i) h->dt_cur[i] = dt[i];
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  for (Group &g : h->groups) {
    Tiling *ts[2 * kMaxCand] = {&g.rhs[0], &g.rhs[1], &g.grad[0], &g.grad[1]};
    for (Tiling *t : ts) {
      for (size_t s = 0; s < g.members.size(); ++s) t->h_inst[s].dt = dt[g.members[s]];
      HIPCHECK(h, hipMemcpy(t->d_inst, t->h_inst.data(), sizeof(InstDesc) * t->h_inst.size(), hipMemcpyHostToDevice));
    }
  }
  return LDG_OK;
}

// One RK4 step of every group (eager launches).  With ke, each group's
// kinetic energies of u^{n+1} go to row *d_step of the KE history.  In the
// mixed path without a closure, a step
// that follows another one in the same call (in32) reads its stage-1 input
// as the FP32 fluctuations the previous step's stage 4 left in ub (out32);
// ub is free then (stage 4 reads ua) and is only rewritten by stage 2.  The
// kernels round an FP64 stage-1 input to the same FP32 values, so results do
// not depend on how steps are split into calls or graph replays.
int rk4_step(ldg_handle *h, bool ke, bool in32, bool out32) {
  static const double b[4] = {1 / 6.0, 1 / 3.0, 1 / 3.0, 1 / 6.0};
  static const double a[3] = {0.5, 0.5, 1.0};
  // with several order groups each runs on its own stream between a fork
  // from and a join back to the handle's stream; the launch helpers use
  // h->stream, which points at the group's stream meanwhile
  const bool fork = !h->gstreams.empty();
  hipStream_t const main = h->stream;
  struct Restore {
    ldg_handle *h;
    hipStream_t s;
    ~Restore() { h->stream = s; }
  } restore{h, main};
  if (fork) HIPCHECK(h, hipEventRecord(h->ev_fork, main));
  int first = 0;
  for (size_t gi = 0; gi < h->groups.size(); ++gi) {
    Group &g = h->groups[gi];
    if (fork) {
      h->stream = h->gstreams[gi];
      HIPCHECK(h, hipStreamWaitEvent(h->stream, h->ev_fork, 0));
    }
    const bool carry = h->mixed && g.cl.net == nullptr;
    StageIn U = carry && in32 ? StageIn{g.ub, true} : state_in(g.u);
    void *bufs[2] = {g.ua, g.ub};
    for (int s = 0; s < 4; ++s) {
      // (a closure evaluated once per step reads the nodal gradient in
      // stage 1 only)
      int rc = launch_grad(h, g, U, -1, g.cl.net && g.cl.per_step && s > 0 ? 0 : -1);
      if (rc) return rc;
      StageArgs st = empty_stage();
      st.u0 = g.u;
      st.acc = g.acc;
      st.acc_init = (s == 0);
      st.cb = b[s];
      if (s < 3) {
        if (h->mixed) st.unext32 = (float *)bufs[s & 1];
        else st.unext = (double *)bufs[s & 1];
        st.ca = a[s];
      }
      if (s == 3 && ke) st.ke_part = g.ke_part;
      if (s == 3 && carry && out32) st.next32 = (float *)g.ub;
      rc = closure_stage(h, g, U, st, s);
      if (rc) return rc;
      rc = launch_rhs(h, g, U, st);
      if (rc) return rc;
      U = StageIn{bufs[s & 1], h->mixed};
    }
    if (ke) {
      const Tiling &t = g.rhs[g.crhs];
      hipLaunchKernelGGL(ldg_finalize_kernel, dim3(g.members.size()), dim3(NT), 0, h->stream,
                         (const double *)g.ke_part, 1, (const InstDesc *)t.d_inst, (int)h->inst.size(),
                         (const double *)g.d_scale, h->ke_hist + first, (const int *)h->d_step);
      HIPCHECK(h, hipGetLastError());
    }
    if (fork) {
      HIPCHECK(h, hipEventRecord(h->ev_join[gi], h->stream));
      h->stream = main;
    }
    std::swap(g.u, g.acc);
    first += (int)g.members.size();
  }
  if (fork)
    for (hipEvent_t ev : h->ev_join) HIPCHECK(h, hipStreamWaitEvent(main, ev, 0));
  if (ke) {
    hipLaunchKernelGGL(ldg_advance_kernel, dim3(1), dim3(1), 0, h->stream, h->d_step);
    HIPCHECK(h, hipGetLastError());
  }
  return LDG_OK;
}

void drop_graph(ldg_handle *h) {
  for (auto &g : h->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g = ldg_handle::GraphSlot();
  }
}

// The graph for the current arrangement of the arrays, built on first use:
// two eager steps recorded on the stream (the u/acc roles are the same after
// two steps, so the graph can be replayed any number of times).  The first
// replay of a call reads the FP64 state (first); later replays read the FP32
// stage-1 input the previous replay left, like eager steps after the first.
int get_graph(ldg_handle *h, bool ke, bool first, hipGraphExec_t *out) {
  const double *u = h->groups[0].u;
  const int NS = (int)(sizeof(h->graphs) / sizeof(h->graphs[0]));
  int slot = -1;
  for (int i = 0; i < NS; ++i) {
    const auto &g = h->graphs[i];
    if (g.exec && g.u == u && g.ke == ke && g.first == first) { *out = g.exec; return LDG_OK; }
  }
  for (int i = 0; i < NS && slot < 0; ++i)
    if (!h->graphs[i].exec) slot = i;
  // replace round-robin, but never the other variant of this arrangement:
  // a call holds both at once
  for (int k = 0; k < NS && slot < 0; ++k) {
    const int i = h->graph_next++ % NS;
    if (!(h->graphs[i].u == u && h->graphs[i].ke == ke)) slot = i;
  }
  auto &gs = h->graphs[slot];
  if (gs.exec) (void)hipGraphExecDestroy(gs.exec);
  gs = ldg_handle::GraphSlot();
  hipGraph_t graph = nullptr;
  HIPCHECK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = rk4_step(h, ke, !first, true);
  if (!rc) rc = rk4_step(h, ke, true, true);
  hipError_t e = hipStreamEndCapture(h->stream, &graph);
  if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
  if (e != hipSuccess) return fail(h, LDG_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(h, LDG_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  // capturing recorded the launches without running them; the u/acc roles
  // were swapped twice, i.e. they are back where they started
  gs.exec = exec;
  gs.ke = ke;
  gs.first = first;
  gs.u = u;
  *out = exec;
  return LDG_OK;
}

// Instance index at each position of the group-by-group numbering used by
// the device-side reductions.
std::vector<int> group_order(const ldg_handle *h) {
  std::vector<int> order;
  for (const Group &g : h->groups) order.insert(order.end(), g.members.begin(), g.members.end());
  return order;
}

int step_impl(ldg_handle *h, const double *dt, int nsteps, double *ke) {
  if (nsteps < 0) return fail(h, LDG_EINVAL, "bad step arguments");
  for (size_t i = 0; i < h->inst.size(); ++i)
    if (!(dt[i] > 0)) return fail(h, LDG_EINVAL, "bad step arguments");
  const size_t ni = h->inst.size();
  int rc = set_dt(h, dt);
  if (rc) return rc;
  // chunks of at most kKeHistSteps steps when the KE history is recorded:
  // the device history row counter restarts and the rows are copied out
  // after each chunk
  const int zero = 0;
  int done = 0;
  while (done < nsteps) {
    const int n = ke ? std::min<int>((int)(h->ke_hist_cap / ni), nsteps - done) : nsteps - done;
    if (ke) HIPCHECK(h, hipMemcpyAsync(h->d_step, &zero, sizeof(int), hipMemcpyHostToDevice, h->stream));
    int k = 0;
    if (h->use_graph && n >= 2) {
      // the first replay of a call reads the FP64 state, later ones the FP32
      // stage-1 input the step before left
      hipGraphExec_t first = nullptr, later = nullptr;
      if (done == 0) rc = get_graph(h, ke != nullptr, true, &first);
      if (!rc && (done > 0 || n >= 4)) rc = get_graph(h, ke != nullptr, false, &later);
      if (rc) return rc;
      for (; k + 2 <= n; k += 2) HIPCHECK(h, hipGraphLaunch(done + k ? later : first, h->stream));
    }
    for (; k < n; ++k) {
      rc = rk4_step(h, ke != nullptr, done + k > 0, done + k + 1 < nsteps);
      if (rc) return rc;
    }
    if (ke)
      HIPCHECK(h, hipMemcpyAsync(ke + (size_t)done * ni, h->ke_hist, sizeof(double) * n * ni, hipMemcpyDeviceToHost,
                                 h->stream));
    done += n;
  }
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  if (ke && ni > 1) {
    // the device history is numbered group by group; reorder each row to
    // instance order
    const std::vector<int> order = group_order(h);
    std::vector<double> row(ni);
    for (int s = 0; s < nsteps; ++s) {
      double *r = ke + (size_t)s * ni;
      for (size_t pos = 0; pos < ni; ++pos) row[order[pos]] = r[pos];
      std::copy(row.begin(), row.end(), r);
    }
  }
  if (ke)
    for (size_t i = 0; i < (size_t)nsteps * ni; ++i)
      if (!std::isfinite(ke[i])) return fail(h, LDG_ENONFINITE, "kinetic energy became non-finite");
  return LDG_OK;
}

// Position of instance i in its group's arrays.
size_t inst_offset(const ldg_handle *h, int i) {
  const Inst &in = h->inst[i];
  const Group &g = h->groups[in.group];
  size_t off = 0;
  for (int s = 0; s < in.slot; ++s) off += (size_t)h->inst[g.members[s]].K;
  return off * g.n * g.n * g.n * NSTATE;
}

size_t inst_values(const ldg_handle *h, int i) {
  const Inst &in = h->inst[i];
  const int n = in.p + 1;
  return (size_t)in.K * n * n * n * NSTATE;
}

ldg_instance default_instance(const ldg_config &c) {
  ldg_instance in;
  std::memset(&in, 0, sizeof(in));
  in.p = c.p; in.nx = c.nx; in.ny = c.ny; in.nz = c.nz;
  in.Re = c.Re; in.M0 = c.M0; in.V0 = c.V0;
  return in;
}

// HIP events destroyed on every way out of a timing function (ldg_autotune,
// ldg_kernel_times).
template <int K>
struct EventSet {
  hipEvent_t e[K] = {};
  int n = 0;
  ~EventSet() {
    for (int i = 0; i < n; ++i) (void)hipEventDestroy(e[i]);
  }
  int create(ldg_handle *h) {
    for (; n < K; ++n) HIPCHECK(h, hipEventCreate(&e[n]));
    return LDG_OK;
  }
};

}  // namespace ldg_host

using namespace ldg_host;

extern "C" {

int ldg_config_default(ldg_config *c) {
  if (!c) return LDG_EINVAL;
  std::memset(c, 0, sizeof(*c));
  c->p = 2;
  c->nx = c->ny = c->nz = 10;
  for (int l = 0; l < 3; ++l) { c->xmin[l] = -M_PI; c->xmax[l] = M_PI; }
  c->L = 1.0;
  c->Re = 1600.0; c->Pr = 0.71; c->gamma = 1.4; c->M0 = 0.1; c->rho0 = 1.0; c->V0 = 1.0; c->R = 8.314;
  c->precision = LDG_FP64;
  return LDG_OK;
}

const char *ldg_last_error(const ldg_handle *h) { return h ? h->err.c_str() : g_create_error; }

int ldg_create(const ldg_config *cfg, ldg_handle **out) {
  if (!cfg || !out) return fail(nullptr, LDG_EINVAL, "null argument");
  const ldg_instance in = default_instance(*cfg);
  return create_impl(cfg, &in, 1, out);
}

int ldg_ensemble_create(const ldg_config *base, const ldg_instance *inst, int n, ldg_handle **out) {
  return create_impl(base, inst, n, out);
}

int ldg_ensemble_sample(unsigned long long seed, int n, ldg_instance *out) {
  if (!out || n < 0) return LDG_EINVAL;
  std::mt19937_64 rng(seed);
  auto uni = [&]() { return (rng() >> 11) * (1.0 / 9007199254740992.0); };   // [0, 1)
  const int sizes[4] = {8, 12, 16, 24};
  for (int i = 0; i < n; ++i) {
    ldg_instance &in = out[i];
    std::memset(&in, 0, sizeof(in));
    const int ne = sizes[rng() % 4];
    in.nx = in.ny = in.nz = ne;
    in.p = 2 + (int)(rng() % 3);
    in.Re = 400.0 * std::exp(uni() * std::log(3200.0 / 400.0));
    in.M0 = 0.05 + 0.25 * uni();
    const double u1 = 1.0 - uni(), u2 = uni();                  // Box-Muller, u1 in (0, 1]
    const double gauss = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    in.V0 = 1.0 + 0.05 * gauss;
    for (int l = 0; l < 3; ++l) in.phase[l] = 2.0 * M_PI * uni();
  }
  return LDG_OK;
}

void ldg_destroy(ldg_handle *h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graph(h);
  for (Group &g : h->groups) {
    double *ptrs[] = {g.u, g.acc, g.k, g.ke_part, g.diag_part, g.d_scale};
    for (double *p : ptrs) if (p) (void)hipFree(p);
    if (g.ua) (void)hipFree(g.ua);
    if (g.ub) (void)hipFree(g.ub);
    if (g.q) (void)hipFree(g.q);
    if (g.ff) (void)hipFree(g.ff);
    if (g.ops) (void)hipFree(g.ops);
    if (g.sops) (void)hipFree(g.sops);
    for (int c = 0; c < kMaxCand; ++c) { free_tiling(g.rhs[c]); free_tiling(g.grad[c]); }
    closure_release(g);
  }
  if (h->ke_hist) (void)hipFree(h->ke_hist);
  if (h->scal) (void)hipFree(h->scal);
  if (h->d_step) (void)hipFree(h->d_step);
  for (hipStream_t gs : h->gstreams) if (gs) (void)hipStreamDestroy(gs);
  for (hipEvent_t ev : h->ev_join) if (ev) (void)hipEventDestroy(ev);
  if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int ldg_sizes(const ldg_handle *h, int *dofs, int *nelem) {
  if (!h) return LDG_EINVAL;
  const int n = h->inst[0].p + 1;
  if (dofs) *dofs = n * n * n;
  if (nelem) *nelem = (int)h->inst[0].K;
  return LDG_OK;
}

int ldg_instance_count(const ldg_handle *h, int *n) {
  if (!h || !n) return LDG_EINVAL;
  *n = (int)h->inst.size();
  return LDG_OK;
}

int ldg_instance_sizes(const ldg_handle *h, int i, int *p, int *dofs, int *nelem) {
  if (!h || i < 0 || i >= (int)h->inst.size()) return LDG_EINVAL;
  const int n = h->inst[i].p + 1;
  if (p) *p = h->inst[i].p;
  if (dofs) *dofs = n * n * n;
  if (nelem) *nelem = (int)h->inst[i].K;
  return LDG_OK;
}

int ldg_initial_condition(ldg_handle *h) {
  if (!h) return LDG_EINVAL;
  for (size_t i = 0; i < h->inst.size(); ++i) {
    const Inst &in = h->inst[i];
    Group &g = h->groups[in.group];
    TGVArgs t;
    std::memset(&t, 0, sizeof(t));
    t.n = g.n; t.nx = in.nx; t.ny = in.ny; t.nz = in.nz; t.nelem = in.K;
    for (int l = 0; l < 3; ++l) { t.xmin[l] = h->cfg.xmin[l]; t.xmax[l] = h->cfg.xmax[l]; t.phase[l] = in.phase[l]; }
    for (int k = 0; k < g.n; ++k) t.r[k] = g.basis.r[k];
    const double c0 = in.V0 / in.M0;
    const double p0 = h->cfg.rho0 * c0 * c0 / h->cfg.gamma;     // solver.cpp:40-43
    const double T0 = p0 / (h->cfg.R * h->cfg.rho0);
    t.L = h->cfg.L; t.p0 = p0; t.rho0 = h->cfg.rho0; t.V0 = in.V0; t.RT0 = h->cfg.R * T0; t.gamma = h->cfg.gamma;
    hipLaunchKernelGGL(ldg_tgv_kernel, dim3(2048), dim3(256), 0, h->stream, g.u + inst_offset(h, (int)i), t);
    HIPCHECK(h, hipGetLastError());
  }
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  return LDG_OK;
}

int ldg_instance_set_state(ldg_handle *h, int i, const double *u) {
  if (!h || !u || i < 0 || i >= (int)h->inst.size()) return LDG_EINVAL;
  Group &g = h->groups[h->inst[i].group];
  HIPCHECK(h, hipMemcpyAsync(g.u + inst_offset(h, i), u, inst_values(h, i) * sizeof(double), hipMemcpyHostToDevice,
                             h->stream));
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  return LDG_OK;
}

int ldg_instance_get_state(ldg_handle *h, int i, double *u) {
  if (!h || !u || i < 0 || i >= (int)h->inst.size()) return LDG_EINVAL;
  Group &g = h->groups[h->inst[i].group];
  HIPCHECK(h, hipMemcpyAsync(u, g.u + inst_offset(h, i), inst_values(h, i) * sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  return LDG_OK;
}

int ldg_set_state(ldg_handle *h, const double *u) { return ldg_instance_set_state(h, 0, u); }
int ldg_get_state(ldg_handle *h, double *u) { return ldg_instance_get_state(h, 0, u); }

int ldg_state_device_ptr(ldg_handle *h, double **dptr) {
  if (!h || !dptr) return LDG_EINVAL;
  *dptr = h->groups[h->inst[0].group].u + inst_offset(h, 0);
  return LDG_OK;
}

int ldg_instance_time_step(const ldg_handle *h, int i, double cfl, double *dt) {
  if (!h || !dt || i < 0 || i >= (int)h->inst.size()) return LDG_EINVAL;
  const Inst &in = h->inst[i];
  const double m = std::fmin(std::fmin(in.dx[0], in.dx[1]), in.dx[2]);
  *dt = cfl * m / (in.V0 * (2 * in.p + 1));     // solver.cpp:47
  return LDG_OK;
}

double ldg_time_step(const ldg_handle *h, double cfl) {
  double dt = -1.0;
  if (!h || ldg_instance_time_step(h, 0, cfl, &dt) != LDG_OK) return -1.0;
  return dt;
}

int ldg_rhs(ldg_handle *h, const double *u, double *k, double *Du) {
  if (!h || !u || !k) return LDG_EINVAL;
  if (h->inst.size() != 1) return fail(h, LDG_EINVAL, "ldg_rhs needs a single-instance handle");
  Group &g = h->groups[0];
  const size_t bytes = g.nU * sizeof(double);
  // the FP64 running-sum array is free between steps: it takes the input
  HIPCHECK(h, hipMemcpyAsync(g.acc, u, bytes, hipMemcpyHostToDevice, h->stream));
  int rc = launch_grad(h, g, state_in(g.acc), -1, Du ? 1 : -1);
  if (rc) return rc;
  StageArgs st = empty_stage();
  st.kout = g.k;
  rc = closure_stage(h, g, state_in(g.acc), st);
  if (rc) return rc;
  rc = launch_rhs(h, g, state_in(g.acc), st);
  if (rc) return rc;
  HIPCHECK(h, hipMemcpyAsync(k, g.k, bytes, hipMemcpyDeviceToHost, h->stream));
  if (Du) {
    std::vector<double> qh(g.nU * 3);
    if (h->mixed) {
      std::vector<float> qf(g.nU * 3);
      HIPCHECK(h, hipMemcpyAsync(qf.data(), g.q, g.nU * 3 * sizeof(float), hipMemcpyDeviceToHost, h->stream));
      HIPCHECK(h, hipStreamSynchronize(h->stream));
      for (size_t i = 0; i < qf.size(); ++i) qh[i] = qf[i];
    } else {
      HIPCHECK(h, hipMemcpyAsync(qh.data(), g.q, bytes * 3, hipMemcpyDeviceToHost, h->stream));
      HIPCHECK(h, hipStreamSynchronize(h->stream));
    }
    // (node, s + 5 l, element) -> Dus(node, s, element, l)
    const size_t dofs = (size_t)g.n * g.n * g.n;
    const long K = g.K;
    for (long e = 0; e < K; ++e)
      for (int l = 0; l < 3; ++l)
        for (int s = 0; s < NSTATE; ++s)
          std::memcpy(Du + dofs * (s + NSTATE * (e + (size_t)K * l)), qh.data() + dofs * ((size_t)e * NGRAD + 5 * l + s),
                      dofs * sizeof(double));
  }
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  return LDG_OK;
}

int ldg_step(ldg_handle *h, double dt, int nsteps, double *ke) {
  if (!h) return fail(h, LDG_EINVAL, "bad step arguments");
  if (h->inst.size() != 1) return fail(h, LDG_EINVAL, "ldg_step needs a single-instance handle (see ldg_ensemble_step)");
  return step_impl(h, &dt, nsteps, ke);
}

int ldg_ensemble_step(ldg_handle *h, const double *dt, int nsteps, double *ke) {
  if (!h || !dt) return fail(h, LDG_EINVAL, "bad step arguments");
  return step_impl(h, dt, nsteps, ke);
}

int ldg_set_graph(ldg_handle *h, int enable) {
  if (!h) return LDG_EINVAL;
  h->use_graph = enable != 0;
  if (!h->use_graph) drop_graph(h);
  return LDG_OK;
}

int ldg_instance_diagnostics(ldg_handle *h, int i, double out[3]) {
  if (!h || !out || i < 0 || i >= (int)h->inst.size()) return LDG_EINVAL;
  const int ni = (int)h->inst.size();
  // one zero-length RK epilogue per group (acc = u + 0 k) gives the nodal KE
  // of u; the Gauss-point sums come from the volume phase of the same launch
  int first = 0;
  for (Group &g : h->groups) {
    int rc = launch_grad(h, g, state_in(g.u));
    if (rc) return rc;
    StageArgs st = empty_stage();
    st.u0 = g.u;
    st.acc = g.acc;
    st.acc_init = 1;
    st.cb = 0.0;
    st.ke_part = g.ke_part;
    st.diag_part = g.diag_part;
    rc = launch_rhs(h, g, state_in(g.u), st);
    if (rc) return rc;
    const Tiling &t = g.rhs[g.crhs];
    hipLaunchKernelGGL(ldg_finalize_kernel, dim3(g.members.size()), dim3(NT), 0, h->stream, (const double *)g.ke_part,
                       1, (const InstDesc *)t.d_inst, ni, (const double *)g.d_scale, h->scal + first,
                       (const int *)nullptr);
    hipLaunchKernelGGL(ldg_finalize_kernel, dim3(g.members.size()), dim3(NT), 0, h->stream,
                       (const double *)g.diag_part, 2, (const InstDesc *)t.d_inst, ni, (const double *)g.d_scale,
                       h->scal + ni + 2 * first, (const int *)nullptr);
    HIPCHECK(h, hipGetLastError());
    first += (int)g.members.size();
  }
  std::vector<double> all(3 * ni);
  HIPCHECK(h, hipMemcpyAsync(all.data(), h->scal, 3 * ni * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  // instances are numbered group by group in the arrays above
  const std::vector<int> order = group_order(h);
  for (int pos = 0; pos < ni; ++pos) {
    if (order[pos] == i) {
      out[0] = all[pos];
      out[1] = all[ni + 2 * pos];
      out[2] = all[ni + 2 * pos + 1];
    }
  }
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(out[k])) return fail(h, LDG_ENONFINITE, "diagnostic is non-finite");
  return LDG_OK;
}

int ldg_diagnostics(ldg_handle *h, double out[3]) {
  if (!h || !out) return LDG_EINVAL;
  return ldg_instance_diagnostics(h, 0, out);
}

int ldg_autotune(ldg_handle *h, int *chosen) {
  if (!h) return LDG_EINVAL;
  drop_graph(h);
  EventSet<2> evs;
  if (int rc = evs.create(h)) return rc;
  hipEvent_t e0 = evs.e[0], e1 = evs.e[1];
  int gi = 0;
  for (Group &g : h->groups) {
    // time the kernels on the stage input they read in a run: in the mixed
    // path without a closure that is the FP32 fluctuation of the state,
    // written into ub (free between calls) by one rhs launch with a_s = 0
    const bool f32 = h->mixed && g.cl.net == nullptr;
    if (f32) {
      StageArgs st = empty_stage();
      st.u0 = g.u;
      st.unext32 = (float *)g.ub;
      st.ca = 0.0;
      const int rc = launch_rhs(h, g, state_in(g.u), st, g.crhs);
      if (rc) return rc;
    }
    const StageIn in = f32 ? StageIn{g.ub, true} : state_in(g.u);
    for (int kind = 0; kind < 2; ++kind) {     // 0: gradient, 1: right-hand side
      // rounds that alternate the candidates, each timed over two launches
      // after an untimed one; a candidate's time is its best round (the
      // clock drifts within a call, and alternating keeps it from favouring
      // whichever candidate ran later)
      float best[kMaxCand];
      for (int c = 0; c < kMaxCand; ++c) best[c] = 1e30f;
      for (int round = 0; round < 3; ++round) {
        for (int c = 0; c < kMaxCand; ++c) {
          if (c > 0 && cand_e(g.n, kind == 1, c) == cand_e(g.n, kind == 1, 0)) continue;
          float ms = 0.f;
          for (int rep = 0; rep < 3; ++rep) {    // first launch warms up, not timed
            if (rep == 1) HIPCHECK(h, hipEventRecord(e0, h->stream));
            int rc;
            if (kind == 0) {
              rc = launch_grad(h, g, in, c);
            } else {
              StageArgs st = empty_stage();
              st.kout = g.k;
              rc = launch_rhs(h, g, in, st, c);
            }
            if (rc) return rc;
          }
          HIPCHECK(h, hipEventRecord(e1, h->stream));
          HIPCHECK(h, hipEventSynchronize(e1));
          HIPCHECK(h, hipEventElapsedTime(&ms, e0, e1));
          best[c] = std::min(best[c], ms);
        }
      }
      int bestc = 0;
      for (int c = 1; c < kMaxCand; ++c)
        if (best[c] < best[bestc]) bestc = c;
      if (kind == 0) g.cgrad = bestc; else g.crhs = bestc;
    }
    if (chosen) {
      chosen[3 * gi] = g.n - 1;
      chosen[3 * gi + 1] = cand_e(g.n, false, g.cgrad);
      chosen[3 * gi + 2] = cand_e(g.n, true, g.crhs);
    }
    ++gi;
  }
  return LDG_OK;
}

int ldg_kernel_times(ldg_handle *h, double dt, int nsteps, double ms[2]) {
  if (!h || !ms || nsteps < 1) return LDG_EINVAL;
  if (h->inst.size() != 1) return fail(h, LDG_EINVAL, "ldg_kernel_times needs a single-instance handle");
  static const double b[4] = {1 / 6.0, 1 / 3.0, 1 / 3.0, 1 / 6.0};
  static const double a[3] = {0.5, 0.5, 1.0};
  int rc = set_dt(h, &dt);
  if (rc) return rc;
  Group &g = h->groups[0];
  EventSet<3> evs;
  if (int c = evs.create(h)) return c;
  hipEvent_t *ev = evs.e;
  double tot[2] = {0, 0};
  int count = 0;
  // the same stage inputs as ldg_step: from the second step on, stage 1
  // reads the FP32 copy of u^n the previous stage 4 wrote (mixed path)
  const bool carry = h->mixed && g.cl.net == nullptr;
  for (int it = 0; it < nsteps; ++it) {
    StageIn U = carry && it > 0 ? StageIn{g.ub, true} : state_in(g.u);
    void *bufs[2] = {g.ua, g.ub};
    for (int s = 0; s < 4; ++s) {
      HIPCHECK(h, hipEventRecord(ev[0], h->stream));
      rc = launch_grad(h, g, U);
      if (rc) return rc;
      HIPCHECK(h, hipEventRecord(ev[1], h->stream));
      StageArgs st = empty_stage();
      st.u0 = g.u; st.acc = g.acc; st.acc_init = (s == 0); st.cb = b[s];
      if (s < 3) {
        if (h->mixed) st.unext32 = (float *)bufs[s & 1];
        else st.unext = (double *)bufs[s & 1];
        st.ca = a[s];
      }
      if (s == 3 && carry && it + 1 < nsteps) st.next32 = (float *)g.ub;
      rc = launch_rhs(h, g, U, st);
      if (rc) return rc;
      HIPCHECK(h, hipEventRecord(ev[2], h->stream));
      HIPCHECK(h, hipEventSynchronize(ev[2]));
      float t0 = 0, t1 = 0;
      HIPCHECK(h, hipEventElapsedTime(&t0, ev[0], ev[1]));
      HIPCHECK(h, hipEventElapsedTime(&t1, ev[1], ev[2]));
      tot[0] += t0; tot[1] += t1; ++count;
      U = StageIn{bufs[s & 1], h->mixed};
    }
    std::swap(g.u, g.acc);
  }
  ms[0] = tot[0] / count;
  ms[1] = tot[1] / count;
  return LDG_OK;
}

}  // extern "C"

// ------------------------------------------------- allocation-free right-hand side
//
// ldg_rhs_noalloc keeps everything beyond its input and output in one fixed
// device array per GPU: the 1D operators, the two instance tables (gradient
// and right-hand-side tilings) and the face fluxes.  Host code never
// uses the array's name as an address: its device address is looked up once
// per device with hipGetSymbolAddress and kept in g_scratch_ptr.
static __device__ __align__(256) unsigned char scratch[LDG_NOALLOC_SCRATCH_BYTES];

namespace ldg_host {

constexpr size_t kScrOps = 0, kScrInstGrad = 4096, kScrInstRhs = 6144, kScrFF = 8192;
// the rhs kernel's operator image of the precision in use follows the
// operators in their slot
template <int N>
constexpr size_t kScrImg = kScrOps + (sizeof(Ops<N, double>) + 15) / 16 * 16;
static_assert(kScrImg<6> + SOpsImage<6>::D <= kScrInstGrad && SOpsImage<6>::F <= SOpsImage<6>::D,
              "operators do not fit their scratch slot");
static_assert(sizeof(InstDesc) <= kScrInstRhs - kScrInstGrad, "instance table does not fit its scratch slot");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (looked up on the first call
// on the device), or nullptr with the error recorded.
unsigned char *scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) {
    set_global_error("ldg_rhs_noalloc: device index out of range");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      char msg[256];
      std::snprintf(msg, sizeof(msg), "ldg_rhs_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
      set_global_error(msg);
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
  }
  return g_scratch_ptr[dev];
}

// The 1D operators of every order, built on the host once per process.
struct OpsTables {
  Ops<2, double> o2;
  Ops<3, double> o3;
  Ops<4, double> o4;
  Ops<5, double> o5;
  Ops<6, double> o6;
  bool ok = true;
};
const OpsTables &ops_tables() {
  static const OpsTables t = [] {
    OpsTables r;
    double a = 0.0;
    r.o2 = make_ops<2>(make_basis(1), a); r.ok = r.ok && a < 1e-9;
    r.o3 = make_ops<3>(make_basis(2), a); r.ok = r.ok && a < 1e-9;
    r.o4 = make_ops<4>(make_basis(3), a); r.ok = r.ok && a < 1e-9;
    r.o5 = make_ops<5>(make_basis(4), a); r.ok = r.ok && a < 1e-9;
    r.o6 = make_ops<6>(make_basis(5), a); r.ok = r.ok && a < 1e-9;
    return r;
  }();
  return t;
}
template <int N> const Ops<N, double> &ops_of(const OpsTables &t);
template <> const Ops<2, double> &ops_of<2>(const OpsTables &t) { return t.o2; }
template <> const Ops<3, double> &ops_of<3>(const OpsTables &t) { return t.o3; }
template <> const Ops<4, double> &ops_of<4>(const OpsTables &t) { return t.o4; }
template <> const Ops<5, double> &ops_of<5>(const OpsTables &t) { return t.o5; }
template <> const Ops<6, double> &ops_of<6>(const OpsTables &t) { return t.o6; }

// The gradient and right-hand-side kernels of ldg_rhs (default tilings, FP64
// input) with every intermediate in scr.  The host copies of the operators
// and instance tables are read by the copies before the stream is
// synchronised (the caller synchronises before returning).
template <int N>
hipError_t noalloc_launch(const ldg_config &cfg, const Inst &in, unsigned char *scr, const double *u, double *k,
                          hipStream_t s, InstDesc dsc[2], unsigned char *img) {
  constexpr int EG = Cand<N>::grad[0], ER = Cand<N>::rhs[0];
  const Ops<N, double> &o = ops_of<N>(ops_tables());
  if (inst_desc(cfg, in, EG, 0, 0, dsc[0]) != LDG_OK || inst_desc(cfg, in, ER, 0, 0, dsc[1]) != LDG_OK)
    return hipErrorInvalidValue;
  // the image of the precision in use only, at kScrImg; the kernel is given
  // the base from which that image's offset lands there
  const double a[3] = {dsc[1].ialpha[0], dsc[1].ialpha[1], dsc[1].ialpha[2]};
  const bool mixed = cfg.precision == LDG_MIXED;
  if (mixed) fill_sops<N, float>(*reinterpret_cast<SOps<N, float> *>(img), o, a);
  else fill_sops<N, double>(*reinterpret_cast<SOps<N, double> *>(img), o, a);
  const size_t ib = mixed ? sizeof(SOps<N, float>) : sizeof(SOps<N, double>);
  const unsigned char *simg =
      scr + kScrImg<N> - (mixed ? SOpsImage<N>::template offset<float>() : SOpsImage<N>::template offset<double>());
  hipError_t e = hipMemcpyAsync(scr + kScrOps, &o, sizeof(o), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(scr + kScrImg<N>, img, ib, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(scr + kScrInstGrad, &dsc[0], sizeof(InstDesc), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(scr + kScrInstRhs, &dsc[1], sizeof(InstDesc), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  const Ops<N, double> *op = reinterpret_cast<const Ops<N, double> *>(scr + kScrOps);
  const LaunchArgs lg{reinterpret_cast<const InstDesc *>(scr + kScrInstGrad), nullptr, dsc[0].nblocks,
                      (dsc[0].nblocks + 7) / 8};
  const LaunchArgs lr{reinterpret_cast<const InstDesc *>(scr + kScrInstRhs), nullptr, dsc[1].nblocks,
                      (dsc[1].nblocks + 7) / 8};
  StageArgs st = empty_stage();
  st.kout = k;
  // the gradient kernel writes no nodal gradient here (WQ = false): the
  // scratch holds the face fluxes only
  if (cfg.precision == LDG_MIXED) {
    float *ff = reinterpret_cast<float *>(scr + kScrFF);
    hipLaunchKernelGGL((ldg_grad_kernel<N, EG, float, double, false>), dim3(8 * lg.per_xcd), dim3(NT), 0, s, u,
                       nullptr, ff, lg, op);
    hipLaunchKernelGGL((ldg_rhs_kernel<N, ER, float, double>), dim3(8 * lr.per_xcd), dim3(NT), 0, s, u,
                       (const float *)ff, lr, simg, st);
  } else {
    double *ff = reinterpret_cast<double *>(scr + kScrFF);
    hipLaunchKernelGGL((ldg_grad_kernel<N, EG, double, double, false>), dim3(8 * lg.per_xcd), dim3(NT), 0, s, u,
                       nullptr, ff, lg, op);
    hipLaunchKernelGGL((ldg_rhs_kernel<N, ER, double, double>), dim3(8 * lr.per_xcd), dim3(NT), 0, s, u,
                       (const double *)ff, lr, simg, st);
  }
  return hipGetLastError();
}

int noalloc_fail(int code, const char *msg) {
  set_global_error(msg);
  return code;
}

}  // namespace ldg_host

extern "C" int ldg_rhs_noalloc(const ldg_config *cfg, const double *u, double *k, void *stream) {
  using namespace ldg_host;
  if (!cfg || !u || !k) return noalloc_fail(LDG_EINVAL, "ldg_rhs_noalloc: null argument");
  const ldg_instance li = default_instance(*cfg);
  if (validate_instance(*cfg, li) != LDG_OK) return LDG_EINVAL;   // message recorded
  Inst in;
  in.p = li.p; in.nx = li.nx; in.ny = li.ny; in.nz = li.nz;
  in.Re = li.Re; in.M0 = li.M0; in.V0 = li.V0;
  const int ns[3] = {in.nx, in.ny, in.nz};
  for (int l = 0; l < 3; ++l) {
    in.phase[l] = 0.0;
    in.dx[l] = (cfg->xmax[l] - cfg->xmin[l]) / ns[l];
  }
  in.K = (long)in.nx * in.ny * in.nz;
  in.group = 0;
  in.slot = 0;
  const int n = in.p + 1;
  // the face fluxes (6 chunks of 5 n^2 values per element)
  const int wb = cfg->precision == LDG_MIXED ? 4 : 8;
  const double qbytes = (double)in.K * 6 * ff_chunk(n, wb) * wb;
  if (kScrFF + qbytes > (double)LDG_NOALLOC_SCRATCH_BYTES) {
    char msg[256];
    std::snprintf(msg, sizeof(msg),
                  "ldg_rhs_noalloc: %d x %d x %d elements at p = %d need %.0f bytes of device scratch; the limit is "
                  "%u (LDG_NOALLOC_SCRATCH_BYTES)", in.nx, in.ny, in.nz, in.p, kScrFF + qbytes,
                  (unsigned)LDG_NOALLOC_SCRATCH_BYTES);
    return noalloc_fail(LDG_EINVAL, msg);
  }
  if (!ops_tables().ok) return noalloc_fail(LDG_EINVAL, "ldg_rhs_noalloc: 1D operators are not centro-symmetric");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return noalloc_fail(LDG_EDEVICE, "ldg_rhs_noalloc: no HIP device");
  unsigned char *scr = scratch_on(dev);
  if (!scr) return LDG_EDEVICE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  InstDesc dsc[2];
  alignas(16) unsigned char img[SOpsImage<6>::D];   // (read by the copies until the synchronisation below)
  hipError_t e = hipErrorInvalidValue;
  switch (n) {
    case 2: e = noalloc_launch<2>(*cfg, in, scr, u, k, s, dsc, img); break;
    case 3: e = noalloc_launch<3>(*cfg, in, scr, u, k, s, dsc, img); break;
    case 4: e = noalloc_launch<4>(*cfg, in, scr, u, k, s, dsc, img); break;
    case 5: e = noalloc_launch<5>(*cfg, in, scr, u, k, s, dsc, img); break;
    case 6: e = noalloc_launch<6>(*cfg, in, scr, u, k, s, dsc, img); break;
  }
  const hipError_t es = hipStreamSynchronize(s);
  if (e == hipSuccess) e = es;
  if (e != hipSuccess) {
    char msg[256];
    std::snprintf(msg, sizeof(msg), "ldg_rhs_noalloc: %s", hipGetErrorString(e));
    return noalloc_fail(LDG_EDEVICE, msg);
  }
  return LDG_OK;
}
