// closure.hip -- learned per-element closure for an under-resolved run
// (C-ABI part of libldg.so, declared in include/ldg.h).  No reference
// counterpart: the reference has no closure.
//
// Training data come from runs of the method itself: a fine (high order)
// state u_f is L2-projected to the coarse order, P u_f, and the closure
// learns the difference between the projected fine right-hand side and the
// coarse right-hand side of the projected state,
//   target = P rhs_f(u_f) - rhs_c(P u_f),
// from the coarse element's state and LDG gradient.  At run time the
// correction is added to the coarse right-hand side inside the RK4 stage
// epilogue, element by element, unless its normalised RMS exceeds a bound
// (then that element keeps the plain right-hand side).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "closure_kernels.hpp"
#include "handle.hpp"

using namespace ldg_host;
using namespace ldg_mlp;

struct ldg_closure {
  int p = 0, n = 0, dofs = 0, nin = 0, nout = 0, hidden = 0;
  long nparam = 0;
  float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
  float *norm = nullptr;            // in_mean[nin], in_scale[nin] (1/std), out_mean[nout], out_std[nout]
  unsigned short *Wb = nullptr;     // padded BF16 weight copies for the fused inference kernel
  int k1p = 0, noutp = 0;
  long adam_t = 0;
  long cap = 0, count = 0;
  float *Xd = nullptr, *Td = nullptr;
  float *Wd = nullptr;              // per-sample loss weight (1 unless weighting by collection)
  std::vector<int> group_of;        // per sample: the collection (snapshot) it came from
  int ncollect = 0;
  int weighting = 0;                // 0: uniform; 1: every collection has the same total weight per unit target energy
  bool normalized = false;
  int wsB = 0;                      // training workspace rows
  float *Xb = nullptr, *Tb = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *Y = nullptr,
        *dY = nullptr, *dH = nullptr, *dZ = nullptr, *Wbat = nullptr;
  long *idx = nullptr;
  double *part = nullptr;
  hipStream_t stream = nullptr;
  std::string err;
};

namespace {

std::string g_closure_error;
constexpr int kLossBlocks = 256;

int cfail(ldg_closure *c, int code, const std::string &msg) {
  if (c) c->err = msg; else g_closure_error = msg;
  return code;
}

#define CCHECK(c, call)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return cfail((c), LDG_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// parameter blocks: W1 (hidden x nin), b1, W2 (hidden x hidden), b2,
// W3 (nout x hidden), b3
struct Offs {
  long W1, b1, W2, b2, W3, b3;
};
Offs offsets(const ldg_closure *c) {
  Offs o;
  o.W1 = 0;
  o.b1 = o.W1 + (long)c->hidden * c->nin;
  o.W2 = o.b1 + c->hidden;
  o.b2 = o.W2 + (long)c->hidden * c->hidden;
  o.W3 = o.b2 + c->hidden;
  o.b3 = o.W3 + (long)c->nout * c->hidden;
  return o;
}

dim3 gemm_grid(int M, int N) { return dim3((N + GT - 1) / GT, (M + GT - 1) / GT); }

template <bool TA, bool TB, int EPI>
void gemm(hipStream_t s, int M, int N, int K, const float *A, int lda, const float *B, int ldb, const float *bias,
          float *Z, float *C, int ldc) {
  if (M <= 0 || N <= 0) return;
  hipLaunchKernelGGL((mlp_gemm<TA, TB, EPI>), gemm_grid(M, N), dim3(256), 0, s, M, N, K, A, lda, B, ldb, bias, Z, C,
                     ldc);
}

int grid_for(long n, int bs = 256) { return (int)std::min<long>((n + bs - 1) / bs, 65535); }

// Forward pass of rows [0, B) of X into Y (Z1/Z2 kept when not null).
void forward(const ldg_closure *c, hipStream_t s, int B, const float *X, float *Z1, float *H1, float *Z2, float *H2,
             float *Y) {
  const Offs o = offsets(c);
  const float *P = c->params;
  gemm<false, true, EPI_BIAS_SILU>(s, B, c->hidden, c->nin, X, c->nin, P + o.W1, c->nin, P + o.b1, Z1, H1, c->hidden);
  gemm<false, true, EPI_BIAS_SILU>(s, B, c->hidden, c->hidden, H1, c->hidden, P + o.W2, c->hidden, P + o.b2, Z2, H2,
                                   c->hidden);
  gemm<false, true, EPI_BIAS>(s, B, c->nout, c->hidden, H2, c->hidden, P + o.W3, c->hidden, P + o.b3, nullptr, Y,
                              c->nout);
}

// Gradient of the mean squared error over rows [0, B) into c->grads; loss
// partials into c->part.  Uses the training workspace.
void backward(ldg_closure *c, hipStream_t s, int B, const float *X, const float *T, const float *rw) {
  const Offs o = offsets(c);
  const float *P = c->params;
  float *G = c->grads;
  forward(c, s, B, X, c->Z1, c->H1, c->Z2, c->H2, c->Y);
  const long ny = (long)B * c->nout;
  hipLaunchKernelGGL(mlp_loss_grad, dim3(kLossBlocks), dim3(256), 0, s, (const float *)c->Y, T, ny, c->nout, rw,
                     2.0f / (float)ny, c->dY, c->part);
  gemm<true, false, EPI_NONE>(s, c->nout, c->hidden, B, c->dY, c->nout, c->H2, c->hidden, nullptr, nullptr, G + o.W3,
                              c->hidden);
  hipLaunchKernelGGL(mlp_colsum, dim3((c->nout + 255) / 256), dim3(256), 0, s, (const float *)c->dY, B, c->nout,
                     G + o.b3);
  gemm<false, false, EPI_NONE>(s, B, c->hidden, c->nout, c->dY, c->nout, P + o.W3, c->hidden, nullptr, nullptr, c->dH,
                               c->hidden);
  const long nh = (long)B * c->hidden;
  hipLaunchKernelGGL(mlp_silu_back, dim3(grid_for(nh)), dim3(256), 0, s, (const float *)c->dH, (const float *)c->Z2,
                     nh, c->dZ);
  gemm<true, false, EPI_NONE>(s, c->hidden, c->hidden, B, c->dZ, c->hidden, c->H1, c->hidden, nullptr, nullptr,
                              G + o.W2, c->hidden);
  hipLaunchKernelGGL(mlp_colsum, dim3((c->hidden + 255) / 256), dim3(256), 0, s, (const float *)c->dZ, B, c->hidden,
                     G + o.b2);
  gemm<false, false, EPI_NONE>(s, B, c->hidden, c->hidden, c->dZ, c->hidden, P + o.W2, c->hidden, nullptr, nullptr,
                               c->dH, c->hidden);
  hipLaunchKernelGGL(mlp_silu_back, dim3(grid_for(nh)), dim3(256), 0, s, (const float *)c->dH, (const float *)c->Z1,
                     nh, c->dZ);
  gemm<true, false, EPI_NONE>(s, c->hidden, c->nin, B, c->dZ, c->hidden, X, c->nin, nullptr, nullptr, G + o.W1,
                              c->nin);
  hipLaunchKernelGGL(mlp_colsum, dim3((c->hidden + 255) / 256), dim3(256), 0, s, (const float *)c->dZ, B, c->hidden,
                     G + o.b1);
}

int ensure_workspace(ldg_closure *c, int B) {
  if (B <= c->wsB) return LDG_OK;
  float **bufs[] = {&c->Xb, &c->Tb, &c->Z1, &c->H1, &c->Z2, &c->H2, &c->Y, &c->dY, &c->dH, &c->dZ, &c->Wbat};
  for (float **b : bufs) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
  if (c->idx) (void)hipFree(c->idx);
  c->idx = nullptr;
  c->wsB = 0;
  const size_t r = (size_t)B;
  const size_t widths[] = {(size_t)c->nin, (size_t)c->nout, (size_t)c->hidden, (size_t)c->hidden, (size_t)c->hidden,
                           (size_t)c->hidden, (size_t)c->nout, (size_t)c->nout, (size_t)c->hidden, (size_t)c->hidden, 1};
  for (int i = 0; i < 11; ++i)
    if (hipMalloc((void **)bufs[i], r * widths[i] * sizeof(float)) != hipSuccess)
      return cfail(c, LDG_ENOMEM, "device allocation failed (closure workspace)");
  if (hipMalloc((void **)&c->idx, r * sizeof(long)) != hipSuccess)
    return cfail(c, LDG_ENOMEM, "device allocation failed (closure workspace)");
  c->wsB = B;
  return LDG_OK;
}

// Refreshes the BF16 weight copies used by the fused inference kernel.
int pack_bf16(ldg_closure *c) {
  const Offs o = offsets(c);
  unsigned short *W1b = c->Wb, *W2b = W1b + (long)c->hidden * c->k1p, *W3b = W2b + (long)c->hidden * c->hidden;
  const long n = (long)c->hidden * c->k1p + (long)c->hidden * c->hidden + (long)c->noutp * c->hidden;
  hipLaunchKernelGGL(mlp_pack_bf16, dim3(grid_for(n)), dim3(256), 0, c->stream, (const float *)(c->params + o.W1),
                     (const float *)(c->params + o.W2), (const float *)(c->params + o.W3), c->nin, c->hidden, c->nout,
                     c->k1p, c->noutp, W1b, W2b, W3b);
  CCHECK(c, hipGetLastError());
  CCHECK(c, hipStreamSynchronize(c->stream));
  return LDG_OK;
}

bool fused_ok(const ldg_closure *c) { return c->n == 3 && c->hidden == 256; }

// 1D L2 projection from order pf to order pc on the reference element:
// P = M_c^{-1} B, B(i, j) = integral of l^c_i l^f_j, both integrals exact
// (Gauss rules of pc+1 and pf+1 points).
std::vector<double> projection_1d(int pf, int pc) {
  const Basis1D F = make_basis(pf), Cb = make_basis(pc);
  const int nf = pf + 1, nc = pc + 1;
  // coarse Lagrange basis at the fine Gauss points
  std::vector<double> Lc(nc * nf);
  for (int q = 0; q < nf; ++q)
    for (int i = 0; i < nc; ++i) {
      double v = 1.0;
      for (int j = 0; j < nc; ++j)
        if (j != i) v *= (F.g[q] - Cb.r[j]) / (Cb.r[i] - Cb.r[j]);
      Lc[i * nf + q] = v;
    }
  std::vector<double> Bm(nc * nf, 0.0);
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nf; ++j)
      for (int q = 0; q < nf; ++q) Bm[i * nf + j] += F.w[q] * Lc[i * nf + q] * F.at(F.I, q, j);
  std::vector<double> P(nc * nf, 0.0);
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nf; ++j)
      for (int k = 0; k < nc; ++k) P[i * nf + j] += Cb.at(Cb.M1inv, i, k) * Bm[k * nf + j];
  return P;
}

int project_device(hipStream_t s, int pf, int pc, long K, const double *src, double *dst, std::string &err) {
  const std::vector<double> P = projection_1d(pf, pc);
  double *dP = nullptr;
  if (hipMalloc((void **)&dP, P.size() * sizeof(double)) != hipSuccess) { err = "device allocation failed"; return LDG_ENOMEM; }
  hipError_t e = hipMemcpyAsync(dP, P.data(), P.size() * sizeof(double), hipMemcpyHostToDevice, s);
  const int nc = pc + 1;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(mlp_project, dim3(grid_for(K * 5 * nc * nc * nc)), dim3(256), 0, s, pf + 1, nc, K,
                       (const double *)dP, src, dst);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(dP);
  if (e != hipSuccess) { err = std::string("projection: ") + hipGetErrorString(e); return LDG_EDEVICE; }
  return LDG_OK;
}

// Background (rho0, E0) of instance in: the state minus it is what the
// closure sees.
void background(const ldg_handle *h, const Inst &in, double &rho0, double &E0) {
  const double c0 = in.V0 / in.M0;
  const double p0 = h->cfg.rho0 * c0 * c0 / h->cfg.gamma;
  rho0 = h->cfg.rho0;
  E0 = p0 / (h->cfg.gamma - 1.0);
}

// features of one instance's elements into X (rows of the instance)
template <typename TQ, typename UT>
void launch_features(hipStream_t s, const UT *U, const TQ *Q, long K, int dofs, double rho0, double E0,
                     const float *norm, int nin, float *X) {
  hipLaunchKernelGGL((mlp_features<TQ, UT>), dim3(grid_for(K * nin)), dim3(256), 0, s, U, Q, K, dofs, rho0, E0, norm,
                     norm + nin, X);
}

// U: the group's stage input (state, or FP32 fluctuations: nothing to subtract)
void features_group(ldg_handle *h, Group &g, StageIn U, const float *norm, int nin, float *X) {
  const int dofs = g.n * g.n * g.n;
  long off = 0;
  for (int idx : g.members) {
    const Inst &in = h->inst[idx];
    double rho0, E0;
    background(h, in, rho0, E0);
    if (U.fluct) rho0 = E0 = 0.0;
    const long uo = off * 5 * dofs;
    if (h->mixed && U.fluct)
      launch_features<float, float>(h->stream, (const float *)U.p + uo, (const float *)g.q + off * 15 * dofs, in.K,
                                    dofs, rho0, E0, norm, nin, X + off * nin);
    else if (h->mixed)
      launch_features<float, double>(h->stream, (const double *)U.p + uo, (const float *)g.q + off * 15 * dofs, in.K,
                                     dofs, rho0, E0, norm, nin, X + off * nin);
    else
      launch_features<double, double>(h->stream, (const double *)U.p + uo, (const double *)g.q + off * 15 * dofs,
                                      in.K, dofs, rho0, E0, norm, nin, X + off * nin);
    off += in.K;
  }
}

template <typename TQ, typename UT>
void launch_infer(hipStream_t s, const ldg_closure *c, const UT *U, const TQ *Q, long K, double rho0, double E0,
                  double bound, int dissipative, double rho_add, double *corr, unsigned char *gate) {
  const Offs o = offsets(c);
  const unsigned short *W1b = c->Wb, *W2b = W1b + (long)c->hidden * c->k1p, *W3b = W2b + (long)c->hidden * c->hidden;
  hipLaunchKernelGGL((mlp_infer<TQ, UT, 540, 256, 135>), dim3((unsigned)((K + FR - 1) / FR)), dim3(256), 0, s, U, Q,
                     K, rho0, E0, (const float *)c->norm, W1b, W2b, W3b, (const float *)(c->params + o.b1),
                     (const float *)(c->params + o.b2), (const float *)(c->params + o.b3), bound, dissipative,
                     rho_add, corr, gate);
}

void free_closure(ldg_closure *c) {
  float *bufs[] = {c->params, c->grads, c->m, c->v, c->norm, c->Xd, c->Td, c->Wd, c->Xb, c->Tb, c->Z1,
                   c->H1, c->Z2, c->H2, c->Y, c->dY, c->dH, c->dZ, c->Wbat};
  for (float *b : bufs)
    if (b) (void)hipFree(b);
  if (c->idx) (void)hipFree(c->idx);
  if (c->part) (void)hipFree(c->part);
  if (c->Wb) (void)hipFree(c->Wb);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // namespace

namespace ldg_host {

int closure_stage(ldg_handle *h, Group &g, StageIn U, StageArgs &st, int stage) {
  GroupClosure &cl = g.cl;
  if (!cl.net) return LDG_OK;
  if (cl.per_step && stage > 0) {   // stage 1's correction and gate, held for the step
    st.corr = cl.corr;
    st.gate = cl.gate;
    return LDG_OK;
  }
  const ldg_closure *c = cl.net;
  const int dofs = g.n * g.n * g.n;
  if (fused_ok(c)) {
    // one fused launch per instance (each has its own background state)
    long off = 0;
    for (int idx : g.members) {
      const Inst &in = h->inst[idx];
      double rho0, E0;
      background(h, in, rho0, E0);
      const double rho_add = U.fluct ? rho0 : 0.0;     // density = input + rho_add
      if (U.fluct) rho0 = E0 = 0.0;
      const long uo = off * 5 * dofs;
      if (h->mixed && U.fluct)
        launch_infer<float, float>(h->stream, c, (const float *)U.p + uo, (const float *)g.q + off * 15 * dofs, in.K,
                                   rho0, E0, cl.bound, cl.dissipative, rho_add, cl.corr + off * c->nout, cl.gate + off);
      else if (h->mixed)
        launch_infer<float, double>(h->stream, c, (const double *)U.p + uo, (const float *)g.q + off * 15 * dofs,
                                    in.K, rho0, E0, cl.bound, cl.dissipative, rho_add, cl.corr + off * c->nout,
                                    cl.gate + off);
      else
        launch_infer<double, double>(h->stream, c, (const double *)U.p + uo, (const double *)g.q + off * 15 * dofs,
                                     in.K, rho0, E0, cl.bound, cl.dissipative, rho_add, cl.corr + off * c->nout,
                                     cl.gate + off);
      off += in.K;
    }
  } else {
    features_group(h, g, U, c->norm, c->nin, cl.X);
    forward(c, h->stream, (int)g.K, cl.X, nullptr, cl.H1, nullptr, cl.H2, cl.Y);
    // the dissipative gate needs the density (rho0 is shared by all instances)
    const double rho_add = U.fluct ? h->cfg.rho0 : 0.0;
    const dim3 grid((unsigned)std::min<long>(g.K, 65535)), block(64);
    const float *om = c->norm + 2 * c->nin, *os = c->norm + 2 * c->nin + c->nout;
    if (U.fluct)
      hipLaunchKernelGGL((mlp_correction<float>), grid, block, 0, h->stream, (const float *)cl.Y, g.K, c->nout, om, os,
                         cl.bound, cl.dissipative, (const float *)U.p, rho_add, cl.corr, cl.gate);
    else
      hipLaunchKernelGGL((mlp_correction<double>), grid, block, 0, h->stream, (const float *)cl.Y, g.K, c->nout, om,
                         os, cl.bound, cl.dissipative, (const double *)U.p, rho_add, cl.corr, cl.gate);
  }
  HIPCHECK(h, hipGetLastError());
  st.corr = cl.corr;
  st.gate = cl.gate;
  return LDG_OK;
}

void closure_release(Group &g) {
  GroupClosure &cl = g.cl;
  float *f[] = {cl.X, cl.H1, cl.H2, cl.Y};
  for (float *p : f)
    if (p) (void)hipFree(p);
  if (cl.corr) (void)hipFree(cl.corr);
  if (cl.gate) (void)hipFree(cl.gate);
  cl = GroupClosure();
}

}  // namespace ldg_host

extern "C" {

const char *ldg_closure_last_error(const ldg_closure *c) { return c ? c->err.c_str() : g_closure_error.c_str(); }

int ldg_closure_create(int p, int hidden, long capacity, unsigned long long seed, ldg_closure **out) {
  if (!out || p < 1 || p > 5 || hidden < 1 || hidden > 4096 || capacity < 0)
    return cfail(nullptr, LDG_EINVAL, "bad closure arguments");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return cfail(nullptr, LDG_EDEVICE, "no HIP device");
  ldg_closure *c = new ldg_closure();
  c->p = p;
  c->n = p + 1;
  c->dofs = c->n * c->n * c->n;
  c->nin = 20 * c->dofs;
  c->nout = 5 * c->dofs;
  c->hidden = hidden;
  const Offs o = offsets(c);
  c->nparam = o.b3 + c->nout;
  c->cap = capacity;
  c->k1p = (c->nin + 15) / 16 * 16;
  c->noutp = (c->nout + 31) / 32 * 32;
  auto bad = [&](int code, const char *msg) { g_closure_error = msg; free_closure(c); return code; };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bad(LDG_EDEVICE, "hipStreamCreate failed");
  const size_t pb = (size_t)c->nparam * sizeof(float);
  if (hipMalloc((void **)&c->params, pb) != hipSuccess || hipMalloc((void **)&c->grads, pb) != hipSuccess ||
      hipMalloc((void **)&c->m, pb) != hipSuccess || hipMalloc((void **)&c->v, pb) != hipSuccess ||
      hipMalloc((void **)&c->norm, sizeof(float) * 2 * (c->nin + c->nout)) != hipSuccess ||
      hipMalloc((void **)&c->part, sizeof(double) * kLossBlocks) != hipSuccess ||
      hipMalloc((void **)&c->Wb, sizeof(unsigned short) * ((size_t)c->hidden * c->k1p + (size_t)c->hidden * c->hidden +
                                                           (size_t)c->noutp * c->hidden)) != hipSuccess)
    return bad(LDG_ENOMEM, "device allocation failed (closure parameters)");
  if (capacity > 0 && (hipMalloc((void **)&c->Xd, (size_t)capacity * c->nin * sizeof(float)) != hipSuccess ||
                       hipMalloc((void **)&c->Td, (size_t)capacity * c->nout * sizeof(float)) != hipSuccess ||
                       hipMalloc((void **)&c->Wd, (size_t)capacity * sizeof(float)) != hipSuccess))
    return bad(LDG_ENOMEM, "device allocation failed (closure dataset)");
  // PyTorch-style uniform initialisation, bound 1/sqrt(fan_in); the output
  // layer scaled by 0.1 so that training starts near a zero correction
  std::vector<float> hp(c->nparam);
  std::mt19937_64 rng(seed);
  auto fill = [&](long off, long count, int fan_in, float scale) {
    const double bnd = scale / std::sqrt((double)fan_in);
    for (long i = 0; i < count; ++i) hp[off + i] = (float)((2.0 * ((rng() >> 11) * (1.0 / 9007199254740992.0)) - 1.0) * bnd);
  };
  fill(o.W1, o.b1 - o.W1, c->nin, 1.0f);
  fill(o.b1, c->hidden, c->nin, 1.0f);
  fill(o.W2, o.b2 - o.W2, c->hidden, 1.0f);
  fill(o.b2, c->hidden, c->hidden, 1.0f);
  fill(o.W3, o.b3 - o.W3, c->hidden, 0.1f);
  fill(o.b3, c->nout, c->hidden, 0.1f);
  std::vector<float> hn(2 * (c->nin + c->nout));
  for (int j = 0; j < c->nin; ++j) { hn[j] = 0.0f; hn[c->nin + j] = 1.0f; }
  for (int j = 0; j < c->nout; ++j) { hn[2 * c->nin + j] = 0.0f; hn[2 * c->nin + c->nout + j] = 1.0f; }
  if (hipMemcpy(c->params, hp.data(), pb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->norm, hn.data(), hn.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(c->m, 0, pb) != hipSuccess || hipMemset(c->v, 0, pb) != hipSuccess || pack_bf16(c) != LDG_OK)
    return bad(LDG_EDEVICE, "closure initialisation failed");
  *out = c;
  return LDG_OK;
}

void ldg_closure_destroy(ldg_closure *c) {
  if (!c) return;
  (void)hipStreamSynchronize(c->stream);
  free_closure(c);
}

int ldg_closure_sizes(const ldg_closure *c, int *nin, int *nout, int *hidden, long *nparam) {
  if (!c) return LDG_EINVAL;
  if (nin) *nin = c->nin;
  if (nout) *nout = c->nout;
  if (hidden) *hidden = c->hidden;
  if (nparam) *nparam = c->nparam;
  return LDG_OK;
}

int ldg_closure_get(ldg_closure *c, float *params, float *norm) {
  if (!c) return LDG_EINVAL;
  if (params) CCHECK(c, hipMemcpy(params, c->params, c->nparam * sizeof(float), hipMemcpyDeviceToHost));
  if (norm) CCHECK(c, hipMemcpy(norm, c->norm, 2 * (c->nin + c->nout) * sizeof(float), hipMemcpyDeviceToHost));
  return LDG_OK;
}

int ldg_closure_set(ldg_closure *c, const float *params, const float *norm) {
  if (!c) return LDG_EINVAL;
  if (params) {
    for (long i = 0; i < c->nparam; ++i)
      if (!std::isfinite(params[i])) return cfail(c, LDG_EINVAL, "non-finite parameter");
    CCHECK(c, hipMemcpy(c->params, params, c->nparam * sizeof(float), hipMemcpyHostToDevice));
    c->adam_t = 0;
    CCHECK(c, hipMemset(c->m, 0, c->nparam * sizeof(float)));
    CCHECK(c, hipMemset(c->v, 0, c->nparam * sizeof(float)));
    int rc = pack_bf16(c);
    if (rc) return rc;
  }
  if (norm) {
    const int nn = 2 * (c->nin + c->nout);
    for (int i = 0; i < nn; ++i)
      if (!std::isfinite(norm[i])) return cfail(c, LDG_EINVAL, "non-finite normalisation");
    CCHECK(c, hipMemcpy(c->norm, norm, nn * sizeof(float), hipMemcpyHostToDevice));
  }
  return LDG_OK;
}

int ldg_closure_forward(ldg_closure *c, const float *X, int B, float *Y) {
  if (!c || !X || !Y || B < 1) return LDG_EINVAL;
  int rc = ensure_workspace(c, B);
  if (rc) return rc;
  CCHECK(c, hipMemcpyAsync(c->Xb, X, (size_t)B * c->nin * sizeof(float), hipMemcpyHostToDevice, c->stream));
  forward(c, c->stream, B, c->Xb, nullptr, c->H1, nullptr, c->H2, c->Y);
  CCHECK(c, hipGetLastError());
  CCHECK(c, hipMemcpyAsync(Y, c->Y, (size_t)B * c->nout * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  CCHECK(c, hipStreamSynchronize(c->stream));
  return LDG_OK;
}

int ldg_closure_gradient(ldg_closure *c, const float *X, const float *T, int B, float *grad, double *loss) {
  if (!c || !X || !T || B < 1) return LDG_EINVAL;
  int rc = ensure_workspace(c, B);
  if (rc) return rc;
  CCHECK(c, hipMemcpyAsync(c->Xb, X, (size_t)B * c->nin * sizeof(float), hipMemcpyHostToDevice, c->stream));
  CCHECK(c, hipMemcpyAsync(c->Tb, T, (size_t)B * c->nout * sizeof(float), hipMemcpyHostToDevice, c->stream));
  backward(c, c->stream, B, c->Xb, c->Tb, nullptr);
  CCHECK(c, hipGetLastError());
  if (grad) CCHECK(c, hipMemcpyAsync(grad, c->grads, c->nparam * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  std::vector<double> part(kLossBlocks);
  CCHECK(c, hipMemcpyAsync(part.data(), c->part, kLossBlocks * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CCHECK(c, hipStreamSynchronize(c->stream));
  if (loss) *loss = std::accumulate(part.begin(), part.end(), 0.0) / ((double)B * c->nout);
  return LDG_OK;
}

int ldg_project_state(ldg_handle *fine, int i, ldg_handle *coarse, int j) {
  if (!fine || !coarse || i < 0 || j < 0 || i >= (int)fine->inst.size() || j >= (int)coarse->inst.size())
    return LDG_EINVAL;
  const Inst &a = fine->inst[i], &b = coarse->inst[j];
  if (a.nx != b.nx || a.ny != b.ny || a.nz != b.nz || b.p > a.p)
    return fail(coarse, LDG_EINVAL, "projection needs the same mesh and a coarse order <= the fine order");
  HIPCHECK(fine, hipStreamSynchronize(fine->stream));
  std::string err;
  const int rc = project_device(coarse->stream, a.p, b.p, a.K, fine->groups[a.group].u + inst_offset(fine, i),
                                coarse->groups[b.group].u + inst_offset(coarse, j), err);
  if (rc) return fail(coarse, rc, err);
  return LDG_OK;
}

int ldg_closure_collect(ldg_closure *c, ldg_handle *fine, ldg_handle *coarse, double keep, unsigned long long seed) {
  if (!c || !fine || !coarse || !(keep > 0) || keep > 1) return LDG_EINVAL;
  const int ni = (int)fine->inst.size();
  if ((int)coarse->inst.size() != ni) return cfail(c, LDG_EINVAL, "fine and coarse handles differ in instance count");
  for (int i = 0; i < ni; ++i) {
    const Inst &a = fine->inst[i], &b = coarse->inst[i];
    if (b.p != c->p || a.p < b.p || a.nx != b.nx || a.ny != b.ny || a.nz != b.nz)
      return cfail(c, LDG_EINVAL, "instance pairs need the same mesh, the closure's order on the coarse side");
  }
  if (c->normalized) return cfail(c, LDG_EINVAL, "the dataset was already normalised; collect before fitting");
  // fine right-hand sides k_f = rhs(u_f) of every instance (plain, no closure)
  for (Group &g : fine->groups) {
    if (launch_grad(fine, g, state_in(g.u))) return cfail(c, LDG_EDEVICE, fine->err);
    StageArgs st = empty_stage();
    st.kout = g.k;
    if (launch_rhs(fine, g, state_in(g.u), st)) return cfail(c, LDG_EDEVICE, fine->err);
  }
  CCHECK(c, hipStreamSynchronize(fine->stream));
  // project u_f -> coarse u, k_f -> coarse acc (FP64 scratch for P k_f)
  for (int i = 0; i < ni; ++i) {
    const Inst &a = fine->inst[i], &b = coarse->inst[i];
    const Group &gf = fine->groups[a.group];
    Group &gc = coarse->groups[b.group];
    std::string err;
    int rc = project_device(coarse->stream, a.p, b.p, a.K, gf.u + inst_offset(fine, i), gc.u + inst_offset(coarse, i), err);
    if (!rc) rc = project_device(coarse->stream, a.p, b.p, a.K, gf.k + inst_offset(fine, i), gc.acc + inst_offset(coarse, i), err);
    if (rc) return cfail(c, rc, err);
  }
  // coarse right-hand side of the projected state, features and targets
  std::mt19937_64 rng(seed);
  for (Group &g : coarse->groups) {
    if (launch_grad(coarse, g, state_in(g.u), -1, 1)) return cfail(c, LDG_EDEVICE, coarse->err);
    StageArgs st = empty_stage();
    st.kout = g.k;
    if (launch_rhs(coarse, g, state_in(g.u), st)) return cfail(c, LDG_EDEVICE, coarse->err);
    float *Xf = nullptr, *Tf = nullptr;
    long *didx = nullptr;
    if (hipMalloc((void **)&Xf, (size_t)g.K * c->nin * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&Tf, (size_t)g.K * c->nout * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&didx, (size_t)g.K * sizeof(long)) != hipSuccess) {
      if (Xf) (void)hipFree(Xf);
      if (Tf) (void)hipFree(Tf);
      return cfail(c, LDG_ENOMEM, "device allocation failed (collection)");
    }
    // raw features: identity normalisation
    std::vector<float> ident(2 * c->nin);
    for (int j = 0; j < c->nin; ++j) { ident[j] = 0.0f; ident[c->nin + j] = 1.0f; }
    float *dident = nullptr;
    hipError_t e = hipMalloc((void **)&dident, ident.size() * sizeof(float));
    if (e == hipSuccess) e = hipMemcpyAsync(dident, ident.data(), ident.size() * sizeof(float), hipMemcpyHostToDevice, coarse->stream);
    if (e == hipSuccess) {
      features_group(coarse, g, state_in(g.u), dident, c->nin, Xf);
      hipLaunchKernelGGL(mlp_targets, dim3(grid_for((long)g.nU)), dim3(256), 0, coarse->stream, (const double *)g.acc,
                         (const double *)g.k, (long)g.nU, Tf);
      e = hipGetLastError();
    }
    // subsample: round(keep K) elements without replacement
    long take = std::min<long>((long)std::llround(keep * g.K), c->cap - c->count);
    if (e == hipSuccess && take > 0) {
      std::vector<long> perm(g.K);
      std::iota(perm.begin(), perm.end(), 0L);
      for (long t = 0; t < take; ++t) {
        const long r = t + (long)(rng() % (unsigned long long)(g.K - t));
        std::swap(perm[t], perm[r]);
      }
      std::sort(perm.begin(), perm.begin() + take);
      e = hipMemcpyAsync(didx, perm.data(), take * sizeof(long), hipMemcpyHostToDevice, coarse->stream);
      if (e == hipSuccess) {
        hipLaunchKernelGGL(mlp_gather, dim3(grid_for(take * c->nin)), dim3(256), 0, coarse->stream, (const float *)Xf,
                           c->nin, (const long *)didx, (int)take, c->Xd + (size_t)c->count * c->nin);
        hipLaunchKernelGGL(mlp_gather, dim3(grid_for(take * c->nout)), dim3(256), 0, coarse->stream, (const float *)Tf,
                           c->nout, (const long *)didx, (int)take, c->Td + (size_t)c->count * c->nout);
        e = hipGetLastError();
      }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(coarse->stream);
    (void)hipFree(Xf);
    (void)hipFree(Tf);
    (void)hipFree(didx);
    if (dident) (void)hipFree(dident);
    if (e != hipSuccess) return cfail(c, LDG_EDEVICE, std::string("collection: ") + hipGetErrorString(e));
    if (take > 0) {
      c->group_of.insert(c->group_of.end(), take, c->ncollect);
      c->count += take;
    }
  }
  ++c->ncollect;
  return LDG_OK;
}

int ldg_closure_add_samples(ldg_closure *c, const float *X, const float *T, long n) {
  if (!c || !X || !T || n < 0) return LDG_EINVAL;
  if (c->normalized) return cfail(c, LDG_EINVAL, "the dataset was already normalised; add samples before fitting");
  if (c->count + n > c->cap) return cfail(c, LDG_EINVAL, "dataset capacity exceeded");
  CCHECK(c, hipMemcpy(c->Xd + (size_t)c->count * c->nin, X, (size_t)n * c->nin * sizeof(float), hipMemcpyHostToDevice));
  CCHECK(c, hipMemcpy(c->Td + (size_t)c->count * c->nout, T, (size_t)n * c->nout * sizeof(float), hipMemcpyHostToDevice));
  c->group_of.insert(c->group_of.end(), n, c->ncollect++);
  c->count += n;
  return LDG_OK;
}

int ldg_closure_set_weighting(ldg_closure *c, int mode) {
  if (!c || mode < 0 || mode > 1) return LDG_EINVAL;
  if (c->normalized) return cfail(c, LDG_EINVAL, "choose the weighting before fitting the normalisation");
  c->weighting = mode;
  return LDG_OK;
}

int ldg_closure_get_weights(ldg_closure *c, long first, long n, float *w) {
  if (!c || !w || first < 0 || n < 0 || first + n > c->count) return LDG_EINVAL;
  if (!c->normalized) return cfail(c, LDG_EINVAL, "weights are set when the normalisation is fitted");
  CCHECK(c, hipMemcpy(w, c->Wd + first, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
  return LDG_OK;
}

int ldg_closure_samples(const ldg_closure *c, long *n) {
  if (!c || !n) return LDG_EINVAL;
  *n = c->count;
  return LDG_OK;
}

int ldg_closure_get_samples(ldg_closure *c, long first, long n, float *X, float *T) {
  if (!c || first < 0 || n < 0 || first + n > c->count) return LDG_EINVAL;
  if (X) CCHECK(c, hipMemcpy(X, c->Xd + first * c->nin, (size_t)n * c->nin * sizeof(float), hipMemcpyDeviceToHost));
  if (T) CCHECK(c, hipMemcpy(T, c->Td + first * c->nout, (size_t)n * c->nout * sizeof(float), hipMemcpyDeviceToHost));
  return LDG_OK;
}

int ldg_closure_fit_normalization(ldg_closure *c) {
  if (!c || c->count < 2) return cfail(c, LDG_EINVAL, "need at least 2 samples");
  if (c->normalized) return cfail(c, LDG_EINVAL, "already normalised");
  std::vector<float> hn(2 * (c->nin + c->nout));
  const int widths[2] = {c->nin, c->nout};
  float *data[2] = {c->Xd, c->Td};
  const long chunk = 4096;
  const long nch = (c->count + chunk - 1) / chunk;
  for (int t = 0; t < 2; ++t) {
    const int w = widths[t];
    double *part = nullptr;
    CCHECK(c, hipMalloc((void **)&part, sizeof(double) * 2 * w * nch));
    hipLaunchKernelGGL(mlp_moments, dim3((w + 63) / 64, (unsigned)nch), dim3(64), 0, c->stream, (const float *)data[t],
                       c->count, w, chunk, part);
    std::vector<double> hp(2 * w * nch);
    hipError_t e = hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(part);
    if (e != hipSuccess) return cfail(c, LDG_EDEVICE, "moments failed");
    float *mean = hn.data() + (t == 0 ? 0 : 2 * c->nin), *sc = mean + w;
    for (int j = 0; j < w; ++j) {
      double s = 0, s2 = 0;
      for (long k = 0; k < nch; ++k) { s += hp[(k * 2 + 0) * w + j]; s2 += hp[(k * 2 + 1) * w + j]; }
      const double mu = s / c->count;
      const double var = std::max(0.0, s2 / c->count - mu * mu);
      double sd = std::sqrt(var);
      if (!(sd > 1e-30)) sd = 1.0;
      mean[j] = (float)mu;
      sc[j] = t == 0 ? (float)(1.0 / sd) : (float)sd;   // inputs store 1/std, outputs std
    }
  }
  CCHECK(c, hipMemcpy(c->norm, hn.data(), hn.size() * sizeof(float), hipMemcpyHostToDevice));
  // normalise the dataset in place: X scaled by 1/std, T by 1/std
  std::vector<float> oinv(c->nout);
  for (int j = 0; j < c->nout; ++j) oinv[j] = 1.0f / hn[2 * c->nin + c->nout + j];
  float *doinv = nullptr;
  CCHECK(c, hipMalloc((void **)&doinv, c->nout * sizeof(float)));
  hipError_t e = hipMemcpy(doinv, oinv.data(), c->nout * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(mlp_normalize, dim3(grid_for(c->count * c->nin)), dim3(256), 0, c->stream, c->Xd, c->count,
                       c->nin, (const float *)c->norm, (const float *)(c->norm + c->nin));
    hipLaunchKernelGGL(mlp_normalize, dim3(grid_for(c->count * c->nout)), dim3(256), 0, c->stream, c->Td, c->count,
                       c->nout, (const float *)(c->norm + 2 * c->nin), (const float *)doinv);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(doinv);
  if (e != hipSuccess) return cfail(c, LDG_EDEVICE, "normalisation failed");
  // loss weights: uniform, or 1 / (mean normalised target energy of the
  // sample's collection), scaled to mean 1 over the samples
  std::vector<float> w(c->count, 1.0f);
  if (c->weighting == 1) {
    float *ms = nullptr;
    CCHECK(c, hipMalloc((void **)&ms, c->count * sizeof(float)));
    hipLaunchKernelGGL(mlp_rowmeansq, dim3(grid_for(c->count)), dim3(256), 0, c->stream, (const float *)c->Td,
                       c->count, c->nout, ms);
    std::vector<float> hm(c->count);
    e = hipMemcpyAsync(hm.data(), ms, c->count * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(ms);
    if (e != hipSuccess) return cfail(c, LDG_EDEVICE, "weights failed");
    std::vector<double> gs(c->ncollect, 0.0), gn(c->ncollect, 0.0);
    for (long i = 0; i < c->count; ++i) { gs[c->group_of[i]] += hm[i]; gn[c->group_of[i]] += 1.0; }
    double tot = 0.0;
    for (long i = 0; i < c->count; ++i) {
      const double m = gs[c->group_of[i]] / gn[c->group_of[i]];
      w[i] = (float)(m > 0 ? 1.0 / m : 1.0);
      tot += w[i];
    }
    for (long i = 0; i < c->count; ++i) w[i] = (float)(w[i] * (c->count / tot));
  }
  CCHECK(c, hipMemcpy(c->Wd, w.data(), c->count * sizeof(float), hipMemcpyHostToDevice));
  c->normalized = true;
  return LDG_OK;
}

int ldg_closure_train(ldg_closure *c, int epochs, int batch, double lr, double val_fraction, unsigned long long seed,
                      double *losses) {
  if (!c || epochs < 1 || batch < 1 || !(lr > 0) || val_fraction < 0 || val_fraction >= 1) return LDG_EINVAL;
  const long nval = (long)std::floor(val_fraction * c->count);
  const long ntrain = c->count - nval;
  if (ntrain < batch) return cfail(c, LDG_EINVAL, "fewer training samples than one batch");
  int rc = ensure_workspace(c, batch);
  if (rc) return rc;
  // a seeded random hold-out of nval samples; the rest is reshuffled every epoch
  std::mt19937_64 rng(seed);
  std::vector<long> all(c->count);
  std::iota(all.begin(), all.end(), 0L);
  std::shuffle(all.begin(), all.end(), rng);
  std::vector<long> perm(all.begin(), all.begin() + ntrain);
  long *didx = nullptr, *dval = nullptr;
  CCHECK(c, hipMalloc((void **)&didx, std::max<long>(ntrain, 1) * sizeof(long)));
  if (hipMalloc((void **)&dval, std::max<long>(nval, 1) * sizeof(long)) != hipSuccess) {
    (void)hipFree(didx);
    return cfail(c, LDG_ENOMEM, "device allocation failed (training indices)");
  }
  hipError_t e = nval ? hipMemcpy(dval, all.data() + ntrain, nval * sizeof(long), hipMemcpyHostToDevice) : hipSuccess;
  const float *wsrc = c->normalized ? c->Wd : nullptr;     // per-sample weights (set with the normalisation)
  std::vector<double> part(kLossBlocks);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  auto gather = [&](const long *idx, int rows) {
    hipLaunchKernelGGL(mlp_gather, dim3(grid_for((long)rows * c->nin)), dim3(256), 0, c->stream,
                       (const float *)c->Xd, c->nin, idx, rows, c->Xb);
    hipLaunchKernelGGL(mlp_gather, dim3(grid_for((long)rows * c->nout)), dim3(256), 0, c->stream,
                       (const float *)c->Td, c->nout, idx, rows, c->Tb);
    if (wsrc)
      hipLaunchKernelGGL(mlp_gather, dim3(grid_for(rows)), dim3(256), 0, c->stream, wsrc, 1, idx, rows, c->Wbat);
  };
  for (int ep = 0; ep < epochs && e == hipSuccess; ++ep) {
    std::shuffle(perm.begin(), perm.end(), rng);
    e = hipMemcpy(didx, perm.data(), ntrain * sizeof(long), hipMemcpyHostToDevice);
    double tl = 0.0;
    long nb = 0;
    for (long b0 = 0; b0 + batch <= ntrain && e == hipSuccess; b0 += batch, ++nb) {
      gather(didx + b0, batch);
      backward(c, c->stream, batch, c->Xb, c->Tb, wsrc ? c->Wbat : nullptr);
      ++c->adam_t;
      const float bc1 = 1.0f - (float)std::pow(b1, (double)c->adam_t), bc2 = 1.0f - (float)std::pow(b2, (double)c->adam_t);
      hipLaunchKernelGGL(mlp_adam, dim3(grid_for(c->nparam)), dim3(256), 0, c->stream, c->params,
                         (const float *)c->grads, c->m, c->v, c->nparam, (float)lr, b1, b2, eps, bc1, bc2);
      e = hipMemcpyAsync(part.data(), c->part, kLossBlocks * sizeof(double), hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      tl += std::accumulate(part.begin(), part.end(), 0.0) / ((double)batch * c->nout);
    }
    double vl = 0.0;
    long vn = 0;
    for (long v0 = 0; v0 < nval && e == hipSuccess; v0 += batch) {
      const int rows = (int)std::min<long>(batch, nval - v0);
      gather(dval + v0, rows);
      forward(c, c->stream, rows, c->Xb, nullptr, c->H1, nullptr, c->H2, c->Y);
      hipLaunchKernelGGL(mlp_loss_grad, dim3(kLossBlocks), dim3(256), 0, c->stream, (const float *)c->Y,
                         (const float *)c->Tb, (long)rows * c->nout, c->nout,
                         (const float *)(wsrc ? c->Wbat : nullptr), 0.0f, (float *)nullptr, c->part);
      e = hipMemcpyAsync(part.data(), c->part, kLossBlocks * sizeof(double), hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      vl += std::accumulate(part.begin(), part.end(), 0.0);
      vn += (long)rows * c->nout;
    }
    if (losses) {
      losses[2 * ep] = nb ? tl / nb : 0.0;
      losses[2 * ep + 1] = vn ? vl / vn : 0.0;
    }
  }
  (void)hipFree(didx);
  (void)hipFree(dval);
  if (e != hipSuccess) return cfail(c, LDG_EDEVICE, std::string("training: ") + hipGetErrorString(e));
  return pack_bf16(c);
}

int ldg_set_closure(ldg_handle *h, const ldg_closure *c, double bound, int flags) {
  const int both = LDG_CLOSURE_DISSIPATIVE | LDG_CLOSURE_ENERGIZING;
  if (!h || (c && !(bound >= 0)) || (flags & ~(both | LDG_CLOSURE_PER_STEP)) || (flags & both) == both)
    return LDG_EINVAL;
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  drop_graph(h);
  for (Group &g : h->groups) {
    closure_release(g);
    if (!c || g.n != c->n) continue;
    GroupClosure &cl = g.cl;
    const size_t K = (size_t)g.K;
    if (hipMalloc((void **)&cl.X, K * c->nin * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&cl.H1, K * c->hidden * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&cl.H2, K * c->hidden * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&cl.Y, K * c->nout * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&cl.corr, K * c->nout * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&cl.gate, K) != hipSuccess) {
      closure_release(g);
      return fail(h, LDG_ENOMEM, "device allocation failed (closure buffers)");
    }
    cl.net = c;
    cl.bound = bound;
    // the energy gate's sign: +1 keeps corrections that do not raise the
    // element's kinetic energy, -1 those that do not lower it, 0 none
    cl.dissipative = (flags & LDG_CLOSURE_DISSIPATIVE) ? 1 : (flags & LDG_CLOSURE_ENERGIZING) ? -1 : 0;
    cl.per_step = (flags & LDG_CLOSURE_PER_STEP) != 0;
  }
  return LDG_OK;
}

int ldg_closure_gated(ldg_handle *h, long *off, long *total) {
  if (!h) return LDG_EINVAL;
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  long o = 0, t = 0;
  for (Group &g : h->groups) {
    if (!g.cl.net) continue;
    std::vector<unsigned char> gh(g.K);
    HIPCHECK(h, hipMemcpy(gh.data(), g.cl.gate, g.K, hipMemcpyDeviceToHost));
    for (unsigned char x : gh) o += (x == 0);
    t += g.K;
  }
  if (off) *off = o;
  if (total) *total = t;
  return LDG_OK;
}

}  // extern "C"
