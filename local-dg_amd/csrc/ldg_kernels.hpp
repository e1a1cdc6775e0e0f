// ldg_kernels.hpp -- HIP kernels of the LDG right-hand side for gfx950.
//
// One RK4 stage is two launches over the whole periodic mesh:
//   ldg_grad_kernel  auxiliary gradient Q_l = (1/alpha_l)(Dn U + L0 jump_l)
//                    (Solver::rk4Rhs first half, solver.cpp:490-508) on the
//                    faces, the Roe + one-sided viscous face fluxes of both
//                    elements of every face (solver.cpp:643-698, 750-814,
//                    891-921, 1194-1374), lifted to the face nodes
//   ldg_rhs_kernel   the volume gradient and volume integrals, minus the
//                    lifted face terms, inverse mass and the fused RK4
//                    stage epilogue (solver.cpp:510-535 and 410-434)
// Both stage a block of E consecutive elements in LDS, work in 1D "lines"
// of N = p+1 values owned by one thread (in-place sum-factorised sweeps,
// with the even/odd split of the centro-symmetric 1D operators, symop.hpp),
// and read the periodic neighbours' face slices straight from HBM (the
// on-device form of the single-rank halo exchange, solver.cpp:929-998).
//
// Precision: the state, the RK4 combination, the inverse mass and every
// reduction are FP64.  The interpolation, flux evaluation and test-function
// contractions run in the compute type T (double for the plain path, float
// for the mixed path), on fluctuations about the constant background
// (rho0, 0, 0, 0, E0 = p0/(gamma-1)): the background is subtracted in FP64
// when U is loaded (exact, both values are close), so the O(1/M0^2) part of
// the energy never enters the T-precision arithmetic.  A constant state is a
// discrete steady state, so the fluctuation form changes nothing but
// rounding.
//
// Data in HBM:
//   u0, acc, unext : FP64 (node, state, element) -- the reference layout
//   U (stage input): the same, or in the mixed path (RK stages 2-4) FP32
//                    fluctuations about the background in that layout
//   FF             : T, per element [direction][side][state][face node]:
//                    the lifted face fluxes (gradient kernel -> rhs kernel)
//   Q              : TQ (node, s + 5*l, element) -- the nodal gradient, 15
//                    fields per element, written only when asked for
#pragma once

#include <hip/hip_runtime.h>

#include "symop.hpp"

namespace ldg {

constexpr int NSTATE = 5;
constexpr int NGRAD = 15;
constexpr int NT = 256;          // threads per workgroup

// 1D operators (row-major [row][col]), FP64, in device memory.
template <int N, typename T>
struct Ops {
  T I[N][N];      // I[q][i]  = l_i(g_q)
  T D[N][N];      // D[q][i]  = l_i'(g_q)
  T Dn[N][N];     // Dn[i][j] = l_j'(r_i)
  T Minv[N][N];   // inverse 1D mass matrix
  T L0[N];        // Minv[:, 0]: lift of the jump at the left face
  T IL0[N];       // I L0: that lift at the Gauss points
  T w[N];         // Gauss weights
  // the sweep operators in even/odd split form: I, I^T, D^T, Minv, D
  SymMat<N, T> sI, sIt, sDt, sMi, sD;
};

struct MeshArgs {
  int nx, ny, nz;
  long nelem;
  int nblocks;           // logical workgroups (ceil(nelem / E))
  int per_xcd;           // ceil(nblocks / 8)
  double ialpha[3];      // 2 / dx_l
};

struct PhysArgs {
  double gamma, gm1, mu, kappa;   // kappa = mu gamma / Pr acting on grad(e)
  double rho0, E0, H0;            // background: E0 = p0/(gamma-1), H0 = E0 + p0
};

struct StageArgs {
  const double *u0;   // u^n (RK base state)
  double *acc;        // running u^n + dt sum b_s k_s
  double *unext;      // next stage input u^n + dt a_s k_s
  float *unext32;     // or, in the mixed path, that input as FP32 fluctuations about the background
  float *next32;      // stage 4 (mixed): the new running sum u^{n+1} as FP32 fluctuations, the
                      // next step's stage-1 input when that step follows in the same call
  double *kout;       // raw k (parity / single RHS)
  double cb, ca;      // RK4 coefficients b_s, a_s (times each instance's dt)
  int acc_init;       // 1: acc = u0 + dt b k   0: acc += dt b k
  double *ke_part;    // per-workgroup partial sums of the reference KE of acc
  double *diag_part;  // per-workgroup partial sums (Gauss KE, enstrophy)
  const double *corr;          // learned closure: k += corr where gate[element] != 0
  const unsigned char *gate;
};

// One solver instance inside a launch.  A launch covers the elements of all
// instances of one order (an ensemble); a single solver is a launch of one
// instance.  Instances are stored back to back in the launch's arrays.
// n / d for 0 <= n < 2^31 by one 32x32->64-bit multiply and a shift
// (Granlund and Montgomery, "Division by invariant integers using
// multiplication", 1994): with l = ceil(log2 d), s = 31 + l and
// m = ceil(2^s / d) < 2^32, floor(n m / 2^s) = floor(n / d) because the
// error n (m d - 2^s) / (d 2^s) < n / 2^s <= 2^-l <= 1/d.  Built on the host;
// with uniform operands the device evaluates it on the scalar unit, where a
// plain division by a run-time value would take a dozen vector instructions.
struct FastDiv {
  unsigned m;
  int s;
  __host__ __device__ inline int div(int n) const {
    return (int)(((unsigned long long)(unsigned)n * m) >> s);
  }
};
inline FastDiv make_fastdiv(int d) {
  int l = 0;
  while ((1LL << l) < d) ++l;
  FastDiv f;
  f.s = 31 + l;
  f.m = (unsigned)(((1ULL << f.s) + (unsigned long long)d - 1) / (unsigned long long)d);
  return f;
}

struct InstDesc {
  int nx, ny, nz;
  int nblocks;           // its logical workgroups (whole bricks, see block_elements)
  long nelem;            // elements of this instance
  long elem_off;         // its first element in the launch's arrays
  int block_off;         // its first logical workgroup
  int nbx;               // element blocks per x-row: ceil(nx / E)
  int tb[3];             // brick of tb[0] x tb[1] x tb[2] blocks (x blocks, y rows, z planes)
  int nbrick[3];         // bricks per direction
  FastDiv div_bs, div_nb0, div_nb1, div_tb0, div_tb1;   // by tb0 tb1 tb2, nbrick[0], nbrick[1], tb[0], tb[1]
  double ialpha[3];      // 2 / dx_l
  double dt;             // its time step
  PhysArgs ph;
};

// Elements of an instance's logical workgroup lbl (0-based within the
// instance): blocks are E consecutive elements of one x-row, and blocks are
// visited brick by brick (tb blocks per brick, bricks x-fastest), so that the
// workgroups resident at the same time on an XCD cover a compact region and
// their face neighbours in y and z are mostly in that XCD's L2.  Returns the
// coordinates of the first element and the count, 0 for a padding block of
// a partial brick.
__host__ __device__ inline void block_elements(const InstDesc &d, int E, int lbl, int ec[3], int &ne) {
  const int bs = d.tb[0] * d.tb[1] * d.tb[2];
  const int brick = d.div_bs.div(lbl), off = lbl - brick * bs;
  const int t = d.div_nb0.div(brick), Bx = brick - t * d.nbrick[0];
  const int Bz = d.div_nb1.div(t), By = t - Bz * d.nbrick[1];
  const int u = d.div_tb0.div(off), ox = off - u * d.tb[0];
  const int oz = d.div_tb1.div(u), oy = u - oz * d.tb[1];
  const int bx = Bx * d.tb[0] + ox, ey = By * d.tb[1] + oy, ez = Bz * d.tb[2] + oz;
  ec[0] = ec[1] = ec[2] = 0;
  ne = 0;
  if (bx >= d.nbx || ey >= d.ny || ez >= d.nz) return;
  ec[0] = bx * E; ec[1] = ey; ec[2] = ez;
  ne = (d.nx - ec[0]) < E ? (d.nx - ec[0]) : E;
}

struct LaunchArgs {
  const InstDesc *inst;
  const int *blk_inst;   // logical workgroup -> instance (nullptr: one instance)
  int nblocks;           // logical workgroups of the launch
  int per_xcd;           // ceil(nblocks / 8)
};

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of element blocks so that face neighbours in x and y
// share that XCD's L2.  Returns -1 for padding workgroups.
__device__ inline int logical_block(const LaunchArgs &la) {
  const int b = blockIdx.x;
  const int lb = (b & 7) * la.per_xcd + (b >> 3);
  return lb < la.nblocks ? lb : -1;
}

// The instance of logical workgroup lb, its mesh, and the block's first
// element (coordinates ec and index kl, local to the instance) and element
// count.
template <int E>
__device__ inline const InstDesc &block_instance(const LaunchArgs &la, int lb, MeshArgs &m, int ec[3], long &kl,
                                                 int &ne) {
  const InstDesc &d = la.inst[la.blk_inst ? la.blk_inst[lb] : 0];
  m.nx = d.nx; m.ny = d.ny; m.nz = d.nz;
  m.nelem = d.nelem;
  m.ialpha[0] = d.ialpha[0]; m.ialpha[1] = d.ialpha[1]; m.ialpha[2] = d.ialpha[2];
  block_elements(d, E, lb - d.block_off, ec, ne);
  kl = ec[0] + (long)d.nx * (ec[1] + (long)d.ny * ec[2]);
  return d;
}

// c + dir (dir = +-1) on a periodic axis of n elements
__device__ inline int wrap_step(int c, int dir, int n) {
  c += dir;
  return c < 0 ? c + n : (c >= n ? c - n : c);
}

// Face neighbours of the block's elements (periodic; the reference builds
// the same table as eToE, mesh.cpp:201-315, with ghost slots on the faces),
// computed once per workgroup from the first element's coordinates ec:
// nb[e][f] for faces f = -x,+x,-y,+y,-z,+z.
__device__ inline void fill_neighbors(const MeshArgs &m, const int ec[3], long off, int ne, long (*nb)[6], int tid) {
  if (tid < ne * 6) {
    const int e = tid / 6, f = tid % 6, dim = f >> 1, dir = (f & 1) ? 1 : -1;
    int x = ec[0] + e, y = ec[1], z = ec[2];
    if (dim == 0) x = wrap_step(x, dir, m.nx);
    else if (dim == 1) y = wrap_step(y, dir, m.ny);
    else z = wrap_step(z, dir, m.nz);
    nb[e][f] = off + x + (long)m.nx * (y + (long)m.ny * z);
  }
}

// node index of position i along direction d in the line with face
// coordinates (a, b): x-lines (iy, iz), y-lines (ix, iz), z-lines (ix, iy)
// -- the same (a, b) order as the reference's face map (mesh.cpp:428-501).
template <int N>
__device__ inline int line_base(int d, int a, int b) {
  return d == 0 ? (N * a + N * N * b) : d == 1 ? (a + N * N * b) : (a + N * b);
}
template <int N>
__device__ inline int line_stride(int d) { return d == 0 ? 1 : d == 1 ? N : N * N; }

// ---------------------------------------------------------------------
// Pointwise physics in fluctuation form, textbook forms of
// solver.cpp:1040-1374.  A state w = (rho - rho0, rho u, rho v, rho w,
// E - E0); p' = p - p0 = (gamma-1)(w4 - |m|^2/(2 rho)) and
// E + p = H0 + w4 + p'.
// ---------------------------------------------------------------------
// Reciprocal and square root: correctly rounded in FP64; in FP32 (the
// mixed path, whose error budget is set by FP32 rounding anyway) the
// hardware approximations v_rcp_f32 / v_sqrt_f32 (1 ulp).
__device__ inline double rcp(double x) { return 1.0 / x; }
__device__ inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double sqrt_(double x) { return sqrt(x); }
__device__ inline float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }

template <typename T>
struct Bg {
  T rho0, E0, H0, gm1, mu, kappa;
  __device__ inline explicit Bg(const PhysArgs &p)
      : rho0(T(p.rho0)), E0(T(p.E0)), H0(T(p.H0)), gm1(T(p.gm1)), mu(T(p.mu)), kappa(T(p.kappa)) {}
};

// Velocity gradient gv[i][l] and internal-energy gradient ge[l] from the
// state (density rho = 1/ir, velocity v, total energy E) and the conserved
// gradients g[s + 5 l].
template <typename T>
__device__ inline void prim_grads(T ir, T E, const T v[3], const T *g, T gv[3][3], T ge[3]) {
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    const T drho = g[5 * l];
#pragma unroll
    for (int i = 0; i < 3; ++i) gv[i][l] = (g[1 + i + 5 * l] - v[i] * drho) * ir;
    const T dEr = (g[4 + 5 * l] - E * ir * drho) * ir;
    ge[l] = dEr - (v[0] * gv[0][l] + v[1] * gv[1][l] + v[2] * gv[2][l]);
  }
}

// Volume flux F[s + 5 l] = Fc_l(u) - Fv_l(u, g) (fluxC minus fluxV,
// solver.cpp:1040-1188).  The momentum flux carries p' = p - p0: a constant
// pressure contributes exactly nothing to the discrete residual (the Gauss
// rules integrate d(phi)/dx and the face traces of phi exactly).
template <typename T>
__device__ inline void volume_flux(const T w[5], const T *g, const Bg<T> &b, T F[15]) {
  const T rho = b.rho0 + w[0];
  const T ir = rcp(rho);
  const T v[3] = {w[1] * ir, w[2] * ir, w[3] * ir};
  const T pf = b.gm1 * (w[4] - T(0.5) * (w[1] * v[0] + w[2] * v[1] + w[3] * v[2]));
  const T hv = b.H0 + w[4] + pf;   // E + p
  T gv[3][3], ge[3];
  prim_grads<T>(ir, b.E0 + w[4], v, g, gv, ge);
  const T div = gv[0][0] + gv[1][1] + gv[2][2];
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    T tau[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) tau[i] = b.mu * (gv[i][l] + gv[l][i]) - (i == l ? b.mu * T(2.0 / 3.0) * div : T(0));
    F[5 * l] = w[1 + l];
#pragma unroll
    for (int i = 0; i < 3; ++i) F[1 + i + 5 * l] = w[1 + i] * v[l] + (i == l ? pf : T(0)) - tau[i];
    F[4 + 5 * l] = hv * v[l] - (v[0] * tau[0] + v[1] * tau[1] + v[2] * tau[2] + b.kappa * ge[l]);
  }
}

// Fluxes of one face, for both elements that share it.  The face has the
// element K on its - side and the element N on its + side (normal +e_D out
// of K); wK, wN are the two traces and q the gradient trace chosen by the
// one-sided switch, which is N's own on this face for both sides (the
// switch takes the + side element's gradient: the neighbour's for K, whose
// normal sum is positive, and N's own for N).  Returns K's flux
//   outK = Roe(uN, uK, +e_D) - Fv_D(uK, q)
// (numericalFluxC, solver.cpp:1194-1374; numericalFluxV, solver.cpp:891-921)
// and N's flux on the same face, whose normal is -e_D,
//   outN = Roe(uK, uN, -e_D) + Fv_D(uN, q) = -Roe(uN, uK, +e_D) + Fv_D(uN, q):
// the Roe flux is evaluated once and enters the two sides with opposite
// signs (F*(a, b, n) = -F*(b, a, -n): the Roe average is symmetric in the
// two states and |A(-n)| = |A(n)|), while the viscous flux is evaluated on
// each side's own trace, as the reference does.
template <typename T, int D>
__device__ inline void visc_face(const T w[5], T ir, const T v[3], const T *q, const Bg<T> &b, T fv[4]) {
  T gv[3][3], ge[3];
  prim_grads<T>(ir, b.E0 + w[4], v, q, gv, ge);
  const T div = gv[0][0] + gv[1][1] + gv[2][2];
  T tau[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) tau[i] = b.mu * (gv[i][D] + gv[D][i]) - (i == D ? b.mu * T(2.0 / 3.0) * div : T(0));
#pragma unroll
  for (int i = 0; i < 3; ++i) fv[i] = tau[i];
  fv[3] = v[0] * tau[0] + v[1] * tau[1] + v[2] * tau[2] + b.kappa * ge[D];
}

template <typename T, int D>
__device__ inline void face_flux_pair(const T wK[5], const T wN[5], const T *q, const Bg<T> &b, T outK[5],
                                      T outN[5]) {
  const T rN = b.rho0 + wN[0], rK = b.rho0 + wK[0];
  const T irN = rcp(rN), irK = rcp(rK);
  const T vN[3] = {wN[1] * irN, wN[2] * irN, wN[3] * irN};
  const T vK[3] = {wK[1] * irK, wK[2] * irK, wK[3] * irK};
  const T pN = b.gm1 * (wN[4] - T(0.5) * (wN[1] * vN[0] + wN[2] * vN[1] + wN[3] * vN[2]));   // p - p0
  const T pK = b.gm1 * (wK[4] - T(0.5) * (wK[1] * vK[0] + wK[2] * vK[1] + wK[3] * vK[2]));
  const T hN = b.H0 + wN[4] + pN, hK = b.H0 + wK[4] + pK;                                   // E + p
  const T HN = hN * irN, HK = hK * irK;
  const T qN = vN[D], qK = vK[D];
  const T s = sqrt_(rN * irK);
  const T iden = rcp(s + T(1));
  T vt[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) vt[i] = (s * vN[i] + vK[i]) * iden;
  const T ht = (s * HN + HK) * iden;
  const T q2 = vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2];
  const T qt = vt[D];
  const T c2 = b.gm1 * (ht - T(0.5) * q2);
  const T c = sqrt_(c2), ic = rcp(c);
  const T l0 = fabs(qt), lp = fabs(qt + c), lm = fabs(qt - c);
  T dd[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) dd[k] = wN[k] - wK[k];
  const T dp = b.gm1 * (T(0.5) * q2 * dd[0] - (vt[0] * dd[1] + vt[1] * dd[2] + vt[2] * dd[3]) + dd[4]);
  const T dV = dd[1 + D] - qt * dd[0];
  const T s1 = T(0.5) * (lp + lm) - l0, s2 = T(0.5) * (lp - lm);
  const T A = s1 * dp * ic * ic + s2 * dV * ic;
  const T B = s2 * dp * ic + s1 * dV;
  T R[5];
  R[0] = T(0.5) * (rN * qN + rK * qK) - T(0.5) * (l0 * dd[0] + A);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const T pn = (i == D) ? (pN + pK) : T(0);
    const T bn = (i == D) ? B : T(0);
    R[1 + i] = T(0.5) * (wN[1 + i] * qN + wK[1 + i] * qK + pn) - T(0.5) * (l0 * dd[1 + i] + A * vt[i] + bn);
  }
  R[4] = T(0.5) * (hN * qN + hK * qK) - T(0.5) * (l0 * dd[4] + A * ht + B * qt);
  T fK[4], fN[4];
  visc_face<T, D>(wK, irK, vK, q, b, fK);
  visc_face<T, D>(wN, irN, vN, q, b, fN);
  outK[0] = R[0];
  outN[0] = -R[0];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    outK[1 + i] = R[1 + i] - fK[i];
    outN[1 + i] = fN[i] - R[1 + i];
  }
}

// ---------------------------------------------------------------------
// Block-wide deterministic sum of one double per thread -> every thread: a
// fixed butterfly inside each wave (every lane ends with the same wave
// total), then the NT/64 wave totals in a fixed order through sw[NT/64].
// A caller that sums again with the same sw puts a barrier in between.
// ---------------------------------------------------------------------
__device__ inline double block_sum(double v, double *sw) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = v;
  __syncthreads();
  static_assert(NT == 256, "four wave totals");
  return (sw[0] + sw[1]) + (sw[2] + sw[3]);
}

// ---------------------------------------------------------------------
// Kernel A: auxiliary gradient (solver.cpp:490-508, restated in strong
// form: -M1^{-1} S1 + face lifts = Dn + M1^{-1} e_0 e_0^T, exactly), and
// the face fluxes of the elements' negative faces.
//
// Gradient: computed in the gradient's own type TQ (FP64, or FP32 in the
// mixed path) from the fluctuation of the state about the background.
// One thread per node computes all 15 components and stores them straight
// to HBM (consecutive threads -> consecutive nodes, coalesced).
//
// Face fluxes: the face between an element N and its -d neighbour K needs
// K's trace (already staged here for the gradient's jump term), N's trace
// and N's gradient trace (the one-sided switch picks it for both sides), so
// the workgroup that owns N evaluates that face once for both elements:
// interpolation of the two states and the 11 gradient components the
// viscous flux reads to the face Gauss points (two sweeps), then Roe +
// viscous fluxes (face_flux_pair), weighted for the lift
// (w_a w_b 2/dx_d), written to FF[element][d][side][state][face point] with
// side 0 = N's own flux on its -d face and side 1 = K's flux on its +d face.
// The right-hand-side kernel only lifts them.  In type TQ, on fluctuations
// about the background, like the rest of the flux arithmetic.
// ---------------------------------------------------------------------
// Gradient components that the viscous flux across a face of direction D
// reads (numericalFluxV, solver.cpp:891-921, with the face normal +-e_D:
// the stress row sigma_D. and the heat flux along D): the derivatives of all
// five conserved variables along D, and along each other direction l those
// of rho, m_D and m_l -- 11 of the 15.  The other four enter the
// reference's face flux multiplied by a zero normal component, so only
// these are carried to the face points.  Slot j of direction D holds
// component k = s + 5 l.
constexpr int NFG = 11;
__host__ __device__ constexpr int face_grad(int D, int j) {
  return j < NSTATE ? j + NSTATE * D
                    : ((j - NSTATE) % 3 == 0 ? 0 : (j - NSTATE) % 3 == 1 ? 1 + D : 1 + (D + 1 + (j - NSTATE) / 3) % 3) +
                          NSTATE * ((D + 1 + (j - NSTATE) / 3) % 3);
}

// Occupancy target of the gradient kernel (waves per SIMD), which sets its
// register budget: 5 with an FP32 gradient up to p = 4 (the face flux needs
// more registers than the 80 of 6 waves), 4 with FP64 and at p = 5; never
// more than its LDS footprint allows (one workgroup = one wave per SIMD,
// 160 KiB of LDS per CU).
template <int N, int E, typename TQ>
struct GradWaves {
  static constexpr int NFF = 2 * NSTATE + NFG;
  static constexpr int lds = (int)sizeof(TQ) * (E * NSTATE * N * N * N + E * 3 * NFF * N * N + N * (N + 1) + N) +
                             (int)sizeof(SymMat<N, TQ>) + 8 * 6 * E;
  static constexpr int by_lds = (160 * 1024) / lds;
  static constexpr int want = sizeof(TQ) == 4 && N <= 5 ? 5 : 4;
  static constexpr int value = want < by_lds ? want : by_lds;
};

// Face-flux array: per element, 3 directions x 2 sides, each a chunk of 5
// states x N^2 face nodes.
__host__ __device__ constexpr int ff_chunk(int n, int /*bytes*/) { return NSTATE * n * n; }
template <int N, typename T>
__host__ __device__ constexpr int ff_per_element() { return 6 * ff_chunk(N, (int)sizeof(T)); }

// 16-byte vector of T
template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = float4; };
template <> struct Vec16<double> { using type = double2; };

// Face fluxes of direction D at one face point: fp points at the point's
// slot of field 0 (fields NF apart: own trace 0-4, neighbour's trace 5-9,
// the gradient components face_grad(D, j) 10-20).
template <typename C, int D, int NF>
__device__ inline void face_point_flux(const C *fp, const Bg<C> &bg, C outK[5], C outN[5]) {
  C wN[5], wK[5], q[NGRAD];
#pragma unroll
  for (int s = 0; s < NSTATE; ++s) { wN[s] = fp[s * NF]; wK[s] = fp[(NSTATE + s) * NF]; }
#pragma unroll
  for (int k = 0; k < NGRAD; ++k) q[k] = C(0);   // not read by the direction-D flux
#pragma unroll
  for (int j = 0; j < NFG; ++j) q[face_grad(D, j)] = fp[(2 * NSTATE + j) * NF];
  face_flux_pair<C, D>(wK, wN, q, bg, outK, outN);
}

// Gradient components of one -D face node for the face flux (the sums of
// the full gradient at that node): along D the line through the node,
// along the two other directions l lines in the face plane, each with the
// lift of the jump on its -l face (the neighbour traces staged in sf); the
// node's own trace with them.  su, sf: the element's state and face slots.
template <int N, int D, int NFF, typename C>
__device__ inline void face_node_grad(const C *__restrict__ su, C *__restrict__ sf, const C (*__restrict__ sD)[N + 1],
                                      const double *ia, unsigned ab) {
  constexpr int NN = N * N * N, NF = N * N;
  constexpr unsigned ST[3] = {1, N, NF};   // memory stride along x, y, z
  const unsigned a = ab % N, b = ab / N;
  // the node (0, a, b), (a, 0, b) or (a, b, 0)
  const unsigned c[3] = {D == 0 ? 0u : a, D == 0 ? a : D == 1 ? 0u : b, D == 2 ? 0u : b};
  const unsigned node = c[0] + N * c[1] + NF * c[2];
  C *fs = sf + D * NFF * NF + ab;
#pragma unroll
  for (int s = 0; s < NSTATE; ++s) {
    const C *us = su + s * NN + node;
    C q = sD[0][N] * sf[(D * NFF + NSTATE + s) * NF + ab];
#pragma unroll
    for (int j = 0; j < N; ++j) q += sD[0][j] * us[j * ST[D]];
    fs[s * NF] = us[0];
    fs[(2 * NSTATE + s) * NF] = q * C(ia[D]);
  }
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int l = (D + 1 + t) % 3;
    const unsigned il = c[l];
    // the node's coordinates on the -l face
    const unsigned abl = l == 0 ? c[1] + N * c[2] : l == 1 ? c[0] + N * c[2] : c[0] + N * c[1];
    const unsigned start = node - il * ST[l];
    const C *row = sD[il];
#pragma unroll
    for (int idx = 0; idx < 3; ++idx) {
      const int s = idx == 0 ? 0 : idx == 1 ? 1 + D : 1 + l;
      const C *us = su + s * NN + start;
      C q = row[N] * sf[(l * NFF + NSTATE + s) * NF + abl];
#pragma unroll
      for (int j = 0; j < N; ++j) q += row[j] * us[j * ST[l]];
      fs[(2 * NSTATE + NSTATE + 3 * t + idx) * NF] = q * C(ia[l]);
    }
  }
}

// WQ: also compute the gradient at every node and write it to Q (for the
// closure's features and ldg_rhs's gradient output); the time-stepping path
// needs only the face nodes' components, the right-hand-side kernel
// computes the volume gradient at the Gauss points itself.
template <int N, int E, typename TQ, typename UT, bool WQ>
__global__ void __launch_bounds__(NT, (GradWaves<N, E, TQ>::value))
ldg_grad_kernel(const UT *__restrict__ U, TQ *__restrict__ Q, TQ *__restrict__ FF, LaunchArgs la,
                const Ops<N, double> *__restrict__ opp) {
  constexpr int NN = N * N * N, NF = N * N;
  // face fields of each element's -d face: own U (0-4), the -d neighbour's U
  // (5-9), the gradient components of direction d (10-20)
  constexpr int NFF = 2 * NSTATE + NFG;
  // arithmetic type: FP32 with an FP32 gradient (mixed), FP64 otherwise
  using C = TQ;
  __shared__ C su[E][NSTATE][NN];
  __shared__ C sf[E][3][NFF][NF];   // -d face of each element: nodal traces, then Gauss-point values
  __shared__ C sD[N][N + 1];
  __shared__ C sw[N];
  __shared__ SymMat<N, C> sI, sIt;
  __shared__ long snb[E][6];
  const int lb = logical_block(la);
  if (lb < 0) return;
  MeshArgs m;
  long kl;
  int ne, ec[3];
  const InstDesc &d = block_instance<E>(la, lb, m, ec, kl, ne);
  if (ne == 0) return;                // padding block of a partial brick
  const long k0 = d.elem_off + kl;    // first element of the block in the arrays
  const unsigned tid = threadIdx.x;
  fill_neighbors(m, ec, d.elem_off, ne, snb, tid);

  // Q_l(i) = (Dn u)(i) + L0(i) (u_0 - u_nbr): the own trace u_0 is the line's
  // first node, so its lift joins column 0 of Dn and only the neighbour's
  // trace enters separately, with weight -L0(i) in column N
  if (tid < N * (N + 1)) {
    const int i = tid / (N + 1), j = tid % (N + 1);
    sD[i][j] = C(j < N ? opp->Dn[i][j] + (j == 0 ? opp->L0[i] : 0.0) : -opp->L0[i]);
  }
  {
    constexpr int MS = sizeof(SymMat<N, C>) / sizeof(C);   // coefficients of I, one per thread
    if (tid >= 64 && tid < 64 + MS)
      reinterpret_cast<C *>(&sI)[tid - 64] = C(reinterpret_cast<const double *>(&opp->sI)[tid - 64]);
    if (tid >= 192 && tid < 192 + MS)
      reinterpret_cast<C *>(&sIt)[tid - 192] = C(reinterpret_cast<const double *>(&opp->sIt)[tid - 192]);
  }
  if (tid >= 128 && tid < 128 + N) sw[tid - 128] = C(opp->w[tid - 128]);
  __syncthreads();   // neighbour table

  // the block's state (one node per thread, its five variables) and the -d
  // neighbours' traces on their positive faces (one face point per thread)
  // into LDS as fluctuations about the background, all of a thread's loads
  // issued before the first is stored.  The loads are unconditional (an
  // index past the block is clamped to its last item) and converted only
  // when stored: a load under a branch is waited for at the branch's end.
  // An FP64 input in the mixed path (RK stage 1) is rounded to the FP32
  // fluctuation the other stages read, so a step's result does not depend
  // on which form its input came in; the FP64 path differentiates the
  // fluctuation too (the same gradient: Dn annihilates constants).
  {
    constexpr int UL = (E * NN + NT - 1) / NT;
    constexpr int JL = (E * 3 * NF + NT - 1) / NT;
    const unsigned tot = ne * NN, totj = ne * 3 * NF;
    UT v[UL][NSTATE], vj[JL][NSTATE];
    unsigned so[UL], sj[JL];
#pragma unroll
    for (int it = 0; it < UL; ++it) {
      const unsigned w = tid + it * NT < tot ? tid + it * NT : tot - 1;
      const unsigned e = w / NN;
      so[it] = w + e * ((NSTATE - 1) * NN);   // su[e][0][node]
      const UT *src = U + k0 * (NSTATE * NN) + (e * (NSTATE * NN) + (w - e * NN));
#pragma unroll
      for (int s = 0; s < NSTATE; ++s) v[it][s] = src[s * NN];
    }
#pragma unroll
    for (int it = 0; it < JL; ++it) {
      const unsigned w = tid + it * NT < totj ? tid + it * NT : totj - 1;
      const unsigned ed = w / NF, ab = w - ed * NF;   // ed = e * 3 + d
      const unsigned e = ed / 3, dd = ed - 3 * e;
      const unsigned b = ab / N, a = ab - b * N;
      sj[it] = ed * (NFF * NF) + NSTATE * NF + ab;    // sf[e][dd][NSTATE][ab]
      // node (N-1, a, b), (a, N-1, b) or (a, b, N-1) of the -d neighbour: its +d face
      const unsigned node = (dd == 0 ? N : 1) * a + (dd == 2 ? N : NF) * b + (dd == 0 ? 1 : dd == 1 ? N : NF) * (N - 1);
      const UT *src = U + snb[e][2 * dd] * (NSTATE * NN) + node;
#pragma unroll
      for (int s = 0; s < NSTATE; ++s) vj[it][s] = src[s * NN];
    }
    const double brho = d.ph.rho0, bE = d.ph.E0;
    auto fluct = [&](UT x, int s) -> C {
      if (sizeof(UT) == 4) return C(x);   // already the FP32 fluctuation
      const double f = (double)x - (s == 0 ? brho : s == 4 ? bE : 0.0);
      return C(f);
    };
    C *const suf = &su[0][0][0];
    C *const sff = &sf[0][0][0][0];
#pragma unroll
    for (int it = 0; it < UL; ++it) {
      if (tid + it * NT < tot) {
#pragma unroll
        for (int s = 0; s < NSTATE; ++s) suf[so[it] + s * NN] = fluct(v[it][s], s);
      }
    }
#pragma unroll
    for (int it = 0; it < JL; ++it) {
      if (tid + it * NT < totj) {
#pragma unroll
        for (int s = 0; s < NSTATE; ++s) sff[sj[it] + s * NF] = fluct(vj[it][s], s);
      }
    }
  }
  __syncthreads();

  // gradient, one node per thread; the node's own trace on the element's -d
  // faces and the face components of its gradient go to the face slots (a
  // node on an edge or corner belongs to several faces)
  auto node_grad = [&](unsigned w) {
    const unsigned e = w / NN, node = w - e * NN;
    const unsigned iz = node / NF, rem = node - iz * NF;
    const unsigned iy = rem / N, ix = rem - iy * N;
    C rx[N + 1], ry[N + 1], rz[N + 1];
#pragma unroll
    for (int j = 0; j <= N; ++j) { rx[j] = sD[ix][j]; ry[j] = sD[iy][j]; rz[j] = sD[iz][j]; }
    TQ *Qe = Q + (k0 + e) * NGRAD * NN + node;
    C g[NGRAD], ox[NSTATE], oy[NSTATE], oz[NSTATE];
#pragma unroll
    for (int s = 0; s < NSTATE; ++s) {
      const C *us = su[e][s];
      C qx = rx[N] * sf[e][0][NSTATE + s][iy + N * iz];
      C qy = ry[N] * sf[e][1][NSTATE + s][ix + N * iz];
      C qz = rz[N] * sf[e][2][NSTATE + s][ix + N * iy];
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const C ux = us[j + N * iy + NF * iz], uy = us[ix + N * j + NF * iz], uz = us[ix + N * iy + NF * j];
        if (j == 0) { ox[s] = ux; oy[s] = uy; oz[s] = uz; }   // the line's first node: the -d face trace
        qx += rx[j] * ux;
        qy += ry[j] * uy;
        qz += rz[j] * uz;
      }
      g[0 * NSTATE + s] = qx * C(m.ialpha[0]);
      g[1 * NSTATE + s] = qy * C(m.ialpha[1]);
      g[2 * NSTATE + s] = qz * C(m.ialpha[2]);
      Qe[(0 * NSTATE + s) * NN] = g[0 * NSTATE + s];
      Qe[(1 * NSTATE + s) * NN] = g[1 * NSTATE + s];
      Qe[(2 * NSTATE + s) * NN] = g[2 * NSTATE + s];
    }
    if (ix == 0) {
      C *fs = &sf[e][0][0][iy + N * iz];
#pragma unroll
      for (int s = 0; s < NSTATE; ++s) fs[s * NF] = ox[s];
#pragma unroll
      for (int j = 0; j < NFG; ++j) fs[(2 * NSTATE + j) * NF] = g[face_grad(0, j)];
    }
    if (iy == 0) {
      C *fs = &sf[e][1][0][ix + N * iz];
#pragma unroll
      for (int s = 0; s < NSTATE; ++s) fs[s * NF] = oy[s];
#pragma unroll
      for (int j = 0; j < NFG; ++j) fs[(2 * NSTATE + j) * NF] = g[face_grad(1, j)];
    }
    if (iz == 0) {
      C *fs = &sf[e][2][0][ix + N * iy];
#pragma unroll
      for (int s = 0; s < NSTATE; ++s) fs[s * NF] = oz[s];
#pragma unroll
      for (int j = 0; j < NFG; ++j) fs[(2 * NSTATE + j) * NF] = g[face_grad(2, j)];
    }
  };
  if constexpr (WQ) {
    if constexpr (E * NN <= NT) {
      if (tid < (unsigned)ne * NN) node_grad(tid);   // one pass (every order's default tiling)
    } else {
      for (unsigned w = tid; w < (unsigned)ne * NN; w += NT) node_grad(w);
    }
  } else {
    // the face nodes only; the points of one direction padded to whole
    // waves, as in the flux phase below
    constexpr unsigned PD = (E * NF + 63) / 64 * 64;
    constexpr unsigned FSTEP = 3 * PD <= NT ? 0x40000000u : NT;
    for (unsigned w = tid; w < 3 * PD; w += FSTEP) {
      const unsigned dd = w / PD, r = w - dd * PD;
      if (r >= (unsigned)ne * NF) continue;
      const unsigned e = r / NF, ab = r - e * NF;
      if (dd == 0) face_node_grad<N, 0, NFF, C>(su[e][0], sf[e][0][0], sD, m.ialpha, ab);
      else if (dd == 1) face_node_grad<N, 1, NFF, C>(su[e][0], sf[e][0][0], sD, m.ialpha, ab);
      else face_node_grad<N, 2, NFF, C>(su[e][0], sf[e][0][0], sD, m.ialpha, ab);
    }
  }
  __syncthreads();

  // interpolate the 21 face fields to the face Gauss points: along a, then
  // along b, in place; face coordinates (a, b) = (y, z), (x, z), (x, y) for
  // d = 0, 1, 2.  A line is (element, direction, field, position): w = fl N
  // + x with fl = (e 3 + d) NFF + f, whose face slot starts at fl NF in sf;
  // the two sweeps share the decomposition, done once per thread.
  {
    constexpr unsigned SW = 3 * NFF * N;              // lines per element
    constexpr int SL = (E * SW + NT - 1) / NT;        // lines per thread
    C *const sff = &sf[0][0][0][0];
    unsigned pa[SL], pb[SL];                          // line starts of the two sweeps, ~0 past the block
#pragma unroll
    for (int it = 0; it < SL; ++it) {
      const unsigned w = tid + it * NT, fl = w / N, x = w - fl * N;
      const bool in = w < (unsigned)ne * SW;
      pa[it] = in ? fl * NF + N * x : ~0u;
      pb[it] = fl * NF + x;
    }
    SymReg<N, C> I;
    I.load(sI);
#pragma unroll
    for (int it = 0; it < SL; ++it) {
      if (pa[it] == ~0u) continue;
      C v[N], o[N];
#pragma unroll
      for (int a = 0; a < N; ++a) v[a] = sff[pa[it] + a];
      sym_apply<N, 1, C>(I.m, v, o);
#pragma unroll
      for (int q = 0; q < N; ++q) sff[pa[it] + q] = o[q];
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < SL; ++it) {
      if (pa[it] == ~0u) continue;
      C v[N], o[N];
#pragma unroll
      for (int b = 0; b < N; ++b) v[b] = sff[pb[it] + N * b];
      sym_apply<N, 1, C>(I.m, v, o);
#pragma unroll
      for (int q = 0; q < N; ++q) sff[pb[it] + N * q] = o[q];
    }
  }
  __syncthreads();

  // fluxes of both sides at every face Gauss point, weighted for the lift,
  // into the point's slots 0-9 (read only by the thread of that point).
  // The points of one direction are padded to whole waves, so that every
  // wave evaluates one direction's flux (the direction is a template
  // parameter: a wave holding points of several directions would run each
  // variant with part of its lanes)
  {
    const Bg<C> bg(d.ph);
    constexpr unsigned PD = (E * NF + 63) / 64 * 64;
    constexpr unsigned FSTEP = 3 * PD <= NT ? 0x40000000u : NT;   // one pass when the points fit
    for (unsigned w = tid; w < 3 * PD; w += FSTEP) {
      const unsigned dd = w / PD, r = w - dd * PD;
      if (r >= (unsigned)ne * NF) continue;
      const unsigned e = r / NF, ab = r - e * NF;
      C *fp = &sf[e][dd][0][ab];
      C outK[5], outN[5];
      if (dd == 0) face_point_flux<C, 0, NF>(fp, bg, outK, outN);
      else if (dd == 1) face_point_flux<C, 1, NF>(fp, bg, outK, outN);
      else face_point_flux<C, 2, NF>(fp, bg, outK, outN);
      // (a select, not an index: a run-time index into m would put it in scratch)
      const double ia = dd == 0 ? m.ialpha[0] : dd == 1 ? m.ialpha[1] : m.ialpha[2];
      const C wt = sw[ab % N] * sw[ab / N] * C(ia);
#pragma unroll
      for (int s = 0; s < NSTATE; ++s) {
        fp[s * NF] = outN[s] * wt;                   // side 0: this element, its -d face
        fp[(NSTATE + s) * NF] = outK[s] * wt;        // side 1: the -d neighbour, its +d face
      }
    }
  }
  __syncthreads();

  // lift to the face nodes, I^T along a then along b, and write FF: per
  // element and direction the 10 flux fields (side, state) are lines
  // (field f, x) of N values; the b sweep stores straight to HBM
  // (FF[element][2 d + side][state][node]), so the right-hand-side kernel
  // only adds them
  {
    constexpr unsigned LL = 2 * NSTATE * N;           // lines per (element, direction)
    constexpr int LI = (E * 3 * LL + NT - 1) / NT;
    constexpr int CH = ff_chunk(N, (int)sizeof(TQ));
    C *const sff = &sf[0][0][0][0];
    unsigned lb_[LI], lf[LI], lx[LI];                 // (e 3 + d), field, x; lb_ = ~0 past the block
#pragma unroll
    for (int it = 0; it < LI; ++it) {
      const unsigned w = tid + it * NT, ed = w / LL, r = w - ed * LL, f = r / N;
      lb_[it] = w < (unsigned)ne * 3 * LL ? ed : ~0u;
      lf[it] = f;
      lx[it] = r - f * N;
    }
    SymReg<N, C> It;
    It.load(sIt);
#pragma unroll
    for (int it = 0; it < LI; ++it) {
      if (lb_[it] == ~0u) continue;
      C *p = sff + (lb_[it] * NFF + lf[it]) * NF + N * lx[it];
      C v[N], o[N];
#pragma unroll
      for (int a = 0; a < N; ++a) v[a] = p[a];
      sym_apply<N, 1, C>(It.m, v, o);
#pragma unroll
      for (int a = 0; a < N; ++a) p[a] = o[a];
    }
    __syncthreads();
    TQ *const ffb = FF + k0 * ff_per_element<N, TQ>();
#pragma unroll
    for (int it = 0; it < LI; ++it) {
      if (lb_[it] == ~0u) continue;
      const unsigned ed = lb_[it], f = lf[it], x = lx[it];
      const C *p = sff + (ed * NFF + f) * NF + x;
      C v[N], o[N];
#pragma unroll
      for (int b = 0; b < N; ++b) v[b] = p[N * b];
      sym_apply<N, 1, C>(It.m, v, o);
      // element ed / 3, chunk 2 d + side, state f mod 5, node x + N b
      const unsigned e = ed / 3, dd = ed - 3 * e, side = f / NSTATE;
      TQ *dst = ffb + (e * ff_per_element<N, TQ>() + (2 * dd + side) * CH + (f - NSTATE * side) * NF + x);
#pragma unroll
      for (int b = 0; b < N; ++b) dst[N * b] = o[b];
    }
  }
}

// ---------------------------------------------------------------------
// Kernel B: volume gradient and integrals, the face terms, inverse mass,
// RK epilogue.
//
// LDS per element: a 20-field volume buffer of type T (in FP32 with room
// for the two jump face sets after it), which last holds the FP64
// inverse-mass sweeps (5 FP64 fields fit in the space of 10 FP32 fields).
// The face terms come from HBM into registers (the gradient kernel lifted
// them to the face nodes).
// ---------------------------------------------------------------------
template <int N, typename T>
struct RhsLds {
  static constexpr int NN = N * N * N, NF = N * N;
  static constexpr int NV = NSTATE + NGRAD;          // 20 volume fields
  static constexpr int SV = NV * NN;
  static constexpr int CH = ff_chunk(N, (int)sizeof(T));
  static constexpr int FE = 6 * CH;                  // face fluxes of one element (3 directions, 2 sides)
  static constexpr int SJ = NSTATE * NF;             // one face field set: the y or z jumps
  static constexpr int VW = 16 / (int)sizeof(T);     // values per 16-byte word
  // the y and z jump sets: in FP64 inside the Q_z slot (which the volume
  // stage fills last; less LDS, so more workgroups fit a CU), in FP32 after
  // the volume fields (LDS does not limit the FP32 kernel's occupancy, and
  // the Q_z slot would cost a barrier)
  static constexpr bool JQZ = sizeof(T) == sizeof(double);
  static_assert(2 * SJ <= NSTATE * NN, "the jump sets fit the Q_z slot");
  static constexpr int SJ0 = JQZ ? 3 * NSTATE * NN : SV;   // start of the jump sets
  static constexpr int BUF = ((JQZ ? SV : SV + 2 * SJ) + VW - 1) / VW * VW;   // per element (16-byte multiple)
  // occupancy target (waves per SIMD) handed to the register allocator
  static constexpr int WAVES = sizeof(T) == 4 ? (N <= 5 ? 4 : 2) : (N <= 2 ? 4 : N <= 5 ? 3 : 2);
};

// Operators in LDS: the compute-type copies used by the sweeps and the FP64
// inverse mass.  Each phase loads the matrix it needs into a register array
// before its loop (LDS loads produce VGPRs: the coefficients neither occupy
// SGPRs nor spill).
template <int N, typename T>
struct SOps {
  SymMat<N, T> I, It, Dt;   // interpolation, its transpose, D^T (compute type)
  SymMat<N, T> Ds[3];       // D = I Dn times 1/alpha_l, l = x, y, z
  T il0[3][N];              // I L0 times 1/alpha_l
  SymMat<N, double> Mi;     // inverse mass (FP64)
  T w[N];                   // Gauss weights (compute type and FP64)
  double wd[N];
};

// The rhs kernel's operators for one instance, built on the host from the
// FP64 operators and the instance's a_l = 1/alpha_l = 2/dx_l: I, I^T, D^T
// and D a_l (l = x, y, z) in type T, each coefficient rounded once; the
// FP64 inverse mass; (I L0) a_l; the Gauss weights.  The kernel copies the
// image of its instance into LDS (no conversion or scaling on the device).
template <int N, typename T>
void fill_sops(SOps<N, T> &so, const Ops<N, double> &op, const double a[3]) {
  constexpr int MS = sizeof(SymMat<N, T>) / sizeof(T);   // coefficients per operator
  static_assert(sizeof(SymMat<N, T>) == MS * sizeof(T) && sizeof(SymMat<N, double>) == MS * sizeof(double),
                "operators are plain coefficient arrays");
  using SO = SOps<N, T>;
  static_assert(offsetof(SO, Ds) == 3 * sizeof(SymMat<N, T>), "I, I^T, D^T, D_x, D_y, D_z are contiguous");
  const SymMat<N, double> *src[6] = {&op.sI, &op.sIt, &op.sDt, &op.sD, &op.sD, &op.sD};
  const double f[6] = {1.0, 1.0, 1.0, a[0], a[1], a[2]};
  T *dst = reinterpret_cast<T *>(&so.I);
  for (int o = 0; o < 6; ++o)
    for (int i = 0; i < MS; ++i) dst[o * MS + i] = T(reinterpret_cast<const double *>(src[o])[i] * f[o]);
  so.Mi = op.sMi;
  for (int l = 0; l < 3; ++l)
    for (int q = 0; q < N; ++q) so.il0[l][q] = T(op.IL0[q] * a[l]);
  for (int q = 0; q < N; ++q) {
    so.w[q] = T(op.w[q]);
    so.wd[q] = op.w[q];
  }
}

// Per instance, the FP32 image then the FP64 image, each padded to 16 bytes.
template <int N>
struct SOpsImage {
  static constexpr int F = (int)((sizeof(SOps<N, float>) + 15) / 16 * 16);
  static constexpr int D = (int)((sizeof(SOps<N, double>) + 15) / 16 * 16);
  static constexpr int BYTES = F + D;
  template <typename T>
  static constexpr int offset() { return sizeof(T) == sizeof(double) ? F : 0; }
};

template <int N, typename T>
struct Mat {
  T v[N][N];
  __device__ inline void load(const T (&src)[N][N]) {
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int b = 0; b < N; ++b) v[a][b] = src[a][b];
  }
};

template <int N, int E, typename T, typename UT>
__global__ void __launch_bounds__(NT, (RhsLds<N, T>::WAVES))
ldg_rhs_kernel(const UT *__restrict__ U, const T *__restrict__ FF, LaunchArgs la, const unsigned char *__restrict__ simg,
               StageArgs st) {
  using L = RhsLds<N, T>;
  constexpr int NN = L::NN, NF = L::NF, BUF = L::BUF;
  __shared__ __align__(16) T sbuf[E][BUF];
  __shared__ double sred[NT / 64];
  __shared__ double sbg[NSTATE];      // the background state (rho0, 0, 0, 0, E0)
  __shared__ SOps<N, T> sops;
  __shared__ long snb[E][6];
  const int lb = logical_block(la);
  if (lb < 0) return;
  MeshArgs m;
  long kl;
  int ne, ec[3];
  const InstDesc &d = block_instance<E>(la, lb, m, ec, kl, ne);
  if (ne == 0) {                      // padding block of a partial brick: empty partial sums
    if (threadIdx.x == 0) {
      if (st.ke_part) st.ke_part[lb] = 0.0;
      if (st.diag_part) { st.diag_part[2 * lb] = 0.0; st.diag_part[2 * lb + 1] = 0.0; }
    }
    return;
  }
  const long k0 = d.elem_off + kl;    // first element of the block in the arrays
  const PhysArgs ph = d.ph;
  // the stage input is the conserved state (UT = double) or its FP32
  // fluctuation about the background (UT = float, written by the previous
  // stage's epilogue)
  constexpr bool FLUCT = sizeof(UT) == 4;
  const double brho = FLUCT ? 0.0 : ph.rho0, bE = FLUCT ? 0.0 : ph.E0;
  const double dtb = d.dt * st.cb, dta = d.dt * st.ca;
  const int tid = threadIdx.x;
  const SOps<N, T> *so = &sops;
  const Bg<T> bg(ph);
  // FP64 view of an element's volume buffer, used by the inverse-mass
  // sweeps and the epilogue
  auto md = [&](int e) -> double * { return reinterpret_cast<double *>(&sbuf[e][0]); };

  // The volume lines of the x sweep stage (below): item i = (e, s, ab) of
  // the block's ne x 5 x N^2 x-lines, in wave-aligned ranges of KP items per
  // kind.  Their HBM loads are issued here, so that they land while the
  // face terms are lifted.  Kind 0 computes I u and the y and z jumps, kind
  // 1 D u and the x jump; when both kinds fit the workgroup side by side
  // (one element per workgroup at p = 4) they run on separate waves,
  // otherwise every item does both.
  constexpr int XL = NSTATE * NF;                    // lines of one direction per element
  constexpr unsigned KP = (E * XL + 63) / 64 * 64;   // items of one kind, padded to whole waves
  constexpr bool XSPLIT = 2 * KP <= NT;
  constexpr int XI = XSPLIT ? 1 : (int)((KP + NT - 1) / NT);
  const unsigned totx = ne * XL;
  UT vo[XI][N], vx[XI], vy[XI][N], vz[XI][N];
  // the item's face terms (lifted to the nodes by the gradient kernel): the
  // -x face at the line's first node, the +x face at its last, and the
  // line's nodes on a -y/+y and a -z/+z face where it lies on one
  T fa[XI], fb[XI], fy[XI][N], fz[XI][N];
  // Line items (e, s, ab) of the sweep phases below run in kinds of KP
  // items (a kind = one family of lines, e.g. the z-lines of F_x).  When KP
  // divides the workgroup (p = 4: E = 2 fills it, E = 1 half of it) a
  // thread's item is the same in every pass and phase: it is decomposed once
  // here, and kind_loop gives the thread the kinds (pass * NT / KP +
  // tid / KP) of its wave-uniform share.
  constexpr bool HOIST = KP <= NT && NT % KP == 0;
  struct LineIt {
    unsigned e, s, ab;
  };
  auto line_at = [&](unsigned i) -> LineIt {
    LineIt L;
    L.e = i / XL;
    const unsigned r = i - L.e * XL;
    L.s = r / NF;
    L.ab = r - L.s * NF;
    return L;
  };
  const unsigned own_i = HOIST ? (unsigned)tid % KP : 0u;
  const LineIt own = line_at(own_i < totx ? own_i : 0u);
  // body(kind, item, pass): pass = the loop pass the item came in (kind-0
  // items come in passes 0, 1, ... in order)
  auto kind_loop = [&](const unsigned K, auto &&body) {
    if constexpr (HOIST) {
      constexpr unsigned PER = NT / KP;
#pragma unroll
      for (unsigned pass = 0; pass < (K + PER - 1) / PER; ++pass) {
        const unsigned kind = pass * PER + (unsigned)tid / KP;
        if (kind < K && own_i < totx) body(kind, own, pass);
      }
    } else {
#pragma unroll
      for (unsigned pass = 0; pass < (K * KP + NT - 1) / NT; ++pass) {
        const unsigned w = tid + pass * NT;
        const unsigned kind = w / KP, i = w - kind * KP;
        if (kind < K && i < totx) body(kind, line_at(i), pass);
      }
    }
  };
  unsigned xi[XI];                                   // the item, or ~0
  auto x_item = [&](int it) -> unsigned {
    const unsigned w = tid + it * NT, i = XSPLIT ? w % KP : w;
    return i < totx ? i : ~0u;
  };
#pragma unroll
  for (int it = 0; it < XI; ++it) {
    xi[it] = x_item(it);
    const unsigned i = xi[it] == ~0u ? totx - 1 : xi[it];
    const unsigned e = i / XL, r = i - e * XL;       // r = s NF + ab: the line starts at N r in its element
    const UT *src = U + k0 * (NSTATE * NN) + (e * (NSTATE * NN) + N * r);
#pragma unroll
    for (int q = 0; q < N; ++q) vo[it][q] = src[q];
  }

  fill_neighbors(m, ec, d.elem_off, ne, snb, tid);
  {
    // the instance's operator image (8-byte words) into LDS
    static_assert(sizeof(SOps<N, T>) % 8 == 0, "whole 8-byte words");
    constexpr int W8 = sizeof(SOps<N, T>) / 8;
    const int ii = la.blk_inst ? la.blk_inst[lb] : 0;
    const uint2 *src = reinterpret_cast<const uint2 *>(simg + (long)ii * SOpsImage<N>::BYTES +
                                                       SOpsImage<N>::template offset<T>());
    for (int w = tid; w < W8; w += NT) reinterpret_cast<uint2 *>(&sops)[w] = src[w];
  }
  if (tid < NSTATE) sbg[tid] = tid == 0 ? ph.rho0 : tid == NSTATE - 1 ? ph.E0 : 0.0;
  __syncthreads();
  // the neighbours' traces the jumps need: the -x neighbour's node at the
  // line's start, and the -y (-z) neighbour's x-line along the -y (-z) face;
  // the latter loaded by every item (from the line itself where the line is
  // not on that face: unconditional loads are not waited for at a branch)
#pragma unroll
  for (int it = 0; it < XI; ++it) {
    const unsigned i = xi[it] == ~0u ? totx - 1 : xi[it];
    const unsigned e = i / XL, r = i - e * XL, s = r / NF, ab = r - s * NF;
    const unsigned iy = ab % N, iz = ab / N;
    const UT *own = U + k0 * (NSTATE * NN) + (e * (NSTATE * NN) + N * r);
    vx[it] = U[snb[e][0] * (NSTATE * NN) + (N * r + (N - 1))];
    const UT *ny = iy == 0 ? U + snb[e][2] * (NSTATE * NN) + (s * NN + N * (N - 1) + NF * iz) : own;
    const UT *nz = iz == 0 ? U + snb[e][4] * (NSTATE * NN) + (s * NN + NF * (N - 1) + N * iy) : own;
#pragma unroll
    for (int q = 0; q < N; ++q) {
      vy[it][q] = ny[q];
      vz[it][q] = nz[q];
    }
    // FF[element][2 d + side][state][node]: side 0 of the element's own
    // record for its -d faces, side 1 of the +d neighbour's for its +d faces
    constexpr int CH = L::CH, FE = L::FE;
    const T *const ffo = FF + (k0 + e) * FE + s * NF;
    fa[it] = ffo[ab];
    fb[it] = FF[snb[e][1] * FE + (CH + s * NF + ab)];
    const T *fyp = iy == N - 1 ? FF + snb[e][3] * FE + (3 * CH + s * NF + N * iz) : ffo + 2 * CH + N * iz;
    const T *fzp = iz == N - 1 ? FF + snb[e][5] * FE + (5 * CH + s * NF + N * iy) : ffo + 4 * CH + N * iy;
#pragma unroll
    for (int q = 0; q < N; ++q) {
      fy[it][q] = fyp[q];
      fz[it][q] = fzp[q];
    }
  }

  // ================= volume =================
  // The state and its gradient at the volume Gauss points, from the state
  // alone.  The gradient kernel's Q_l = (1/alpha_l) (Dn u + L0 J_l) along
  // direction l (J_l: the jump u - u_nbr on the element's -l face, the only
  // face whose neighbour trace enters the one-sided LDG gradient),
  // interpolated to the Gauss points, factorises as
  //   Q_l = (1/alpha_l) [ D_l (x) I (x) I u + (I L0)_l (x) (I (x) I) J_l ]
  // with D = I Dn, the basis derivatives at the Gauss points (exact: the
  // derivative of a basis polynomial lies in their span).  Sweeps:
  //   x (lines from HBM):  X = I u,  Q_x <- D u + IL0 J_x;  J_y, J_z <- I J
  //   y:  XY = I X,  Q_y <- D X + IL0 (I J_y);  Q_x <- I Q_x;  J_z <- I J_z
  //   z:  u_g = I XY,  Q_z <- D XY + IL0 (I I J_z);  Q_x, Q_y <- I Q
  // (45 field sweeps instead of 60 for interpolating a stored gradient, and
  // no gradient traffic).  Fields of the volume buffer: u 0-4, Q_l 5+5l..;
  // the y and z jumps after them.  In type T on fluctuations about the
  // background, as the face terms.  A line item (e, s, ab) names the x-line
  // (y, z) = (a, b), the y-line (x, z) = (a, b) and the z-line (x, y) =
  // (a, b) of field s; each stage runs its kinds of lines in wave-aligned
  // ranges of KP items.
  auto sjy = [&](int e) -> T * { return &sbuf[e][L::SJ0]; };
  auto sjz = [&](int e) -> T * { return &sbuf[e][L::SJ0 + L::SJ]; };
  {
    SymReg<N, T> I, Dx;
    I.load(so->I);
    Dx.load(so->Ds[0]);
    T l0[N];
#pragma unroll
    for (int q = 0; q < N; ++q) l0[q] = so->il0[0][q];
#pragma unroll
    for (int it = 0; it < XI; ++it) {
      if (xi[it] == ~0u) continue;
      const unsigned i = xi[it];
      const unsigned e = i / XL, r = i - e * XL, s = r / NF, ab = r - s * NF;
      const unsigned iy = ab % N, iz = ab / N;
      const double bgs = s == 0 ? brho : s == 4 ? bE : 0.0;
      auto fl = [&](UT x) -> T { return FLUCT ? T(x) : T((double)x - bgs); };
      const bool k0s = !XSPLIT || tid < KP, k1s = !XSPLIT || tid >= KP;   // kinds of this item
      T u[N], o[N];
#pragma unroll
      for (int q = 0; q < N; ++q) u[q] = fl(vo[it][q]);
      T *px = &sbuf[e][N * r];
      if (k0s) {
        sym_apply<N, 1, T>(I.m, u, o);
#pragma unroll
        for (int q = 0; q < N; ++q) px[q] = o[q];
        if (iy == 0) {
          T j[N];
#pragma unroll
          for (int q = 0; q < N; ++q) j[q] = u[q] - fl(vy[it][q]);
          sym_apply<N, 1, T>(I.m, j, o);
#pragma unroll
          for (int q = 0; q < N; ++q) sjy(e)[s * NF + q + N * iz] = o[q];
        }
        if (iz == 0) {
          T j[N];
#pragma unroll
          for (int q = 0; q < N; ++q) j[q] = u[q] - fl(vz[it][q]);
          sym_apply<N, 1, T>(I.m, j, o);
#pragma unroll
          for (int q = 0; q < N; ++q) sjz(e)[s * NF + q + N * iy] = o[q];
        }
      }
      if (k1s) {
        const T jx = u[0] - fl(vx[it]);
        sym_apply<N, -1, T>(Dx.m, u, o);
#pragma unroll
        for (int q = 0; q < N; ++q) px[NSTATE * NN + q] = o[q] + l0[q] * jx;
      }
    }
  }
  __syncthreads();
  // y sweeps: kind 0 I X (in place) and D X + IL0 J_y into the Q_y slot;
  // kind 1 I Q_x (in place); then I J_z along y
  {
    SymReg<N, T> I, Dy;
    I.load(so->I);
    Dy.load(so->Ds[1]);
    T l0[N];
#pragma unroll
    for (int q = 0; q < N; ++q) l0[q] = so->il0[1][q];
    kind_loop(2, [&](unsigned kind, const LineIt &L, unsigned pass) {
      const unsigned a = L.ab % N, b = L.ab / N;
      T *p = &sbuf[L.e][(kind == 0 ? L.s : NSTATE + L.s) * NN + a + NF * b];
      T v[N], o[N];
#pragma unroll
      for (int q = 0; q < N; ++q) v[q] = p[N * q];
      sym_apply<N, 1, T>(I.m, v, o);
#pragma unroll
      for (int q = 0; q < N; ++q) p[N * q] = o[q];
      if (kind == 0) {
        const T jy = sjy(L.e)[L.s * NF + L.ab];
        sym_apply<N, -1, T>(Dy.m, v, o);
        T *pq = p + 2 * NSTATE * NN;
#pragma unroll
        for (int q = 0; q < N; ++q) pq[N * q] = o[q] + l0[q] * jy;
      }
    });
    for (unsigned w = tid; w < (unsigned)ne * NSTATE * N; w += NT) {
      const unsigned e = w / (NSTATE * N), r = w - e * (NSTATE * N);
      T *p = sjz(e) + (r / N) * NF + r % N;
      T v[N], o[N];
#pragma unroll
      for (int q = 0; q < N; ++q) v[q] = p[N * q];
      sym_apply<N, 1, T>(I.m, v, o);
#pragma unroll
      for (int q = 0; q < N; ++q) p[N * q] = o[q];
    }
  }
  __syncthreads();
  // z sweeps: kind 0 I XY (in place) and D XY + IL0 J_z into the Q_z slot;
  // kinds 1 and 2 I Q_x and I Q_y (in place).  Where the z jumps share the
  // Q_z slot (FP64), each kind-0 item reads its jump before a barrier that
  // precedes the first Q_z store.
  {
    SymReg<N, T> I, Dz;
    I.load(so->I);
    Dz.load(so->Ds[2]);
    T l0[N];
#pragma unroll
    for (int q = 0; q < N; ++q) l0[q] = so->il0[2][q];
    constexpr int ZI = HOIST ? 1 : (int)((KP + NT - 1) / NT);   // kind-0 items per thread
    T jzr[ZI];
#pragma unroll
    for (int it = 0; it < ZI; ++it) {
      const unsigned i = HOIST ? own_i : tid + it * NT;
      const LineIt L = HOIST ? own : line_at(i < totx ? i : 0u);
      jzr[it] = (HOIST ? (unsigned)tid < KP : i < KP) && i < totx ? sjz(L.e)[L.s * NF + L.ab] : T(0);
    }
    if constexpr (L::JQZ) __syncthreads();
    kind_loop(3, [&](unsigned kind, const LineIt &L, unsigned pass) {
      T *p = &sbuf[L.e][(kind * NSTATE + L.s) * NN + L.ab];
      T v[N], o[N];
#pragma unroll
      for (int q = 0; q < N; ++q) v[q] = p[NF * q];
      sym_apply<N, 1, T>(I.m, v, o);
#pragma unroll
      for (int q = 0; q < N; ++q) p[NF * q] = o[q];
      if (kind == 0) {
        // this item's jump (kind-0 items come in passes 0 .. ZI-1)
        const T jz = jzr[HOIST ? 0 : pass];
        sym_apply<N, -1, T>(Dz.m, v, o);
        T *pq = p + 3 * NSTATE * NN;
#pragma unroll
        for (int q = 0; q < N; ++q) pq[NF * q] = o[q] + l0[q] * jz;
      }
    });
  }
  __syncthreads();

  // pointwise volume flux, weighted; optional Gauss diagnostics (FP64 sums)
  double dke = 0.0, dens = 0.0;
  for (int w = tid; w < ne * NN; w += NT) {
    const int e = w / NN, q = w - e * NN;
    T *pv = &sbuf[e][q];
    T u[5], g[15], F[15];
#pragma unroll
    for (int s = 0; s < 5; ++s) u[s] = pv[s * NN];
#pragma unroll
    for (int k = 0; k < 15; ++k) g[k] = pv[(NSTATE + k) * NN];
    if (st.diag_part) {
      const double rho = ph.rho0 + (double)u[0];
      const double ir = 1.0 / rho;
      const double v[3] = {(double)u[1] * ir, (double)u[2] * ir, (double)u[3] * ir};
      double gd[15], gv[3][3], ge[3];
#pragma unroll
      for (int k = 0; k < 15; ++k) gd[k] = g[k];
      prim_grads<double>(ir, ph.E0 + (double)u[4], v, gd, gv, ge);
      const double w1 = gv[2][1] - gv[1][2], w2 = gv[0][2] - gv[2][0], w3 = gv[1][0] - gv[0][1];
      const double wq = so->wd[q % N] * so->wd[(q / N) % N] * so->wd[q / NF];
      dke += wq * 0.5 * ((double)u[1] * v[0] + (double)u[2] * v[1] + (double)u[3] * v[2]);
      dens += wq * 0.5 * rho * (w1 * w1 + w2 * w2 + w3 * w3);
    }
    volume_flux<T>(u, g, bg, F);
    const T wq = so->w[q % N] * so->w[(q / N) % N] * so->w[q / NF];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const T sc = wq * T(m.ialpha[l]);
#pragma unroll
      for (int s = 0; s < 5; ++s) pv[(NSTATE + 5 * l + s) * NN] = F[s + 5 * l] * sc;
    }
  }
  if (st.diag_part) {
    const double a = block_sum(dke, sred);
    __syncthreads();
    const double b = block_sum(dens, sred);
    if (tid == 0) { st.diag_part[2 * lb] = a; st.diag_part[2 * lb + 1] = b; }
  }
  __syncthreads();

  // test-function contraction: z sweep (kinds 0, 1: I^T on F_x, F_y; kind
  // 2: D^T on F_z)
  {
    SymReg<N, T> It, Dt;
    It.load(so->It);
    Dt.load(so->Dt);
    kind_loop(3, [&](unsigned kind, const LineIt &L, unsigned pass) {
      T *p = &sbuf[L.e][(NSTATE + kind * NSTATE + L.s) * NN + L.ab];
      T v[N], o[N];
#pragma unroll
      for (int q = 0; q < N; ++q) v[q] = p[q * NF];
      if (kind < 2) sym_apply<N, 1, T>(It.m, v, o);
      else sym_apply<N, -1, T>(Dt.m, v, o);
#pragma unroll
      for (int i = 0; i < N; ++i) p[i * NF] = o[i];
    });
  }
  __syncthreads();
  // y sweep: F_x <- I^T F_x ; G <- D^T F_y + I^T F_z (stored in the F_y slot)
  {
    SymReg<N, T> It;
    It.load(so->It);
    kind_loop(1, [&](unsigned, const LineIt &L, unsigned) {
      const unsigned base = (L.ab % N) + NF * (L.ab / N);
      T *px = &sbuf[L.e][(NSTATE + L.s) * NN + base];
      T *py = px + NSTATE * NN;
      const T *pz = px + 2 * NSTATE * NN;
      T vx[N], vy[N], vz[N], ox[N], og[N];
#pragma unroll
      for (int q = 0; q < N; ++q) { vx[q] = px[q * N]; vy[q] = py[q * N]; vz[q] = pz[q * N]; }
      sym_apply<N, 1, T>(It.m, vx, ox);
      sym_apply2<N, -1, 1, T>(so->Dt, vy, It.m, vz, og);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        px[i * N] = ox[i];
        py[i * N] = og[i];
      }
    });
  }
  __syncthreads();
  // x sweep: r = D^T F_x + I^T G + face residual (T), then M1^{-1} along x
  // in FP64.  The FP64 result goes to the start of the element's buffer
  // (5 FP64 fields), which overlaps the F_x fields still being read by other
  // threads when T = float: every thread keeps its results in registers
  // until all reads are done.
  {
    constexpr int XR = HOIST ? 1 : (E * XL + NT - 1) / NT;
    double res[XR][N];
    bool has[XR];
    SymReg<N, T> It;
    It.load(so->It);
    // volume minus the face terms of the line's nodes (the gradient
    // kernel's lifted face fluxes, loaded with the x-stage lines: item it)
    auto xline = [&](const LineIt &L, int it, double (&out)[N]) {
      const unsigned base = N * L.ab;
      const T *px = &sbuf[L.e][(NSTATE + L.s) * NN + base];
      const T *pg = px + NSTATE * NN;
      T vx[N], vg[N], o[N];
      double rr[N];
#pragma unroll
      for (int q = 0; q < N; ++q) { vx[q] = px[q]; vg[q] = pg[q]; }
      sym_apply2<N, -1, 1, T>(so->Dt, vx, It.m, vg, o);
      o[0] -= fa[it];
      o[N - 1] -= fb[it];
      const unsigned iy = L.ab % N, iz = L.ab / N;
      if (iy == 0 || iy == N - 1) {
#pragma unroll
        for (int i = 0; i < N; ++i) o[i] -= fy[it][i];
      }
      if (iz == 0 || iz == N - 1) {
#pragma unroll
        for (int i = 0; i < N; ++i) o[i] -= fz[it][i];
      }
#pragma unroll
      for (int i = 0; i < N; ++i) rr[i] = (double)o[i];
      sym_apply<N, 1, double>(so->Mi, rr, out);
    };
    if constexpr (HOIST) {
      has[0] = tid < (int)KP && own_i < totx;
      if (has[0]) xline(own, 0, res[0]);
    } else {
#pragma unroll
      for (int it = 0; it < XR; ++it) {
        const unsigned w = tid + it * NT;
        has[it] = w < totx;
        if (has[it]) xline(line_at(w), it, res[it]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < XR; ++it) {
      if (!has[it]) continue;
      const LineIt L = HOIST ? own : line_at(tid + it * NT);
      double *po = md(L.e) + L.s * NN + N * L.ab;
#pragma unroll
      for (int i = 0; i < N; ++i) po[i] = res[it][i];
    }
  }
  __syncthreads();
  // FP64 layout: the RK epilogue's HBM operands (u^n and the running sum)
  // are copied straight into the part of the volume buffer that is free from
  // here on (after the 5 FP64 result fields) by LDS-direct loads, so their
  // latency is covered by the inverse-mass sweeps below (1.5 % faster at
  // 64^3, p = 4).  The FP32 layout has room for one of the two only; there
  // they are loaded into registers at the same point instead.
  const bool need_u0 = st.acc || st.unext || st.unext32;
  const bool need_acc = st.acc && !st.acc_init;
  constexpr bool PF = sizeof(T) == sizeof(double);
  auto pf_u0 = [&](int e) -> double * { return md(e) + NSTATE * NN; };
  auto pf_acc = [&](int e) -> double * { return md(e) + 2 * NSTATE * NN; };
  if (PF) {
    constexpr int DW = 2 * NSTATE * NN;    // dwords per element and array
    const int lane = tid & 63, wave = tid >> 6;
    for (int e = 0; e < ne; ++e) {
      const long gofs_e = (k0 + e) * NSTATE * NN;
      for (int c = wave; c * 64 < DW; c += NT / 64) {
        const int i = c * 64 + lane;
        if (i < DW) {
          if (need_u0)
            __builtin_amdgcn_global_load_lds(reinterpret_cast<const unsigned *>(st.u0 + gofs_e) + i,
                                             (__attribute__((address_space(3))) void *)(reinterpret_cast<unsigned *>(pf_u0(e)) + c * 64),
                                             4, 0, 0);
          if (need_acc)
            __builtin_amdgcn_global_load_lds(reinterpret_cast<const unsigned *>(st.acc + gofs_e) + i,
                                             (__attribute__((address_space(3))) void *)(reinterpret_cast<unsigned *>(pf_acc(e)) + c * 64),
                                             4, 0, 0);
        }
      }
    }
  }
  // Epilogue items.  The block's values are contiguous in every (node,
  // state, element) array: group g = 5 e + s (element e, state s) is words
  // g N^3 .. g N^3 + N^3 - 1 past the block's base gofs (uniform base,
  // 32-bit offset), and words e BUFD + s N^3 + q of the FP64 view of the
  // volume buffer.  Item (it, thread t) is word w = t + it NT.  When N^3 is
  // a multiple of the wave (N = 4) a wave's items lie in one group, which
  // is then computed as a wave-uniform value (no per-lane division).
  constexpr int EPI = (E * NSTATE * NN + NT - 1) / NT;
  constexpr bool GS = NN % 64 == 0;
  constexpr unsigned BUFD = L::BUF * sizeof(T) / sizeof(double);   // an element's buffer, in doubles
  const unsigned tot_epi = ne * NSTATE * NN;
  const long gofs = k0 * NSTATE * NN;
  struct EpiItem {
    bool ok;
    unsigned w, li, s;   // word past gofs, word of the FP64 view, state
  };
  auto epi_item = [&](int it) -> EpiItem {
    EpiItem x;
    unsigned g, q;
    if constexpr (GS) {
      g = (unsigned)__builtin_amdgcn_readfirstlane((int)((tid + it * NT) / NN));
      q = (unsigned)tid % NN;
    } else {
      g = (tid + it * NT) / NN;
      q = tid + it * NT - g * NN;
    }
    const unsigned e = g / NSTATE;
    x.w = g * NN + q;
    x.ok = x.w < tot_epi;
    x.s = g - e * NSTATE;
    x.li = x.w + e * (BUFD - NSTATE * NN);
    return x;
  };
  double ru[EPI], ra[EPI];
  {
    const double *const u0b = st.u0 + gofs, *const accb = st.acc + gofs;
#pragma unroll
    for (int it = 0; it < EPI; ++it) {
      const EpiItem x = epi_item(it);
      ru[it] = 0.0;
      ra[it] = 0.0;
      if (!PF && x.ok) {
        if (need_u0) ru[it] = u0b[x.w];
        if (need_acc) ra[it] = accb[x.w];
      }
    }
  }
  // M1^{-1} along y, then along z (FP64)
#pragma unroll
  for (int dd = 1; dd < 3; ++dd) {
    {
      SymReg<N, double> Mi;
      Mi.load(so->Mi);
      kind_loop(1, [&](unsigned, const LineIt &L, unsigned) {
        constexpr unsigned LS[3] = {1, N, N * N};
        const unsigned a = L.ab % N, b = L.ab / N;
        double *p = md(L.e) + L.s * NN + (dd == 1 ? a + NF * b : a + N * b);
        double v[N], o[N];
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = p[i * LS[dd]];
        sym_apply<N, 1, double>(Mi.m, v, o);
#pragma unroll
        for (int i = 0; i < N; ++i) p[i * LS[dd]] = o[i];
      });
    }
    __syncthreads();
  }

  // the LDS-direct loads above have landed once this wave's memory counter
  // is zero; the barrier makes every wave's copies visible
  if (PF) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }

  // ---- epilogue (FP64): k, RK4 accumulation and next stage input.  All of
  // a thread's HBM reads are issued before any is used.
  {
    double *const mb = md(0);
    double kv[EPI];
#pragma unroll
    for (int it = 0; it < EPI; ++it) {
      const EpiItem x = epi_item(it);
      kv[it] = 0.0;
      if (x.ok) {
        if (PF && need_u0) ru[it] = mb[x.li + NSTATE * NN];
        if (PF && need_acc) ra[it] = mb[x.li + 2 * NSTATE * NN];
        kv[it] = mb[x.li];
        if (st.corr && st.gate[k0 + x.w / (NSTATE * NN)]) kv[it] += st.corr[gofs + x.w];
      }
    }
    if (st.ke_part) __syncthreads();   // every k read before the new state reuses the buffer
    double *const koutb = st.kout + gofs, *const accb = st.acc + gofs, *const unextb = st.unext + gofs;
    float *const next32b = st.next32 + gofs, *const unext32b = st.unext32 + gofs;
#pragma unroll
    for (int it = 0; it < EPI; ++it) {
      const EpiItem x = epi_item(it);
      if (x.ok) {
        const unsigned w = x.w;
        const double k = kv[it];
        // background of the item's state variable (the FP32 copies are fluctuations)
        const double bgv = sbg[x.s];
        if (st.kout) koutb[w] = k;
        if (st.acc) {
          const double a = (st.acc_init ? ru[it] : ra[it]) + dtb * k;
          accb[w] = a;
          if (st.ke_part) mb[x.li] = a;
          if (st.next32) next32b[w] = (float)(a - bgv);
        }
        if (st.unext) unextb[w] = ru[it] + dta * k;
        if (st.unext32) unext32b[w] = (float)(ru[it] + dta * k - bgv);
      }
    }
  }
  if (st.ke_part) {
    __syncthreads();
    double kel = 0.0;
    for (int w = tid; w < ne * NN; w += NT) {
      const int e = w / NN, q = w - e * NN;
      const double *a = md(e);
      const double rho = a[q], m1 = a[NN + q], m2 = a[2 * NN + q], m3 = a[3 * NN + q];
      const double wq = so->wd[q % N] * so->wd[(q / N) % N] * so->wd[q / NF];
      kel += wq * (m1 * m1 + m2 * m2 + m3 * m3) / (2.0 * rho);
    }
    const double t = block_sum(kel, sred);
    if (tid == 0) st.ke_part[lb] = t;
  }
}

// Ordered per-instance sums of per-workgroup partials, one workgroup per
// instance: out[(row * row_stride + i) * stride + j] = scale_i * (sum over
// the instance's blocks b of part[b * stride + j]), row = *step (0 if null).
static __global__ void __launch_bounds__(NT) ldg_finalize_kernel(const double *__restrict__ part, int stride,
                                                            const InstDesc *__restrict__ inst, int row_stride,
                                                            const double *__restrict__ scale, double *__restrict__ out,
                                                            const int *__restrict__ step) {
  __shared__ double sred[NT / 64];
  const int i = blockIdx.x;
  const int b0 = inst[i].block_off, b1 = b0 + inst[i].nblocks;
  const long row = step ? *step : 0;
  constexpr int A = 32;   // accumulators per thread: loads in flight (one workgroup reads every partial)
  for (int j = 0; j < stride; ++j) {
    // thread t sums blocks b0 + t + k NT into accumulator k mod A; the
    // order depends only on b - b0
    double acc[A];
#pragma unroll
    for (int k = 0; k < A; ++k) acc[k] = 0.0;
    for (int base = b0 + threadIdx.x; base < b1; base += A * NT) {
      double x[A];
#pragma unroll
      for (int k = 0; k < A; ++k) {
        const int b = base + k * NT;
        x[k] = b < b1 ? part[(long)b * stride + j] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < A; ++k) acc[k] += x[k];
    }
#pragma unroll
    for (int h = A / 2; h > 0; h /= 2)
#pragma unroll
      for (int k = 0; k < h; ++k) acc[k] += acc[k + h];
    const double v = acc[0];
    const double t = block_sum(v, sred);
    if (threadIdx.x == 0) out[(row * row_stride + i) * stride + j] = t * scale[i];
    __syncthreads();
  }
}

// Advances the history row written by ldg_finalize_kernel (one thread).
static __global__ void ldg_advance_kernel(int *step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

// Taylor-Green initial state at the CGL nodes (solver.cpp:194-236).
struct TGVArgs {
  int n, nx, ny, nz;
  long nelem;
  double xmin[3], xmax[3];
  double r[8];
  double L, p0, rho0, V0, RT0, gamma;
  double phase[3];       // the field is evaluated at x + phase (0 for the reference case)
};

static __global__ void ldg_tgv_kernel(double *__restrict__ u, TGVArgs t) {
  const int dofs = t.n * t.n * t.n;
  const long total = t.nelem * dofs;
  for (long w = blockIdx.x * (long)blockDim.x + threadIdx.x; w < total; w += (long)gridDim.x * blockDim.x) {
    const long k = w / dofs;
    const int iN = (int)(w - k * dofs);
    const long ie[3] = {k % t.nx, (k / t.nx) % t.ny, k / ((long)t.nx * t.ny)};
    const int in[3] = {iN % t.n, (iN / t.n) % t.n, iN / (t.n * t.n)};
    const int ns[3] = {t.nx, t.ny, t.nz};
    double X[3];
    for (int l = 0; l < 3; ++l) {
      const double lo = ie[l] * (t.xmax[l] - t.xmin[l]) / ns[l] + t.xmin[l];
      const double hi = (ie[l] + 1) * (t.xmax[l] - t.xmin[l]) / ns[l] + t.xmin[l];
      X[l] = lo + .5 * (t.r[in[l]] + 1.0) * (hi - lo);
    }
    const double x = X[0] + t.phase[0], y = X[1] + t.phase[1], z = X[2] + t.phase[2], L = t.L;
    const double pr = t.p0 + (t.rho0 * t.V0 * t.V0 / 16.0) * (cos(2 * x / L) + cos(2 * y / L)) * (cos(2 * z / L) + 2);
    const double rho = pr / t.RT0;
    const double m1 = t.rho0 * t.V0 * sin(x / L) * cos(y / L) * cos(z / L);
    const double m2 = -t.rho0 * t.V0 * cos(x / L) * sin(y / L) * cos(z / L);
    const double m3 = 0.0;
    double *uk = u + k * NSTATE * dofs;
    uk[iN] = rho;
    uk[iN + dofs] = m1;
    uk[iN + 2 * dofs] = m2;
    uk[iN + 3 * dofs] = m3;
    uk[iN + 4 * dofs] = pr / (t.gamma - 1) + (m1 * m1 + m2 * m2 + m3 * m3) / (2.0 * rho);
  }
}

}  // namespace ldg
