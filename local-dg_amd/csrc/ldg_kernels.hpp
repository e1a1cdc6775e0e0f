// ldg_kernels.hpp -- HIP kernels of the LDG right-hand side for gfx950.
//
// One RK4 stage is two launches over the whole periodic mesh:
//   ldg_grad_kernel  auxiliary gradient Q_l = (1/alpha_l)(Dn U + L0 jump_l)
//                    (Solver::rk4Rhs first half, solver.cpp:490-508)
//   ldg_rhs_kernel   face Roe + one-sided viscous fluxes, volume integrals,
//                    inverse mass and the fused RK4 stage epilogue
//                    (solver.cpp:510-535 and 410-434)
// Both stage a block of E consecutive elements in LDS, work in 1D "lines"
// of N = p+1 values owned by one thread (in-place sum-factorised sweeps),
// and read the periodic neighbours' face slices straight from HBM (the
// on-device form of the single-rank halo exchange, solver.cpp:929-998).
//
// Data in HBM (all FP64):
//   U, u0, acc, unext : (node, state, element)  -- the reference layout
//   Q                 : (node, s + 5*l, element) -- 15 gradient fields per element
#pragma once

#include <hip/hip_runtime.h>

namespace ldg {

constexpr int NSTATE = 5;
constexpr int NGRAD = 15;
constexpr int NT = 256;          // threads per workgroup

// 1D operators (row-major [row][col]); passed by value as kernel arguments
// so that fully unrolled loops read them as uniform scalars.
template <int N, typename T>
struct Ops {
  T I[N][N];      // I[q][i]  = l_i(g_q)
  T D[N][N];      // D[q][i]  = l_i'(g_q)
  T Dn[N][N];     // Dn[i][j] = l_j'(r_i)
  T Minv[N][N];   // inverse 1D mass matrix
  T L0[N];        // Minv[:, 0]: lift of the jump at the left face
  T w[N];         // Gauss weights
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
  double pref;                    // background pressure p0 = rho0 c0^2 / gamma
};

struct StageArgs {
  const double *u0;   // u^n (RK base state)
  double *acc;        // running u^n + dt sum b_s k_s
  double *unext;      // next stage input u^n + dt a_s k_s
  double *kout;       // raw k (parity / single RHS)
  double dtb, dta;
  int acc_init;       // 1: acc = u0 + dtb k   0: acc += dtb k
  double *ke_part;    // per-workgroup partial sums of the reference KE of acc
  double *diag_part;  // per-workgroup partial sums (Gauss KE, enstrophy)
};

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of element blocks so that face neighbours in x and y
// share that XCD's L2.  Returns -1 for padding workgroups.
__device__ inline int logical_block(const MeshArgs &m) {
  const int b = blockIdx.x;
  const int lb = (b & 7) * m.per_xcd + (b >> 3);
  return lb < m.nblocks ? lb : -1;
}

__device__ inline long neighbor(const MeshArgs &m, long K, int d, int dir) {
  long ex = K % m.nx, ey = (K / m.nx) % m.ny, ez = K / ((long)m.nx * m.ny);
  if (d == 0) ex = (ex + dir + m.nx) % m.nx;
  else if (d == 1) ey = (ey + dir + m.ny) % m.ny;
  else ez = (ez + dir + m.nz) % m.nz;
  return ex + m.nx * (ey + (long)m.ny * ez);
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
// Pointwise physics (FP64), textbook forms of solver.cpp:1040-1374
// ---------------------------------------------------------------------

// Velocity gradient gv[i][l] and internal-energy gradient ge[l] from the
// state and the conserved gradients g[s + 5 l].
template <typename T>
__device__ inline void prim_grads(const T u[5], T ir, const T v[3], const T *g, T gv[3][3], T ge[3]) {
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    const T drho = g[5 * l];
#pragma unroll
    for (int i = 0; i < 3; ++i) gv[i][l] = (g[1 + i + 5 * l] - v[i] * drho) * ir;
    const T dEr = (g[4 + 5 * l] - u[4] * ir * drho) * ir;
    ge[l] = dEr - (v[0] * gv[0][l] + v[1] * gv[1][l] + v[2] * gv[2][l]);
  }
}

// Volume flux F[s + 5 l] = Fc_l(u) - Fv_l(u, g) (fluxC minus fluxV,
// solver.cpp:1040-1188).  The momentum flux carries p - pref instead of p:
// a constant pressure contributes exactly nothing to the discrete residual
// (the Gauss rules integrate d(phi)/dx and the face traces of phi exactly),
// so this only removes the round-off of the O(1/M0^2) background.
template <typename T>
__device__ inline void volume_flux(const T u[5], const T *g, const PhysArgs &ph, T F[15]) {
  const T ir = T(1) / u[0];
  const T v[3] = {u[1] * ir, u[2] * ir, u[3] * ir};
  const T pr = T(ph.gm1) * (u[4] - T(0.5) * (u[1] * v[0] + u[2] * v[1] + u[3] * v[2]));
  const T pf = pr - T(ph.pref);
  T gv[3][3], ge[3];
  prim_grads<T>(u, ir, v, g, gv, ge);
  const T mu = T(ph.mu), kap = T(ph.kappa);
  const T div = gv[0][0] + gv[1][1] + gv[2][2];
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    T tau[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) tau[i] = mu * (gv[i][l] + gv[l][i]) - (i == l ? mu * T(2.0 / 3.0) * div : T(0));
    F[5 * l] = u[1 + l];
#pragma unroll
    for (int i = 0; i < 3; ++i) F[1 + i + 5 * l] = u[1 + i] * v[l] + (i == l ? pf : T(0)) - tau[i];
    F[4 + 5 * l] = (u[4] + pr) * v[l] - (v[0] * tau[0] + v[1] * tau[1] + v[2] * tau[2] + kap * ge[l]);
  }
}

// Face flux for element K on the face with outward normal sgn*e_d:
// Roe(uN, uK, n) (numericalFluxC, solver.cpp:1194-1374) plus the one-sided
// viscous flux -Fv_d(uK, q) sgn (numericalFluxV, solver.cpp:891-921; q is
// the neighbour's gradient on positive faces, the element's own otherwise).
template <typename T>
__device__ inline void face_flux(const T uK[5], const T uN[5], const T *q, int d, T sgn,
                                 const PhysArgs &ph, T out[5]) {
  const T gm1 = T(ph.gm1);
  const T irN = T(1) / uN[0], irK = T(1) / uK[0];
  const T vN[3] = {uN[1] * irN, uN[2] * irN, uN[3] * irN};
  const T vK[3] = {uK[1] * irK, uK[2] * irK, uK[3] * irK};
  const T pN = gm1 * (uN[4] - T(0.5) * (uN[1] * vN[0] + uN[2] * vN[1] + uN[3] * vN[2]));
  const T pK = gm1 * (uK[4] - T(0.5) * (uK[1] * vK[0] + uK[2] * vK[1] + uK[3] * vK[2]));
  const T HN = (uN[4] + pN) * irN, HK = (uK[4] + pK) * irK;
  const T qN = sgn * vN[d], qK = sgn * vK[d];
  const T s = sqrt(uN[0] * irK);
  const T iden = T(1) / (s + T(1));
  T vt[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) vt[i] = (s * vN[i] + vK[i]) * iden;
  const T ht = (s * HN + HK) * iden;
  const T q2 = vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2];
  const T qt = sgn * vt[d];
  const T c2 = gm1 * (ht - T(0.5) * q2);
  const T c = sqrt(c2), ic = T(1) / c;
  const T l0 = fabs(qt), lp = fabs(qt + c), lm = fabs(qt - c);
  T dd[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) dd[k] = uN[k] - uK[k];
  const T dp = gm1 * (T(0.5) * q2 * dd[0] - (vt[0] * dd[1] + vt[1] * dd[2] + vt[2] * dd[3]) + dd[4]);
  const T dV = sgn * dd[1 + d] - qt * dd[0];
  const T s1 = T(0.5) * (lp + lm) - l0, s2 = T(0.5) * (lp - lm);
  const T A = s1 * dp * ic * ic + s2 * dV * ic;
  const T B = s2 * dp * ic + s1 * dV;
  out[0] = T(0.5) * (uN[0] * qN + uK[0] * qK) - T(0.5) * (l0 * dd[0] + A);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const T pn = (i == d) ? sgn * ((pN - T(ph.pref)) + (pK - T(ph.pref))) : T(0);   // see volume_flux
    const T bn = (i == d) ? sgn * B : T(0);
    out[1 + i] = T(0.5) * (uN[1 + i] * qN + uK[1 + i] * qK + pn) - T(0.5) * (l0 * dd[1 + i] + A * vt[i] + bn);
  }
  out[4] = T(0.5) * ((uN[4] + pN) * qN + (uK[4] + pK) * qK) - T(0.5) * (l0 * dd[4] + A * ht + B * qt);

  // viscous part: only Fv_d is needed since n = sgn e_d
  T gv[3][3], ge[3];
  prim_grads<T>(uK, irK, vK, q, gv, ge);
  const T mu = T(ph.mu), kap = T(ph.kappa);
  const T div = gv[0][0] + gv[1][1] + gv[2][2];
  T tau[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) tau[i] = mu * (gv[i][d] + gv[d][i]) - (i == d ? mu * T(2.0 / 3.0) * div : T(0));
#pragma unroll
  for (int i = 0; i < 3; ++i) out[1 + i] -= sgn * tau[i];
  out[4] -= sgn * (vK[0] * tau[0] + vK[1] * tau[1] + vK[2] * tau[2] + kap * ge[d]);
}

// ---------------------------------------------------------------------
// Block-wide deterministic sum of one double per thread -> thread 0.
// ---------------------------------------------------------------------
__device__ inline double block_sum(double v, double *sred) {
  sred[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sred[threadIdx.x] += sred[threadIdx.x + s];
    __syncthreads();
  }
  return sred[0];
}

// ---------------------------------------------------------------------
// Kernel A: auxiliary gradient (solver.cpp:490-508, restated in strong
// form: -M1^{-1} S1 + face lifts = Dn + M1^{-1} e_0 e_0^T, exactly).
// One thread per node computes all 15 components and stores them straight
// to HBM (consecutive threads -> consecutive nodes, coalesced).
// ---------------------------------------------------------------------
template <int N, int E>
__global__ void __launch_bounds__(NT) ldg_grad_kernel(const double *__restrict__ U, double *__restrict__ Q,
                                                        MeshArgs m, Ops<N, double> op) {
  constexpr int NN = N * N * N, NF = N * N;
  __shared__ double su[E][NSTATE][NN];
  __shared__ double sj[E][3][NSTATE][NF];
  __shared__ double sD[N][N + 1];
  const int lb = logical_block(m);
  if (lb < 0) return;
  const long k0 = (long)lb * E;
  const int ne = (int)((m.nelem - k0) < E ? (m.nelem - k0) : E);
  const int tid = threadIdx.x;

  if (tid < N * (N + 1)) {
    const int i = tid / (N + 1), j = tid % (N + 1);
    sD[i][j] = j < N ? op.Dn[i][j] : op.L0[i];
  }
  const double *Ub = U + k0 * NSTATE * NN;
  for (int w = tid; w < ne * NSTATE * NN; w += NT) (&su[0][0][0])[w] = Ub[w];
  __syncthreads();

  // jump on the negative face of direction d: own trace minus the -d
  // neighbour's trace on its positive face
  for (int w = tid; w < ne * 3 * NSTATE * NF; w += NT) {
    const int e = w / (3 * NSTATE * NF);
    const int r = w - e * (3 * NSTATE * NF);
    const int d = r / (NSTATE * NF);
    const int s = (r / NF) % NSTATE;
    const int ab = r % NF;
    const int a = ab % N, b = ab / N;
    const long Kn = neighbor(m, k0 + e, d, -1);
    const int base = line_base<N>(d, a, b), st = line_stride<N>(d);
    sj[e][d][s][ab] = su[e][s][base] - U[(Kn * NSTATE + s) * NN + base + (N - 1) * st];
  }
  __syncthreads();

  for (int w = tid; w < ne * NN; w += NT) {
    const int e = w / NN, node = w - e * NN;
    const int ix = node % N, iy = (node / N) % N, iz = node / NF;
    double rx[N + 1], ry[N + 1], rz[N + 1];
#pragma unroll
    for (int j = 0; j <= N; ++j) { rx[j] = sD[ix][j]; ry[j] = sD[iy][j]; rz[j] = sD[iz][j]; }
    double *Qe = Q + (k0 + e) * NGRAD * NN + node;
#pragma unroll
    for (int s = 0; s < NSTATE; ++s) {
      const double *us = su[e][s];
      double qx = rx[N] * sj[e][0][s][iy + N * iz];
      double qy = ry[N] * sj[e][1][s][ix + N * iz];
      double qz = rz[N] * sj[e][2][s][ix + N * iy];
#pragma unroll
      for (int j = 0; j < N; ++j) {
        qx += rx[j] * us[j + N * iy + NF * iz];
        qy += ry[j] * us[ix + N * j + NF * iz];
        qz += rz[j] * us[ix + N * iy + NF * j];
      }
      Qe[(0 * NSTATE + s) * NN] = qx * m.ialpha[0];
      Qe[(1 * NSTATE + s) * NN] = qy * m.ialpha[1];
      Qe[(2 * NSTATE + s) * NN] = qz * m.ialpha[2];
    }
  }
}

// ---------------------------------------------------------------------
// Kernel B: face fluxes + volume integrals + inverse mass + RK epilogue.
//
// LDS per element: a 20-field volume buffer (which first serves as the
// face buffer of one direction: 2 sides x 25 fields x N^2 points) and the
// face residual (5 x N^3).
// ---------------------------------------------------------------------
template <int N>
struct RhsLds {
  static constexpr int NN = N * N * N, NF = N * N;
  static constexpr int NV = NSTATE + NGRAD;          // 20 volume fields
  static constexpr int NFF = 25;                     // face fields: uK(5), uN(5), q~(15)
  static constexpr int SV = NV * NN;
  static constexpr int FB = 2 * NFF * NF;
  static constexpr int BUF = SV > FB ? SV : FB;      // per element
};

template <int N, int E>
__global__ void __launch_bounds__(NT) ldg_rhs_kernel(const double *__restrict__ U, const double *__restrict__ Q,
                                                       MeshArgs m, Ops<N, double> op, PhysArgs ph, StageArgs st) {
  using L = RhsLds<N>;
  constexpr int NN = L::NN, NF = L::NF, NV = L::NV, NFF = L::NFF, BUF = L::BUF;
  __shared__ double sbuf[E][BUF];
  __shared__ double sr[E][NSTATE][NN];
  __shared__ double sred[NT];
  const int lb = logical_block(m);
  if (lb < 0) return;
  const long k0 = (long)lb * E;
  const int ne = (int)((m.nelem - k0) < E ? (m.nelem - k0) : E);
  const int tid = threadIdx.x;

  for (int w = tid; w < ne * NSTATE * NN; w += NT) (&sr[0][0][0])[w] = 0.0;

  // ================= faces, one direction at a time =================
  for (int d = 0; d < 3; ++d) {
    const int st_d = line_stride<N>(d);
    const int ast = d == 0 ? N : 1;          // memory stride of the face coordinate a
    // gather nodal face data straight from HBM and interpolate along a:
    // fb[side][f][qa + N b]; f = uK(0..4), uN(5..9), q~(10..24)
    for (int w = tid; w < ne * 2 * NFF * N; w += NT) {
      const int e = w / (2 * NFF * N);
      const int r = w - e * (2 * NFF * N);
      const int side = r / (NFF * N);
      const int f = (r / N) % NFF;
      const int b = r % N;
      const long K = k0 + e;
      const int own = side ? (N - 1) * st_d : 0;
      const int opp = side ? 0 : (N - 1) * st_d;
      const double *src;
      if (f < NSTATE) {
        src = U + (K * NSTATE + f) * NN + own;
      } else if (f < 2 * NSTATE) {
        src = U + (neighbor(m, K, d, side ? 1 : -1) * NSTATE + (f - NSTATE)) * NN + opp;
      } else if (side == 0) {
        src = Q + (K * NGRAD + (f - 2 * NSTATE)) * NN;                        // own, i_d = 0
      } else {
        src = Q + (neighbor(m, K, d, 1) * NGRAD + (f - 2 * NSTATE)) * NN;     // neighbour's i_d = 0
      }
      src += line_base<N>(d, 0, b);
      double v[N];
#pragma unroll
      for (int a = 0; a < N; ++a) v[a] = src[a * ast];
      double *dst = &sbuf[e][(side * NFF + f) * NF + N * b];
#pragma unroll
      for (int q = 0; q < N; ++q) {
        double acc = 0.0;
#pragma unroll
        for (int a = 0; a < N; ++a) acc += op.I[q][a] * v[a];
        dst[q] = acc;
      }
    }
    __syncthreads();
    // interpolate along b (in place)
    for (int w = tid; w < ne * 2 * NFF * N; w += NT) {
      const int e = w / (2 * NFF * N);
      const int r = w - e * (2 * NFF * N);
      const int row = r / N;                  // side * NFF + f
      const int qa = r % N;
      double *p = &sbuf[e][row * NF + qa];
      double v[N];
#pragma unroll
      for (int b = 0; b < N; ++b) v[b] = p[b * N];
#pragma unroll
      for (int q = 0; q < N; ++q) {
        double acc = 0.0;
#pragma unroll
        for (int b = 0; b < N; ++b) acc += op.I[q][b] * v[b];
        p[q * N] = acc;
      }
    }
    __syncthreads();
    // flux at each face Gauss point, weighted for the lift
    for (int w = tid; w < ne * 2 * NF; w += NT) {
      const int e = w / (2 * NF);
      const int r = w - e * (2 * NF);
      const int side = r / NF;
      const int ab = r % NF;
      const double *fb = &sbuf[e][side * NFF * NF + ab];
      double uK[5], uN[5], qq[15], out[5];
#pragma unroll
      for (int s = 0; s < 5; ++s) { uK[s] = fb[s * NF]; uN[s] = fb[(5 + s) * NF]; }
#pragma unroll
      for (int k = 0; k < 15; ++k) qq[k] = fb[(10 + k) * NF];
      face_flux<double>(uK, uN, qq, d, side ? 1.0 : -1.0, ph, out);
      const double wt = op.w[ab % N] * op.w[ab / N] * m.ialpha[d];
      double *o = &sbuf[e][side * NFF * NF + ab];
#pragma unroll
      for (int s = 0; s < 5; ++s) o[s * NF] = out[s] * wt;
    }
    __syncthreads();
    // lift: I^T along a (in place)
    for (int w = tid; w < ne * 2 * NSTATE * N; w += NT) {
      const int e = w / (2 * NSTATE * N);
      const int r = w - e * (2 * NSTATE * N);
      const int side = r / (NSTATE * N);
      const int s = (r / N) % NSTATE;
      const int qb = r % N;
      double *p = &sbuf[e][(side * NFF + s) * NF + N * qb];
      double v[N];
#pragma unroll
      for (int q = 0; q < N; ++q) v[q] = p[q];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < N; ++q) acc += op.I[q][i] * v[q];
        p[i] = acc;
      }
    }
    __syncthreads();
    // I^T along b, subtracted into the face nodes of the residual
    for (int w = tid; w < ne * 2 * NSTATE * N; w += NT) {
      const int e = w / (2 * NSTATE * N);
      const int r = w - e * (2 * NSTATE * N);
      const int side = r / (NSTATE * N);
      const int s = (r / N) % NSTATE;
      const int a = r % N;
      const double *p = &sbuf[e][(side * NFF + s) * NF + a];
      double v[N];
#pragma unroll
      for (int q = 0; q < N; ++q) v[q] = p[q * N];
      const int off = side ? (N - 1) * st_d : 0;
#pragma unroll
      for (int b = 0; b < N; ++b) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < N; ++q) acc += op.I[q][b] * v[q];
        sr[e][s][line_base<N>(d, a, b) + off] -= acc;
      }
    }
    __syncthreads();
  }

  // ================= volume =================
  // load U and Q x-lines from HBM and interpolate along x
  for (int w = tid; w < ne * NV * NF; w += NT) {
    const int e = w / (NV * NF);
    const int r = w - e * (NV * NF);
    const int f = r / NF;
    const int yz = r % NF;
    const long K = k0 + e;
    const double *src = f < NSTATE ? U + (K * NSTATE + f) * NN + N * yz
                                   : Q + (K * NGRAD + (f - NSTATE)) * NN + N * yz;
    double v[N];
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = src[i];
    double *dst = &sbuf[e][f * NN + N * yz];
#pragma unroll
    for (int q = 0; q < N; ++q) {
      double acc = 0.0;
#pragma unroll
      for (int i = 0; i < N; ++i) acc += op.I[q][i] * v[i];
      dst[q] = acc;
    }
  }
  __syncthreads();
  // y then z sweeps (in place)
#pragma unroll
  for (int dd = 1; dd < 3; ++dd) {
    for (int w = tid; w < ne * NV * NF; w += NT) {
      const int e = w / (NV * NF);
      const int r = w - e * (NV * NF);
      const int f = r / NF;
      const int ab = r % NF;
      const int ls = line_stride<N>(dd);
      double *p = &sbuf[e][f * NN + line_base<N>(dd, ab % N, ab / N)];
      double v[N];
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] = p[i * ls];
#pragma unroll
      for (int q = 0; q < N; ++q) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) acc += op.I[q][i] * v[i];
        p[q * ls] = acc;
      }
    }
    __syncthreads();
  }

  // pointwise volume flux, weighted; optional Gauss diagnostics
  double dke = 0.0, dens = 0.0;
  for (int w = tid; w < ne * NN; w += NT) {
    const int e = w / NN, q = w - e * NN;
    double *pv = &sbuf[e][q];
    double u[5], g[15], F[15];
#pragma unroll
    for (int s = 0; s < 5; ++s) u[s] = pv[s * NN];
#pragma unroll
    for (int k = 0; k < 15; ++k) g[k] = pv[(NSTATE + k) * NN];
    const double wq = op.w[q % N] * op.w[(q / N) % N] * op.w[q / NF];
    if (st.diag_part) {
      const double ir = 1.0 / u[0];
      const double v[3] = {u[1] * ir, u[2] * ir, u[3] * ir};
      double gv[3][3], ge[3];
      prim_grads<double>(u, ir, v, g, gv, ge);
      const double w1 = gv[2][1] - gv[1][2], w2 = gv[0][2] - gv[2][0], w3 = gv[1][0] - gv[0][1];
      dke += wq * 0.5 * (u[1] * v[0] + u[2] * v[1] + u[3] * v[2]);
      dens += wq * 0.5 * u[0] * (w1 * w1 + w2 * w2 + w3 * w3);
    }
    volume_flux<double>(u, g, ph, F);
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const double sc = wq * m.ialpha[l];
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

  // test-function contraction: z sweep (I^T on F_x, F_y; D^T on F_z)
  for (int w = tid; w < ne * NGRAD * NF; w += NT) {
    const int e = w / (NGRAD * NF);
    const int r = w - e * (NGRAD * NF);
    const int f = r / NF;            // 0..14 = s + 5 l
    const int ab = r % NF;
    double *p = &sbuf[e][(NSTATE + f) * NN + ab];
    double v[N];
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = p[q * NF];
    if (f >= 10) {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < N; ++q) acc += op.D[q][i] * v[q];
        p[i * NF] = acc;
      }
    } else {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < N; ++q) acc += op.I[q][i] * v[q];
        p[i * NF] = acc;
      }
    }
  }
  __syncthreads();
  // y sweep: F_x <- I^T F_x ; G <- D^T F_y + I^T F_z (stored in the F_y slot)
  for (int w = tid; w < ne * NSTATE * NF; w += NT) {
    const int e = w / (NSTATE * NF);
    const int r = w - e * (NSTATE * NF);
    const int s = r / NF;
    const int ab = r % NF;
    const int base = (ab % N) + NF * (ab / N);
    double *px = &sbuf[e][(NSTATE + s) * NN + base];
    double *py = &sbuf[e][(NSTATE + 5 + s) * NN + base];
    const double *pz = &sbuf[e][(NSTATE + 10 + s) * NN + base];
    double vx[N], vy[N], vz[N];
#pragma unroll
    for (int q = 0; q < N; ++q) { vx[q] = px[q * N]; vy[q] = py[q * N]; vz[q] = pz[q * N]; }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double ax = 0.0, ag = 0.0;
#pragma unroll
      for (int q = 0; q < N; ++q) {
        ax += op.I[q][i] * vx[q];
        ag += op.D[q][i] * vy[q] + op.I[q][i] * vz[q];
      }
      px[i * N] = ax;
      py[i * N] = ag;
    }
  }
  __syncthreads();
  // x sweep: r = D^T F_x + I^T G + face residual, then M1^{-1} along x
  for (int w = tid; w < ne * NSTATE * NF; w += NT) {
    const int e = w / (NSTATE * NF);
    const int r = w - e * (NSTATE * NF);
    const int s = r / NF;
    const int ab = r % NF;
    const int base = N * ab;
    const double *px = &sbuf[e][(NSTATE + s) * NN + base];
    const double *pg = &sbuf[e][(NSTATE + 5 + s) * NN + base];
    double vx[N], vg[N], rr[N];
#pragma unroll
    for (int q = 0; q < N; ++q) { vx[q] = px[q]; vg[q] = pg[q]; }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double acc = sr[e][s][base + i];
#pragma unroll
      for (int q = 0; q < N; ++q) acc += op.D[q][i] * vx[q] + op.I[q][i] * vg[q];
      rr[i] = acc;
    }
    double *po = &sbuf[e][s * NN + base];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) acc += op.Minv[i][j] * rr[j];
      po[i] = acc;
    }
  }
  __syncthreads();
  // M1^{-1} along y, then along z
#pragma unroll
  for (int dd = 1; dd < 3; ++dd) {
    for (int w = tid; w < ne * NSTATE * NF; w += NT) {
      const int e = w / (NSTATE * NF);
      const int r = w - e * (NSTATE * NF);
      const int s = r / NF;
      const int ab = r % NF;
      const int ls = line_stride<N>(dd);
      double *p = &sbuf[e][s * NN + line_base<N>(dd, ab % N, ab / N)];
      double v[N];
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] = p[i * ls];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) acc += op.Minv[i][j] * v[j];
        p[i * ls] = acc;
      }
    }
    __syncthreads();
  }

  // ---- epilogue: k, RK4 accumulation and next stage input
  const long gofs = k0 * NSTATE * NN;
  for (int w = tid; w < ne * NSTATE * NN; w += NT) {
    const int e = w / (NSTATE * NN), r = w - e * (NSTATE * NN);
    const double k = sbuf[e][r];
    const long gi = gofs + w;
    if (st.kout) st.kout[gi] = k;
    if (st.acc || st.unext) {
      const double u0 = st.u0[gi];
      if (st.acc) {
        const double a = st.acc_init ? u0 + st.dtb * k : st.acc[gi] + st.dtb * k;
        st.acc[gi] = a;
        if (st.ke_part) sbuf[e][NSTATE * NN + r] = a;   // fields 5..9 are free again
      }
      if (st.unext) st.unext[gi] = u0 + st.dta * k;
    }
  }
  if (st.ke_part) {
    __syncthreads();
    double kel = 0.0;
    for (int w = tid; w < ne * NN; w += NT) {
      const int e = w / NN, q = w - e * NN;
      const double *a = &sbuf[e][NSTATE * NN];
      const double rho = a[q], m1 = a[NN + q], m2 = a[2 * NN + q], m3 = a[3 * NN + q];
      const double wq = op.w[q % N] * op.w[(q / N) % N] * op.w[q / NF];
      kel += wq * (m1 * m1 + m2 * m2 + m3 * m3) / (2.0 * rho);
    }
    const double t = block_sum(kel, sred);
    if (tid == 0) st.ke_part[lb] = t;
  }
}


// Ordered sum of per-workgroup partials: out[j] = scale * sum_b part[b*stride + j].
__global__ void __launch_bounds__(NT) ldg_finalize_kernel(const double *__restrict__ part, int nparts, int stride,
                                                            double scale, double *__restrict__ out) {
  __shared__ double sred[NT];
  for (int j = 0; j < stride; ++j) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nparts; b += NT) v += part[(long)b * stride + j];
    const double t = block_sum(v, sred);
    if (threadIdx.x == 0) out[j] = t * scale;
    __syncthreads();
  }
}

// Taylor-Green initial state at the CGL nodes (solver.cpp:194-236).
struct TGVArgs {
  int n, nx, ny, nz;
  long nelem;
  double xmin[3], xmax[3];
  double r[8];
  double L, p0, rho0, V0, RT0, gamma;
};

__global__ void ldg_tgv_kernel(double *__restrict__ u, TGVArgs t) {
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
    const double x = X[0], y = X[1], z = X[2], L = t.L;
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
