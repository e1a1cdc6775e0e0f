// Centro-symmetric 1D operators applied by the even/odd split.
//
// The 1D operators of the tensor-product element live on node sets that
// are symmetric about the element centre (Chebyshev-Gauss-Lobatto nodes,
// Gauss points), so every operator A used by the sweeps satisfies
//   A[N-1-q][N-1-j] = S A[q][j],   S = +1 (interpolation, its transpose, the
//                                  inverse mass) or -1 (differentiation).
// With e_j = v_j + v_{N-1-j} and o_j = v_j - v_{N-1-j} (j < H = N/2) and, for
// odd N, the middle value v_m (m = H):
//   (A v)_q = sum_j P[q][j] e_j + C[q] v_m + sum_j R[q][j] o_j
//   P[q][j] = (A[q][j] + A[q][N-1-j]) / 2,  R[q][j] = (A[q][j] - A[q][N-1-j]) / 2,
//   C[q] = A[q][m];
// the even part (P, C) changes sign between rows q and N-1-q when S = -1, the
// odd part (R) when S = +1.  Splitting into the part K that keeps its sign
// and the part F that flips,
//   (A v)_q = K_q + F_q,   (A v)_{N-1-q} = K_q - F_q,   (A v)_m = K_m,
// so only rows q < ceil(N/2) are evaluated: N = 5 takes 21 operations per
// line instead of 25, N = 6 30 instead of 36, and a sum of two operators
// (D^T a + I^T b, the fused contraction sweeps) 38 instead of 50 at N = 5.
// At N = 3, 4 the operation count is about the same, but the operator needs
// half the coefficients (registers and LDS loads): measured rhs -15 % at
// N = 4 and -4 % at N = 3 (mixed).  At N = 2 the split costs more and the
// plain product is used.  This is the even-odd decomposition of
// spectral-element codes (Solomonoff, "A fast algorithm for spectral
// differentiation", J. Comput. Phys. 98, 1992; Kopriva, "Implementing
// Spectral Methods for Partial Differential Equations", 2009).
#pragma once

#include <hip/hip_runtime.h>

namespace ldg {

// Rows q < Q = ceil(N/2) of A in split form, or A itself for N = 2, where
// the plain product is cheaper.
template <int N, typename T>
struct SymMat {
  static constexpr int H = N / 2, Q = (N + 1) / 2;
  static constexpr bool SPLIT = N >= 3;
  static constexpr int QP = SPLIT ? Q : 1, HP = SPLIT ? H : 1, NA = SPLIT ? 1 : N;
  T P[QP][HP];   // split form (SPLIT)
  T R[QP][HP];
  T C[QP];
  T A[NA][NA];   // plain form (!SPLIT)
};

// Split form of A (row-major N x N), exact for a centro-symmetric A; the
// caller symmetrises A first (see symmetrize).
template <int N, typename T, typename Src>
__host__ __device__ inline void make_symmat(const Src (&A)[N][N], SymMat<N, T> &M) {
  constexpr int H = SymMat<N, T>::H, Q = SymMat<N, T>::Q;
  if constexpr (SymMat<N, T>::SPLIT) {
    for (int q = 0; q < Q; ++q) {
      for (int j = 0; j < H; ++j) {
        M.P[q][j] = T((A[q][j] + A[q][N - 1 - j]) * Src(0.5));
        M.R[q][j] = T((A[q][j] - A[q][N - 1 - j]) * Src(0.5));
      }
      M.C[q] = (N & 1) ? T(A[q][H]) : T(0);
    }
    M.A[0][0] = T(0);
  } else {
    for (int q = 0; q < N; ++q)
      for (int j = 0; j < N; ++j) M.A[q][j] = T(A[q][j]);
    M.P[0][0] = M.R[0][0] = M.C[0] = T(0);
  }
}

// The same operator in another precision.
template <int N, typename T, typename Src>
__host__ __device__ inline void convert_symmat(SymMat<N, T> &d, const SymMat<N, Src> &s) {
  using M = SymMat<N, T>;
  for (int q = 0; q < M::QP; ++q) {
    for (int j = 0; j < M::HP; ++j) { d.P[q][j] = T(s.P[q][j]); d.R[q][j] = T(s.R[q][j]); }
    d.C[q] = T(s.C[q]);
  }
  for (int q = 0; q < M::NA; ++q)
    for (int j = 0; j < M::NA; ++j) d.A[q][j] = T(s.A[q][j]);
}

// Replace A by its centro-symmetric (S = +1) or anti-centro-symmetric
// (S = -1) part: A[q][j] <- (A[q][j] + S A[N-1-q][N-1-j]) / 2, which only
// removes round-off of the operator's construction.
template <int N>
inline void symmetrize(double (&A)[N][N], int S) {
  double B[N][N];
  for (int q = 0; q < N; ++q)
    for (int j = 0; j < N; ++j) B[q][j] = 0.5 * (A[q][j] + S * A[N - 1 - q][N - 1 - j]);
  for (int q = 0; q < N; ++q)
    for (int j = 0; j < N; ++j) A[q][j] = B[q][j];
}

// Coefficient registers of one operator (loaded once per phase from LDS).
template <int N, typename T>
struct SymReg {
  SymMat<N, T> m;
  __host__ __device__ inline void load(const SymMat<N, T> &src) { convert_symmat<N, T, T>(m, src); }
};

namespace detail {
// even and odd parts of the line v
template <int N, typename T>
__host__ __device__ inline void split_line(const T (&v)[N], T (&e)[N / 2 > 0 ? N / 2 : 1], T (&o)[N / 2 > 0 ? N / 2 : 1]) {
#pragma unroll
  for (int j = 0; j < N / 2; ++j) {
    e[j] = v[j] + v[N - 1 - j];
    o[j] = v[j] - v[N - 1 - j];
  }
}
// even part E_q and odd part O_q of row q, accumulated straight into k (the
// part that keeps its sign in row N-1-q) and f (the part that flips)
template <int N, int S, typename T>
__host__ __device__ inline void row_parts(const SymMat<N, T> &M, int q, const T *e, const T *o, T vm, T &k, T &f) {
  constexpr int H = N / 2;
  T &E = S == 1 ? k : f;
  T &O = S == 1 ? f : k;
  if (N & 1) E += M.C[q] * vm;
#pragma unroll
  for (int j = 0; j < H; ++j) E += M.P[q][j] * e[j];
#pragma unroll
  for (int j = 0; j < H; ++j) O += M.R[q][j] * o[j];
}
// middle row (odd N): only the kept part is non-zero
template <int N, int S, typename T>
__host__ __device__ inline T mid_part(const SymMat<N, T> &M, const T *e, const T *o, T vm) {
  constexpr int H = N / 2, m = N / 2;
  T k = T(0);
  if (S == 1) {
    k = M.C[m] * vm;
#pragma unroll
    for (int j = 0; j < H; ++j) k += M.P[m][j] * e[j];
  } else {
#pragma unroll
    for (int j = 0; j < H; ++j) k += M.R[m][j] * o[j];
  }
  return k;
}
}  // namespace detail

// out = A v   (out may not alias v)
template <int N, int S, typename T>
__host__ __device__ inline void sym_apply(const SymMat<N, T> &M, const T (&v)[N], T (&out)[N]) {
  if constexpr (!SymMat<N, T>::SPLIT) {
#pragma unroll
    for (int q = 0; q < N; ++q) {
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < N; ++j) acc += M.A[q][j] * v[j];
      out[q] = acc;
    }
  } else {
    constexpr int H = N / 2;
    T e[H], o[H];
    detail::split_line<N, T>(v, e, o);
    const T vm = (N & 1) ? v[H] : T(0);
#pragma unroll
    for (int q = 0; q < H; ++q) {
      T k = T(0), f = T(0);
      detail::row_parts<N, S, T>(M, q, e, o, vm, k, f);
      out[q] = k + f;
      out[N - 1 - q] = k - f;
    }
    if constexpr ((N & 1) != 0) out[H] = detail::mid_part<N, S, T>(M, e, o, vm);
  }
}

// out = A1 v1 + A2 v2   (the fused contraction sweeps)
template <int N, int S1, int S2, typename T>
__host__ __device__ inline void sym_apply2(const SymMat<N, T> &M1, const T (&v1)[N], const SymMat<N, T> &M2,
                                           const T (&v2)[N], T (&out)[N]) {
  if constexpr (!SymMat<N, T>::SPLIT) {
#pragma unroll
    for (int q = 0; q < N; ++q) {
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < N; ++j) acc += M1.A[q][j] * v1[j] + M2.A[q][j] * v2[j];
      out[q] = acc;
    }
  } else {
    constexpr int H = N / 2;
    T e1[H], o1[H], e2[H], o2[H];
    detail::split_line<N, T>(v1, e1, o1);
    detail::split_line<N, T>(v2, e2, o2);
    const T m1 = (N & 1) ? v1[H] : T(0), m2 = (N & 1) ? v2[H] : T(0);
#pragma unroll
    for (int q = 0; q < H; ++q) {
      T k = T(0), f = T(0);
      detail::row_parts<N, S1, T>(M1, q, e1, o1, m1, k, f);
      detail::row_parts<N, S2, T>(M2, q, e2, o2, m2, k, f);
      out[q] = k + f;
      out[N - 1 - q] = k - f;
    }
    if constexpr ((N & 1) != 0)
      out[H] = detail::mid_part<N, S1, T>(M1, e1, o1, m1) + detail::mid_part<N, S2, T>(M2, e2, o2, m2);
  }
}

}  // namespace ldg
