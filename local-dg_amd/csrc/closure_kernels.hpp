// closure_kernels.hpp -- device code of the learned per-element closure:
// L2 projection between orders, feature/target assembly, a BF16-input MFMA
// GEMM with fused bias + SiLU, and the elementwise pieces of the MLP's
// backward pass and Adam.
//
// The network maps one element's state and LDG gradient (5 + 15 fields at
// the (p+1)^3 nodes; 540 values at p = 2) to a correction of its
// right-hand side (5 fields at the nodes; 135 values):
//   H1 = silu(X W1^T + b1), H2 = silu(H1 W2^T + b2), Y = H2 W3^T + b3.
// Weights are kept in FP32 and enter the matrix cores rounded to BF16, with
// FP32 accumulation.
#pragma once

#include <hip/hip_runtime.h>

namespace ldg_mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even (inputs here are finite).
__device__ inline unsigned short to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }

constexpr int GT = 64;       // output tile (rows and columns) per workgroup
constexpr int GK = 32;       // K chunk staged in LDS
constexpr int GPAD = 8;      // LDS row padding (BF16 elements)
enum { EPI_NONE = 0, EPI_BIAS = 1, EPI_BIAS_SILU = 2 };

// C[m][n] = epi( sum_k A(m,k) B(k,n) ), FP32 in memory, BF16 x BF16 -> FP32
// on v_mfma_f32_32x32x16_bf16.  A(m,k) = TA ? A[k*lda+m] : A[m*lda+k];
// B(k,n) = TB ? B[n*ldb+k] : B[k*ldb+n].  Epilogue: + bias[n]; with SiLU the
// pre-activation goes to Z (if not null) and silu() to C.  Grid:
// (ceil(N/64), ceil(M/64)), 256 threads; four waves own 32x32 quadrants.
// Every global access is guarded by the matrix bounds (out-of-range
// operands are staged as zeros).
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) mlp_gemm(int M, int N, int K, const float *__restrict__ A, int lda,
                                                const float *__restrict__ B, int ldb, const float *__restrict__ bias,
                                                float *__restrict__ Z, float *__restrict__ C, int ldc) {
  __shared__ __align__(16) unsigned short As[GT][GK + GPAD];   // [m][k]
  __shared__ __align__(16) unsigned short Bs[GT][GK + GPAD];   // [n][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  floatx16 acc;
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = 0.0f;

  for (int k0 = 0; k0 < K; k0 += GK) {
    if (!TA) {
      const int r = tid >> 2, c = (tid & 3) * 8, m = m0 + r;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int k = k0 + c + i;
        As[r][c + i] = to_bf16((m < M && k < K) ? A[(long)m * lda + k] : 0.0f);
      }
    } else {
      const int kk = tid >> 3, c = (tid & 7) * 8, k = k0 + kk;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int m = m0 + c + i;
        As[c + i][kk] = to_bf16((m < M && k < K) ? A[(long)k * lda + m] : 0.0f);
      }
    }
    if (TB) {
      const int r = tid >> 2, c = (tid & 3) * 8, n = n0 + r;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int k = k0 + c + i;
        Bs[r][c + i] = to_bf16((n < N && k < K) ? B[(long)n * ldb + k] : 0.0f);
      }
    } else {
      const int kk = tid >> 3, c = (tid & 7) * 8, k = k0 + kk;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int n = n0 + c + i;
        Bs[c + i][kk] = to_bf16((n < N && k < K) ? B[(long)k * ldb + n] : 0.0f);
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; kk += 16) {
      // lane l supplies row (l mod 32) of A / column (l mod 32) of B and the
      // 8 consecutive k of half (l / 32) of the 16-deep step
      const int kb = kk + (lane >> 5) * 8;
      const u16x8 a = *reinterpret_cast<const u16x8 *>(&As[wm + (lane & 31)][kb]);
      const u16x8 b = *reinterpret_cast<const u16x8 *>(&Bs[wn + (lane & 31)][kb]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                    acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // accumulator j of lane l holds row 8 (j/4) + 4 (l/32) + j%4, column l%32
  const int n = n0 + wn + (lane & 31);
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int m = m0 + wm + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
    if (m < M && n < N) {
      float v = acc[j];
      if (EPI >= EPI_BIAS) v += bias[n];
      if (EPI == EPI_BIAS_SILU) {
        if (Z) Z[(long)m * ldc + n] = v;
        v = silu(v);
      }
      C[(long)m * ldc + n] = v;
    }
  }
}

// Features of elements [0, K) of one instance: X[e][j] = (x_j - mean_j) *
// istd_j with x = (state minus the background (rho0, 0, 0, 0, E0): nU
// values, then the gradient: nQ values), both in their per-element layout.
template <typename TQ, typename UT>
__global__ void mlp_features(const UT *__restrict__ U, const TQ *__restrict__ Q, long K, int dofs,
                             double rho0, double E0, const float *__restrict__ mean, const float *__restrict__ istd,
                             float *__restrict__ X) {
  const int nU = 5 * dofs, nin = 20 * dofs;
  const long total = K * nin;
  for (long w = blockIdx.x * (long)blockDim.x + threadIdx.x; w < total; w += (long)gridDim.x * blockDim.x) {
    const long e = w / nin;
    const int j = (int)(w - e * nin);
    double x;
    if (j < nU) {
      const int s = j / dofs;
      x = (double)U[e * nU + j] - (s == 0 ? rho0 : s == 4 ? E0 : 0.0);
    } else {
      x = (double)Q[e * (long)(15 * dofs) + (j - nU)];
    }
    X[w] = (float)((x - (double)mean[j]) * (double)istd[j]);
  }
}

// Training targets of elements [0, K): T = Pk - k (projected fine
// right-hand side minus the coarse right-hand side of the projected state).
__global__ void mlp_targets(const double *__restrict__ Pk, const double *__restrict__ k, long n,
                            float *__restrict__ T) {
  for (long w = blockIdx.x * (long)blockDim.x + threadIdx.x; w < n; w += (long)gridDim.x * blockDim.x)
    T[w] = (float)(Pk[w] - k[w]);
}

// Rate of change of the element's nodal kinetic energy sum_j |m_j|^2 /
// (2 rho_j) caused by a correction c of (rho, m) (one wave; lanes over the
// nodes; every lane returns the sum).
// The correction is c = y * ostd + omean (y: the normalised output row).
// u: the element's state, or its fluctuation (then rho_add = rho0).
template <typename UT>
__device__ inline double ke_production(const UT *__restrict__ u, double rho_add, int dofs, const float *y,
                                       const float *__restrict__ ostd, const float *__restrict__ omean, int lane) {
  auto c = [&](int k) { return (double)y[k] * (double)ostd[k] + (double)omean[k]; };
  double pr = 0.0;
  for (int j = lane; j < dofs; j += 64) {
    const double ir = 1.0 / ((double)u[j] + rho_add);
    const double v0 = (double)u[dofs + j] * ir, v1 = (double)u[2 * dofs + j] * ir, v2 = (double)u[3 * dofs + j] * ir;
    pr += v0 * c(dofs + j) + v1 * c(2 * dofs + j) + v2 * c(3 * dofs + j) - 0.5 * (v0 * v0 + v1 * v1 + v2 * v2) * c(j);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) pr += __shfl_xor(pr, o, 64);
  return pr;
}

// Closure correction of each element (one 64-thread block per element):
// corr = Y * ostd + omean, applied (gate = 1) only if the RMS of the
// normalised output Y is at most `bound` and, with `dissipative` = 1, if it
// does not raise the element's kinetic energy (= -1: does not lower it);
// else gate = 0 (plain RHS).
template <typename UT>
__global__ void __launch_bounds__(64) mlp_correction(const float *__restrict__ Y, long K, int nout,
                                                     const float *__restrict__ omean, const float *__restrict__ ostd,
                                                     double bound, int dissipative, const UT *__restrict__ U,
                                                     double rho_add, double *__restrict__ corr,
                                                     unsigned char *__restrict__ gate) {
  for (long e = blockIdx.x; e < K; e += gridDim.x) {
    const float *y = Y + e * nout;
    float ss = 0.0f;
    for (int j = threadIdx.x; j < nout; j += 64) ss += y[j] * y[j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o, 64);
    bool on = sqrt((double)ss / nout) <= bound;
    if (on && dissipative)
      on = dissipative * ke_production(U + e * nout, rho_add, nout / 5, y, ostd, omean, threadIdx.x) <= 0.0;
    for (int j = threadIdx.x; j < nout; j += 64)
      corr[e * nout + j] = on ? (double)y[j] * (double)ostd[j] + (double)omean[j] : 0.0;
    if (threadIdx.x == 0) gate[e] = on ? 1 : 0;
  }
}

// dY = scale rw (Y - T); per-block partial sums of rw (Y - T)^2 (FP64),
// with rw the weight of the row (rows of `width` values; 1 if rw is null).
__global__ void __launch_bounds__(256) mlp_loss_grad(const float *__restrict__ Y, const float *__restrict__ T,
                                                     long n, int width, const float *__restrict__ rw, float scale,
                                                     float *__restrict__ dY, double *__restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (long w = blockIdx.x * 256L + threadIdx.x; w < n; w += (long)gridDim.x * 256) {
    const float d = Y[w] - T[w];
    const float r = rw ? rw[w / width] : 1.0f;
    if (dY) dY[w] = scale * r * d;
    s += (double)r * d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// dZ = dH * silu'(Z), silu'(z) = s (1 + z (1 - s)), s = sigmoid(z).
__global__ void mlp_silu_back(const float *__restrict__ dH, const float *__restrict__ Z, long n,
                              float *__restrict__ dZ) {
  for (long w = blockIdx.x * (long)blockDim.x + threadIdx.x; w < n; w += (long)gridDim.x * blockDim.x) {
    const float z = Z[w];
    const float s = 1.0f / (1.0f + __expf(-z));
    dZ[w] = dH[w] * s * (1.0f + z * (1.0f - s));
  }
}

// Column sums of a rows x cols row-major matrix (one thread per column,
// fixed order: deterministic).
__global__ void mlp_colsum(const float *__restrict__ D, int rows, int cols, float *__restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.0f;
  for (int r = 0; r < rows; ++r) s += D[(long)r * cols + c];
  out[c] = s;
}

// Adam (bias-corrected), elementwise over all parameters.
__global__ void mlp_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                         float *__restrict__ v, long n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
  for (long w = blockIdx.x * (long)blockDim.x + threadIdx.x; w < n; w += (long)gridDim.x * blockDim.x) {
    const float gi = g[w];
    const float mi = b1 * m[w] + (1.0f - b1) * gi;
    const float vi = b2 * v[w] + (1.0f - b2) * gi * gi;
    m[w] = mi;
    v[w] = vi;
    p[w] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  }
}

// dst[b][:] = src[idx[b]][:] (rows of `width` floats).
__global__ void mlp_gather(const float *__restrict__ src, int width, const long *__restrict__ idx, int rows,
                           float *__restrict__ dst) {
  const long total = (long)rows * width;
  for (long w = blockIdx.x * (long)blockDim.x + threadIdx.x; w < total; w += (long)gridDim.x * blockDim.x) {
    const long b = w / width;
    const int j = (int)(w - b * width);
    dst[w] = src[idx[b] * width + j];
  }
}

// out[r] = mean over the row of D[r][j]^2 (one thread per row).
__global__ void mlp_rowmeansq(const float *__restrict__ D, long rows, int width, float *__restrict__ out) {
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < rows; r += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < width; ++j) s += (double)D[r * width + j] * D[r * width + j];
    out[r] = (float)(s / width);
  }
}

// Per-feature sums over a chunk of rows (FP64), for mean and spread:
// part[(chunk * 2 + 0) * width + j] = sum x, [... + 1] = sum x^2.
__global__ void mlp_moments(const float *__restrict__ D, long rows, int width, long chunk, double *__restrict__ part) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= width) return;
  const long r0 = blockIdx.y * chunk;
  const long r1 = r0 + chunk < rows ? r0 + chunk : rows;
  double s = 0.0, s2 = 0.0;
  for (long r = r0; r < r1; ++r) {
    const double x = D[r * width + j];
    s += x;
    s2 += x * x;
  }
  part[(blockIdx.y * 2L + 0) * width + j] = s;
  part[(blockIdx.y * 2L + 1) * width + j] = s2;
}

// D[r][j] = (D[r][j] - mean[j]) * scale[j], in place.
__global__ void mlp_normalize(float *__restrict__ D, long rows, int width, const float *__restrict__ mean,
                              const float *__restrict__ scale) {
  const long total = rows * width;
  for (long w = blockIdx.x * (long)blockDim.x + threadIdx.x; w < total; w += (long)gridDim.x * blockDim.x) {
    const int j = (int)(w % width);
    D[w] = (D[w] - mean[j]) * scale[j];
  }
}

// BF16 copies of the weight matrices for the fused inference kernel, rows
// padded to a multiple of 16 columns (W1) and the output matrix to a
// multiple of 32 rows (W3), padding zero:
//   W1b [hid][k1p], W2b [hid][hid], W3b [noutp][hid].
__global__ void mlp_pack_bf16(const float *__restrict__ W1, const float *__restrict__ W2,
                              const float *__restrict__ W3, int nin, int hid, int nout, int k1p, int noutp,
                              unsigned short *__restrict__ W1b, unsigned short *__restrict__ W2b,
                              unsigned short *__restrict__ W3b) {
  const long n1 = (long)hid * k1p, n2 = (long)hid * hid, n3 = (long)noutp * hid;
  for (long w = blockIdx.x * (long)blockDim.x + threadIdx.x; w < n1 + n2 + n3; w += (long)gridDim.x * blockDim.x) {
    if (w < n1) {
      const long r = w / k1p;
      const int c = (int)(w - r * k1p);
      W1b[w] = c < nin ? to_bf16(W1[r * nin + c]) : 0;
    } else if (w < n1 + n2) {
      W2b[w - n1] = to_bf16(W2[w - n1]);
    } else {
      const long v = w - n1 - n2;
      const long r = v / hid;
      W3b[v] = r < nout ? to_bf16(W3[v]) : 0;
    }
  }
}

// Fused inference for p = 2 elements (27 nodes: NIN = 540 inputs, NOUT =
// 135 outputs) and HID = 256: one workgroup evaluates the network on FR = 32
// elements of one instance entirely on chip -- features from the state and
// gradient (background removed, normalised, rounded to BF16) into LDS, three
// MFMA layers with hidden activations kept in LDS, then the gate and the
// correction.  Numerically identical to mlp_features + 3 x mlp_gemm +
// mlp_correction (same roundings, same MFMA sequence over k).
constexpr int FR = 32;

// One wave: acc[t] += A (32 rows from LDS, row stride LDA) x W_t^T over KS
// steps of 16, with W_t the 32 weight rows starting at w[t] (BF16, global).
// Weight vectors are prefetched PF steps ahead so that the L2 latency is
// covered by the matrix-core work of the steps in between.
template <int KS, int NTILE, int LDA>
__device__ inline void wave_layer(const unsigned short *__restrict__ A, const unsigned short *const (&w)[NTILE],
                                  int lr, int kh, floatx16 (&acc)[NTILE]) {
  constexpr int PF = 4;
  u16x8 buf[PF][NTILE];
#pragma unroll
  for (int p = 0; p < PF; ++p)
#pragma unroll
    for (int t = 0; t < NTILE; ++t)
      if (p < KS) buf[p][t] = *reinterpret_cast<const u16x8 *>(w[t] + 16 * p + kh);
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    u16x8 cur[NTILE];
#pragma unroll
    for (int t = 0; t < NTILE; ++t) {
      cur[t] = buf[s % PF][t];
      if (s + PF < KS) buf[s % PF][t] = *reinterpret_cast<const u16x8 *>(w[t] + 16 * (s + PF) + kh);
    }
    const u16x8 av = *reinterpret_cast<const u16x8 *>(A + lr * LDA + 16 * s + kh);
#pragma unroll
    for (int t = 0; t < NTILE; ++t)
      acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, cur[t]),
                                                       acc[t], 0, 0, 0);
  }
}

template <typename TQ, typename UT, int NIN, int HID, int NOUT>
__global__ void __launch_bounds__(256) mlp_infer(const UT *__restrict__ U, const TQ *__restrict__ Q, long K,
                                                 double rho0, double E0, const float *__restrict__ norm,
                                                 const unsigned short *__restrict__ W1b,
                                                 const unsigned short *__restrict__ W2b,
                                                 const unsigned short *__restrict__ W3b, const float *__restrict__ b1,
                                                 const float *__restrict__ b2, const float *__restrict__ b3,
                                                 double bound, int dissipative, double rho_add,
                                                 double *__restrict__ corr, unsigned char *__restrict__ gate) {
  constexpr int DOFS = NIN / 20, NU = 5 * DOFS, NQ = 15 * DOFS;
  constexpr int K1P = (NIN + 15) / 16 * 16, NOUTP = (NOUT + 31) / 32 * 32;
  constexpr int XS = K1P + 8, HS = HID + 8;
  static_assert(HID == 256, "four waves of 64 hidden columns");
  static_assert(XS >= HS, "the second hidden layer reuses the feature buffer");
  __shared__ __align__(16) unsigned short Xs[FR][XS];
  __shared__ __align__(16) unsigned short Hs[FR][HS];
  __shared__ float Ys[FR][NOUT + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long e0 = (long)blockIdx.x * FR;
  const int rows = (int)((K - e0) < FR ? (K - e0) : FR);
  const float *mean = norm, *scale = norm + NIN, *omean = norm + 2 * NIN, *ostd = norm + 2 * NIN + NOUT;

  // normalisation to LDS, then the block's state and gradient rows, which
  // are contiguous in HBM: each thread issues a batch of loads before using
  // any of them
  __shared__ float smean[NIN], sscale[NIN];
  for (int j = tid; j < NIN; j += 256) { smean[j] = mean[j]; sscale[j] = scale[j]; }
  for (int w = tid; w < FR * (K1P - NIN); w += 256) Xs[w / (K1P - NIN)][NIN + w % (K1P - NIN)] = 0;
  __syncthreads();
  {
    constexpr int B = 8;
    const UT *Ub = U + e0 * NU;
    for (int base = 0; base < FR * NU; base += 256 * B) {
      double v[B];
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const int idx = base + i * 256 + tid;
        v[i] = idx < rows * NU ? (double)Ub[idx] : 0.0;
      }
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const int idx = base + i * 256 + tid;
        if (idx < FR * NU) {
          const int r = idx / NU, j = idx - r * NU, s = j / DOFS;
          const double x = v[i] - (s == 0 ? rho0 : s == 4 ? E0 : 0.0);
          Xs[r][j] = to_bf16(r < rows ? (float)((x - (double)smean[j]) * (double)sscale[j]) : 0.0f);
        }
      }
    }
    const TQ *Qb = Q + e0 * NQ;
    for (int base = 0; base < FR * NQ; base += 256 * B) {
      double v[B];
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const int idx = base + i * 256 + tid;
        v[i] = idx < rows * NQ ? (double)Qb[idx] : 0.0;
      }
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const int idx = base + i * 256 + tid;
        if (idx < FR * NQ) {
          const int r = idx / NQ, j = NU + (idx - r * NQ);
          Xs[r][j] = to_bf16(r < rows ? (float)((v[i] - (double)smean[j]) * (double)sscale[j]) : 0.0f);
        }
      }
    }
  }
  __syncthreads();

  const int lr = lane & 31, kh = (lane >> 5) * 8;
  // layer 1: wave w owns hidden columns [64 w, 64 w + 64)
  {
    floatx16 a[2];
#pragma unroll
    for (int j = 0; j < 16; ++j) { a[0][j] = 0.0f; a[1][j] = 0.0f; }
    const unsigned short *const w[2] = {W1b + (long)(64 * wave + lr) * K1P, W1b + (long)(64 * wave + 32 + lr) * K1P};
    wave_layer<K1P / 16, 2, XS>(&Xs[0][0], w, lr, kh, a);
    const int c0 = 64 * wave + lr;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int r = 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
      Hs[r][c0] = to_bf16(silu(a[0][j] + b1[c0]));
      Hs[r][c0 + 32] = to_bf16(silu(a[1][j] + b1[c0 + 32]));
    }
  }
  __syncthreads();
  // layer 2 into the feature buffer
  {
    floatx16 a[2];
#pragma unroll
    for (int j = 0; j < 16; ++j) { a[0][j] = 0.0f; a[1][j] = 0.0f; }
    const unsigned short *const w[2] = {W2b + (long)(64 * wave + lr) * HID, W2b + (long)(64 * wave + 32 + lr) * HID};
    wave_layer<HID / 16, 2, HS>(&Hs[0][0], w, lr, kh, a);
    const int c0 = 64 * wave + lr;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int r = 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
      Xs[r][c0] = to_bf16(silu(a[0][j] + b2[c0]));
      Xs[r][c0 + 32] = to_bf16(silu(a[1][j] + b2[c0 + 32]));
    }
  }
  __syncthreads();
  // layer 3: output tiles of 32 columns, tile t on wave t mod 4
  for (int t = wave; t < NOUTP / 32; t += 4) {
    floatx16 a[1];
#pragma unroll
    for (int j = 0; j < 16; ++j) a[0][j] = 0.0f;
    const unsigned short *const w[1] = {W3b + (long)(32 * t + lr) * HID};
    wave_layer<HID / 16, 1, XS>(&Xs[0][0], w, lr, kh, a);
    const int c = 32 * t + lr;
    if (c < NOUT) {
#pragma unroll
      for (int j = 0; j < 16; ++j) Ys[8 * (j >> 2) + 4 * (lane >> 5) + (j & 3)][c] = a[0][j] + b3[c];
    }
  }
  __syncthreads();
  // gate and correction, one wave per row
  for (int r = wave; r < rows; r += 4) {
    float ss = 0.0f;
    for (int j = lane; j < NOUT; j += 64) ss += Ys[r][j] * Ys[r][j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o, 64);
    bool on = sqrt((double)ss / NOUT) <= bound;
    const long e = e0 + r;
    if (on && dissipative)
      on = dissipative * ke_production(U + e * NU, rho_add, DOFS, &Ys[r][0], ostd, omean, lane) <= 0.0;
    for (int j = lane; j < NOUT; j += 64)
      corr[e * NOUT + j] = on ? (double)Ys[r][j] * (double)ostd[j] + (double)omean[j] : 0.0;
    if (lane == 0) gate[e] = on ? 1 : 0;
  }
}

// L2 projection between orders, element by element: dst (nc^3 nodes per
// field) = (P (x) P (x) P) src (nf^3 nodes), P = nc x nf row-major, for 5
// fields per element.  One thread per (element, field, coarse node).
__global__ void mlp_project(int nf, int nc, long K, const double *__restrict__ P, const double *__restrict__ src,
                            double *__restrict__ dst) {
  const int df = nf * nf * nf, dc = nc * nc * nc;
  const long total = K * 5 * dc;
  for (long w = blockIdx.x * (long)blockDim.x + threadIdx.x; w < total; w += (long)gridDim.x * blockDim.x) {
    const long es = w / dc;                 // element * 5 + field
    const int node = (int)(w - es * dc);
    const int i = node % nc, j = (node / nc) % nc, k = node / (nc * nc);
    const double *s = src + es * df;
    double acc = 0.0;
    for (int c = 0; c < nf; ++c) {
      double accb = 0.0;
      for (int b = 0; b < nf; ++b) {
        double acca = 0.0;
        for (int a = 0; a < nf; ++a) acca += P[i * nf + a] * s[a + nf * (b + nf * c)];
        accb += P[j * nf + b] * acca;
      }
      acc += P[k * nf + c] * accb;
    }
    dst[w] = acc;
  }
}

}  // namespace ldg_mlp
