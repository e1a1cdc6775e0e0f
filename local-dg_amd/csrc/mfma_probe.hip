// mfma_probe.hip -- the rhs kernel's y/z test-function contraction at p = 4
// in three forms, to measure whether matrix cores pay for it (a measurement,
// not part of the solver).
//
// Per element (N = 5 nodes per direction, node x + 5 y + 25 z), from the 15
// weighted flux fields F_x, F_y, F_z (5 states each):
//   O_x = (I^T (x) I^T) F_x                    over (y, z), every x and state
//   O_g = (D^T (x) I^T) F_y + (I^T (x) D^T) F_z
// which is what ldg_rhs_kernel's z and y contraction sweeps compute
// (solver.cpp:701-741, 817-864, restated in sum-factorised form).  As in
// the solver, the results overwrite F_x and F_y in the element's LDS buffer.
//   variant 0: vector ALU, the solver's two in-place sum-factorised sweeps
//              with the even/odd split
//   variant 1: v_mfma_f32_32x32x2_f32 on the dense 25x25 slab operators
//              (FP32 inputs)
//   variant 2: v_mfma_f32_32x32x16_f16 on the slab operators with every
//              operand split into FP16 hi + lo halves and three products
//              (hi hi + hi lo + lo hi), FP32 accumulation.  The solver's
//              fluxes exceed the FP16 range at low Mach numbers, so each
//              column is first scaled by a power of two that puts its
//              largest magnitude below 2^14 (and the result scaled back).
// Every workgroup stages E = 4 elements in one LDS buffer and repeats the
// contraction `reps` times on it in place, so that the difference of two
// repetition counts measures the contraction and not the HBM transfer.
// Each repetition reads what the previous one wrote: nothing of it can be
// moved out of the loop.  Column c of the slab GEMMs is (element, state, x)
// = (c / 25, (c / 5) % 5, c % 5); row / k index i = y + 5 z.
#include <hip/hip_runtime.h>

#include "symop.hpp"

namespace ldg_probe {

using ldg::SymMat;
using ldg::SymReg;
using ldg::sym_apply;
using ldg::sym_apply2;

constexpr int N = 5, NF = 25, NN = 125, NS = 5, E = 4, NT = 256;
constexpr int COLS = E * NS * N;   // 100 slab columns per operator family
constexpr int KX = 26, KG = 50;    // K of O_x (25, padded to the FP32 step of 2) and of O_g

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));

struct ProbeOps {
  SymMat<N, float> It, Dt;      // the solver's even/odd split operators (variant 0)
  float Ax[32][KX];             // O_x slab operator, rows i' (25 used), k = i (25 used)
  float Ag[32][KG];             // [A_y | A_z]
  _Float16 Axh[32][32], Axl[32][32];   // hi / lo halves, k padded to 32
  _Float16 Agh[32][64], Agl[32][64];   // k padded to 64
};

// the block's 15 x 125 values per element into LDS (in range: blocks are whole)
__device__ inline void stage(const float *__restrict__ F, long e0, float *sF, int tid) {
  const float *src = F + e0 * (15 * NN);
  for (int w = tid; w < E * 15 * NN; w += NT) sF[w] = src[w];
}

// O_x, O_g (the first 10 fields of every element) to HBM
__device__ inline void store(float *__restrict__ O, long e0, const float *sF, int tid) {
  float *dst = O + e0 * (10 * NN);
  for (int w = tid; w < E * 10 * NN; w += NT) {
    const int e = w / (10 * NN), r = w - e * 10 * NN;
    dst[w] = sF[e * 15 * NN + r];
  }
}

// variant 0: z sweeps in place (I^T on F_x, F_y; D^T on F_z), then y sweeps
// in place (I^T on F_x; D^T F_y + I^T F_z into the F_y slot), as the
// solver does
__global__ void __launch_bounds__(NT) probe_valu(const float *__restrict__ F, float *__restrict__ O,
                                                 const ProbeOps *__restrict__ ops, int reps) {
  __shared__ float sF[E * 15 * NN];
  __shared__ SymMat<N, float> sIt, sDt;
  const int tid = threadIdx.x;
  const long e0 = (long)blockIdx.x * E;
  if (tid == 0) { sIt = ops->It; sDt = ops->Dt; }
  stage(F, e0, sF, tid);
  __syncthreads();
  SymReg<N, float> It, Dt;
  It.load(sIt);
  Dt.load(sDt);
  for (int rep = 0; rep < reps; ++rep) {
    // z: line (e, f, ab) along z, stride NF
    for (int w = tid; w < E * 15 * NF; w += NT) {
      const int e = w / (15 * NF), r = w - e * 15 * NF, f = r / NF, ab = r - f * NF;
      float *p = &sF[(e * 15 + f) * NN + ab];
      float v[N], o[N];
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] = p[i * NF];
      if (f < 10) sym_apply<N, 1, float>(It.m, v, o);
      else sym_apply<N, -1, float>(Dt.m, v, o);
#pragma unroll
      for (int i = 0; i < N; ++i) p[i * NF] = o[i];
    }
    __syncthreads();
    // y: line (e, s, x, z) along y, stride N
    for (int w = tid; w < E * NS * NF; w += NT) {
      const int e = w / (NS * NF), r = w - e * NS * NF, s = r / NF, ab = r - s * NF;
      const int base = ab % N + NF * (ab / N);
      float *px = &sF[(e * 15 + s) * NN + base], *py = px + NS * NN;
      const float *pz = px + 2 * NS * NN;
      float vx[N], vy[N], vz[N], ox[N], og[N];
#pragma unroll
      for (int i = 0; i < N; ++i) { vx[i] = px[i * N]; vy[i] = py[i * N]; vz[i] = pz[i * N]; }
      sym_apply<N, 1, float>(It.m, vx, ox);
      sym_apply2<N, -1, 1, float>(Dt.m, vy, It.m, vz, og);
#pragma unroll
      for (int i = 0; i < N; ++i) { px[i * N] = ox[i]; py[i * N] = og[i]; }
    }
    __syncthreads();
  }
  store(O, e0, sF, tid);
}

// the column c's element, state and x: its slab value i is field f at
// node x + 5 i
__device__ inline int col_base(int c, int f0) {
  const int cc = c < COLS ? c : COLS - 1;   // columns past the block repeat the last (discarded)
  const int e = cc / (NS * N), r = cc - e * NS * N, s = r / N, x = r - s * N;
  return (e * 15 + f0 + s) * NN + x;
}

// output rows i' of acc (row 8 (j/4) + 4 (lane/32) + j%4, column lane%32)
// into field f0 of the column, times 2^sh
__device__ inline void put(const floatx16 &acc, int c, int f0, float *sF, int lane, int sh) {
  if (c >= COLS) return;
  const int b = col_base(c, f0);
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int i = 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
    if (i < NF) sF[b + N * i] = __builtin_amdgcn_ldexpf(acc[j], sh);
  }
}

// variant 1: wave w owns slab columns 32 w .. 32 w + 31 of both families
__global__ void __launch_bounds__(NT) probe_mfma32(const float *__restrict__ F, float *__restrict__ O,
                                                   const ProbeOps *__restrict__ ops, int reps) {
  __shared__ float sF[E * 15 * NN];
  __shared__ float sAx[32][KX], sAg[32][KG];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long e0 = (long)blockIdx.x * E;
  for (int w = tid; w < 32 * KX; w += NT) (&sAx[0][0])[w] = (&ops->Ax[0][0])[w];
  for (int w = tid; w < 32 * KG; w += NT) (&sAg[0][0])[w] = (&ops->Ag[0][0])[w];
  stage(F, e0, sF, tid);
  __syncthreads();
  const int c = 32 * wave + (lane & 31), h = lane >> 5, row = lane & 31;
  const int bx = col_base(c, 0), by = col_base(c, NS), bz = col_base(c, 2 * NS);
  for (int rep = 0; rep < reps; ++rep) {
    floatx16 ax, ag;
#pragma unroll
    for (int j = 0; j < 16; ++j) { ax[j] = 0.0f; ag[j] = 0.0f; }
#pragma unroll
    for (int k0 = 0; k0 < KX; k0 += 2) {
      const int k = k0 + h;                              // < 26; column 25 of the operator is zero
      const float b = k < NF ? sF[bx + N * k] : 0.0f;
      ax = __builtin_amdgcn_mfma_f32_32x32x2f32(sAx[row][k], b, ax, 0, 0, 0);
    }
#pragma unroll
    for (int k0 = 0; k0 < KG; k0 += 2) {
      const int k = k0 + h;                              // F_y for k < 25, F_z after
      const float b = k < NF ? sF[by + N * k] : sF[bz + N * (k - NF)];
      ag = __builtin_amdgcn_mfma_f32_32x32x2f32(sAg[row][k], b, ag, 0, 0, 0);
    }
    // the column's reads are done before its writes (one wave owns it)
    put(ax, c, 0, sF, lane, 0);
    put(ag, c, NS, sF, lane, 0);
    __syncthreads();
  }
  store(O, e0, sF, tid);
}

// the power of two that brings a column whose largest magnitude is m below
// 2^14 (m = f 2^x, f in [0.5, 1): scale by 2^(14 - x)); bounded so that
// tiny columns stay finite
__device__ inline int col_shift(float m) {
  const int x = __builtin_amdgcn_frexp_expf(m);
  const int s = 14 - x;
  return s < 100 ? s : 100;
}

// v (already scaled) as FP16 hi + lo
__device__ inline void split(float v, _Float16 &hi, _Float16 &lo) {
  hi = (_Float16)v;
  lo = (_Float16)(v - (float)hi);
}

// variant 2: as variant 1 with FP16 hi/lo operands on v_mfma_f32_32x32x16_f16;
// lane l supplies row l%32 of A / column l%32 of B and the 8 consecutive k
// of half l/32 of each 16-deep step
__global__ void __launch_bounds__(NT) probe_mfma16(const float *__restrict__ F, float *__restrict__ O,
                                                   const ProbeOps *__restrict__ ops, int reps) {
  __shared__ float sF[E * 15 * NN];
  __shared__ __align__(16) _Float16 sAxh[32][32], sAxl[32][32], sAgh[32][64], sAgl[32][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long e0 = (long)blockIdx.x * E;
  for (int w = tid; w < 32 * 32; w += NT) {
    (&sAxh[0][0])[w] = (&ops->Axh[0][0])[w];
    (&sAxl[0][0])[w] = (&ops->Axl[0][0])[w];
  }
  for (int w = tid; w < 32 * 64; w += NT) {
    (&sAgh[0][0])[w] = (&ops->Agh[0][0])[w];
    (&sAgl[0][0])[w] = (&ops->Agl[0][0])[w];
  }
  stage(F, e0, sF, tid);
  __syncthreads();
  const int c = 32 * wave + (lane & 31), h = lane >> 5, row = lane & 31;
  const int bx = col_base(c, 0), by = col_base(c, NS), bz = col_base(c, 2 * NS);
  for (int rep = 0; rep < reps; ++rep) {
    // the lane's share of its column: k = k0 + 8 h + t of each 16-deep step
    float vx[2][8], vg[4][8];
    float mx = 0.0f, mg = 0.0f;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int k = 16 * s + 8 * h + t;
        vx[s][t] = k < NF ? sF[bx + N * k] : 0.0f;
        mx = fmaxf(mx, fabsf(vx[s][t]));
      }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int k = 16 * s + 8 * h + t;
        vg[s][t] = k < NF ? sF[by + N * k] : k < 2 * NF ? sF[bz + N * (k - NF)] : 0.0f;
        mg = fmaxf(mg, fabsf(vg[s][t]));
      }
    // the column's two halves are lanes l and l ^ 32
    mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
    mg = fmaxf(mg, __shfl_xor(mg, 32, 64));
    const int shx = col_shift(mx), shg = col_shift(mg);
    floatx16 ax, ag;
#pragma unroll
    for (int j = 0; j < 16; ++j) { ax[j] = 0.0f; ag[j] = 0.0f; }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      half8 bh, bl;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        _Float16 hi, lo;
        split(__builtin_amdgcn_ldexpf(vx[s][t], shx), hi, lo);
        bh[t] = hi;
        bl[t] = lo;
      }
      const half8 ah = *reinterpret_cast<const half8 *>(&sAxh[row][16 * s + 8 * h]);
      const half8 al = *reinterpret_cast<const half8 *>(&sAxl[row][16 * s + 8 * h]);
      ax = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, ax, 0, 0, 0);
      ax = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, ax, 0, 0, 0);
      ax = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, ax, 0, 0, 0);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      half8 bh, bl;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        _Float16 hi, lo;
        split(__builtin_amdgcn_ldexpf(vg[s][t], shg), hi, lo);
        bh[t] = hi;
        bl[t] = lo;
      }
      const half8 ah = *reinterpret_cast<const half8 *>(&sAgh[row][16 * s + 8 * h]);
      const half8 al = *reinterpret_cast<const half8 *>(&sAgl[row][16 * s + 8 * h]);
      ag = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, ag, 0, 0, 0);
      ag = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, ag, 0, 0, 0);
      ag = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, ag, 0, 0, 0);
    }
    put(ax, c, 0, sF, lane, -shx);
    put(ag, c, NS, sF, lane, -shg);
    __syncthreads();
  }
  store(O, e0, sF, tid);
}

}  // namespace ldg_probe

namespace {

int launch(int variant, int reps, const float *F, float *O, long K, const ldg_probe::ProbeOps *op, hipStream_t s) {
  using namespace ldg_probe;
  const dim3 grid((unsigned)(K / E)), block(NT);
  if (variant == 0) hipLaunchKernelGGL(probe_valu, grid, block, 0, s, F, O, op, reps);
  else if (variant == 1) hipLaunchKernelGGL(probe_mfma32, grid, block, 0, s, F, O, op, reps);
  else hipLaunchKernelGGL(probe_mfma16, grid, block, 0, s, F, O, op, reps);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace

extern "C" {

// Size of the operator block the caller fills (struct ProbeOps).
int ldg_probe_ops_bytes(void) { return (int)sizeof(ldg_probe::ProbeOps); }

// Host in, host out: copies F (K elements of 15 x 125 floats; K a multiple
// of 4) and the operator block to the device, launches variant 0/1/2 with
// `reps` repetitions per workgroup once untimed and `launches` times timed
// (HIP events on a stream of its own), returns the shortest launch in
// *best_ms and the last launch's output in O (K x 10 x 125 floats).
// Allocates and frees its own device memory.  0 on success, -1 for bad
// arguments, -2 for a HIP error.
int ldg_probe_measure(int variant, int reps, int launches, const float *F, float *O, long K, const void *ops,
                      float *best_ms) {
  using namespace ldg_probe;
  if (K <= 0 || K % E != 0 || reps < 1 || launches < 1 || variant < 0 || variant > 2 || !F || !O || !ops || !best_ms)
    return -1;
  const size_t fb = (size_t)K * 15 * NN * sizeof(float), ob = (size_t)K * 10 * NN * sizeof(float);
  float *dF = nullptr, *dO = nullptr;
  ProbeOps *dops = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  int rc = -2;
  if (hipMalloc(&dF, fb) != hipSuccess || hipMalloc(&dO, ob) != hipSuccess ||
      hipMalloc(&dops, sizeof(ProbeOps)) != hipSuccess || hipStreamCreate(&s) != hipSuccess ||
      hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
    goto done;
  if (hipMemcpy(dF, F, fb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dops, ops, sizeof(ProbeOps), hipMemcpyHostToDevice) != hipSuccess)
    goto done;
  if (launch(variant, reps, dF, dO, K, dops, s) != 0) goto done;
  *best_ms = -1.0f;
  for (int i = 0; i < launches; ++i) {
    float ms = 0.0f;
    if (hipEventRecord(a, s) != hipSuccess || launch(variant, reps, dF, dO, K, dops, s) != 0 ||
        hipEventRecord(b, s) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
        hipEventElapsedTime(&ms, a, b) != hipSuccess)
      goto done;
    if (*best_ms < 0.0f || ms < *best_ms) *best_ms = ms;
  }
  if (hipMemcpy(O, dO, ob, hipMemcpyDeviceToHost) != hipSuccess) goto done;
  rc = 0;
done:
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (s) (void)hipStreamDestroy(s);
  if (dF) (void)hipFree(dF);
  if (dO) (void)hipFree(dO);
  if (dops) (void)hipFree(dops);
  return rc;
}

}  // extern "C"
