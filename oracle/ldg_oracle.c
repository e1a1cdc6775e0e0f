/*
 * ldg_oracle.c -- CPU restatement of the Local-DG explicit right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: a plain-C, FP64,
 * dense-operator restatement of NoseKnowsAll/Local-DG's LDG right-hand side,
 * RK4 driver and kinetic-energy diagnostic, written from reading the
 * reference sources (it contains none of them).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
 * library built from it, and only as the checker or as the timed CPU
 * baseline.  The product path (libldg.so) never links or calls it.
 *
 * Every function cites the reference file:line it follows.  Where this
 * restatement deliberately differs from the reference it says so:
 *   - LAPACK/BLAS calls (dgesv, dsytrf/dsytrs, dsyev, dgemm, dscal) are
 *     replaced by small hand-written routines with the same mathematical
 *     result (LU with partial pivoting, Cholesky for the SPD mass matrix,
 *     cyclic Jacobi for the symmetric Golub-Welsch matrix, plain loops).
 *   - The MPI halo exchange on one rank (self-messages into ghost slots,
 *     solver.cpp:929-998, mesh.cpp:212-314, MPIUtil.cpp:15-34) is restated as
 *     its net effect: the neighbour trace of face f of element K is the
 *     neighbour N = periodic(K, f)'s trace on face f^1.
 *   - The pointwise fluxes (fluxC, fluxV, numericalFluxC) are written from
 *     the textbook formulas they expand (Euler flux, Newtonian viscous
 *     flux with Fourier heat flux in internal energy, Roe flux without an
 *     entropy fix), not from the reference's generated expressions.
 *
 * Arithmetic type.  The working type `real` is double by default.  Built
 * with -DORC_LONG_DOUBLE it is the x87 80-bit long double (64-bit
 * significand, unit round-off 5.4e-20): the same algorithm with every
 * operator, flux, product and sum carried 11 bits further.  That build is
 * the "exact" answer against which both the FP64 oracle and the GPU's FP64
 * path measure their rounding error.  The interface takes and returns
 * double in either build; inputs (states, domain, physics constants, dt)
 * are the same doubles in both, so both solve the same discrete problem.
 *
 * The dense products and triangular solves run over register tiles of a
 * few rows and columns, each entry with exactly the operation order of the
 * plain column-by-column loop, so the tiling changes speed, not a single bit
 * (tests/test_oracle_build.py compares against the untiled loops).
 *
 * Parity pins (see tests/): Gauss/CGL nodes vs closed forms and numpy;
 * operator Kronecker structure and exactness on polynomials; fluxes vs
 * independent textbook implementations and the supersonic-upwind limit of
 * Roe; full RHS vs an independent numpy weak-form quadrature; invariants
 * (constant state, mass conservation, axis-permutation and periodic-shift
 * equivariance); the FP64 build vs the long double build.  The enstrophy and
 * Gauss-point KE functions have no counterpart in the reference (they are
 * this project's own definitions) and are pinned against closed-form
 * Taylor-Green values.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <omp.h>

#ifdef ORC_LONG_DOUBLE
typedef long double real;
#define R_SQRT sqrtl
#define R_COS cosl
#define R_SIN sinl
#define R_FABS fabsl
#define R_CEIL ceill
#define R_POW powl
#define R_PI 3.141592653589793238462643383279502884L
#else
typedef double real;
#define R_SQRT sqrt
#define R_COS cos
#define R_SIN sin
#define R_FABS fabs
#define R_CEIL ceil
#define R_POW pow
#define R_PI M_PI
#endif

#define NS 5    /* conserved states: rho, rho u, rho v, rho w, rho E */
#define DIM 3
#define NFACES 6
#define OB 8    /* columns per block in the dense kernels */

typedef struct {
  int p, n, dofs, nFN, nQ2D, nQ3D;
  int nx, ny, nz, K;
  real xmin[3], xmax[3], dx[3], alpha[3], J;
  /* physics (solver.cpp:30-43) */
  real Re, Pr, gamma, M0, rho0, V0, L, R, c0, p0, T0;
  /* operators (solver.cpp:74-192) */
  real *xg, *wg;            /* 1D Gauss nodes / weights (n) */
  real *wQ2D, *wQ3D;        /* tensor weights */
  real *Interp2D;           /* (nQ2D, nFN) */
  real *Interp3D;           /* (nQ3D, dofs) */
  real *Mchol;              /* Cholesky factor of Mel (dofs, dofs), lower */
  real *Mel;                /* assembled mass matrix (dofs, dofs) */
  real *Sels;               /* (dofs, dofs, 3) */
  real *Kels;               /* (nQ3D, dofs, 3) */
  real *Kels2D;             /* (nQ2D, nFN, 3) */
  int *efToN;               /* (nFN, 6) */
  real *coords;             /* (3, dofs, K) */
  /* work arrays */
  real *uI2, *uI3, *DuI2, *DuI3; /* interpolated traces / volume values */
} oracle_t;

static real *ralloc(size_t n) { return malloc(sizeof(real) * n); }

static real *to_real(const double *x, size_t n) {
  real *r = calloc(n ? n : 1, sizeof(real));
  if (r) for (size_t i = 0; i < n; ++i) r[i] = x[i];
  return r;
}

static void to_double(const real *r, double *x, size_t n) {
  for (size_t i = 0; i < n; ++i) x[i] = (double)r[i];
}

/* ------------------------------------------------------------------ */
/* Small dense linear algebra (column-major, restating the BLAS/LAPACK  */
/* calls of solver.cpp / dgMath.cpp)                                    */
/* ------------------------------------------------------------------ */

/* C(m,n) = alpha * op(A) * B + beta * C, op(A) = A (m,k) or A^T (A is k,m).
   Restates cblas_dgemm with NoTrans/Trans on A and NoTrans on B.  Column j
   of C is c_i += a_{i,pp} (alpha b_pp) over pp in order (NoTrans), or
   c_i += alpha sum_pp a_{pp,i} b_pp with the sum in pp order (Trans).
   The loops below run over register tiles (MB rows x OB columns for
   NoTrans, MI rows x OB columns for Trans), each entry of C with exactly
   that sequence of roundings. */
#ifdef ORC_UNTILED
/* The same product column by column, without tiles (the form the tiled
   kernels are checked against, bitwise). */
static void gemm(int transA, int m, int n, int k, real alpha,
                 const real *A, int lda, const real *B, int ldb,
                 real beta, real *C, int ldc) {
  long j;
#pragma omp parallel for schedule(static)
  for (j = 0; j < n; ++j) {
    real *c = C + (size_t)j * ldc;
    const real *b = B + (size_t)j * ldb;
    if (beta == 0.0) { for (int i = 0; i < m; ++i) c[i] = 0.0; }
    else if (beta != 1.0) { for (int i = 0; i < m; ++i) c[i] *= beta; }
    if (!transA) {
      for (int pp = 0; pp < k; ++pp) {
        const real bp = alpha * b[pp];
        const real *a = A + (size_t)pp * lda;
        for (int i = 0; i < m; ++i) c[i] += a[i] * bp;
      }
    } else {
      for (int i = 0; i < m; ++i) {
        const real *a = A + (size_t)i * lda;
        real s = 0.0;
        for (int pp = 0; pp < k; ++pp) s += a[pp] * b[pp];
        c[i] += alpha * s;
      }
    }
  }
}
#else
#define MB 16   /* rows of C per NoTrans tile */
#define MI 8    /* rows of C per Trans tile */

/* acc[jj][ii] += a(ii, pp) * bt[pp][jj] over pp = 0..k-1 in order, for one
   MB x OB tile; a(ii, pp) = a[pp * lda + ii]. */
static void tile_nt(int k, const real *a, int lda, const real *bt, real acc[OB][MB]) {
#if defined(ORC_LONG_DOUBLE) || !defined(__GNUC__)
  for (int pp = 0; pp < k; ++pp)
    for (int jj = 0; jj < OB; ++jj) {
      const real bp = bt[pp * OB + jj];
      for (int ii = 0; ii < MB; ++ii) acc[jj][ii] += a[(size_t)pp * lda + ii] * bp;
    }
#else
  typedef double v8 __attribute__((vector_size(64)));
  v8 c0[OB], c1[OB];
  for (int jj = 0; jj < OB; ++jj) { memcpy(&c0[jj], &acc[jj][0], 64); memcpy(&c1[jj], &acc[jj][8], 64); }
  for (int pp = 0; pp < k; ++pp) {
    v8 a0, a1;
    memcpy(&a0, a + (size_t)pp * lda, 64);
    memcpy(&a1, a + (size_t)pp * lda + 8, 64);
    for (int jj = 0; jj < OB; ++jj) {
      const double bp = bt[pp * OB + jj];
      c0[jj] += a0 * bp;
      c1[jj] += a1 * bp;
    }
  }
  for (int jj = 0; jj < OB; ++jj) { memcpy(&acc[jj][0], &c0[jj], 64); memcpy(&acc[jj][8], &c1[jj], 64); }
#endif
}

/* s[ii][jj] += sum_pp a(ii, pp) * bt[pp][jj], the sum in pp order, for one
   MI x OB tile; a(ii, pp) = a[ii * k + pp]. */
static void tile_t(int k, const real *a, const real *bt, real s[MI][OB]) {
#if defined(ORC_LONG_DOUBLE) || !defined(__GNUC__)
  for (int pp = 0; pp < k; ++pp)
    for (int ii = 0; ii < MI; ++ii) {
      const real ap = a[(size_t)ii * k + pp];
      for (int jj = 0; jj < OB; ++jj) s[ii][jj] += ap * bt[pp * OB + jj];
    }
#else
  typedef double v8 __attribute__((vector_size(64)));
  v8 c[MI];
  for (int ii = 0; ii < MI; ++ii) memcpy(&c[ii], &s[ii][0], 64);
  for (int pp = 0; pp < k; ++pp) {
    v8 b;
    memcpy(&b, bt + pp * OB, 64);
    for (int ii = 0; ii < MI; ++ii) c[ii] += a[(size_t)ii * k + pp] * b;
  }
  for (int ii = 0; ii < MI; ++ii) memcpy(&s[ii][0], &c[ii], 64);
#endif
}
static void gemm(int transA, int m, int n, int k, real alpha,
                 const real *A, int lda, const real *B, int ldb,
                 real beta, real *C, int ldc) {
  /* A copied with its row count padded to a multiple of the tile height, so
     that every tile reads whole rows (the padding rows are zero and their
     results are discarded) */
  int mp = transA ? (m + MI - 1) / MI * MI : (m + MB - 1) / MB * MB;
  real *Ap = calloc((size_t)mp * k, sizeof(real));
  if (!transA) {
    for (int pp = 0; pp < k; ++pp)
      for (int i = 0; i < m; ++i) Ap[(size_t)pp * mp + i] = A[(size_t)pp * lda + i];
  } else {   /* Ap(pp, i) = A(pp, i), stored with i slowest: row i of op(A) contiguous */
    for (int i = 0; i < m; ++i)
      for (int pp = 0; pp < k; ++pp) Ap[(size_t)i * k + pp] = A[(size_t)i * lda + pp];
  }
  long jb;
  long nblk = (n + OB - 1) / OB;
#pragma omp parallel for schedule(static)
  for (jb = 0; jb < nblk; ++jb) {
    int j0 = (int)(jb * OB), nj = (n - j0 < OB) ? n - j0 : OB;
    real *bt = malloc(sizeof(real) * (size_t)k * OB);   /* B block transposed: bt[pp * OB + jj] */
    for (int jj = 0; jj < nj; ++jj) {
      real *c = C + (size_t)(j0 + jj) * ldc;
      if (beta == 0.0) { for (int i = 0; i < m; ++i) c[i] = 0.0; }
      else if (beta != 1.0) { for (int i = 0; i < m; ++i) c[i] *= beta; }
    }
    if (!transA) {
      for (int pp = 0; pp < k; ++pp)
        for (int jj = 0; jj < OB; ++jj)
          bt[pp * OB + jj] = (jj < nj) ? alpha * B[(size_t)(j0 + jj) * ldb + pp] : 0.0;
      for (int i0 = 0; i0 < m; i0 += MB) {
        real acc[OB][MB];
        int mi = (m - i0 < MB) ? m - i0 : MB;
        for (int jj = 0; jj < OB; ++jj)
          for (int ii = 0; ii < MB; ++ii)
            acc[jj][ii] = (jj < nj && ii < mi) ? C[(size_t)(j0 + jj) * ldc + i0 + ii] : 0.0;
        tile_nt(k, Ap + i0, mp, bt, acc);
        for (int jj = 0; jj < nj; ++jj)
          for (int ii = 0; ii < mi; ++ii) C[(size_t)(j0 + jj) * ldc + i0 + ii] = acc[jj][ii];
      }
    } else {
      for (int pp = 0; pp < k; ++pp)
        for (int jj = 0; jj < OB; ++jj)
          bt[pp * OB + jj] = (jj < nj) ? B[(size_t)(j0 + jj) * ldb + pp] : 0.0;
      for (int i0 = 0; i0 < m; i0 += MI) {
        real s[MI][OB];
        int mi = (m - i0 < MI) ? m - i0 : MI;
        for (int ii = 0; ii < MI; ++ii)
          for (int jj = 0; jj < OB; ++jj) s[ii][jj] = 0.0;
        tile_t(k, Ap + (size_t)i0 * k, bt, s);
        for (int ii = 0; ii < mi; ++ii)
          for (int jj = 0; jj < nj; ++jj) C[(size_t)(j0 + jj) * ldc + i0 + ii] += alpha * s[ii][jj];
      }
    }
    free(bt);
  }
  free(Ap);
}
#endif

/* Solve A X = B in place (B <- A^{-1} B); A (n,n) is destroyed.
   Gaussian elimination with partial pivoting: restates LAPACKE_dgesv
   (solver.cpp:110-111, dgMath.cpp:561-562, 601-602). Returns 0 or -1. */
static int lu_solve(int n, real *A, int nrhs, real *B) {
  for (int c = 0; c < n; ++c) {
    int piv = c; real best = R_FABS(A[c + (size_t)c * n]);
    for (int r = c + 1; r < n; ++r) {
      real v = R_FABS(A[r + (size_t)c * n]);
      if (v > best) { best = v; piv = r; }
    }
    if (best == 0.0) return -1;
    if (piv != c) {
      for (int j = 0; j < n; ++j) { real t = A[c + (size_t)j*n]; A[c + (size_t)j*n] = A[piv + (size_t)j*n]; A[piv + (size_t)j*n] = t; }
      for (int j = 0; j < nrhs; ++j) { real t = B[c + (size_t)j*n]; B[c + (size_t)j*n] = B[piv + (size_t)j*n]; B[piv + (size_t)j*n] = t; }
    }
    real d = A[c + (size_t)c * n];
    for (int r = c + 1; r < n; ++r) {
      real f = A[r + (size_t)c * n] / d;
      if (f == 0.0) continue;
      for (int j = c; j < n; ++j) A[r + (size_t)j*n] -= f * A[c + (size_t)j*n];
      for (int j = 0; j < nrhs; ++j) B[r + (size_t)j*n] -= f * B[c + (size_t)j*n];
    }
  }
  for (int j = 0; j < nrhs; ++j) {
    real *b = B + (size_t)j * n;
    for (int r = n - 1; r >= 0; --r) {
      real s = b[r];
      for (int c = r + 1; c < n; ++c) s -= A[r + (size_t)c * n] * b[c];
      b[r] = s / A[r + (size_t)r * n];
    }
  }
  return 0;
}

/* Cholesky factor of an SPD matrix, lower triangle kept in L.
   Replaces the Bunch-Kaufman LDL^T of LAPACKE_dsytrf (solver.cpp:155-156):
   the mass matrix is SPD, so both factorizations solve the same system. */
static int cholesky(int n, const real *A, real *L) {
  memset(L, 0, sizeof(real) * (size_t)n * n);
  for (int j = 0; j < n; ++j) {
    real s = A[j + (size_t)j * n];
    for (int k = 0; k < j; ++k) s -= L[j + (size_t)k * n] * L[j + (size_t)k * n];
    if (s <= 0.0) return -1;
    real d = R_SQRT(s);
    L[j + (size_t)j * n] = d;
    for (int i = j + 1; i < n; ++i) {
      real t = A[i + (size_t)j * n];
      for (int k = 0; k < j; ++k) t -= L[i + (size_t)k * n] * L[j + (size_t)k * n];
      L[i + (size_t)j * n] = t / d;
    }
  }
  return 0;
}

/* B <- Mel^{-1} B for nrhs columns, using the Cholesky factor: forward
   substitution with L, then backward with L^T, each row
   b_i = (b_i - sum_k l_ik b_k) / l_ii with the sum in k order.
   Restates LAPACKE_dsytrs (solver.cpp:505-506, 532-533).  Columns are
   solved CB at a time (independent chains for the vector units), each in
   exactly that order. */
#ifdef ORC_UNTILED
static void chol_solve(int n, const real *L, long nrhs, real *B) {
  long j;
#pragma omp parallel for schedule(static)
  for (j = 0; j < nrhs; ++j) {
    real *b = B + (size_t)j * n;
    for (int i = 0; i < n; ++i) {
      real s = b[i];
      const real *li = L + i;
      for (int k = 0; k < i; ++k) s -= li[(size_t)k * n] * b[k];
      b[i] = s / L[i + (size_t)i * n];
    }
    for (int i = n - 1; i >= 0; --i) {
      real s = b[i];
      const real *col = L + (size_t)i * n;
      for (int k = i + 1; k < n; ++k) s -= col[k] * b[k];
      b[i] = s / col[i];
    }
  }
}
#else
#define CB 32

/* s[jj] -= l[t * ls] * bt[t * CB + jj] for t = 0..cnt-1 in order. */
static void row_update(int cnt, const real *l, long ls, const real *bt, real s[CB]) {
  const long bs = CB;
#if defined(ORC_LONG_DOUBLE) || !defined(__GNUC__)
  for (int t = 0; t < cnt; ++t) {
    const real lv = l[t * ls];
    for (int jj = 0; jj < CB; ++jj) s[jj] -= lv * bt[t * bs + jj];
  }
#else
  typedef double v8 __attribute__((vector_size(64)));
  v8 c0, c1, c2, c3;
  memcpy(&c0, s, 64); memcpy(&c1, s + 8, 64); memcpy(&c2, s + 16, 64); memcpy(&c3, s + 24, 64);
  for (int t = 0; t < cnt; ++t) {
    const double lv = l[t * ls];
    const double *b = bt + t * bs;
    v8 b0, b1, b2, b3;
    memcpy(&b0, b, 64); memcpy(&b1, b + 8, 64); memcpy(&b2, b + 16, 64); memcpy(&b3, b + 24, 64);
    c0 -= lv * b0; c1 -= lv * b1; c2 -= lv * b2; c3 -= lv * b3;
  }
  memcpy(s, &c0, 64); memcpy(s + 8, &c1, 64); memcpy(s + 16, &c2, 64); memcpy(s + 24, &c3, 64);
#endif
}
static void chol_solve(int n, const real *L, long nrhs, real *B) {
  long jb;
  long nblk = (nrhs + CB - 1) / CB;
#pragma omp parallel for schedule(static)
  for (jb = 0; jb < nblk; ++jb) {
    long j0 = jb * CB;
    int nj = (nrhs - j0 < CB) ? (int)(nrhs - j0) : CB;
    real *bt = malloc(sizeof(real) * (size_t)n * CB);   /* bt[i * CB + jj] */
    real s[CB];
    for (int i = 0; i < n; ++i)
      for (int jj = 0; jj < CB; ++jj)
        bt[i * CB + jj] = (jj < nj) ? B[(size_t)(j0 + jj) * n + i] : 0.0;
    for (int i = 0; i < n; ++i) {        /* forward: L y = b */
      for (int jj = 0; jj < CB; ++jj) s[jj] = bt[i * CB + jj];
      row_update(i, L + i, n, bt, s);
      const real d = L[i + (size_t)i * n];
      for (int jj = 0; jj < CB; ++jj) bt[i * CB + jj] = s[jj] / d;
    }
    for (int i = n - 1; i >= 0; --i) {   /* backward: L^T x = y */
      for (int jj = 0; jj < CB; ++jj) s[jj] = bt[i * CB + jj];
      const real *col = L + (size_t)i * n;
      row_update(n - 1 - i, col + i + 1, 1, bt + (size_t)(i + 1) * CB, s);
      const real d = col[i];
      for (int jj = 0; jj < CB; ++jj) bt[i * CB + jj] = s[jj] / d;
    }
    for (int jj = 0; jj < nj; ++jj)
      for (int i = 0; i < n; ++i) B[(size_t)(j0 + jj) * n + i] = bt[i * CB + jj];
    free(bt);
  }
}
#endif

/* Eigen-decomposition of a small symmetric matrix by cyclic Jacobi
   rotations.  A (n,n) is overwritten; eigenvalues in ev, eigenvectors as
   columns of V.  Stands in for LAPACKE_dsyev in gaussQuad
   (dgMath.cpp:105). */
static void jacobi_eig(int n, real *A, real *ev, real *V) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) V[i + j * n] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    real off = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) off += A[i + j * n] * A[i + j * n];
    if (off < 1e-300) break;
    for (int pp = 0; pp < n; ++pp) {
      for (int q = pp + 1; q < n; ++q) {
        real apq = A[pp + q * n];
        if (R_FABS(apq) < 1e-300) continue;
        real app = A[pp + pp * n], aqq = A[q + q * n];
        real theta = (aqq - app) / (2.0 * apq);
        real t = (theta >= 0 ? 1.0 : -1.0) / (R_FABS(theta) + R_SQRT(theta * theta + 1.0));
        real c = 1.0 / R_SQRT(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  /* A <- A G */
          real akp = A[k + pp * n], akq = A[k + q * n];
          A[k + pp * n] = c * akp - s * akq;
          A[k + q * n] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  /* A <- G^T A */
          real apk = A[pp + k * n], aqk = A[q + k * n];
          A[pp + k * n] = c * apk - s * aqk;
          A[q + k * n] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          real vkp = V[k + pp * n], vkq = V[k + q * n];
          V[k + pp * n] = c * vkp - s * vkq;
          V[k + q * n] = s * vkp + c * vkq;
        }
      }
    }
  }
  for (int i = 0; i < n; ++i) ev[i] = A[i + i * n];
}

/* ------------------------------------------------------------------ */
/* dgMath restatement                                                   */
/* ------------------------------------------------------------------ */

/* Chebyshev-Gauss-Lobatto nodes, ascending: cos((p-i) pi / p)
   (dgMath.cpp:37-45). */
static void chebyshev(int p, real *x) {
  for (int i = 0; i <= p; ++i) x[i] = R_COS((p - i) * R_PI / p);
}

/* Golub-Welsch Gauss-Legendre rule with n = ceil((p+1)/2) points on [-1,1]
   (dgMath.cpp:91-114): eigenvalues of the Jacobi matrix with off-diagonal
   (i+1)/sqrt(4(i+1)^2-1) are the nodes; w_i = 2 v_{0,i}^2.  Ascending. */
static int gauss_quad(int p, real *x, real *w) {
  int n = (int)R_CEIL((p + 1) / 2.0);
  real A[64], V[64], ev[8];
  memset(A, 0, sizeof(A));
  for (int i = 0; i < n - 1; ++i) {
    real t = (i + 1) / R_SQRT(4.0 * (i + 1) * (i + 1) - 1.0);
    A[i + (i + 1) * n] = t; A[(i + 1) + i * n] = t;
  }
  jacobi_eig(n, A, ev, V);
  int idx[8];
  for (int i = 0; i < n; ++i) idx[i] = i;
  for (int i = 0; i < n; ++i)            /* sort ascending (dsyev order) */
    for (int j = i + 1; j < n; ++j)
      if (ev[idx[j]] < ev[idx[i]]) { int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
  for (int i = 0; i < n; ++i) {
    x[i] = ev[idx[i]];
    real v0 = V[0 + idx[i] * n];
    w[i] = 2.0 * v0 * v0;
  }
  return n;
}

/* Legendre polynomials P_0..P_p at x by the three-term recurrence
   (dgMath.cpp:177-209); out(i, ip) column-major with leading dim nx. */
static void legendre(int p, int nx, const real *x, real *P) {
  for (int i = 0; i < nx; ++i) P[i] = 1.0;
  if (p == 0) return;
  for (int i = 0; i < nx; ++i) P[i + nx] = x[i];
  for (int ip = 0; ip < p - 1; ++ip)
    for (int i = 0; i < nx; ++i)
      P[i + (size_t)(ip + 2) * nx] = ((2 * ip + 3) * x[i] * P[i + (size_t)(ip + 1) * nx]
                                      - (ip + 1) * P[i + (size_t)ip * nx]) / (ip + 2);
}

/* Derivatives of the Legendre polynomials (dgMath.cpp:216-248). */
static void dlegendre(int p, int nx, const real *x, const real *P, real *dP) {
  for (int i = 0; i < nx; ++i) dP[i] = 0.0;
  if (p == 0) return;
  for (int i = 0; i < nx; ++i) dP[i + nx] = 1.0;
  for (int ip = 0; ip < p - 1; ++ip)
    for (int i = 0; i < nx; ++i)
      dP[i + (size_t)(ip + 2) * nx] = ((2 * ip + 3) * (x[i] * dP[i + (size_t)(ip + 1) * nx] + P[i + (size_t)(ip + 1) * nx])
                                       - (ip + 1) * dP[i + (size_t)ip * nx]) / (ip + 2);
}

/* Tensor-product Legendre Vandermonde on a tensor grid of n1^D points given
   by 1D coordinates x1 (dgMath.cpp:254-294, 361-411): V(pt, mode) with
   pt = ix + n1 iy (+ n1^2 iz), mode = ipx + n iy (+ n^2 ipz).  If dirDeriv
   >= 0 the factor in that direction is differentiated (dgMath.cpp:417-501). */
static void legendre_nd(int D, int p, int n1, const real *x1, int dirDeriv, real *V) {
  int n = p + 1;
  real *P = ralloc((size_t)n1 * n), *dP = ralloc((size_t)n1 * n);
  legendre(p, n1, x1, P);
  dlegendre(p, n1, x1, P, dP);
  int npts = (D == 2) ? n1 * n1 : n1 * n1 * n1;
  int nmod = (D == 2) ? n * n : n * n * n;
  for (int m = 0; m < nmod; ++m) {
    int mm[3] = { m % n, (m / n) % n, m / (n * n) };
    for (int q = 0; q < npts; ++q) {
      int qq[3] = { q % n1, (q / n1) % n1, q / (n1 * n1) };
      real v = 1.0;
      for (int d = 0; d < D; ++d) {
        const real *T = (d == dirDeriv) ? dP : P;
        v *= T[qq[d] + (size_t)mm[d] * n1];
      }
      V[q + (size_t)m * npts] = v;
    }
  }
  free(P); free(dP);
}

/* Interpolation matrix from the CGL tensor grid to the Gauss tensor grid:
   INTERP = L_to * inv(L_from) (dgMath.cpp:545-615).  If dirDeriv >= 0 the
   result is the derivative of the nodal basis at the Gauss points
   (solver.cpp:116-129, dPhiQ). */
static void interp_matrix(int D, int p, const real *xfrom1, const real *xto1,
                          int dirDeriv, real *out) {
  int n = p + 1;
  int nf = (D == 2) ? n * n : n * n * n;
  real *Lf = ralloc((size_t)nf * nf), *C = calloc((size_t)nf * nf, sizeof(real));
  real *Lt = ralloc((size_t)nf * nf);
  legendre_nd(D, p, n, xfrom1, -1, Lf);
  for (int i = 0; i < nf; ++i) C[i + (size_t)i * nf] = 1.0;
  lu_solve(nf, Lf, nf, C);
  legendre_nd(D, p, n, xto1, dirDeriv, Lt);
  gemm(0, nf, nf, nf, 1.0, Lt, nf, C, nf, 0.0, out, nf);
  free(Lf); free(C); free(Lt);
}

/* ------------------------------------------------------------------ */
/* Mesh restatement (mesh.cpp:51-355, 389-501)                          */
/* ------------------------------------------------------------------ */

/* Element id ix + nx iy + nx ny iz (mesh.cpp:162).  Periodic neighbour
   across face f in order -x,+x,-y,+y,-z,+z (mesh.cpp:212-318; the ghost
   slots on one periodic rank hold exactly this neighbour's data). */
static int neighbor(const oracle_t *o, int K, int f) {
  int ix = K % o->nx, iy = (K / o->nx) % o->ny, iz = K / (o->nx * o->ny);
  int d = f / 2, s = (f % 2) ? 1 : -1;
  if (d == 0) ix = (ix + s + o->nx) % o->nx;
  if (d == 1) iy = (iy + s + o->ny) % o->ny;
  if (d == 2) iz = (iz + s + o->nz) % o->nz;
  return ix + o->nx * (iy + o->ny * iz);
}

/* Face-node map efToN(iFN, f) (mesh.cpp:428-501). */
static void init_face_map(int n, int *ef) {
  for (int b = 0; b < n; ++b)
    for (int a = 0; a < n; ++a) {
      int j = a + n * b;
      ef[j + 0 * n * n] = 0 + n * a + n * n * b;             /* -x: (iy,iz) */
      ef[j + 1 * n * n] = (n - 1) + n * a + n * n * b;       /* +x */
      ef[j + 2 * n * n] = a + 0 + n * n * b;                 /* -y: (ix,iz) */
      ef[j + 3 * n * n] = a + n * (n - 1) + n * n * b;       /* +y */
      ef[j + 4 * n * n] = a + n * b + 0;                     /* -z: (ix,iy) */
      ef[j + 5 * n * n] = a + n * b + n * n * (n - 1);       /* +z */
    }
}

/* Outward normal component l of face f (mesh.cpp:338-353). */
static real normal_comp(int f, int l) {
  if (f / 2 != l) return 0.0;
  return (f % 2) ? 1.0 : -1.0;
}

/* ------------------------------------------------------------------ */
/* Pointwise physics (textbook forms of solver.cpp:1040-1374)           */
/* ------------------------------------------------------------------ */

/* Euler flux F(iS + 5 l) of conserved state u (solver.cpp:1040-1086).
   The reference hard-codes gamma = 1.4 in the pressure and uses p.gamma in
   the energy flux; with gamma = 1.4 both agree with this form. */
static void fluxC(const real *u, real gamma, real *F) {
  real rho = u[0], ir = 1.0 / rho;
  real v[3] = { u[1] * ir, u[2] * ir, u[3] * ir };
  real pr = (gamma - 1.0) * (u[4] - 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
  for (int l = 0; l < 3; ++l) {
    F[0 + 5 * l] = u[1 + l];
    for (int i = 0; i < 3; ++i) F[1 + i + 5 * l] = u[1 + i] * v[l] + (i == l ? pr : 0.0);
    F[4 + 5 * l] = (u[4] + pr) * v[l];
  }
}

/* Navier-Stokes viscous flux Fv(iS + 5 l) from the state u and conserved
   gradients Du(iS + 5 l) (solver.cpp:1089-1188): Newtonian stress with
   mu = 1/Re and Stokes' hypothesis, heat flux (gamma/(Pr Re)) grad(e) with
   e = E/rho - |v|^2/2. */
static void fluxV(const real *u, const real *Du, real Re, real Pr, real gamma, real *F) {
  real rho = u[0], ir = 1.0 / rho;
  real v[3] = { u[1] * ir, u[2] * ir, u[3] * ir };
  real gv[3][3]; /* gv[i][l] = d v_i / d x_l */
  real ge[3];
  for (int l = 0; l < 3; ++l) {
    real drho = Du[0 + 5 * l];
    for (int i = 0; i < 3; ++i) gv[i][l] = (Du[1 + i + 5 * l] - v[i] * drho) * ir;
    real dEr = (Du[4 + 5 * l] - u[4] * ir * drho) * ir;   /* d(E/rho) */
    ge[l] = dEr - (v[0] * gv[0][l] + v[1] * gv[1][l] + v[2] * gv[2][l]);
  }
  real mu = 1.0 / Re, div = gv[0][0] + gv[1][1] + gv[2][2];
  const real two_thirds = (real)2.0 / (real)3.0;
  for (int l = 0; l < 3; ++l) {
    real tau[3];
    for (int i = 0; i < 3; ++i)
      tau[i] = mu * (gv[i][l] + gv[l][i] - (i == l ? two_thirds * div : 0.0));
    F[0 + 5 * l] = 0.0;
    for (int i = 0; i < 3; ++i) F[1 + i + 5 * l] = tau[i];
    F[4 + 5 * l] = v[0] * tau[0] + v[1] * tau[1] + v[2] * tau[2] + mu * gamma / Pr * ge[l];
  }
}

/* Roe flux F*(uN, uK, n) for element K with outward normal n
   (solver.cpp:1194-1374): F* = (F(uN)+F(uK)).n/2 - |A_roe|(uN-uK)/2 with
   Roe averages, eigenvalues un, un+-c and no entropy fix. */
static void roe(const real *uN, const real *uK, const real *n, real gamma, real *out) {
  real gm1 = gamma - 1.0;
  real rN = uN[0], rK = uK[0];
  real vN[3] = { uN[1] / rN, uN[2] / rN, uN[3] / rN };
  real vK[3] = { uK[1] / rK, uK[2] / rK, uK[3] / rK };
  real pN = gm1 * (uN[4] - 0.5 * rN * (vN[0]*vN[0] + vN[1]*vN[1] + vN[2]*vN[2]));
  real pK = gm1 * (uK[4] - 0.5 * rK * (vK[0]*vK[0] + vK[1]*vK[1] + vK[2]*vK[2]));
  real HN = (uN[4] + pN) / rN, HK = (uK[4] + pK) / rK;
  real qN = vN[0]*n[0] + vN[1]*n[1] + vN[2]*n[2];
  real qK = vK[0]*n[0] + vK[1]*n[1] + vK[2]*n[2];
  /* Roe averages with weight sqrt(rN/rK) on the N side */
  real s = R_SQRT(rN / rK), den = s + 1.0;
  real vt[3], ht, q2 = 0.0, qt = 0.0;
  for (int i = 0; i < 3; ++i) { vt[i] = (s * vN[i] + vK[i]) / den; q2 += vt[i] * vt[i]; qt += vt[i] * n[i]; }
  ht = (s * HN + HK) / den;
  real c2 = gm1 * (ht - 0.5 * q2), c = R_SQRT(c2);
  real l0 = R_FABS(qt), lp = R_FABS(qt + c), lm = R_FABS(qt - c);
  real d[5];
  for (int k = 0; k < 5; ++k) d[k] = uN[k] - uK[k];
  real dp = gm1 * (0.5 * q2 * d[0] - (vt[0]*d[1] + vt[1]*d[2] + vt[2]*d[3]) + d[4]); /* linearised pressure jump */
  real dV = n[0]*d[1] + n[1]*d[2] + n[2]*d[3] - qt * d[0];                              /* rho~ * jump of normal velocity */
  real s1 = 0.5 * (lp + lm) - l0, s2 = 0.5 * (lp - lm);
  real a = s1 * dp / c2 + s2 * dV / c;
  real b = s2 * dp / c + s1 * dV;
  out[0] = 0.5 * (rN * qN + rK * qK) - 0.5 * (l0 * d[0] + a);
  for (int i = 0; i < 3; ++i)
    out[1 + i] = 0.5 * (uN[1 + i] * qN + uK[1 + i] * qK + n[i] * (pN + pK))
               - 0.5 * (l0 * d[1 + i] + a * vt[i] + b * n[i]);
  out[4] = 0.5 * ((uN[4] + pN) * qN + (uK[4] + pK) * qK) - 0.5 * (l0 * d[4] + a * ht + b * qt);
}

/* Exported double-precision forms of the three pointwise fluxes (tests). */
void orc_fluxC(const double *u, double gamma, double *F) {
  real ur[5], Fr[15];
  for (int i = 0; i < 5; ++i) ur[i] = u[i];
  fluxC(ur, gamma, Fr);
  to_double(Fr, F, 15);
}

void orc_fluxV(const double *u, const double *Du, double Re, double Pr, double gamma, double *F) {
  real ur[5], Dr[15], Fr[15];
  for (int i = 0; i < 5; ++i) ur[i] = u[i];
  for (int i = 0; i < 15; ++i) Dr[i] = Du[i];
  fluxV(ur, Dr, Re, Pr, gamma, Fr);
  to_double(Fr, F, 15);
}

void orc_roe(const double *uN, const double *uK, const double *n, double gamma, double *out) {
  real a[5], b[5], nr[3], o[5];
  for (int i = 0; i < 5; ++i) { a[i] = uN[i]; b[i] = uK[i]; }
  for (int i = 0; i < 3; ++i) nr[i] = n[i];
  roe(a, b, nr, gamma, o);
  to_double(o, out, 5);
}

/* LDG auxiliary flux: the K-side value if n_l > 0, else the N-side one,
   times n_l (solver.cpp:867-884). */
static real flux_L(real uK, real uN, real nl) { return (nl > 0 ? uK : uN) * nl; }

/* One-sided viscous numerical flux: -sum_l Fv_l(uK, Du*) n_l with
   Du* = DuN if sum(n) > 0 else DuK (solver.cpp:891-921). */
static void flux_V_num(const oracle_t *o, const real *uK, const real *DuN,
                       const real *DuK, const real *n, real *out) {
  real F[15];
  real sum = n[0] + n[1] + n[2];
  fluxV(uK, (sum > 0) ? DuN : DuK, o->Re, o->Pr, o->gamma, F);
  for (int s = 0; s < 5; ++s) {
    out[s] = 0.0;
    for (int l = 0; l < 3; ++l) out[s] -= F[s + 5 * l] * n[l];
  }
}

/* ------------------------------------------------------------------ */
/* Setup: Solver::Solver + precompute* (solver.cpp:7-192)              */
/* ------------------------------------------------------------------ */

void orc_destroy(oracle_t *o);

oracle_t *orc_create(int p, int nx, int ny, int nz, const double *xmin, const double *xmax,
                     double L, const double *phys /* Re,Pr,gamma,M0,rho0,V0,R or NULL */) {
  if (p < 1 || p > 8 || nx < 1 || ny < 1 || nz < 1) return NULL;
  oracle_t *o = calloc(1, sizeof(oracle_t));
  o->p = p; o->n = p + 1;
  int n = o->n;
  o->dofs = n * n * n; o->nFN = n * n;
  o->nx = nx; o->ny = ny; o->nz = nz; o->K = nx * ny * nz;
  for (int d = 0; d < 3; ++d) { o->xmin[d] = xmin[d]; o->xmax[d] = xmax[d]; }
  int ns[3] = { nx, ny, nz };
  for (int d = 0; d < 3; ++d) {
    o->dx[d] = (o->xmax[d] - o->xmin[d]) / ns[d];   /* mesh.cpp:149-151 */
    o->alpha[d] = o->dx[d] / 2.0;                   /* solver.cpp:144-147 */
  }
  o->J = o->alpha[0] * o->alpha[1] * o->alpha[2];
  /* physics constants (solver.cpp:30-43), as doubles in either build */
  double ph[7] = { 1600.0, 0.71, 1.4, 0.1, 1.0, 1.0, 8.314 };
  if (phys) for (int i = 0; i < 7; ++i) ph[i] = phys[i];
  o->Re = ph[0]; o->Pr = ph[1]; o->gamma = ph[2]; o->M0 = ph[3]; o->rho0 = ph[4]; o->V0 = ph[5]; o->R = ph[6];
  o->L = L;
  o->c0 = o->V0 / o->M0;
  o->p0 = o->rho0 * o->c0 * o->c0 / o->gamma;
  o->T0 = o->p0 / (o->R * o->rho0);

  /* quadrature: gaussQuad(2*order) -> order+1 points (solver.cpp:80, 88) */
  o->xg = ralloc(8); o->wg = ralloc(8);
  int nq = gauss_quad(2 * p, o->xg, o->wg);
  o->nQ2D = nq * nq; o->nQ3D = nq * nq * nq;
  o->wQ2D = ralloc(o->nQ2D);
  o->wQ3D = ralloc(o->nQ3D);
  for (int b = 0; b < nq; ++b) for (int a = 0; a < nq; ++a) o->wQ2D[a + nq * b] = o->wg[a] * o->wg[b];
  for (int c = 0; c < nq; ++c) for (int b = 0; b < nq; ++b) for (int a = 0; a < nq; ++a)
    o->wQ3D[a + nq * (b + nq * c)] = o->wg[a] * o->wg[b] * o->wg[c];

  real cheb[16];
  chebyshev(p, cheb);
  int dofs = o->dofs, nQ3D = o->nQ3D, nQ2D = o->nQ2D, nFN = o->nFN;

  /* Interp2D / Interp3D (solver.cpp:74-92) */
  o->Interp2D = ralloc((size_t)nQ2D * nFN);
  o->Interp3D = ralloc((size_t)nQ3D * dofs);
  interp_matrix(2, p, cheb, o->xg, -1, o->Interp2D);
  interp_matrix(3, p, cheb, o->xg, -1, o->Interp3D);

  /* phiQ, dPhiQ (solver.cpp:98-129) */
  real *phiQ = o->Interp3D;                 /* identical construction */
  real *dPhiQ = ralloc((size_t)nQ3D * dofs * 3);
  for (int l = 0; l < 3; ++l) interp_matrix(3, p, cheb, o->xg, l, dPhiQ + (size_t)l * nQ3D * dofs);
  real *wphi = ralloc((size_t)nQ3D * dofs);   /* polyQuad = w * phiQ (solver.cpp:131-140) */
  for (int j = 0; j < dofs; ++j) for (int q = 0; q < nQ3D; ++q) wphi[q + (size_t)j * nQ3D] = phiQ[q + (size_t)j * nQ3D] * o->wQ3D[q];

  /* Mel = J phiQ^T W phiQ (solver.cpp:142-156) */
  o->Mel = ralloc((size_t)dofs * dofs);
  gemm(1, dofs, dofs, nQ3D, o->J, wphi, nQ3D, phiQ, nQ3D, 0.0, o->Mel, dofs);
  o->Mchol = ralloc((size_t)dofs * dofs);
  if (cholesky(dofs, o->Mel, o->Mchol) != 0) { orc_destroy(o); free(dPhiQ); free(wphi); return NULL; }

  /* Sels_l = (J/alpha_l) dPhiQ_l^T W phiQ (solver.cpp:158-165) */
  o->Sels = ralloc((size_t)dofs * dofs * 3);
  for (int l = 0; l < 3; ++l)
    gemm(1, dofs, dofs, nQ3D, o->J / o->alpha[l], dPhiQ + (size_t)l * nQ3D * dofs, nQ3D,
         wphi, nQ3D, 0.0, o->Sels + (size_t)l * dofs * dofs, dofs);

  /* Kels_l(q,i) = dPhiQ_l(q,i) w_q J/alpha_l (solver.cpp:167-176) */
  o->Kels = ralloc((size_t)nQ3D * dofs * 3);
  for (int l = 0; l < 3; ++l)
    for (int i = 0; i < dofs; ++i)
      for (int q = 0; q < nQ3D; ++q)
        o->Kels[q + (size_t)nQ3D * (i + (size_t)dofs * l)] = dPhiQ[q + (size_t)nQ3D * (i + (size_t)dofs * l)] * o->wQ3D[q] * (o->J / o->alpha[l]);

  /* Kels2D_l(r,j) = Interp2D(r,j) w2_r J/alpha_l (solver.cpp:178-190) */
  o->Kels2D = ralloc((size_t)nQ2D * nFN * 3);
  for (int l = 0; l < 3; ++l)
    for (int j = 0; j < nFN; ++j)
      for (int r = 0; r < nQ2D; ++r)
        o->Kels2D[r + (size_t)nQ2D * (j + (size_t)nFN * l)] = o->Interp2D[r + (size_t)nQ2D * j] * o->wQ2D[r] * (o->J / o->alpha[l]);
  free(dPhiQ); free(wphi);

  /* mesh connectivity and node coordinates (mesh.cpp:127-146, 389-418) */
  o->efToN = malloc(sizeof(int) * nFN * 6);
  init_face_map(n, o->efToN);
  o->coords = ralloc(3 * (size_t)dofs * o->K);
  for (int k = 0; k < o->K; ++k) {
    int ie[3] = { k % nx, (k / nx) % ny, k / (nx * ny) };
    for (int iN = 0; iN < dofs; ++iN) {
      int in[3] = { iN % n, (iN / n) % n, iN / (n * n) };
      for (int l = 0; l < 3; ++l) {
        real lo = ie[l] * (o->xmax[l] - o->xmin[l]) / ns[l] + o->xmin[l];        /* vertex of eToV(0,k) */
        real hi = (ie[l] + 1) * (o->xmax[l] - o->xmin[l]) / ns[l] + o->xmin[l];  /* vertex of eToV(7,k) */
        real scale = .5 * (cheb[in[l]] + 1.0);
        o->coords[l + 3 * ((size_t)iN + (size_t)dofs * k)] = lo + scale * (hi - lo);
      }
    }
  }
  size_t K = o->K;
  o->uI2 = ralloc((size_t)nQ2D * NS * 6 * K);
  o->uI3 = ralloc((size_t)nQ3D * NS * K);
  o->DuI2 = ralloc((size_t)nQ2D * NS * 6 * K * 3);
  o->DuI3 = ralloc((size_t)nQ3D * NS * K * 3);
  if (!o->uI2 || !o->uI3 || !o->DuI2 || !o->DuI3) { orc_destroy(o); return NULL; }
  return o;
}

void orc_destroy(oracle_t *o) {
  if (!o) return;
  free(o->xg); free(o->wg); free(o->wQ2D); free(o->wQ3D); free(o->Interp2D); free(o->Interp3D);
  free(o->Mel); free(o->Mchol); free(o->Sels); free(o->Kels); free(o->Kels2D); free(o->efToN);
  free(o->coords); free(o->uI2); free(o->uI3); free(o->DuI2); free(o->DuI3);
  free(o);
}

/* Taylor-Green initial state at the nodes (solver.cpp:194-236). */
void orc_initial_condition(const oracle_t *o, double *u) {
  int dofs = o->dofs;
  for (long k = 0; k < o->K; ++k)
    for (int iN = 0; iN < dofs; ++iN) {
      const real *X = o->coords + 3 * ((size_t)iN + (size_t)dofs * k);
      real x = X[0], y = X[1], z = X[2], L = o->L;
      real pr = o->p0 + (o->rho0 * o->V0 * o->V0 / 16.0) * (R_COS(2 * x / L) + R_COS(2 * y / L)) * (R_COS(2 * z / L) + 2);
      real uk[5];
      uk[0] = pr / (o->R * o->T0);
      uk[1] = o->rho0 * o->V0 * R_SIN(x / L) * R_COS(y / L) * R_COS(z / L);
      uk[2] = -o->rho0 * o->V0 * R_COS(x / L) * R_SIN(y / L) * R_COS(z / L);
      uk[3] = 0 * o->rho0;
      uk[4] = pr / (o->gamma - 1) + (uk[1] * uk[1] + uk[2] * uk[2] + uk[3] * uk[3]) / (2.0 * uk[0]);
      double *ud = u + (size_t)dofs * NS * k;
      for (int s = 0; s < NS; ++s) ud[iN + s * dofs] = (double)uk[s];
    }
}

/* ------------------------------------------------------------------ */
/* RHS: Solver::rk4Rhs (solver.cpp:486-535)                             */
/* ------------------------------------------------------------------ */

/* Solver::interpolate (solver.cpp:541-577): face traces at face Gauss
   points for all 6 faces and volume values at Gauss points, for `dim`
   stacked fields of size (dofs, 5, K). */
static void interpolate(const oracle_t *o, const real *cur, real *I2, real *I3, int dim) {
  int nFN = o->nFN, nQ2D = o->nQ2D, dofs = o->dofs, nQ3D = o->nQ3D;
  size_t K = o->K;
  real *onF = ralloc((size_t)nFN * NS * 6 * K);
  for (int l = 0; l < dim; ++l) {
    long k;
#pragma omp parallel for schedule(static)
    for (k = 0; k < (long)K; ++k)
      for (int f = 0; f < 6; ++f)
        for (int s = 0; s < NS; ++s)
          for (int j = 0; j < nFN; ++j)
            onF[j + (size_t)nFN * (s + NS * (f + 6 * (size_t)k))] =
                cur[o->efToN[j + nFN * f] + (size_t)dofs * (s + NS * ((size_t)k + K * l))];
    gemm(0, nQ2D, NS * 6 * K, nFN, 1.0, o->Interp2D, nQ2D, onF, nFN, 0.0, I2 + (size_t)nQ2D * NS * 6 * K * l, nQ2D);
  }
  gemm(0, nQ3D, NS * K * dim, dofs, 1.0, o->Interp3D, nQ3D, cur, dofs, 0.0, I3, nQ3D);
  free(onF);
}

/* Lift a face flux fStar(nQ2D, 5) with Kels2D(:,:,f/2)^T and add it to the
   face nodes of `res` element k (solver.cpp:619-629, 683-693, 799-809). */
static void lift_add(const oracle_t *o, int f, const real *fStar, real *resk) {
  int nFN = o->nFN, nQ2D = o->nQ2D, dofs = o->dofs;
  const real *K2 = o->Kels2D + (size_t)nQ2D * nFN * (f / 2);
  for (int s = 0; s < NS; ++s)
    for (int j = 0; j < nFN; ++j) {
      real acc = 0.0;
      for (int r = 0; r < nQ2D; ++r) acc += K2[r + (size_t)nQ2D * j] * fStar[r + nQ2D * s];
      resk[o->efToN[j + nFN * f] + (size_t)dofs * s] += acc;
    }
}

/* rk4Rhs: k = Mel^{-1}( Kels^T (fc - fv) - lift(Fc* + Fv*) ), with the
   auxiliary gradient Du_l = Mel^{-1}( lift(uhat n_l) - Sels_l u ).
   Writes k (dofs,5,K) and, if Du != NULL, Du (dofs,5,K,3). */
static void rhs(oracle_t *o, const real *u, real *kout, real *Du_out) {
  int dofs = o->dofs, nQ2D = o->nQ2D, nQ3D = o->nQ3D;
  size_t K = o->K, nU = (size_t)dofs * NS * K;
  real *Dus = Du_out ? Du_out : ralloc(nU * 3);

  interpolate(o, u, o->uI2, o->uI3, 1);                       /* solver.cpp:491 */

  memset(Dus, 0, sizeof(real) * nU * 3);                      /* solver.cpp:495 */
  long k;
#pragma omp parallel for schedule(static)
  for (k = 0; k < (long)K; ++k) {                              /* localDGFlux, solver.cpp:585-635 */
    real fStar[64 * NS];
    for (int l = 0; l < 3; ++l)
      for (int f = 0; f < 6; ++f) {
        real nl = normal_comp(f, l);
        if (nl == 0.0) continue;   /* zero normal: the reference adds a zero flux */
        int N = neighbor(o, (int)k, f), fN = f ^ 1;
        for (int s = 0; s < NS; ++s)
          for (int r = 0; r < nQ2D; ++r) {
            real uK = o->uI2[r + (size_t)nQ2D * (s + NS * (f + 6 * (size_t)k))];
            real uN = o->uI2[r + (size_t)nQ2D * (s + NS * (fN + 6 * (size_t)N))];
            fStar[r + nQ2D * s] = flux_L(uK, uN, nl);
          }
        lift_add(o, f, fStar, Dus + (size_t)dofs * NS * (k + K * l));
      }
  }
  for (int l = 0; l < 3; ++l) {                                /* solver.cpp:497-508 */
    real *Dl = Dus + nU * l;
    gemm(0, dofs, NS * K, dofs, -1.0, o->Sels + (size_t)dofs * dofs * l, dofs, u, dofs, 1.0, Dl, dofs);
    chol_solve(dofs, o->Mchol, NS * K, Dl);
  }

  interpolate(o, Dus, o->DuI2, o->DuI3, 3);                    /* solver.cpp:511 */

  real *res = kout;                                            /* solver.cpp:516-517 */
  memset(res, 0, sizeof(real) * nU);
  size_t s2 = (size_t)nQ2D * NS * 6 * K;                       /* stride of l in DuI2 */
#pragma omp parallel for schedule(static)
  for (k = 0; k < (long)K; ++k) {
    real fStar[64 * NS], uK[5], uN[5], DuK[15], DuN[15], fl[5];
    real *resk = res + (size_t)dofs * NS * k;
    for (int f = 0; f < 6; ++f) {                               /* convectDGFlux, solver.cpp:643-698 */
      int N = neighbor(o, (int)k, f), fN = f ^ 1;
      real nrm[3] = { normal_comp(f, 0), normal_comp(f, 1), normal_comp(f, 2) };
      for (int r = 0; r < nQ2D; ++r) {
        for (int s = 0; s < NS; ++s) {
          uK[s] = o->uI2[r + (size_t)nQ2D * (s + NS * (f + 6 * (size_t)k))];
          uN[s] = o->uI2[r + (size_t)nQ2D * (s + NS * (fN + 6 * (size_t)N))];
        }
        roe(uN, uK, nrm, o->gamma, fl);
        for (int s = 0; s < NS; ++s) fStar[r + nQ2D * s] = fl[s];
      }
      lift_add(o, f, fStar, resk);
    }
    for (int f = 0; f < 6; ++f) {                               /* viscousDGFlux, solver.cpp:750-814 */
      int N = neighbor(o, (int)k, f), fN = f ^ 1;
      real nrm[3] = { normal_comp(f, 0), normal_comp(f, 1), normal_comp(f, 2) };
      for (int r = 0; r < nQ2D; ++r) {
        for (int s = 0; s < NS; ++s) {
          uK[s] = o->uI2[r + (size_t)nQ2D * (s + NS * (f + 6 * (size_t)k))];
          uN[s] = o->uI2[r + (size_t)nQ2D * (s + NS * (fN + 6 * (size_t)N))];
          for (int l = 0; l < 3; ++l) {
            DuK[s + 5 * l] = o->DuI2[r + (size_t)nQ2D * (s + NS * (f + 6 * (size_t)k)) + s2 * l];
            DuN[s + 5 * l] = o->DuI2[r + (size_t)nQ2D * (s + NS * (fN + 6 * (size_t)N)) + s2 * l];
          }
        }
        flux_V_num(o, uK, DuN, DuK, nrm, fl);
        for (int s = 0; s < NS; ++s) fStar[r + nQ2D * s] = fl[s];
      }
      lift_add(o, f, fStar, resk);
    }
    for (int i = 0; i < dofs * NS; ++i) resk[i] = -resk[i];    /* dscal(-1), solver.cpp:523 */
  }

  /* volume fluxes fc - fv at the Gauss points (solver.cpp:701-741, 817-864) */
  real *fcv = ralloc((size_t)nQ3D * NS * K * 3);
  size_t s3 = (size_t)nQ3D * NS * K;
#pragma omp parallel for schedule(static)
  for (k = 0; k < (long)K; ++k) {
    real uq[5], Duq[15], Fc[15], Fv[15];
    for (int q = 0; q < nQ3D; ++q) {
      for (int s = 0; s < NS; ++s) {
        uq[s] = o->uI3[q + (size_t)nQ3D * (s + NS * (size_t)k)];
        for (int l = 0; l < 3; ++l) Duq[s + 5 * l] = o->DuI3[q + (size_t)nQ3D * (s + NS * (size_t)k) + s3 * l];
      }
      fluxC(uq, o->gamma, Fc);
      fluxV(uq, Duq, o->Re, o->Pr, o->gamma, Fv);
      for (int l = 0; l < 3; ++l)
        for (int s = 0; s < NS; ++s)
          fcv[q + (size_t)nQ3D * (s + NS * (size_t)k) + s3 * l] = Fc[s + 5 * l] - Fv[s + 5 * l];
    }
  }
  for (int l = 0; l < 3; ++l)                                  /* residual += Kels_l^T (fc_l - fv_l) */
    gemm(1, dofs, NS * K, nQ3D, 1.0, o->Kels + (size_t)nQ3D * dofs * l, nQ3D, fcv + s3 * l, nQ3D, 1.0, res, dofs);
  free(fcv);

  chol_solve(dofs, o->Mchol, NS * K, res);                     /* solver.cpp:531-533 */
  if (!Du_out) free(Dus);
}

void orc_rhs(oracle_t *o, const double *u, double *kout, double *Du_out) {
  size_t nU = (size_t)o->dofs * NS * o->K;
  real *ur = to_real(u, nU), *kr = ralloc(nU), *Dr = Du_out ? ralloc(nU * 3) : NULL;
  rhs(o, ur, kr, Dr);
  to_double(kr, kout, nU);
  if (Du_out) to_double(Dr, Du_out, nU * 3);
  free(ur); free(kr); free(Dr);
}

/* Solver::computeKE (solver.cpp:1007-1034): sum over elements and nodes of
   J w_q |m|^2/(2 rho) using NODAL values u(iFQ, ., k) weighted by the Gauss
   weights wQ(iFQ) (the reference's definition), normalised by
   rho0 (2 pi L)^3. */
static real ke_sum(const oracle_t *o, const real *u) {
  int dofs = o->dofs;
  real total = 0.0;
  for (long k = 0; k < o->K; ++k) {
    const real *uk = u + (size_t)dofs * NS * k;
    real ke = 0.0;
    for (int q = 0; q < o->nQ3D; ++q)
      ke += o->J * o->wQ3D[q] * (uk[q + dofs] * uk[q + dofs] + uk[q + 2 * dofs] * uk[q + 2 * dofs]
                                 + uk[q + 3 * dofs] * uk[q + 3 * dofs]) / (2 * uk[q]);
    total += ke;
  }
  return total / (o->rho0 * R_POW(2 * R_PI * o->L, 3.0));
}

double orc_ke(const oracle_t *o, const double *u) {
  real *ur = to_real(u, (size_t)o->dofs * NS * o->K);
  double r = (double)ke_sum(o, ur);
  free(ur);
  return r;
}

/* Gauss-point diagnostics (this project's definitions; no reference
   counterpart): out[0] = kinetic energy (1/(rho0 V)) int |m|^2/(2 rho),
   out[1] = enstrophy (1/(rho0 V)) int rho |curl v|^2 / 2, both by the
   element Gauss rule on the interpolated state and LDG gradient. */
void orc_diagnostics(oracle_t *o, const double *u, double *out) {
  int dofs = o->dofs, nQ3D = o->nQ3D;
  size_t K = o->K, nU = (size_t)dofs * NS * K;
  real *ur = to_real(u, nU), *Du = ralloc(nU * 3), *k = ralloc(nU);
  rhs(o, ur, k, Du);              /* leaves uI3 / DuI3 at the Gauss points */
  size_t s3 = (size_t)nQ3D * NS * K;
  real ke = 0.0, ens = 0.0;
  for (size_t e = 0; e < K; ++e) {
    real kel = 0.0, enl = 0.0;
    for (int q = 0; q < nQ3D; ++q) {
      real U[5], G[15];
      for (int s = 0; s < NS; ++s) {
        U[s] = o->uI3[q + (size_t)nQ3D * (s + NS * e)];
        for (int l = 0; l < 3; ++l) G[s + 5 * l] = o->DuI3[q + (size_t)nQ3D * (s + NS * e) + s3 * l];
      }
      real ir = 1.0 / U[0], v[3] = { U[1] * ir, U[2] * ir, U[3] * ir }, gv[3][3];
      for (int i = 0; i < 3; ++i) for (int l = 0; l < 3; ++l) gv[i][l] = (G[1 + i + 5 * l] - v[i] * G[0 + 5 * l]) * ir;
      real w1 = gv[2][1] - gv[1][2], w2 = gv[0][2] - gv[2][0], w3 = gv[1][0] - gv[0][1];
      real wq = o->J * o->wQ3D[q];
      kel += wq * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) * ir * 0.5;
      enl += wq * U[0] * (w1 * w1 + w2 * w2 + w3 * w3) * 0.5;
    }
    ke += kel; ens += enl;
  }
  real V = o->rho0 * R_POW(2 * R_PI * o->L, 3.0);
  out[0] = (double)(ke / V); out[1] = (double)(ens / V);
  free(ur); free(Du); free(k);
}

/* RK4 time stepping (solver.cpp:295-480): classic tableau with stored
   stages; uCurr = u + dt a_{s-1} k_{s-1}; after the four stages
   u += dt b_s k_s in stage order; KE of the new state recorded per step.
   The state is carried in `real` between steps and rounded to double once
   at the end. */
int orc_step(oracle_t *o, double *u, double dt, int nsteps, double *ke) {
  const real b[4] = { (real)1 / 6, (real)1 / 3, (real)1 / 3, (real)1 / 6 };
  const real a[3] = { 0.5, 0.5, 1.0 };
  size_t nU = (size_t)o->dofs * NS * o->K;
  real *ur = to_real(u, nU), *ks = ralloc(nU * 4), *uc = ralloc(nU);
  if (!ur || !ks || !uc) { free(ur); free(ks); free(uc); return -1; }
  for (int it = 0; it < nsteps; ++it) {
    for (int s = 0; s < 4; ++s) {
      if (s == 0) memcpy(uc, ur, sizeof(real) * nU);
      else { for (size_t i = 0; i < nU; ++i) uc[i] = ur[i] + dt * a[s - 1] * ks[i + nU * (s - 1)]; }
      rhs(o, uc, ks + nU * s, NULL);
    }
    for (int s = 0; s < 4; ++s)
      for (size_t i = 0; i < nU; ++i) ur[i] += dt * b[s] * ks[i + nU * s];
    if (ke) ke[it] = (double)ke_sum(o, ur);
  }
  to_double(ur, u, nU);
  free(ur); free(ks); free(uc);
  return 0;
}

/* Reference time step (solver.cpp:44-52) for a CFL factor (0.01 in the
   reference): dt = cfl min(dx)/(V0 (2p+1)). */
double orc_dt(const oracle_t *o, double cfl) {
  real m = o->dx[0] < o->dx[1] ? o->dx[0] : o->dx[1];
  if (o->dx[2] < m) m = o->dx[2];
  return (double)(cfl * m / (o->V0 * (2 * o->p + 1)));
}

/* Accessors used by the tests. */
int orc_dofs(const oracle_t *o) { return o->dofs; }
int orc_nelem(const oracle_t *o) { return o->K; }
int orc_nq(const oracle_t *o) { return o->nQ3D; }
/* OpenMP threads of the library's parallel loops (timing the CPU baseline
   at one thread and at all of them) */
void orc_set_threads(int n) { omp_set_num_threads(n); }
int orc_max_threads(void) { return omp_get_max_threads(); }
/* sizeof(real): 8 for the FP64 build, 16 (80-bit x87) for the long double one */
int orc_real_size(void) { return (int)sizeof(real); }
void orc_get_gauss(const oracle_t *o, double *x, double *w) { to_double(o->xg, x, o->n); to_double(o->wg, w, o->n); }
void orc_get_coords(const oracle_t *o, double *c) { to_double(o->coords, c, 3 * (size_t)o->dofs * o->K); }
/* which: 0 Interp2D, 1 Interp3D, 2 Mel, 3 Sels, 4 Kels, 5 Kels2D */
int orc_get_operator(const oracle_t *o, int which, double *out) {
  size_t sz[6] = { (size_t)o->nQ2D * o->nFN, (size_t)o->nQ3D * o->dofs, (size_t)o->dofs * o->dofs,
                   (size_t)o->dofs * o->dofs * 3, (size_t)o->nQ3D * o->dofs * 3, (size_t)o->nQ2D * o->nFN * 3 };
  const real *src[6] = { o->Interp2D, o->Interp3D, o->Mel, o->Sels, o->Kels, o->Kels2D };
  if (which < 0 || which > 5) return -1;
  to_double(src[which], out, sz[which]);
  return (int)sz[which];
}
void orc_get_physics(const oracle_t *o, double *out) {
  real v[11] = { o->Re, o->Pr, o->gamma, o->M0, o->rho0, o->V0, o->R, o->c0, o->p0, o->T0, o->J };
  to_double(v, out, 11);
}
