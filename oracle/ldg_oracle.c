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
 * Parity pins (see tests/): Gauss/CGL nodes vs closed forms and numpy;
 * operator Kronecker structure and exactness on polynomials; fluxes vs
 * independent textbook implementations and the supersonic-upwind limit of
 * Roe; full RHS vs an independent numpy weak-form quadrature; invariants
 * (constant state, mass conservation, axis-permutation and periodic-shift
 * equivariance).  The enstrophy and Gauss-point KE functions have no
 * counterpart in the reference (they are this project's own definitions)
 * and are pinned against closed-form Taylor-Green values.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define NS 5    /* conserved states: rho, rho u, rho v, rho w, rho E */
#define DIM 3
#define NFACES 6

typedef struct {
  int p, n, dofs, nFN, nQ2D, nQ3D;
  int nx, ny, nz, K;
  double xmin[3], xmax[3], dx[3], alpha[3], J;
  /* physics (solver.cpp:30-43) */
  double Re, Pr, gamma, M0, rho0, V0, L, R, c0, p0, T0;
  /* operators (solver.cpp:74-192) */
  double *xg, *wg;          /* 1D Gauss nodes / weights (n) */
  double *wQ2D, *wQ3D;      /* tensor weights */
  double *Interp2D;         /* (nQ2D, nFN) */
  double *Interp3D;         /* (nQ3D, dofs) */
  double *Mchol;            /* Cholesky factor of Mel (dofs, dofs), lower */
  double *Mel;              /* assembled mass matrix (dofs, dofs) */
  double *Sels;             /* (dofs, dofs, 3) */
  double *Kels;             /* (nQ3D, dofs, 3) */
  double *Kels2D;           /* (nQ2D, nFN, 3) */
  int *efToN;               /* (nFN, 6) */
  double *coords;           /* (3, dofs, K) */
  /* work arrays */
  double *uI2, *uI3, *DuI2, *DuI3; /* interpolated traces / volume values */
} oracle_t;

/* ------------------------------------------------------------------ */
/* Small dense linear algebra (column-major, restating the BLAS/LAPACK  */
/* calls of solver.cpp / dgMath.cpp)                                    */
/* ------------------------------------------------------------------ */

/* C(m,n) = alpha * op(A) * B + beta * C, op(A) = A (m,k) or A^T (A is k,m).
   Restates cblas_dgemm with NoTrans/Trans on A and NoTrans on B. */
static void gemm(int transA, int m, int n, int k, double alpha,
                 const double *A, int lda, const double *B, int ldb,
                 double beta, double *C, int ldc) {
  long j;
#pragma omp parallel for schedule(static)
  for (j = 0; j < n; ++j) {
    double *c = C + (size_t)j * ldc;
    const double *b = B + (size_t)j * ldb;
    if (beta == 0.0) { for (int i = 0; i < m; ++i) c[i] = 0.0; }
    else if (beta != 1.0) { for (int i = 0; i < m; ++i) c[i] *= beta; }
    if (!transA) {
      for (int pp = 0; pp < k; ++pp) {
        const double bp = alpha * b[pp];
        const double *a = A + (size_t)pp * lda;
        for (int i = 0; i < m; ++i) c[i] += a[i] * bp;
      }
    } else {
      for (int i = 0; i < m; ++i) {
        const double *a = A + (size_t)i * lda;
        double s = 0.0;
        for (int pp = 0; pp < k; ++pp) s += a[pp] * b[pp];
        c[i] += alpha * s;
      }
    }
  }
}

/* Solve A X = B in place (B <- A^{-1} B); A (n,n) is destroyed.
   Gaussian elimination with partial pivoting: restates LAPACKE_dgesv
   (solver.cpp:110-111, dgMath.cpp:561-562, 601-602). Returns 0 or -1. */
static int lu_solve(int n, double *A, int nrhs, double *B) {
  for (int c = 0; c < n; ++c) {
    int piv = c; double best = fabs(A[c + (size_t)c * n]);
    for (int r = c + 1; r < n; ++r) {
      double v = fabs(A[r + (size_t)c * n]);
      if (v > best) { best = v; piv = r; }
    }
    if (best == 0.0) return -1;
    if (piv != c) {
      for (int j = 0; j < n; ++j) { double t = A[c + (size_t)j*n]; A[c + (size_t)j*n] = A[piv + (size_t)j*n]; A[piv + (size_t)j*n] = t; }
      for (int j = 0; j < nrhs; ++j) { double t = B[c + (size_t)j*n]; B[c + (size_t)j*n] = B[piv + (size_t)j*n]; B[piv + (size_t)j*n] = t; }
    }
    double d = A[c + (size_t)c * n];
    for (int r = c + 1; r < n; ++r) {
      double f = A[r + (size_t)c * n] / d;
      if (f == 0.0) continue;
      for (int j = c; j < n; ++j) A[r + (size_t)j*n] -= f * A[c + (size_t)j*n];
      for (int j = 0; j < nrhs; ++j) B[r + (size_t)j*n] -= f * B[c + (size_t)j*n];
    }
  }
  for (int j = 0; j < nrhs; ++j) {
    double *b = B + (size_t)j * n;
    for (int r = n - 1; r >= 0; --r) {
      double s = b[r];
      for (int c = r + 1; c < n; ++c) s -= A[r + (size_t)c * n] * b[c];
      b[r] = s / A[r + (size_t)r * n];
    }
  }
  return 0;
}

/* Cholesky factor of an SPD matrix, lower triangle kept in L.
   Replaces the Bunch-Kaufman LDL^T of LAPACKE_dsytrf (solver.cpp:155-156):
   the mass matrix is SPD, so both factorizations solve the same system. */
static int cholesky(int n, const double *A, double *L) {
  memset(L, 0, sizeof(double) * (size_t)n * n);
  for (int j = 0; j < n; ++j) {
    double s = A[j + (size_t)j * n];
    for (int k = 0; k < j; ++k) s -= L[j + (size_t)k * n] * L[j + (size_t)k * n];
    if (s <= 0.0) return -1;
    double d = sqrt(s);
    L[j + (size_t)j * n] = d;
    for (int i = j + 1; i < n; ++i) {
      double t = A[i + (size_t)j * n];
      for (int k = 0; k < j; ++k) t -= L[i + (size_t)k * n] * L[j + (size_t)k * n];
      L[i + (size_t)j * n] = t / d;
    }
  }
  return 0;
}

/* B <- Mel^{-1} B for nrhs columns, using the Cholesky factor.
   Restates LAPACKE_dsytrs (solver.cpp:505-506, 532-533). */
static void chol_solve(int n, const double *L, long nrhs, double *B) {
  long j;
#pragma omp parallel for schedule(static)
  for (j = 0; j < nrhs; ++j) {
    double *b = B + (size_t)j * n;
    for (int i = 0; i < n; ++i) {
      double s = b[i];
      const double *li = L + i;
      for (int k = 0; k < i; ++k) s -= li[(size_t)k * n] * b[k];
      b[i] = s / L[i + (size_t)i * n];
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = b[i];
      const double *col = L + (size_t)i * n;
      for (int k = i + 1; k < n; ++k) s -= col[k] * b[k];
      b[i] = s / col[i];
    }
  }
}

/* Eigen-decomposition of a small symmetric matrix by cyclic Jacobi
   rotations.  A (n,n) is overwritten; eigenvalues in ev, eigenvectors as
   columns of V.  Stands in for LAPACKE_dsyev in gaussQuad
   (dgMath.cpp:105). */
static void jacobi_eig(int n, double *A, double *ev, double *V) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) V[i + j * n] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) off += A[i + j * n] * A[i + j * n];
    if (off < 1e-300) break;
    for (int pp = 0; pp < n; ++pp) {
      for (int q = pp + 1; q < n; ++q) {
        double apq = A[pp + q * n];
        if (fabs(apq) < 1e-300) continue;
        double app = A[pp + pp * n], aqq = A[q + q * n];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  /* A <- A G */
          double akp = A[k + pp * n], akq = A[k + q * n];
          A[k + pp * n] = c * akp - s * akq;
          A[k + q * n] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  /* A <- G^T A */
          double apk = A[pp + k * n], aqk = A[q + k * n];
          A[pp + k * n] = c * apk - s * aqk;
          A[q + k * n] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[k + pp * n], vkq = V[k + q * n];
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
static void chebyshev(int p, double *x) {
  for (int i = 0; i <= p; ++i) x[i] = cos((p - i) * M_PI / p);
}

/* Golub-Welsch Gauss-Legendre rule with n = ceil((p+1)/2) points on [-1,1]
   (dgMath.cpp:91-114): eigenvalues of the Jacobi matrix with off-diagonal
   (i+1)/sqrt(4(i+1)^2-1) are the nodes; w_i = 2 v_{0,i}^2.  Ascending. */
static int gauss_quad(int p, double *x, double *w) {
  int n = (int)ceil((p + 1) / 2.0);
  double A[64], V[64], ev[8];
  memset(A, 0, sizeof(A));
  for (int i = 0; i < n - 1; ++i) {
    double t = (i + 1) / sqrt(4.0 * (i + 1) * (i + 1) - 1.0);
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
    double v0 = V[0 + idx[i] * n];
    w[i] = 2.0 * v0 * v0;
  }
  return n;
}

/* Legendre polynomials P_0..P_p at x by the three-term recurrence
   (dgMath.cpp:177-209); out(i, ip) column-major with leading dim nx. */
static void legendre(int p, int nx, const double *x, double *P) {
  for (int i = 0; i < nx; ++i) P[i] = 1.0;
  if (p == 0) return;
  for (int i = 0; i < nx; ++i) P[i + nx] = x[i];
  for (int ip = 0; ip < p - 1; ++ip)
    for (int i = 0; i < nx; ++i)
      P[i + (size_t)(ip + 2) * nx] = ((2 * ip + 3) * x[i] * P[i + (size_t)(ip + 1) * nx]
                                      - (ip + 1) * P[i + (size_t)ip * nx]) / (ip + 2);
}

/* Derivatives of the Legendre polynomials (dgMath.cpp:216-248). */
static void dlegendre(int p, int nx, const double *x, const double *P, double *dP) {
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
static void legendre_nd(int D, int p, int n1, const double *x1, int dirDeriv, double *V) {
  int n = p + 1;
  double *P = malloc(sizeof(double) * n1 * n), *dP = malloc(sizeof(double) * n1 * n);
  legendre(p, n1, x1, P);
  dlegendre(p, n1, x1, P, dP);
  int npts = (D == 2) ? n1 * n1 : n1 * n1 * n1;
  int nmod = (D == 2) ? n * n : n * n * n;
  for (int m = 0; m < nmod; ++m) {
    int mm[3] = { m % n, (m / n) % n, m / (n * n) };
    for (int q = 0; q < npts; ++q) {
      int qq[3] = { q % n1, (q / n1) % n1, q / (n1 * n1) };
      double v = 1.0;
      for (int d = 0; d < D; ++d) {
        const double *T = (d == dirDeriv) ? dP : P;
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
static void interp_matrix(int D, int p, const double *xfrom1, const double *xto1,
                          int dirDeriv, double *out) {
  int n = p + 1;
  int nf = (D == 2) ? n * n : n * n * n;
  double *Lf = malloc(sizeof(double) * nf * nf), *C = calloc((size_t)nf * nf, sizeof(double));
  double *Lt = malloc(sizeof(double) * nf * nf);
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
static double normal_comp(int f, int l) {
  if (f / 2 != l) return 0.0;
  return (f % 2) ? 1.0 : -1.0;
}

/* ------------------------------------------------------------------ */
/* Pointwise physics (textbook forms of solver.cpp:1040-1374)           */
/* ------------------------------------------------------------------ */

/* Euler flux F(iS + 5 l) of conserved state u (solver.cpp:1040-1086).
   The reference hard-codes gamma = 1.4 in the pressure and uses p.gamma in
   the energy flux; with gamma = 1.4 both agree with this form. */
void orc_fluxC(const double *u, double gamma, double *F) {
  double rho = u[0], ir = 1.0 / rho;
  double v[3] = { u[1] * ir, u[2] * ir, u[3] * ir };
  double pr = (gamma - 1.0) * (u[4] - 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
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
void orc_fluxV(const double *u, const double *Du, double Re, double Pr, double gamma, double *F) {
  double rho = u[0], ir = 1.0 / rho;
  double v[3] = { u[1] * ir, u[2] * ir, u[3] * ir };
  double gv[3][3]; /* gv[i][l] = d v_i / d x_l */
  double ge[3];
  for (int l = 0; l < 3; ++l) {
    double drho = Du[0 + 5 * l];
    for (int i = 0; i < 3; ++i) gv[i][l] = (Du[1 + i + 5 * l] - v[i] * drho) * ir;
    double dEr = (Du[4 + 5 * l] - u[4] * ir * drho) * ir;   /* d(E/rho) */
    ge[l] = dEr - (v[0] * gv[0][l] + v[1] * gv[1][l] + v[2] * gv[2][l]);
  }
  double mu = 1.0 / Re, div = gv[0][0] + gv[1][1] + gv[2][2];
  for (int l = 0; l < 3; ++l) {
    double tau[3];
    for (int i = 0; i < 3; ++i)
      tau[i] = mu * (gv[i][l] + gv[l][i] - (i == l ? 2.0 / 3.0 * div : 0.0));
    F[0 + 5 * l] = 0.0;
    for (int i = 0; i < 3; ++i) F[1 + i + 5 * l] = tau[i];
    F[4 + 5 * l] = v[0] * tau[0] + v[1] * tau[1] + v[2] * tau[2] + mu * gamma / Pr * ge[l];
  }
}

/* Roe flux F*(uN, uK, n) for element K with outward normal n
   (solver.cpp:1194-1374): F* = (F(uN)+F(uK)).n/2 - |A_roe|(uN-uK)/2 with
   Roe averages, eigenvalues un, un+-c and no entropy fix. */
void orc_roe(const double *uN, const double *uK, const double *n, double gamma, double *out) {
  double gm1 = gamma - 1.0;
  double rN = uN[0], rK = uK[0];
  double vN[3] = { uN[1] / rN, uN[2] / rN, uN[3] / rN };
  double vK[3] = { uK[1] / rK, uK[2] / rK, uK[3] / rK };
  double pN = gm1 * (uN[4] - 0.5 * rN * (vN[0]*vN[0] + vN[1]*vN[1] + vN[2]*vN[2]));
  double pK = gm1 * (uK[4] - 0.5 * rK * (vK[0]*vK[0] + vK[1]*vK[1] + vK[2]*vK[2]));
  double HN = (uN[4] + pN) / rN, HK = (uK[4] + pK) / rK;
  double qN = vN[0]*n[0] + vN[1]*n[1] + vN[2]*n[2];
  double qK = vK[0]*n[0] + vK[1]*n[1] + vK[2]*n[2];
  /* Roe averages with weight sqrt(rN/rK) on the N side */
  double s = sqrt(rN / rK), den = s + 1.0;
  double vt[3], ht, q2 = 0.0, qt = 0.0;
  for (int i = 0; i < 3; ++i) { vt[i] = (s * vN[i] + vK[i]) / den; q2 += vt[i] * vt[i]; qt += vt[i] * n[i]; }
  ht = (s * HN + HK) / den;
  double c2 = gm1 * (ht - 0.5 * q2), c = sqrt(c2);
  double l0 = fabs(qt), lp = fabs(qt + c), lm = fabs(qt - c);
  double d[5];
  for (int k = 0; k < 5; ++k) d[k] = uN[k] - uK[k];
  double dp = gm1 * (0.5 * q2 * d[0] - (vt[0]*d[1] + vt[1]*d[2] + vt[2]*d[3]) + d[4]); /* linearised pressure jump */
  double dV = n[0]*d[1] + n[1]*d[2] + n[2]*d[3] - qt * d[0];                              /* rho~ * jump of normal velocity */
  double s1 = 0.5 * (lp + lm) - l0, s2 = 0.5 * (lp - lm);
  double a = s1 * dp / c2 + s2 * dV / c;
  double b = s2 * dp / c + s1 * dV;
  out[0] = 0.5 * (rN * qN + rK * qK) - 0.5 * (l0 * d[0] + a);
  for (int i = 0; i < 3; ++i)
    out[1 + i] = 0.5 * (uN[1 + i] * qN + uK[1 + i] * qK + n[i] * (pN + pK))
               - 0.5 * (l0 * d[1 + i] + a * vt[i] + b * n[i]);
  out[4] = 0.5 * ((uN[4] + pN) * qN + (uK[4] + pK) * qK) - 0.5 * (l0 * d[4] + a * ht + b * qt);
}

/* LDG auxiliary flux: the K-side value if n_l > 0, else the N-side one,
   times n_l (solver.cpp:867-884). */
static double flux_L(double uK, double uN, double nl) { return (nl > 0 ? uK : uN) * nl; }

/* One-sided viscous numerical flux: -sum_l Fv_l(uK, Du*) n_l with
   Du* = DuN if sum(n) > 0 else DuK (solver.cpp:891-921). */
static void flux_V_num(const oracle_t *o, const double *uK, const double *DuN,
                       const double *DuK, const double *n, double *out) {
  double F[15];
  double sum = n[0] + n[1] + n[2];
  orc_fluxV(uK, (sum > 0) ? DuN : DuK, o->Re, o->Pr, o->gamma, F);
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
    o->dx[d] = (xmax[d] - xmin[d]) / ns[d];         /* mesh.cpp:149-151 */
    o->alpha[d] = o->dx[d] / 2.0;                   /* solver.cpp:144-147 */
  }
  o->J = o->alpha[0] * o->alpha[1] * o->alpha[2];
  /* physics constants (solver.cpp:30-43) */
  o->Re = 1600.0; o->Pr = 0.71; o->gamma = 1.4; o->M0 = 0.1; o->rho0 = 1.0; o->V0 = 1.0; o->R = 8.314;
  if (phys) { o->Re = phys[0]; o->Pr = phys[1]; o->gamma = phys[2]; o->M0 = phys[3]; o->rho0 = phys[4]; o->V0 = phys[5]; o->R = phys[6]; }
  o->L = L;
  o->c0 = o->V0 / o->M0;
  o->p0 = o->rho0 * o->c0 * o->c0 / o->gamma;
  o->T0 = o->p0 / (o->R * o->rho0);

  /* quadrature: gaussQuad(2*order) -> order+1 points (solver.cpp:80, 88) */
  o->xg = malloc(sizeof(double) * 8); o->wg = malloc(sizeof(double) * 8);
  int nq = gauss_quad(2 * p, o->xg, o->wg);
  o->nQ2D = nq * nq; o->nQ3D = nq * nq * nq;
  o->wQ2D = malloc(sizeof(double) * o->nQ2D);
  o->wQ3D = malloc(sizeof(double) * o->nQ3D);
  for (int b = 0; b < nq; ++b) for (int a = 0; a < nq; ++a) o->wQ2D[a + nq * b] = o->wg[a] * o->wg[b];
  for (int c = 0; c < nq; ++c) for (int b = 0; b < nq; ++b) for (int a = 0; a < nq; ++a)
    o->wQ3D[a + nq * (b + nq * c)] = o->wg[a] * o->wg[b] * o->wg[c];

  double cheb[16];
  chebyshev(p, cheb);
  int dofs = o->dofs, nQ3D = o->nQ3D, nQ2D = o->nQ2D, nFN = o->nFN;

  /* Interp2D / Interp3D (solver.cpp:74-92) */
  o->Interp2D = malloc(sizeof(double) * nQ2D * nFN);
  o->Interp3D = malloc(sizeof(double) * nQ3D * dofs);
  interp_matrix(2, p, cheb, o->xg, -1, o->Interp2D);
  interp_matrix(3, p, cheb, o->xg, -1, o->Interp3D);

  /* phiQ, dPhiQ (solver.cpp:98-129) */
  double *phiQ = o->Interp3D;                 /* identical construction */
  double *dPhiQ = malloc(sizeof(double) * nQ3D * dofs * 3);
  for (int l = 0; l < 3; ++l) interp_matrix(3, p, cheb, o->xg, l, dPhiQ + (size_t)l * nQ3D * dofs);
  double *wphi = malloc(sizeof(double) * nQ3D * dofs);   /* polyQuad = w * phiQ (solver.cpp:131-140) */
  for (int j = 0; j < dofs; ++j) for (int q = 0; q < nQ3D; ++q) wphi[q + (size_t)j * nQ3D] = phiQ[q + (size_t)j * nQ3D] * o->wQ3D[q];

  /* Mel = J phiQ^T W phiQ (solver.cpp:142-156) */
  o->Mel = malloc(sizeof(double) * dofs * dofs);
  gemm(1, dofs, dofs, nQ3D, o->J, wphi, nQ3D, phiQ, nQ3D, 0.0, o->Mel, dofs);
  o->Mchol = malloc(sizeof(double) * dofs * dofs);
  if (cholesky(dofs, o->Mel, o->Mchol) != 0) { orc_destroy(o); free(dPhiQ); free(wphi); return NULL; }

  /* Sels_l = (J/alpha_l) dPhiQ_l^T W phiQ (solver.cpp:158-165) */
  o->Sels = malloc(sizeof(double) * dofs * dofs * 3);
  for (int l = 0; l < 3; ++l)
    gemm(1, dofs, dofs, nQ3D, o->J / o->alpha[l], dPhiQ + (size_t)l * nQ3D * dofs, nQ3D,
         wphi, nQ3D, 0.0, o->Sels + (size_t)l * dofs * dofs, dofs);

  /* Kels_l(q,i) = dPhiQ_l(q,i) w_q J/alpha_l (solver.cpp:167-176) */
  o->Kels = malloc(sizeof(double) * nQ3D * dofs * 3);
  for (int l = 0; l < 3; ++l)
    for (int i = 0; i < dofs; ++i)
      for (int q = 0; q < nQ3D; ++q)
        o->Kels[q + (size_t)nQ3D * (i + (size_t)dofs * l)] = dPhiQ[q + (size_t)nQ3D * (i + (size_t)dofs * l)] * o->wQ3D[q] * (o->J / o->alpha[l]);

  /* Kels2D_l(r,j) = Interp2D(r,j) w2_r J/alpha_l (solver.cpp:178-190) */
  o->Kels2D = malloc(sizeof(double) * nQ2D * nFN * 3);
  for (int l = 0; l < 3; ++l)
    for (int j = 0; j < nFN; ++j)
      for (int r = 0; r < nQ2D; ++r)
        o->Kels2D[r + (size_t)nQ2D * (j + (size_t)nFN * l)] = o->Interp2D[r + (size_t)nQ2D * j] * o->wQ2D[r] * (o->J / o->alpha[l]);
  free(dPhiQ); free(wphi);

  /* mesh connectivity and node coordinates (mesh.cpp:127-146, 389-418) */
  o->efToN = malloc(sizeof(int) * nFN * 6);
  init_face_map(n, o->efToN);
  o->coords = malloc(sizeof(double) * 3 * (size_t)dofs * o->K);
  for (int k = 0; k < o->K; ++k) {
    int ie[3] = { k % nx, (k / nx) % ny, k / (nx * ny) };
    for (int iN = 0; iN < dofs; ++iN) {
      int in[3] = { iN % n, (iN / n) % n, iN / (n * n) };
      for (int l = 0; l < 3; ++l) {
        double lo = ie[l] * (xmax[l] - xmin[l]) / ns[l] + xmin[l];        /* vertex of eToV(0,k) */
        double hi = (ie[l] + 1) * (xmax[l] - xmin[l]) / ns[l] + xmin[l];  /* vertex of eToV(7,k) */
        double scale = .5 * (cheb[in[l]] + 1.0);
        o->coords[l + 3 * ((size_t)iN + (size_t)dofs * k)] = lo + scale * (hi - lo);
      }
    }
  }
  size_t K = o->K;
  o->uI2 = malloc(sizeof(double) * nQ2D * NS * 6 * K);
  o->uI3 = malloc(sizeof(double) * nQ3D * NS * K);
  o->DuI2 = malloc(sizeof(double) * nQ2D * NS * 6 * K * 3);
  o->DuI3 = malloc(sizeof(double) * nQ3D * NS * K * 3);
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
      const double *X = o->coords + 3 * ((size_t)iN + (size_t)dofs * k);
      double x = X[0], y = X[1], z = X[2], L = o->L;
      double pr = o->p0 + (o->rho0 * o->V0 * o->V0 / 16.0) * (cos(2 * x / L) + cos(2 * y / L)) * (cos(2 * z / L) + 2);
      double *uk = u + (size_t)dofs * NS * k;
      uk[iN + 0 * dofs] = pr / (o->R * o->T0);
      uk[iN + 1 * dofs] = o->rho0 * o->V0 * sin(x / L) * cos(y / L) * cos(z / L);
      uk[iN + 2 * dofs] = -o->rho0 * o->V0 * cos(x / L) * sin(y / L) * cos(z / L);
      uk[iN + 3 * dofs] = 0 * o->rho0;
      uk[iN + 4 * dofs] = pr / (o->gamma - 1) + (uk[iN + 1 * dofs] * uk[iN + 1 * dofs] + uk[iN + 2 * dofs] * uk[iN + 2 * dofs]
                                                + uk[iN + 3 * dofs] * uk[iN + 3 * dofs]) / (2.0 * uk[iN + 0 * dofs]);
    }
}

/* ------------------------------------------------------------------ */
/* RHS: Solver::rk4Rhs (solver.cpp:486-535)                             */
/* ------------------------------------------------------------------ */

/* Solver::interpolate (solver.cpp:541-577): face traces at face Gauss
   points for all 6 faces and volume values at Gauss points, for `dim`
   stacked fields of size (dofs, 5, K). */
static void interpolate(const oracle_t *o, const double *cur, double *I2, double *I3, int dim) {
  int nFN = o->nFN, nQ2D = o->nQ2D, dofs = o->dofs, nQ3D = o->nQ3D;
  size_t K = o->K;
  double *onF = malloc(sizeof(double) * nFN * NS * 6 * K);
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
static void lift_add(const oracle_t *o, int f, const double *fStar, double *resk) {
  int nFN = o->nFN, nQ2D = o->nQ2D, dofs = o->dofs;
  const double *K2 = o->Kels2D + (size_t)nQ2D * nFN * (f / 2);
  for (int s = 0; s < NS; ++s)
    for (int j = 0; j < nFN; ++j) {
      double acc = 0.0;
      for (int r = 0; r < nQ2D; ++r) acc += K2[r + (size_t)nQ2D * j] * fStar[r + nQ2D * s];
      resk[o->efToN[j + nFN * f] + (size_t)dofs * s] += acc;
    }
}

/* rk4Rhs: k = Mel^{-1}( Kels^T (fc - fv) - lift(Fc* + Fv*) ), with the
   auxiliary gradient Du_l = Mel^{-1}( lift(uhat n_l) - Sels_l u ).
   Writes k (dofs,5,K) and, if Du != NULL, Du (dofs,5,K,3). */
void orc_rhs(oracle_t *o, const double *u, double *kout, double *Du_out) {
  int dofs = o->dofs, nQ2D = o->nQ2D, nQ3D = o->nQ3D;
  size_t K = o->K, nU = (size_t)dofs * NS * K;
  double *Dus = Du_out ? Du_out : malloc(sizeof(double) * nU * 3);

  interpolate(o, u, o->uI2, o->uI3, 1);                       /* solver.cpp:491 */

  memset(Dus, 0, sizeof(double) * nU * 3);                    /* solver.cpp:495 */
  long k;
#pragma omp parallel for schedule(static)
  for (k = 0; k < (long)K; ++k) {                              /* localDGFlux, solver.cpp:585-635 */
    double fStar[64 * NS];
    for (int l = 0; l < 3; ++l)
      for (int f = 0; f < 6; ++f) {
        double nl = normal_comp(f, l);
        if (nl == 0.0) continue;   /* zero normal: the reference adds a zero flux */
        int N = neighbor(o, (int)k, f), fN = f ^ 1;
        for (int s = 0; s < NS; ++s)
          for (int r = 0; r < nQ2D; ++r) {
            double uK = o->uI2[r + (size_t)nQ2D * (s + NS * (f + 6 * (size_t)k))];
            double uN = o->uI2[r + (size_t)nQ2D * (s + NS * (fN + 6 * (size_t)N))];
            fStar[r + nQ2D * s] = flux_L(uK, uN, nl);
          }
        lift_add(o, f, fStar, Dus + (size_t)dofs * NS * (k + K * l));
      }
  }
  for (int l = 0; l < 3; ++l) {                                /* solver.cpp:497-508 */
    double *Dl = Dus + nU * l;
    gemm(0, dofs, NS * K, dofs, -1.0, o->Sels + (size_t)dofs * dofs * l, dofs, u, dofs, 1.0, Dl, dofs);
    chol_solve(dofs, o->Mchol, NS * K, Dl);
  }

  interpolate(o, Dus, o->DuI2, o->DuI3, 3);                    /* solver.cpp:511 */

  double *res = kout;                                          /* solver.cpp:516-517 */
  memset(res, 0, sizeof(double) * nU);
  size_t s2 = (size_t)nQ2D * NS * 6 * K;                       /* stride of l in DuI2 */
#pragma omp parallel for schedule(static)
  for (k = 0; k < (long)K; ++k) {
    double fStar[64 * NS], uK[5], uN[5], DuK[15], DuN[15], fl[5];
    double *resk = res + (size_t)dofs * NS * k;
    for (int f = 0; f < 6; ++f) {                               /* convectDGFlux, solver.cpp:643-698 */
      int N = neighbor(o, (int)k, f), fN = f ^ 1;
      double nrm[3] = { normal_comp(f, 0), normal_comp(f, 1), normal_comp(f, 2) };
      for (int r = 0; r < nQ2D; ++r) {
        for (int s = 0; s < NS; ++s) {
          uK[s] = o->uI2[r + (size_t)nQ2D * (s + NS * (f + 6 * (size_t)k))];
          uN[s] = o->uI2[r + (size_t)nQ2D * (s + NS * (fN + 6 * (size_t)N))];
        }
        orc_roe(uN, uK, nrm, o->gamma, fl);
        for (int s = 0; s < NS; ++s) fStar[r + nQ2D * s] = fl[s];
      }
      lift_add(o, f, fStar, resk);
    }
    for (int f = 0; f < 6; ++f) {                               /* viscousDGFlux, solver.cpp:750-814 */
      int N = neighbor(o, (int)k, f), fN = f ^ 1;
      double nrm[3] = { normal_comp(f, 0), normal_comp(f, 1), normal_comp(f, 2) };
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
  double *fcv = malloc(sizeof(double) * (size_t)nQ3D * NS * K * 3);
  size_t s3 = (size_t)nQ3D * NS * K;
#pragma omp parallel for schedule(static)
  for (k = 0; k < (long)K; ++k) {
    double uq[5], Duq[15], Fc[15], Fv[15];
    for (int q = 0; q < nQ3D; ++q) {
      for (int s = 0; s < NS; ++s) {
        uq[s] = o->uI3[q + (size_t)nQ3D * (s + NS * (size_t)k)];
        for (int l = 0; l < 3; ++l) Duq[s + 5 * l] = o->DuI3[q + (size_t)nQ3D * (s + NS * (size_t)k) + s3 * l];
      }
      orc_fluxC(uq, o->gamma, Fc);
      orc_fluxV(uq, Duq, o->Re, o->Pr, o->gamma, Fv);
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

/* Solver::computeKE (solver.cpp:1007-1034): sum over elements and nodes of
   J w_q |m|^2/(2 rho) using NODAL values u(iFQ, ., k) weighted by the Gauss
   weights wQ(iFQ) (the reference's definition), normalised by
   rho0 (2 pi L)^3. */
double orc_ke(const oracle_t *o, const double *u) {
  int dofs = o->dofs;
  double total = 0.0;
  for (long k = 0; k < o->K; ++k) {
    const double *uk = u + (size_t)dofs * NS * k;
    double ke = 0.0;
    for (int q = 0; q < o->nQ3D; ++q)
      ke += o->J * o->wQ3D[q] * (uk[q + dofs] * uk[q + dofs] + uk[q + 2 * dofs] * uk[q + 2 * dofs]
                                 + uk[q + 3 * dofs] * uk[q + 3 * dofs]) / (2 * uk[q]);
    total += ke;
  }
  return total / (o->rho0 * pow(2 * M_PI * o->L, 3.0));
}

/* Gauss-point diagnostics (this project's definitions; no reference
   counterpart): out[0] = kinetic energy (1/(rho0 V)) int |m|^2/(2 rho),
   out[1] = enstrophy (1/(rho0 V)) int rho |curl v|^2 / 2, both by the
   element Gauss rule on the interpolated state and LDG gradient. */
void orc_diagnostics(oracle_t *o, const double *u, double *out) {
  int dofs = o->dofs, nQ3D = o->nQ3D;
  size_t K = o->K, nU = (size_t)dofs * NS * K;
  double *Du = malloc(sizeof(double) * nU * 3), *k = malloc(sizeof(double) * nU);
  orc_rhs(o, u, k, Du);           /* leaves uI3 / DuI3 at the Gauss points */
  size_t s3 = (size_t)nQ3D * NS * K;
  double ke = 0.0, ens = 0.0;
  for (size_t e = 0; e < K; ++e) {
    double kel = 0.0, enl = 0.0;
    for (int q = 0; q < nQ3D; ++q) {
      double U[5], G[15];
      for (int s = 0; s < NS; ++s) {
        U[s] = o->uI3[q + (size_t)nQ3D * (s + NS * e)];
        for (int l = 0; l < 3; ++l) G[s + 5 * l] = o->DuI3[q + (size_t)nQ3D * (s + NS * e) + s3 * l];
      }
      double ir = 1.0 / U[0], v[3] = { U[1] * ir, U[2] * ir, U[3] * ir }, gv[3][3];
      for (int i = 0; i < 3; ++i) for (int l = 0; l < 3; ++l) gv[i][l] = (G[1 + i + 5 * l] - v[i] * G[0 + 5 * l]) * ir;
      double w1 = gv[2][1] - gv[1][2], w2 = gv[0][2] - gv[2][0], w3 = gv[1][0] - gv[0][1];
      double wq = o->J * o->wQ3D[q];
      kel += wq * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) * ir * 0.5;
      enl += wq * U[0] * (w1 * w1 + w2 * w2 + w3 * w3) * 0.5;
    }
    ke += kel; ens += enl;
  }
  double V = o->rho0 * pow(2 * M_PI * o->L, 3.0);
  out[0] = ke / V; out[1] = ens / V;
  free(Du); free(k);
}

/* RK4 time stepping (solver.cpp:295-480): classic tableau with stored
   stages; uCurr = u + dt a_{s-1} k_{s-1}; after the four stages
   u += dt b_s k_s in stage order; KE of the new state recorded per step. */
int orc_step(oracle_t *o, double *u, double dt, int nsteps, double *ke) {
  static const double b[4] = { 1 / 6.0, 1 / 3.0, 1 / 3.0, 1 / 6.0 };
  static const double a[3] = { 0.5, 0.5, 1.0 };
  size_t nU = (size_t)o->dofs * NS * o->K;
  double *ks = malloc(sizeof(double) * nU * 4), *uc = malloc(sizeof(double) * nU);
  if (!ks || !uc) { free(ks); free(uc); return -1; }
  for (int it = 0; it < nsteps; ++it) {
    for (int s = 0; s < 4; ++s) {
      if (s == 0) memcpy(uc, u, sizeof(double) * nU);
      else { for (size_t i = 0; i < nU; ++i) uc[i] = u[i] + dt * a[s - 1] * ks[i + nU * (s - 1)]; }
      orc_rhs(o, uc, ks + nU * s, NULL);
    }
    for (int s = 0; s < 4; ++s)
      for (size_t i = 0; i < nU; ++i) u[i] += dt * b[s] * ks[i + nU * s];
    if (ke) ke[it] = orc_ke(o, u);
  }
  free(ks); free(uc);
  return 0;
}

/* Reference time step (solver.cpp:44-52) for a CFL factor (0.01 in the
   reference): dt = cfl min(dx)/(V0 (2p+1)). */
double orc_dt(const oracle_t *o, double cfl) {
  double m = o->dx[0] < o->dx[1] ? o->dx[0] : o->dx[1];
  if (o->dx[2] < m) m = o->dx[2];
  return cfl * m / (o->V0 * (2 * o->p + 1));
}

/* Accessors used by the tests. */
int orc_dofs(const oracle_t *o) { return o->dofs; }
int orc_nelem(const oracle_t *o) { return o->K; }
int orc_nq(const oracle_t *o) { return o->nQ3D; }
void orc_get_gauss(const oracle_t *o, double *x, double *w) { memcpy(x, o->xg, sizeof(double) * o->n); memcpy(w, o->wg, sizeof(double) * o->n); }
void orc_get_coords(const oracle_t *o, double *c) { memcpy(c, o->coords, sizeof(double) * 3 * (size_t)o->dofs * o->K); }
/* which: 0 Interp2D, 1 Interp3D, 2 Mel, 3 Sels, 4 Kels, 5 Kels2D */
int orc_get_operator(const oracle_t *o, int which, double *out) {
  size_t sz[6] = { (size_t)o->nQ2D * o->nFN, (size_t)o->nQ3D * o->dofs, (size_t)o->dofs * o->dofs,
                   (size_t)o->dofs * o->dofs * 3, (size_t)o->nQ3D * o->dofs * 3, (size_t)o->nQ2D * o->nFN * 3 };
  const double *src[6] = { o->Interp2D, o->Interp3D, o->Mel, o->Sels, o->Kels, o->Kels2D };
  if (which < 0 || which > 5) return -1;
  memcpy(out, src[which], sizeof(double) * sz[which]);
  return (int)sz[which];
}
void orc_get_physics(const oracle_t *o, double *out) {
  out[0] = o->Re; out[1] = o->Pr; out[2] = o->gamma; out[3] = o->M0; out[4] = o->rho0;
  out[5] = o->V0; out[6] = o->R; out[7] = o->c0; out[8] = o->p0; out[9] = o->T0; out[10] = o->J;
}
