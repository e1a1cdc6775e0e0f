// basis.hpp -- 1D operators of the tensor-product nodal DG basis (host, FP64).
//
// The reference builds dense n^3 x n^3 element matrices from Legendre
// Vandermonde solves (solver.cpp:74-192, dgMath.cpp:37-615).  Every one of
// them is a Kronecker product of the 1D pieces built here, so the device
// kernels only need these n x n factors:
//   r      Chebyshev-Gauss-Lobatto nodes, ascending (dgMath.cpp:37-45)
//   g, w   (p+1)-point Gauss-Legendre rule (dgMath.cpp:91-114)
//   I      I(q,i) = l_i(g_q)     Lagrange basis at the Gauss points
//   D      D(q,i) = l_i'(g_q)    its derivative at the Gauss points
//   Dn     Dn(i,j) = l_j'(r_i)   nodal differentiation matrix
//   M1inv  inverse of M1 = I^T W I (1D mass matrix; Mel = J M1(x)M1(x)M1)
// Nodes and weights come from Newton iteration on the Legendre recurrence
// and the barycentric Lagrange formulas, not from eigen-solves.
#pragma once

#include <cmath>
#include <vector>

namespace ldg {

struct Basis1D {
  int n = 0;
  std::vector<double> r, g, w;           // n
  std::vector<double> I, D, Dn, M1inv;   // n*n, row-major [row][col]
  double at(const std::vector<double> &m, int row, int col) const { return m[row * n + col]; }
};

// Legendre P_n and P_n' at x by the three-term recurrence.
inline void legendre_pn(int n, double x, double &P, double &dP) {
  double p0 = 1.0, p1 = x;
  if (n == 0) { P = 1.0; dP = 0.0; return; }
  for (int k = 2; k <= n; ++k) {
    double pk = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
    p0 = p1; p1 = pk;
  }
  P = p1;
  dP = n * (x * p1 - p0) / (x * x - 1.0);
}

inline bool invert(int n, std::vector<double> A, std::vector<double> &Ainv) {
  Ainv.assign(n * n, 0.0);
  for (int i = 0; i < n; ++i) Ainv[i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r) if (std::fabs(A[r * n + c]) > std::fabs(A[piv * n + c])) piv = r;
    if (A[piv * n + c] == 0.0) return false;
    for (int j = 0; j < n; ++j) { std::swap(A[c * n + j], A[piv * n + j]); std::swap(Ainv[c * n + j], Ainv[piv * n + j]); }
    double d = A[c * n + c];
    for (int j = 0; j < n; ++j) { A[c * n + j] /= d; Ainv[c * n + j] /= d; }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      double f = A[r * n + c];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) { A[r * n + j] -= f * A[c * n + j]; Ainv[r * n + j] -= f * Ainv[c * n + j]; }
    }
  }
  return true;
}

inline Basis1D make_basis(int p) {
  Basis1D B;
  const int n = p + 1;
  B.n = n;
  B.r.resize(n); B.g.resize(n); B.w.resize(n);
  for (int i = 0; i < n; ++i) B.r[i] = std::cos((p - i) * M_PI / p);
  // Gauss-Legendre nodes: Newton on P_n from Chebyshev-like guesses, ascending.
  for (int i = 0; i < n; ++i) {
    double x = -std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double P, dP;
      legendre_pn(n, x, P, dP);
      double dx = P / dP;
      x -= dx;
      if (std::fabs(dx) < 1e-17) break;
    }
    double P, dP;
    legendre_pn(n, x, P, dP);
    B.g[i] = x;
    B.w[i] = 2.0 / ((1.0 - x * x) * dP * dP);
  }
  // barycentric weights of the CGL nodes
  std::vector<double> bw(n, 1.0);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k)
      if (k != j) bw[j] /= (B.r[j] - B.r[k]);
  auto lagrange = [&](double x, std::vector<double> &L, std::vector<double> &dL) {
    L.assign(n, 0.0); dL.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
      // l_i(x) = prod_{k!=i} (x - r_k)/(r_i - r_k); derivative by the product rule
      double v = 1.0, d = 0.0;
      for (int k = 0; k < n; ++k) {
        if (k == i) continue;
        double f = (x - B.r[k]) / (B.r[i] - B.r[k]);
        d = d * f + v / (B.r[i] - B.r[k]);
        v *= f;
      }
      L[i] = v; dL[i] = d;
    }
  };
  B.I.resize(n * n); B.D.resize(n * n); B.Dn.resize(n * n);
  std::vector<double> L, dL;
  for (int q = 0; q < n; ++q) {
    lagrange(B.g[q], L, dL);
    for (int i = 0; i < n; ++i) { B.I[q * n + i] = L[i]; B.D[q * n + i] = dL[i]; }
  }
  // nodal differentiation: off-diagonal barycentric formula, diagonal = -row sum
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      double v = (bw[j] / bw[i]) / (B.r[i] - B.r[j]);
      B.Dn[i * n + j] = v;
      s += v;
    }
    B.Dn[i * n + i] = -s;
  }
  std::vector<double> M1(n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int q = 0; q < n; ++q) s += B.I[q * n + i] * B.w[q] * B.I[q * n + j];
      M1[i * n + j] = s;
    }
  invert(n, M1, B.M1inv);
  return B;
}

}  // namespace ldg
