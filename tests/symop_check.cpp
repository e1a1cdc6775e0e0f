// Host check of the even/odd operator application (local-dg_amd/csrc/symop.hpp)
// used by every sum-factorised sweep of the rhs kernel, built with g++ by
// tests/test_symop.py.
//   1. random centro-symmetric (S = +1) and anti-centro-symmetric (S = -1)
//      matrices, N = 2..7: sym_apply and sym_apply2 against the plain
//      matrix-vector product;
//   2. the element's own 1D operators (basis.hpp) for p = 1..5: I, I^T,
//      Minv are centro-symmetric and D^T anti-centro-symmetric to round-off
//      (what the library checks before it symmetrises them), and the split
//      form reproduces the plain product on them.
// Exit status 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <random>

#include "basis.hpp"
#include "symop.hpp"

using namespace ldg;

static int bad = 0;

template <int N>
static double apply_err(const double (&A)[N][N], int S, const double (&B)[N][N], std::mt19937 &g) {
  std::uniform_real_distribution<double> U(-1, 1);
  SymMat<N, double> M, MB;
  make_symmat<N, double>(A, M);
  make_symmat<N, double>(B, MB);
  double v[N], w[N], o[N], o2[N];
  for (int i = 0; i < N; ++i) v[i] = U(g), w[i] = U(g);
  if (S == 1) {
    sym_apply<N, 1, double>(M, v, o);
    sym_apply2<N, 1, 1, double>(M, v, MB, w, o2);
  } else {
    sym_apply<N, -1, double>(M, v, o);
    sym_apply2<N, -1, 1, double>(M, v, MB, w, o2);
  }
  double err = 0, scale = 0;
  for (int q = 0; q < N; ++q) {
    double r = 0, r2 = 0;
    for (int j = 0; j < N; ++j) {
      r += A[q][j] * v[j];
      r2 += A[q][j] * v[j] + B[q][j] * w[j];
      scale = std::fmax(scale, std::fabs(A[q][j]) + std::fabs(B[q][j]));
    }
    err = std::fmax(err, std::fmax(std::fabs(r - o[q]), std::fabs(r2 - o2[q])));
  }
  return err / (scale * N);
}

template <int N>
static void random_case(std::mt19937 &g) {
  std::uniform_real_distribution<double> U(-1, 1);
  for (int S : {1, -1}) {
    double A[N][N], B[N][N];
    for (int q = 0; q < N; ++q)
      for (int j = 0; j < N; ++j) A[q][j] = U(g), B[q][j] = U(g);
    symmetrize<N>(A, S);
    symmetrize<N>(B, 1);
    const double e = apply_err<N>(A, S, B, g);
    std::printf("random N=%d S=%+d rel err %.2e\n", N, S, e);
    if (!(e < 1e-15)) ++bad;
  }
}

template <int N>
static double asym(const double (&A)[N][N], int S) {
  double d = 0;
  for (int q = 0; q < N; ++q)
    for (int j = 0; j < N; ++j) d = std::fmax(d, std::fabs(A[q][j] - S * A[N - 1 - q][N - 1 - j]));
  return d;
}

template <int N>
static void basis_case(std::mt19937 &g) {
  const Basis1D B = make_basis(N - 1);
  double I[N][N], It[N][N], Dt[N][N], Mi[N][N];
  for (int a = 0; a < N; ++a)
    for (int b = 0; b < N; ++b) {
      I[a][b] = B.at(B.I, a, b);
      It[a][b] = B.at(B.I, b, a);
      Dt[a][b] = B.at(B.D, b, a);
      Mi[a][b] = B.at(B.M1inv, a, b);
    }
  const double d = std::fmax(std::fmax(asym<N>(I, 1), asym<N>(It, 1)), std::fmax(asym<N>(Dt, -1), asym<N>(Mi, 1)));
  symmetrize<N>(I, 1);
  symmetrize<N>(It, 1);
  symmetrize<N>(Dt, -1);
  symmetrize<N>(Mi, 1);
  const double e = std::fmax(std::fmax(apply_err<N>(I, 1, It, g), apply_err<N>(Dt, -1, It, g)), apply_err<N>(Mi, 1, I, g));
  std::printf("basis p=%d asymmetry %.2e, split-form rel err %.2e\n", N - 1, d, e);
  if (!(d < 1e-12) || !(e < 1e-15)) ++bad;
}

int main() {
  std::mt19937 g(7);
  random_case<2>(g);
  random_case<3>(g);
  random_case<4>(g);
  random_case<5>(g);
  random_case<6>(g);
  random_case<7>(g);
  basis_case<2>(g);
  basis_case<3>(g);
  basis_case<4>(g);
  basis_case<5>(g);
  basis_case<6>(g);
  std::printf("%s\n", bad ? "FAILED" : "ok");
  return bad;
}
