// Host side of the checkpoint format test (tests/test_checkpoint.py), built
// with g++ against local-dg_amd/host/io.cpp:
//   checkpoint_check write FILE   writes a small checkpoint with known values
//   checkpoint_check read FILE    reads one and prints its header and sums
#include <cstdio>
#include <cstring>

#include "io.hpp"

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  ldg::Checkpoint c;
  if (!std::strcmp(argv[1], "write")) {
    c.p = 1; c.nx = 3; c.ny = 4; c.nz = 5; c.precision = 1;
    c.step = 7; c.timesteps = 11;
    c.dt = 0.125; c.tf = 1.375; c.L = 1.0; c.Re = 1600; c.M0 = 0.1; c.V0 = 1.0; c.Pr = 0.71;
    for (int l = 0; l < 3; ++l) { c.lo[l] = -3.0 - l; c.hi[l] = 3.0 + l; }
    c.u.resize(8 * 5 * 60);
    for (std::size_t i = 0; i < c.u.size(); ++i) c.u[i] = 0.5 * i - 1.0;
    for (int i = 0; i < 7; ++i) c.ke.push_back(1.0 / (i + 1));
    return ldg::writeCheckpoint(argv[2], c) ? 0 : 1;
  }
  if (!ldg::readCheckpoint(argv[2], c)) {
    std::printf("unreadable\n");
    return 1;
  }
  double su = 0, sk = 0;
  for (double v : c.u) su += v;
  for (double v : c.ke) sk += v;
  std::printf("%d %d %d %d %d %ld %ld %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %zu %.17g %zu %.17g\n", c.p, c.nx,
              c.ny, c.nz, c.precision, c.step, c.timesteps, c.dt, c.tf, c.L, c.Re, c.M0, c.V0, c.Pr, c.lo[2], c.hi[1],
              c.u.size(), su, c.ke.size(), sk);
  return 0;
}
