// Host check of the workgroup -> element mapping of the LDG kernels
// (local-dg_amd/csrc/ldg_kernels.hpp), built with hipcc for the host only by
// tests/test_blockmap.py (no GPU needed):
//   1. FastDiv (multiply-shift division by a run-time constant) equals
//      integer division for every divisor 1..4096 on all n < 2^17 and on a
//      million random n < 2^31;
//   2. block_elements, with tiles chosen as the library chooses them, covers
//      every element of a mesh exactly once, each block inside one x-row,
//      for E = 1..32 and meshes from 3^3 to 64 x 9 x 11; padding blocks of
//      partial bricks report no elements;
//   3. the XCD-aware numbering (workgroup b -> (b mod 8) ceil(B/8) + b div 8)
//      reaches every logical block exactly once.
// Exit status 0 when every check holds.
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "ldg_kernels.hpp"

using namespace ldg;

static long bad = 0;

static void fastdiv_checks() {
  std::mt19937 g(5);
  std::uniform_int_distribution<int> big(0, 0x7fffffff);
  for (int d = 1; d <= 4096; ++d) {
    const FastDiv f = make_fastdiv(d);
    for (int n = 0; n < (1 << 17); ++n)
      if (f.div(n) != n / d) ++bad;
    for (int k = 0; k < 250; ++k) {
      const int n = big(g);
      if (f.div(n) != n / d) ++bad;
    }
  }
  for (int k = 0; k < 1000000; ++k) {
    const int d = 1 + big(g) % 100000, n = big(g);
    if (make_fastdiv(d).div(n) != n / d) ++bad;
  }
  std::printf("FastDiv: %ld mismatches\n", bad);
}

// the tiling ldg.hip builds for an instance (build_tiling)
static InstDesc tiling(int nx, int ny, int nz, int E) {
  InstDesc d{};
  d.nx = nx; d.ny = ny; d.nz = nz;
  d.nelem = (long)nx * ny * nz;
  d.nbx = (nx + E - 1) / E;
  d.tb[0] = std::min(std::max(1, 8 / E), d.nbx);
  d.tb[1] = std::min(8, ny);
  d.tb[2] = std::min(4, nz);
  d.nbrick[0] = (d.nbx + d.tb[0] - 1) / d.tb[0];
  d.nbrick[1] = (ny + d.tb[1] - 1) / d.tb[1];
  d.nbrick[2] = (nz + d.tb[2] - 1) / d.tb[2];
  d.nblocks = d.nbrick[0] * d.nbrick[1] * d.nbrick[2] * d.tb[0] * d.tb[1] * d.tb[2];
  d.div_bs = make_fastdiv(d.tb[0] * d.tb[1] * d.tb[2]);
  d.div_nb0 = make_fastdiv(d.nbrick[0]);
  d.div_nb1 = make_fastdiv(d.nbrick[1]);
  d.div_tb0 = make_fastdiv(d.tb[0]);
  d.div_tb1 = make_fastdiv(d.tb[1]);
  return d;
}

static void mapping_checks() {
  long cases = 0, before = bad;
  for (int E : {1, 2, 4, 8, 16, 32})
    for (int nx : {3, 4, 5, 7, 8, 12, 17, 24, 64})
      for (int ny : {3, 5, 8, 9, 16})
        for (int nz : {3, 4, 6, 11}) {
          const InstDesc d = tiling(nx, ny, nz, E);
          std::vector<int> cnt(d.nelem, 0);
          for (int b = 0; b < d.nblocks; ++b) {
            int ec[3], ne;
            block_elements(d, E, b, ec, ne);
            if (ne < 0 || ne > E) { ++bad; continue; }
            if (ne == 0) continue;
            if (ec[0] % E || ec[0] + ne > nx || ec[1] >= ny || ec[2] >= nz || ec[0] < 0 || ec[1] < 0 || ec[2] < 0) {
              ++bad;
              continue;
            }
            if (ne != std::min(E, nx - ec[0])) ++bad;
            for (int e = 0; e < ne; ++e) cnt[ec[0] + e + (long)nx * (ec[1] + (long)ny * ec[2])]++;
          }
          for (int c : cnt)
            if (c != 1) ++bad;
          ++cases;
        }
  std::printf("block_elements: %ld meshes x tilings, %ld failures\n", cases, bad - before);
}

static void xcd_checks() {
  long before = bad;
  for (int B : {1, 7, 8, 9, 63, 64, 65, 1000, 131072, 131071}) {
    const int per = (B + 7) / 8;
    std::vector<int> hit(B, 0);
    for (int b = 0; b < 8 * per; ++b) {
      const int lb = (b & 7) * per + (b >> 3);
      if (lb < B) hit[lb]++;
    }
    for (int h : hit)
      if (h != 1) ++bad;
  }
  std::printf("XCD numbering: %ld failures\n", bad - before);
}

int main() {
  fastdiv_checks();
  mapping_checks();
  xcd_checks();
  std::printf("%s\n", bad ? "FAILED" : "ok");
  return bad ? 1 : 0;
}
