// io.hpp -- Paraview XYZV text snapshots in the reference format
// (io.h:22-42, io.cpp:8-99): header "X, Y, Z, u0, ..., u4", one line per
// node, default ostream precision (6 significant digits), files <name>.<n>.
#pragma once

#include <string>
#include <vector>

#include "array.hpp"

namespace ldg {

bool initXYZVFile(const std::string &filename, const std::string &valuename, int nStates);
bool initXYZVFile(const std::string &filename, int timeseries, const std::string &valuename, int nStates);
// globalCoords (3, dofs, K), arr (dofs, nStates, K)
bool exportToXYZVFile(const std::string &filename, const darray &globalCoords, const darray &arr);
bool exportToXYZVFile(const std::string &filename, int timeseries, const darray &globalCoords, const darray &arr);

// Binary checkpoint of a run (no reference counterpart: the reference only
// dumps an XYZV text file when its wall-clock limit is hit).  Little-endian:
//   char[8] "LDGCKPT1"; int32 p, nx, ny, nz, precision;
//   int64 step, timesteps; float64 dt, tf, L, Re, M0, V0, Pr, lo[3], hi[3];
//   int64 n, float64 u[n]  (the state, (dofs, 5, K) column-major);
//   int64 m, float64 ke[m] (reference KE after steps 1..m)
// The same layout is read and written by ldg_amd.io.
struct Checkpoint {
  int p = 0, nx = 0, ny = 0, nz = 0, precision = 0;
  long step = 0, timesteps = 0;
  double dt = 0, tf = 0, L = 0, Re = 0, M0 = 0, V0 = 0, Pr = 0;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  std::vector<double> u, ke;
};
bool writeCheckpoint(const std::string &filename, const Checkpoint &c);
bool readCheckpoint(const std::string &filename, Checkpoint &c);   // false on I/O or format errors

}  // namespace ldg
