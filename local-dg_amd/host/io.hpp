// io.hpp -- Paraview XYZV text snapshots in the reference format
// (io.h:22-42, io.cpp:8-99): header "X, Y, Z, u0, ..., u4", one line per
// node, default ostream precision (6 significant digits), files <name>.<n>.
#pragma once

#include <string>

#include "array.hpp"

namespace ldg {

bool initXYZVFile(const std::string &filename, const std::string &valuename, int nStates);
bool initXYZVFile(const std::string &filename, int timeseries, const std::string &valuename, int nStates);
// globalCoords (3, dofs, K), arr (dofs, nStates, K)
bool exportToXYZVFile(const std::string &filename, const darray &globalCoords, const darray &arr);
bool exportToXYZVFile(const std::string &filename, int timeseries, const darray &globalCoords, const darray &arr);

}  // namespace ldg
