#include "io.hpp"

#include <fstream>
#include <iostream>
#include <sstream>

namespace ldg {

bool initXYZVFile(const std::string &filename, const std::string &valuename, int nStates) {
  std::ofstream f(filename, std::ios::out);
  if (f.fail()) {
    std::cerr << "ERROR: could not open file " << filename << std::endl;
    return false;
  }
  f << "X, Y, Z, ";
  for (int s = 0; s < nStates - 1; ++s) f << valuename << s << ", ";
  f << valuename << nStates - 1 << std::endl;
  return true;
}

bool initXYZVFile(const std::string &filename, int timeseries, const std::string &valuename, int nStates) {
  std::ostringstream oss;
  oss << filename << "." << timeseries;
  return initXYZVFile(oss.str(), valuename, nStates);
}

bool exportToXYZVFile(const std::string &filename, const darray &X, const darray &arr) {
  std::ofstream f(filename, std::ios::out | std::ios::app);
  if (f.fail()) {
    std::cerr << "ERROR: could not open file " << filename << std::endl;
    return false;
  }
  const std::size_t dofs = arr.size(0), ns = arr.size(1), K = arr.size(2);
  for (std::size_t k = 0; k < K; ++k)
    for (std::size_t i = 0; i < dofs; ++i) {
      for (int l = 0; l < 3; ++l) f << X(l, i, k) << ", ";
      for (std::size_t s = 0; s + 1 < ns; ++s) f << arr(i, s, k) << ", ";
      f << arr(i, ns - 1, k) << "\n";
    }
  return !f.fail();
}

bool exportToXYZVFile(const std::string &filename, int timeseries, const darray &X, const darray &arr) {
  std::ostringstream oss;
  oss << filename << "." << timeseries;
  return exportToXYZVFile(oss.str(), X, arr);
}

}  // namespace ldg
