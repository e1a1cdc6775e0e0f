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

namespace {
const char kMagic[8] = {'L', 'D', 'G', 'C', 'K', 'P', 'T', '1'};

template <typename V>
void put(std::ofstream &f, const V &v) { f.write(reinterpret_cast<const char *>(&v), sizeof(V)); }
template <typename V>
bool get(std::ifstream &f, V &v) { return (bool)f.read(reinterpret_cast<char *>(&v), sizeof(V)); }

void put_vec(std::ofstream &f, const std::vector<double> &v) {
  put(f, (long long)v.size());
  f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
}
bool get_vec(std::ifstream &f, std::vector<double> &v, long long limit) {
  long long n = 0;
  if (!get(f, n) || n < 0 || n > limit) return false;
  v.resize((std::size_t)n);
  return (bool)f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(n * sizeof(double)));
}
}  // namespace

bool writeCheckpoint(const std::string &filename, const Checkpoint &c) {
  std::ofstream f(filename, std::ios::out | std::ios::binary | std::ios::trunc);
  if (f.fail()) {
    std::cerr << "ERROR: could not open file " << filename << std::endl;
    return false;
  }
  f.write(kMagic, 8);
  for (int v : {c.p, c.nx, c.ny, c.nz, c.precision}) put(f, (int)v);
  put(f, (long long)c.step);
  put(f, (long long)c.timesteps);
  for (double v : {c.dt, c.tf, c.L, c.Re, c.M0, c.V0, c.Pr}) put(f, v);
  for (double v : c.lo) put(f, v);
  for (double v : c.hi) put(f, v);
  put_vec(f, c.u);
  put_vec(f, c.ke);
  return !f.fail();
}

bool readCheckpoint(const std::string &filename, Checkpoint &c) {
  std::ifstream f(filename, std::ios::in | std::ios::binary);
  if (f.fail()) return false;
  char magic[8];
  if (!f.read(magic, 8) || std::string(magic, 8) != std::string(kMagic, 8)) return false;
  int iv[5];
  for (int &v : iv) if (!get(f, v)) return false;
  c.p = iv[0]; c.nx = iv[1]; c.ny = iv[2]; c.nz = iv[3]; c.precision = iv[4];
  long long step = 0, timesteps = 0;
  if (!get(f, step) || !get(f, timesteps)) return false;
  c.step = (long)step;
  c.timesteps = (long)timesteps;
  double *dv[] = {&c.dt, &c.tf, &c.L, &c.Re, &c.M0, &c.V0, &c.Pr, &c.lo[0], &c.lo[1], &c.lo[2], &c.hi[0], &c.hi[1], &c.hi[2]};
  for (double *v : dv) if (!get(f, *v)) return false;
  if (c.p < 1 || c.nx < 1 || c.ny < 1 || c.nz < 1 || c.step < 0 || c.step > c.timesteps) return false;
  const long long n = (long long)(c.p + 1) * (c.p + 1) * (c.p + 1) * 5 * c.nx * c.ny * c.nz;
  if (!get_vec(f, c.u, n) || (long long)c.u.size() != n) return false;
  return get_vec(f, c.ke, c.timesteps);
}

}  // namespace ldg
