#include "solver.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <iostream>
#include <stdexcept>
#include <sys/stat.h>

#include "io.hpp"

namespace ldg {

Solver::Solver() : Solver(2, 10, 1.0, 1.0, Mesh()) {}

Solver::Solver(int p, int dtSnaps_, double tf_, double L, const Mesh &mesh_, const SolverOptions &opt_)
    : mesh(mesh_), opt(opt_), tf(tf_), L(L), dtSnaps(dtSnaps_), order(p) {
  ldg_config c;
  ldg_config_default(&c);
  c.p = p;
  c.nx = mesh.nx; c.ny = mesh.ny; c.nz = mesh.nz;
  c.xmin[0] = mesh.botLeft.x; c.xmin[1] = mesh.botLeft.y; c.xmin[2] = mesh.botLeft.z;
  c.xmax[0] = mesh.topRight.x; c.xmax[1] = mesh.topRight.y; c.xmax[2] = mesh.topRight.z;
  c.L = L;
  c.Re = opt.Re; c.Pr = opt.Pr; c.gamma = opt.gamma; c.M0 = opt.M0; c.rho0 = opt.rho0; c.V0 = opt.V0; c.R = opt.R;
  c.precision = opt.precision;
  check(ldg_create(&c, &h), "ldg_create");
  int nel = 0;
  ldg_sizes(h, &dofs_, &nel);
  mesh.mpi.initDatatype((p + 1) * (p + 1));   // solver.cpp:58-59
  mesh.mpi.initFaces(Mesh::N_FACES);
  // dt from the CFL rule, then rounded so that tf is hit exactly (solver.cpp:47-52)
  dt = ldg_time_step(h, opt.cfl);
  timesteps = (long)std::ceil(tf / dt);
  dt = tf / timesteps;
  check(ldg_initial_condition(h), "ldg_initial_condition");
  // node coordinates for the snapshot files (mesh.cpp:389-408)
  const int n = p + 1;
  coords.realloc(3, dofs_, mesh.nElements);
  const int ns[3] = {mesh.nx, mesh.ny, mesh.nz};
  const double lo[3] = {mesh.botLeft.x, mesh.botLeft.y, mesh.botLeft.z};
  const double hi[3] = {mesh.topRight.x, mesh.topRight.y, mesh.topRight.z};
  for (long k = 0; k < mesh.nElements; ++k) {
    const long ie[3] = {k % mesh.nx, (k / mesh.nx) % mesh.ny, k / ((long)mesh.nx * mesh.ny)};
    for (int iN = 0; iN < dofs_; ++iN) {
      const int in[3] = {iN % n, (iN / n) % n, iN / (n * n)};
      for (int l = 0; l < 3; ++l) {
        const double a = ie[l] * (hi[l] - lo[l]) / ns[l] + lo[l];
        const double b = (ie[l] + 1) * (hi[l] - lo[l]) / ns[l] + lo[l];
        coords(l, iN, k) = a + .5 * (std::cos((p - in[l]) * M_PI / p) + 1.0) * (b - a);
      }
    }
  }
}

Solver::~Solver() { ldg_destroy(h); }

void Solver::check(int rc, const char *what) const {
  if (rc != LDG_OK) throw std::runtime_error(std::string(what) + ": " + ldg_last_error(h));
}

const darray &Solver::u() {
  if (uHost.size() != (std::size_t)dofs_ * 5 * mesh.nElements) uHost.realloc(dofs_, 5, mesh.nElements);
  check(ldg_get_state(h, uHost.data()), "ldg_get_state");
  return uHost;
}

void Solver::saveCheckpoint(const std::string &filename) {
  Checkpoint c;
  c.p = order; c.nx = mesh.nx; c.ny = mesh.ny; c.nz = mesh.nz; c.precision = opt.precision;
  c.step = step; c.timesteps = timesteps;
  c.dt = dt; c.tf = tf; c.L = L; c.Re = opt.Re; c.M0 = opt.M0; c.V0 = opt.V0; c.Pr = opt.Pr;
  c.lo[0] = mesh.botLeft.x; c.lo[1] = mesh.botLeft.y; c.lo[2] = mesh.botLeft.z;
  c.hi[0] = mesh.topRight.x; c.hi[1] = mesh.topRight.y; c.hi[2] = mesh.topRight.z;
  const darray &uh = u();
  c.u.assign(uh.data(), uh.data() + uh.size());
  c.ke.assign(kes.begin(), kes.begin() + std::min<std::size_t>(kes.size(), (std::size_t)step));
  if (!writeCheckpoint(filename, c)) throw std::runtime_error("could not write checkpoint " + filename);
}

void Solver::loadCheckpoint(const std::string &filename) {
  Checkpoint c;
  if (!readCheckpoint(filename, c)) throw std::runtime_error("could not read checkpoint " + filename);
  const bool same = c.p == order && c.nx == mesh.nx && c.ny == mesh.ny && c.nz == mesh.nz && c.timesteps == timesteps &&
                    c.precision == opt.precision && c.L == L &&
                    c.dt == dt && c.Re == opt.Re && c.M0 == opt.M0 && c.V0 == opt.V0 && c.Pr == opt.Pr &&
                    c.lo[0] == mesh.botLeft.x && c.hi[0] == mesh.topRight.x && c.lo[1] == mesh.botLeft.y &&
                    c.hi[1] == mesh.topRight.y && c.lo[2] == mesh.botLeft.z && c.hi[2] == mesh.topRight.z;
  if (!same)
    throw std::runtime_error("checkpoint " + filename +
                             " is from a different run (order, mesh, precision, L, physics or dt)");
  check(ldg_set_state(h, c.u.data()), "ldg_set_state");
  step = c.step;
  kes.assign(timesteps, 0.0);
  std::copy(c.ke.begin(), c.ke.end(), kes.begin());
}

void Solver::dgTimeStep() {
  if (step == 0 || kes.size() != (std::size_t)timesteps) kes.resize(timesteps, 0.0);
  if (opt.verbose) std::cout << "Time stepping until tf = " << tf << std::endl;
  if (opt.snapshots) mkdir(opt.outputDir.c_str(), 0755);
  const auto start = std::chrono::steady_clock::now();
  auto elapsed = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count(); };
  double perStep = 0.0;   // measured seconds per step, to stop a call at the wall limit
  long &iStep = step;
  while (iStep < timesteps) {
    if (elapsed() >= opt.wallLimitSeconds) {    // dump state and stop (solver.cpp:349-360)
      if (!opt.checkpointFile.empty()) saveCheckpoint(opt.checkpointFile);
      mkdir(opt.outputDir.c_str(), 0755);
      const std::string name = opt.outputDir + "/xyznavier.tmp";
      if (!initXYZVFile(name, 0, "u", 5) || !exportToXYZVFile(name, 0, coords, u()))
        throw std::runtime_error("could not write " + name);
      break;
    }
    if (opt.snapshots && iStep % dtSnaps == 0) {
      const std::string name = opt.outputDir + "/xyzu.txt";
      if (!initXYZVFile(name, (int)(iStep / dtSnaps), "u", 5) ||
          !exportToXYZVFile(name, (int)(iStep / dtSnaps), coords, u()))
        throw std::runtime_error("could not write " + name);
    }
    // steps up to the next snapshot, but no more than fit before the wall
    // limit at the measured rate, so the limit is checked between steps as in
    // the reference (solver.cpp:349) without one call per step
    long chunk = std::min<long>(timesteps - iStep, dtSnaps - iStep % dtSnaps);
    if (perStep > 0) {
      const double left = opt.wallLimitSeconds - elapsed();
      chunk = std::max<long>(1, std::min<long>(chunk, (long)(left / perStep)));
    } else {
      chunk = 1;
    }
    if (opt.verbose) std::cout << "time = " << iStep * dt << std::endl;
    const double t0 = elapsed();
    check(ldg_step(h, dt, (int)chunk, kes.data() + iStep), "ldg_step");
    perStep = (elapsed() - t0) / chunk;
    iStep += chunk;
  }
  if (iStep == timesteps && !opt.checkpointFile.empty()) saveCheckpoint(opt.checkpointFile);
  // dissipation -(KE[i] - KE[i-2]) / (2 dt) over the steps run (solver.cpp:437-440)
  keDiss.clear();
  for (long i = 2; i < iStep; ++i) keDiss.push_back(-(kes[i] - kes[i - 2]) / (2 * dt));
  if (opt.verbose) {
    std::cout << "Finished time stepping. Time elapsed = " << elapsed() << std::endl;
    std::cout << "Outputting kinetic energy dissipation for Navier-Stokes: " << std::endl;
    for (std::size_t i = 0; i < keDiss.size(); ++i) std::cout << keDiss[i] << (i + 1 < keDiss.size() ? ", " : "\n");
  }
}

}  // namespace ldg
