// host_solver_check.cpp -- drives the C++ host (Mesh, MPIUtil, Solver,
// dgTimeStep, checkpoints) the way the reference's driver block does
// (driver.cpp:26-47: MPIUtil, Mesh{nx, nx, nx, botLeft, topRight, mpi},
// Solver{p, dtSnaps, tf, L, mesh}, dgTimeStep()), and checks the host's
// bookkeeping: a run stopped by the wall-clock limit and continued, a wall
// limit with snapshots off, and checkpoints refused across precisions.
// Linked against the real libldg.so (on a GPU) or the CPU test double
// tests/ldg_abi_stub.cpp.  Prints "ok" and exits 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>

#include "mesh.hpp"
#include "mpiutil.hpp"
#include "solver.hpp"

#define CHECK(cond)                                                            \
  do {                                                                         \
    if (!(cond)) {                                                             \
      std::cerr << "check failed: " #cond " (line " << __LINE__ << ")\n";      \
      std::exit(1);                                                            \
    }                                                                          \
  } while (0)

static bool exists(const std::string &path) {
  struct stat st;
  return stat(path.c_str(), &st) == 0;
}

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  // the reference's driver block, with a short final time
  ldg::MPIUtil mpi{};
  CHECK(mpi.rank == mpi.ROOT && mpi.np == 1);
  for (int f = 0; f < ldg::MPIUtil::N_FACES; ++f) CHECK(mpi.neighbors[f] == mpi.rank && mpi.tags[f] == (f ^ 1));
  const double L = 1.0, size = M_PI * L;
  ldg::Point botLeft{-size, -size, -size}, topRight{size, size, size};
  const int nx = 4, p = 2, dtSnaps = 3;
  const double tf = 0.01;
  ldg::Mesh mesh{nx, nx, nx, botLeft, topRight, mpi};
  CHECK(mesh.nElements == nx * nx * nx);
  CHECK(std::fabs(mesh.minDX - 2 * size / nx) < 1e-15);
  ldg::SolverOptions quiet;
  quiet.verbose = false;
  quiet.snapshots = false;
  {
    ldg::Solver dgSolver{p, dtSnaps, tf, L, mesh, quiet};
    dgSolver.dgTimeStep();
    CHECK(dgSolver.currentStep() == dgSolver.timeSteps());
    CHECK((long)dgSolver.kineticEnergy().size() == dgSolver.timeSteps());
    CHECK(std::fabs(dgSolver.timeStep() * dgSolver.timeSteps() - tf) < 1e-15);
  }
  // stopped by the wall limit before the first step, then continued
  {
    ldg::SolverOptions o = quiet;
    o.wallLimitSeconds = 0;
    o.outputDir = dir + "/wall_out";
    o.checkpointFile = dir + "/wall.ckpt";
    ldg::Solver s{p, dtSnaps, tf, L, mesh, o};
    s.dgTimeStep();
    CHECK(s.currentStep() == 0);
    CHECK(s.kineticEnergy().empty());
    CHECK(exists(o.checkpointFile));                 // checkpoint written with snapshots off
    CHECK(exists(o.outputDir + "/xyznavier.tmp.0"));  // output directory created for the dump
    // the checkpoint written at the stop resumes in a new solver
    ldg::Solver t{p, dtSnaps, tf, L, mesh, quiet};
    t.loadCheckpoint(o.checkpointFile);
    CHECK(t.currentStep() == 0);
    // continue the stopped run, and run the resumed one
    s.setWallLimit(1e9);
    s.dgTimeStep();
    CHECK(s.currentStep() == s.timeSteps());
    const std::vector<double> ke = s.kineticEnergy();
    CHECK((long)ke.size() == s.timeSteps());
    CHECK((long)s.keDissipation().size() == s.timeSteps() - 2);
    t.dgTimeStep();
    CHECK(t.currentStep() == t.timeSteps());
    CHECK(t.kineticEnergy() == ke);
  }
  // a checkpoint of the other precision is refused
  {
    ldg::SolverOptions m = quiet;
    m.precision = LDG_MIXED;
    m.checkpointFile = dir + "/mixed.ckpt";
    ldg::Solver s{p, dtSnaps, tf, L, mesh, m};
    s.dgTimeStep();
    ldg::Solver f{p, dtSnaps, tf, L, mesh, quiet};
    bool refused = false;
    try {
      f.loadCheckpoint(m.checkpointFile);
    } catch (const std::runtime_error &e) {
      refused = std::string(e.what()).find("different run") != std::string::npos;
    }
    CHECK(refused);
  }
  // a mesh held by more than one process is refused
  {
    ldg::MPIUtil two{};
    two.np = 2;
    bool refused = false;
    try {
      ldg::Mesh bad{nx, nx, nx, botLeft, topRight, two};
    } catch (const std::invalid_argument &) {
      refused = true;
    }
    CHECK(refused);
  }
  std::cout << "ok" << std::endl;
  return 0;
}
