// driver.cpp -- command-line driver with the options of the reference's
// driver (driver.cpp:26-47: L, nx, p, tf, dtSnaps; domain [-pi L, pi L]^3)
// plus the ones the reference hard-codes in Solver::Solver.
//
//   ldg_driver [-p 2] [-nx 10] [-L 1] [-tf 10] [-dtSnaps 30] [-cfl 0.01]
//              [-Re 1600] [-M0 0.1] [-mixed] [-nosnap] [-out output] [-quiet]
//              [-checkpoint FILE] [-resume FILE] [-wallLimit SECONDS]
// -checkpoint writes a binary checkpoint when the run ends or hits the wall
// limit; -resume continues a run from one (same options otherwise);
// -wallLimit replaces the 19.5 h limit (the reference's, solver.cpp:351).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "mesh.hpp"
#include "mpiutil.hpp"
#include "solver.hpp"

int main(int argc, char **argv) {
  int p = 2, nx = 10, dtSnaps = 30;
  double L = 1.0, tf = 10.0;
  const char *resume = nullptr;
  ldg::SolverOptions opt;
  for (int i = 1; i < argc; ++i) {
    auto next = [&](const char *flag) -> const char * {
      if (i + 1 >= argc) { std::cerr << "missing value for " << flag << std::endl; std::exit(2); }
      return argv[++i];
    };
    if (!std::strcmp(argv[i], "-p")) p = std::atoi(next("-p"));
    else if (!std::strcmp(argv[i], "-nx")) nx = std::atoi(next("-nx"));
    else if (!std::strcmp(argv[i], "-L")) L = std::atof(next("-L"));
    else if (!std::strcmp(argv[i], "-tf")) tf = std::atof(next("-tf"));
    else if (!std::strcmp(argv[i], "-dtSnaps")) dtSnaps = std::atoi(next("-dtSnaps"));
    else if (!std::strcmp(argv[i], "-cfl")) opt.cfl = std::atof(next("-cfl"));
    else if (!std::strcmp(argv[i], "-Re")) opt.Re = std::atof(next("-Re"));
    else if (!std::strcmp(argv[i], "-M0")) opt.M0 = std::atof(next("-M0"));
    else if (!std::strcmp(argv[i], "-mixed")) opt.precision = LDG_MIXED;
    else if (!std::strcmp(argv[i], "-nosnap")) opt.snapshots = false;
    else if (!std::strcmp(argv[i], "-out")) opt.outputDir = next("-out");
    else if (!std::strcmp(argv[i], "-quiet")) opt.verbose = false;
    else if (!std::strcmp(argv[i], "-checkpoint")) opt.checkpointFile = next("-checkpoint");
    else if (!std::strcmp(argv[i], "-resume")) resume = next("-resume");
    else if (!std::strcmp(argv[i], "-wallLimit")) opt.wallLimitSeconds = std::atof(next("-wallLimit"));
    else { std::cerr << "unknown option " << argv[i] << std::endl; return 2; }
  }
  try {
    ldg::MPIUtil mpi{};
    if (mpi.rank == mpi.ROOT) std::cout << "MPI tasks = " << mpi.np << std::endl;
    const double size = M_PI * L;
    ldg::Mesh mesh{nx, nx, nx, ldg::Point(-size, -size, -size), ldg::Point(size, size, size), mpi};
    std::cout << "p = " << p << "\ntf = " << tf << "\nnx = " << nx << std::endl;
    ldg::Solver solver(p, dtSnaps, tf, L, mesh, opt);
    if (resume) solver.loadCheckpoint(resume);
    solver.dgTimeStep();
  } catch (const std::exception &e) {
    std::cerr << "ERROR: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
