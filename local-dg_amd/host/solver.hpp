// solver.hpp -- the reference's Solver API (solver.h:20-115) on libldg.so.
//
//   Solver(int p, int dtSnaps, double tf, double L, const Mesh& mesh);
//   void dgTimeStep();
//
// Construction builds the operators and the Taylor-Green state on the GPU
// (Solver::Solver / initialCondition, solver.cpp:7-71, 194-236);
// dgTimeStep() runs the RK4 loop to tf on the GPU (solver.cpp:295-453).
// Options the reference hard-codes are gathered in SolverOptions with the
// reference's values as defaults.
#pragma once

#include <string>
#include <vector>

#include "array.hpp"
#include "ldg.h"
#include "mesh.hpp"

namespace ldg {

struct SolverOptions {
  double cfl = 0.01;              // solver.cpp:47
  double Re = 1600.0, Pr = 0.71, gamma = 1.4, M0 = 0.1, rho0 = 1.0, V0 = 1.0, R = 8.314;  // solver.cpp:30-36
  int precision = LDG_FP64;
  bool snapshots = true;          // output/xyzu.txt.<n> every dtSnaps steps (solver.cpp:366-381)
  std::string outputDir = "output";
  double wallLimitSeconds = 19.5 * 60 * 60;   // solver.cpp:351, checked before every step
  bool verbose = true;
  std::string checkpointFile;     // if set: binary checkpoint at the end of dgTimeStep and at the wall limit
};

class Solver {
 public:
  Solver();                                                                   // Solver{2, 10, 1.0, 1.0, Mesh{}}
  Solver(int p, int dtSnaps, double tf, double L, const Mesh &mesh, const SolverOptions &opt = SolverOptions());
  ~Solver();
  Solver(const Solver &) = delete;
  Solver &operator=(const Solver &) = delete;

  void dgTimeStep();

  // binary checkpoint / resume (io.hpp): the state, the step reached and
  // the KE history so far; loadCheckpoint continues the same run (same
  // order, mesh, physics and dt), and dgTimeStep then runs the rest of it
  void saveCheckpoint(const std::string &filename);
  void loadCheckpoint(const std::string &filename);
  long currentStep() const { return step; }
  // replaces the wall-clock limit for later dgTimeStep calls
  void setWallLimit(double seconds) { opt.wallLimitSeconds = seconds; }

  // results and state: the KE history of the steps run so far
  std::vector<double> kineticEnergy() const { return std::vector<double>(kes.begin(), kes.begin() + step); }
  const std::vector<double> &keDissipation() const { return keDiss; }
  double timeStep() const { return dt; }
  long timeSteps() const { return timesteps; }
  int dofs() const { return dofs_; }
  const darray &u();                      // current state (copied from the device)
  const darray &globalCoords() const { return coords; }
  ldg_handle *handle() { return h; }

 private:
  Mesh mesh;
  SolverOptions opt;
  ldg_handle *h = nullptr;
  double tf, dt, L;
  long timesteps, step = 0;
  int dtSnaps, order, dofs_;
  darray uHost, coords;
  std::vector<double> kes, keDiss;   // kes has one entry per time step of the run
  void check(int rc, const char *what) const;
};

}  // namespace ldg
