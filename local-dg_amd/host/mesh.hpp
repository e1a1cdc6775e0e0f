// mesh.hpp -- the reference's Mesh constructor surface (mesh.h:42-187) for a
// uniform periodic box held whole on one GPU.  The MPIUtil argument of the
// reference's constructors (mesh.h:54-57) is the single-rank MPIUtil of
// mpiutil.hpp: the periodic neighbours it describes are gathered on the
// device, so the mesh keeps only its geometry and a copy of it.
#pragma once

#include <iosfwd>

#include "mpiutil.hpp"

namespace ldg {

struct Point {
  double x = 0.0, y = 0.0, z = 0.0;
  Point() = default;
  Point(double x_, double y_, double z_) : x(x_), y(y_), z(z_) {}
};

class Mesh {
 public:
  static const int DIM = 3;
  static const int N_FACES = 6;

  Mesh();                                    // 10^3 elements on [0,1]^3 (mesh.cpp:44-46)
  explicit Mesh(const MPIUtil &mpi);         // the same, mesh.h:55
  Mesh(int nx, int ny, int nz);              // on [0,1]^3
  Mesh(int nx, int ny, int nz, const Point &botLeft, const Point &topRight);
  // the form the reference's driver uses (mesh.h:57, driver.cpp:31)
  Mesh(int nx, int ny, int nz, const Point &botLeft, const Point &topRight, const MPIUtil &mpi);

  MPIUtil mpi;
  int nx, ny, nz;
  long nElements;
  Point botLeft, topRight;
  double minDX, minDY, minDZ;
};

std::ostream &operator<<(std::ostream &out, const Mesh &mesh);

}  // namespace ldg
