// mesh.hpp -- the reference's Mesh constructor surface (mesh.h:42-187) for a
// uniform periodic box held whole on one GPU.  The reference's MPIUtil
// argument has no counterpart: its single-rank periodic neighbours become an
// on-device gather, so only the geometry is kept here.
#pragma once

#include <iosfwd>

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
  Mesh(int nx, int ny, int nz);              // on [0,1]^3
  Mesh(int nx, int ny, int nz, const Point &botLeft, const Point &topRight);

  int nx, ny, nz;
  long nElements;
  Point botLeft, topRight;
  double minDX, minDY, minDZ;
};

std::ostream &operator<<(std::ostream &out, const Mesh &mesh);

}  // namespace ldg
