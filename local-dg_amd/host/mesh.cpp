#include "mesh.hpp"

#include <ostream>
#include <stdexcept>

namespace ldg {

Mesh::Mesh() : Mesh(10, 10, 10) {}

Mesh::Mesh(int nx_, int ny_, int nz_) : Mesh(nx_, ny_, nz_, Point(0, 0, 0), Point(1, 1, 1)) {}

Mesh::Mesh(int nx_, int ny_, int nz_, const Point &bl, const Point &tr)
    : nx(nx_), ny(ny_), nz(nz_), botLeft(bl), topRight(tr) {
  // the reference needs interior elements in every direction (mesh.cpp:102-107)
  if (nx < 3 || ny < 3 || nz < 3) throw std::invalid_argument("Mesh: need at least 3 elements per direction");
  if (!(tr.x > bl.x && tr.y > bl.y && tr.z > bl.z)) throw std::invalid_argument("Mesh: empty domain");
  nElements = (long)nx * ny * nz;
  minDX = (tr.x - bl.x) / nx;   // mesh.cpp:149-151
  minDY = (tr.y - bl.y) / ny;
  minDZ = (tr.z - bl.z) / nz;
}

std::ostream &operator<<(std::ostream &out, const Mesh &m) {
  return out << (long)(m.nx + 1) * (m.ny + 1) * (m.nz + 1) << " vertices connected with " << m.nElements
             << " elements.";
}

}  // namespace ldg
