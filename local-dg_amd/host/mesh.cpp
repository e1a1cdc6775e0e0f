#include "mesh.hpp"

#include <ostream>
#include <stdexcept>

namespace ldg {

Mesh::Mesh() : Mesh(10, 10, 10) {}

Mesh::Mesh(const MPIUtil &mpi_) : Mesh(10, 10, 10, Point(0, 0, 0), Point(1, 1, 1), mpi_) {}

Mesh::Mesh(int nx_, int ny_, int nz_) : Mesh(nx_, ny_, nz_, Point(0, 0, 0), Point(1, 1, 1)) {}

Mesh::Mesh(int nx_, int ny_, int nz_, const Point &bl, const Point &tr)
    : Mesh(nx_, ny_, nz_, bl, tr, MPIUtil()) {}

Mesh::Mesh(int nx_, int ny_, int nz_, const Point &bl, const Point &tr, const MPIUtil &mpi_)
    : mpi(mpi_), nx(nx_), ny(ny_), nz(nz_), botLeft(bl), topRight(tr) {
  // one process holds the whole box: the mesh must be the whole periodic domain
  if (mpi.np != 1) throw std::invalid_argument("Mesh: one process holds the whole mesh (np must be 1)");
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
