// mpiutil.hpp -- the reference's MPIUtil (MPIUtil.h:11-58) for one process
// that holds the whole periodic box on one GPU.
//
// The reference builds a periodic 3D Cartesian communicator (MPIUtil.cpp:4-36);
// on one rank every face neighbour is the rank itself and the halo exchange
// is a self-message into ghost slots.  Here that exchange is the GPU's
// periodic face-neighbour gather, so this type keeps the reference's public
// fields with their single-rank values and links no MPI: a caller written
// against the reference (MPIUtil mpi{}; Mesh mesh{nx, nx, nx, bl, tr, mpi};)
// compiles and runs unchanged.
#pragma once

#include <iostream>
#include <string>

namespace ldg {

class MPIUtil {
 public:
  static const int DIM = 3;
  static const int N_FACES = 2 * DIM;
  static const int ROOT = 0;

  int rank = 0;                 // this process
  int np = 1;                   // one process
  int nps[DIM] = {1, 1, 1};     // process grid
  int coords[DIM] = {0, 0, 0};  // this process's place in it
  int neighbors[N_FACES] = {0, 0, 0, 0, 0, 0};   // -x,+x,-y,+y,-z,+z: itself (periodic)
  int tags[N_FACES] = {1, 0, 3, 2, 5, 4};        // receive tag of face iF is iF^1 (MPIUtil.cpp:31-34)
  int faceMap[N_FACES] = {0, 1, 2, 3, 4, 5};
  int nodesPerFace = 0;

  MPIUtil() = default;

  // MPI_FACE is a contiguous run of nodesPerFace doubles (MPIUtil.cpp:39-43);
  // only its size is kept
  void initDatatype(int n) { nodesPerFace = n; }
  // faceMap(l) = meshDim - N_FACES + l (MPIUtil.cpp:47-52)
  void initFaces(int meshDim) {
    for (int l = 0; l < N_FACES; ++l) faceMap[l] = meshDim - N_FACES + l;
  }
  // one rank: print with the rank prefix (MPIUtil.cpp:55-64)
  void printString(const std::string &s) const { std::cout << rank << ": " << s << std::endl; }
};

}  // namespace ldg
