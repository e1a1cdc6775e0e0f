// array.hpp -- column-major N-d array (first index fastest), the layout
// contract of the reference's array.h (array.h:17-430): u(iN, iS, iK) lives
// at iN + dofs*(iS + 5*iK).  Owns its data; sizes are 64-bit so meshes past
// 2^31 values (large ensembles) still index correctly.
#pragma once

#include <cstddef>
#include <vector>

namespace ldg {

class darray {
 public:
  darray() = default;
  explicit darray(std::size_t s0, std::size_t s1 = 1, std::size_t s2 = 1, std::size_t s3 = 1)
      : dims_{s0, s1, s2, s3}, data_(s0 * s1 * s2 * s3, 0.0) {}
  void realloc(std::size_t s0, std::size_t s1 = 1, std::size_t s2 = 1, std::size_t s3 = 1) {
    dims_[0] = s0; dims_[1] = s1; dims_[2] = s2; dims_[3] = s3;
    data_.assign(s0 * s1 * s2 * s3, 0.0);
  }
  double &operator()(std::size_t i, std::size_t j = 0, std::size_t k = 0, std::size_t l = 0) {
    return data_[i + dims_[0] * (j + dims_[1] * (k + dims_[2] * l))];
  }
  double operator()(std::size_t i, std::size_t j = 0, std::size_t k = 0, std::size_t l = 0) const {
    return data_[i + dims_[0] * (j + dims_[1] * (k + dims_[2] * l))];
  }
  std::size_t size() const { return data_.size(); }
  std::size_t size(int d) const { return dims_[d]; }
  double *data() { return data_.data(); }
  const double *data() const { return data_.data(); }

 private:
  std::size_t dims_[4] = {0, 1, 1, 1};
  std::vector<double> data_;
};

}  // namespace ldg
