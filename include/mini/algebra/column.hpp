// mini::algebra column type of the C++ host surface.
//
// The reference's System exchanges solution and residual columns as
// algebra::DynamicVector<Scalar>, an Eigen column (algebra/eigen.hpp:22-23).
// Eigen is not part of this library, so DynamicVector here is a plain dense
// column with the element-wise operations the temporal solvers use.  Each
// element is updated on its own in index order, so every operation gives the
// IEEE result of the same scalar operation, bit for bit.
#ifndef MINI_ALGEBRA_COLUMN_HPP_
#define MINI_ALGEBRA_COLUMN_HPP_

#include <cassert>
#include <cstddef>
#include <vector>

namespace mini {
namespace algebra {

template <typename Scalar>
class DynamicVector {
 public:
  DynamicVector() = default;
  explicit DynamicVector(std::size_t n) : v_(n, Scalar(0)) {}

  std::size_t size() const { return v_.size(); }
  Scalar *data() { return v_.data(); }
  Scalar const *data() const { return v_.data(); }
  Scalar &operator[](std::size_t i) { return v_[i]; }
  Scalar const &operator[](std::size_t i) const { return v_[i]; }

  DynamicVector &operator+=(DynamicVector const &that) {
    assert(that.size() == size());
    for (std::size_t i = 0; i < v_.size(); ++i) v_[i] += that.v_[i];
    return *this;
  }
  DynamicVector &operator*=(Scalar s) {
    for (auto &x : v_) x *= s;
    return *this;
  }
  DynamicVector &operator/=(Scalar s) {
    for (auto &x : v_) x /= s;
    return *this;
  }
  bool operator==(DynamicVector const &that) const { return v_ == that.v_; }

 private:
  std::vector<Scalar> v_;
};

// Row-major dense matrix; only the product with a column is needed
// (temporal::Constant).
template <typename Scalar>
class DynamicMatrix {
 public:
  DynamicMatrix(std::size_t rows, std::size_t cols) : rows_(rows), cols_(cols), a_(rows * cols, Scalar(0)) {}
  std::size_t rows() const { return rows_; }
  std::size_t cols() const { return cols_; }
  Scalar &operator()(std::size_t i, std::size_t j) { return a_[i * cols_ + j]; }
  Scalar const &operator()(std::size_t i, std::size_t j) const { return a_[i * cols_ + j]; }
  DynamicVector<Scalar> operator*(DynamicVector<Scalar> const &x) const {
    assert(x.size() == cols_);
    DynamicVector<Scalar> y(rows_);
    for (std::size_t i = 0; i < rows_; ++i) {
      Scalar acc = 0;
      for (std::size_t j = 0; j < cols_; ++j) acc += (*this)(i, j) * x[j];
      y[i] = acc;
    }
    return y;
  }

 private:
  std::size_t rows_, cols_;
  std::vector<Scalar> a_;
};

}  // namespace algebra
}  // namespace mini

#endif  // MINI_ALGEBRA_COLUMN_HPP_
