// Test stand-in for green::ndarray::ndarray: owning, row-major, shape().
#pragma once
#include <array>
#include <vector>
namespace green::ndarray {
  template <class T, size_t D>
  class ndarray {
  public:
    template <class... S>
    explicit ndarray(S... s) : shape_{static_cast<size_t>(s)...}, d_((static_cast<size_t>(s) * ...)) {}
    T*                           data() { return d_.data(); }
    const T*                     data() const { return d_.data(); }
    const std::array<size_t, D>& shape() const { return shape_; }
    size_t                       size() const { return d_.size(); }
  private:
    std::array<size_t, D> shape_;
    std::vector<T>        d_;
  };
}  // namespace green::ndarray
