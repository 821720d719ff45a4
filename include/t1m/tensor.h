// t1m::tensor and the scalar concepts: the same interface as the reference
// (include/t1m/t1m.h:13-66 of MatthiasReumann/t1m).  The reference's README
// includes "t1m/tensor.h" (README.md:30) but ships no such header; this one
// exists so that include works.
//
// Difference from the reference: memory_layout::row_major gets defined
// strides (last mode fastest); the reference leaves them uninitialised
// (t1m.h:51-52).
#pragma once

#include <array>
#include <complex>
#include <concepts>
#include <cstddef>
#include <cstdint>
#include <type_traits>

namespace t1m {

template <typename T>
concept TensorScalarArithmetic =
    std::is_same_v<T, float> || std::is_same_v<T, double>;

template <typename T>
concept TensorScalarCompound = std::is_same_v<T, std::complex<float>> ||
                               std::is_same_v<T, std::complex<double>>;

template <typename T>
concept TensorScalar = TensorScalarArithmetic<T> || TensorScalarCompound<T>;

namespace detail {
template <typename T> struct underlying { using type = T; };
template <typename T> struct underlying<std::complex<T>> { using type = T; };
}  // namespace detail

// float / double for real and complex scalars alike.  (The reference's
// std::conditional_t form names T::value_type eagerly, so it only works for
// complex T; t1m.h:25-27.)
template <typename T>
using underlying_scalar_t = typename detail::underlying<T>::type;

enum memory_layout : std::uint8_t { row_major, col_major };

// Non-owning view of a dense tensor; data may be host or device memory.
template <TensorScalar T, std::size_t ndim>
class tensor {
 public:
  using dim_array = std::array<std::size_t, ndim>;
  using stride_array = std::array<std::size_t, ndim>;

  explicit tensor(const dim_array& dims, T* data,
                  memory_layout layout = col_major)
      : data_(data), layout_{layout}, dims_(dims) {}

  [[nodiscard]] constexpr std::size_t rank() const noexcept { return ndim; }
  [[nodiscard]] constexpr dim_array dims() const noexcept { return dims_; }
  [[nodiscard]] constexpr stride_array strides() const {
    stride_array s{};
    if constexpr (ndim > 0) {
      if (layout_ == col_major) {
        std::size_t acc = 1;
        for (std::size_t i = 0; i < ndim; ++i) { s[i] = acc; acc *= dims_[i]; }
      } else {
        std::size_t acc = 1;
        for (std::size_t i = ndim; i-- > 0;) { s[i] = acc; acc *= dims_[i]; }
      }
    }
    return s;
  }
  [[nodiscard]] T* data() noexcept { return data_; }
  [[nodiscard]] const T* data() const noexcept { return data_; }
  [[nodiscard]] constexpr memory_layout layout() const noexcept { return layout_; }

 private:
  T* data_;
  memory_layout layout_;
  dim_array dims_;
};

}  // namespace t1m
