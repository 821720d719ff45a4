// Drop-in replacement for the reference's public header include/t1m/t1m.h
// (MatthiasReumann/t1m): the same t1m::tensor and the same two
// t1m::contract overloads (reference t1m.h:560-568 real, 570-577 complex),
// executed on an MI355X through libt1m_hip.so (include/t1m_hip.h).
//
// Differences a caller can see:
//  * the trailing `const cntx_t* cntx = bli_gks_query_cntx()` parameter (a
//    BLIS context) becomes `const t1m::context* ctx = nullptr` (stream,
//    precision, synchronisation); no caller in the reference passes it;
//  * errors the reference leaves undefined (bad labels, extent mismatch)
//    throw std::invalid_argument; device failures throw std::runtime_error;
//    allocation failures throw std::bad_alloc as in the reference
//    (t1m.h:455-458);
//  * tensors may live in host memory (copied through the device per call,
//    as the reference's examples and tests allocate them) or in device
//    memory (no copies).
// Semantics are the reference's: C += alpha * A . B, beta accepted and
// ignored (t1m.h:497-502, 543-548), alpha fixed to 1 for complex data.
#pragma once

#include <complex>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "t1m/tensor.h"
#include "t1m_hip.h"

namespace t1m {

// Replaces the BLIS cntx_t* of the reference signature.
struct context {
  void* stream = nullptr;                     // hipStream_t
  int precision = T1M_PRECISION_NATIVE;       // t1m_precision
  bool synchronize = true;                    // result visible on return
  bool autotune = false;                      // time tile shapes on first use
  bool deterministic = false;                 // never split K (no atomics)
};

namespace detail {

template <typename T>
constexpr int dtype_of() {
  if constexpr (std::is_same_v<T, float>) return T1M_FLOAT32;
  else if constexpr (std::is_same_v<T, double>) return T1M_FLOAT64;
  else if constexpr (std::is_same_v<T, std::complex<float>>) return T1M_COMPLEX64;
  else return T1M_COMPLEX128;
}

template <typename T, std::size_t ndim>
struct desc {
  int64_t dims[ndim > 0 ? ndim : 1];
  int64_t strides[ndim > 0 ? ndim : 1];
  t1m_tensor t;
  explicit desc(const tensor<T, ndim>& x) {
    const auto d = x.dims();
    const auto s = x.strides();
    for (std::size_t i = 0; i < ndim; ++i) {
      dims[i] = static_cast<int64_t>(d[i]);
      strides[i] = static_cast<int64_t>(s[i]);
    }
    t.dtype = dtype_of<T>();
    t.rank = static_cast<int32_t>(ndim);
    t.dims = dims;
    t.strides = strides;
    t.data = const_cast<void*>(static_cast<const void*>(x.data()));
  }
};

inline void raise(int status) {
  switch (status) {
    case T1M_SUCCESS: return;
    case T1M_ERROR_OUT_OF_MEMORY: throw std::bad_alloc();
    case T1M_ERROR_INVALID_ARGUMENT:
    case T1M_ERROR_UNSUPPORTED: throw std::invalid_argument(t1m_last_error());
    default: throw std::runtime_error(t1m_last_error());
  }
}

template <typename T, std::size_t na, std::size_t nb, std::size_t nc>
void run(double alpha, const tensor<T, na>& a, const std::string& la,
         const tensor<T, nb>& b, const std::string& lb, tensor<T, nc>& c,
         const std::string& lc, const context* ctx) {
  desc<T, na> A(a);
  desc<T, nb> B(b);
  desc<T, nc> C(c);
  t1m_options o;
  t1m_options_init(&o);
  o.alpha_re = alpha;
  if (ctx) {
    o.stream = ctx->stream;
    o.precision = ctx->precision;
    o.synchronize = ctx->synchronize ? 1 : 0;
    o.autotune = ctx->autotune ? 1 : 0;
    o.deterministic = ctx->deterministic ? 1 : 0;
  }
  raise(t1m_contract(&A.t, la.c_str(), &B.t, lb.c_str(), &C.t, lc.c_str(), &o));
}

}  // namespace detail

// Real contraction: C += alpha * A . B (beta is ignored, as in the reference).
template <TensorScalarArithmetic T, std::size_t ndim_a, std::size_t ndim_b,
          std::size_t ndim_c>
void contract(const T alpha, const tensor<T, ndim_a>& a,
              const std::string& labels_a, const tensor<T, ndim_b>& b,
              const std::string& labels_b, const T beta, tensor<T, ndim_c>& c,
              const std::string& labels_c, const context* ctx = nullptr) {
  (void)beta;
  detail::run(static_cast<double>(alpha), a, labels_a, b, labels_b, c,
              labels_c, ctx);
}

// Complex contraction: C += A . B (1M method on the device).
template <TensorScalarCompound T, std::size_t ndim_a, std::size_t ndim_b,
          std::size_t ndim_c, class U = typename T::value_type>
void contract(const tensor<T, ndim_a>& a, const std::string& labels_a,
              const tensor<T, ndim_b>& b, const std::string& labels_b,
              tensor<T, ndim_c>& c, const std::string& labels_c,
              const context* ctx = nullptr) {
  detail::run(1.0, a, labels_a, b, labels_b, c, labels_c, ctx);
}

// Vector-Jacobian product of C = A . B (no counterpart in the reference):
// grad_a += grad_c . conj(B), grad_b += conj(A) . grad_c.
template <TensorScalar T, std::size_t ndim_a, std::size_t ndim_b,
          std::size_t ndim_c>
void contract_vjp(const tensor<T, ndim_a>& a, const std::string& labels_a,
                  const tensor<T, ndim_b>& b, const std::string& labels_b,
                  const tensor<T, ndim_c>& grad_c, const std::string& labels_c,
                  tensor<T, ndim_a>* grad_a, tensor<T, ndim_b>* grad_b,
                  const context* ctx = nullptr) {
  detail::desc<T, ndim_a> A(a);
  detail::desc<T, ndim_b> B(b);
  detail::desc<T, ndim_c> G(grad_c);
  t1m_options o;
  t1m_options_init(&o);
  if (ctx) {
    o.stream = ctx->stream;
    o.precision = ctx->precision;
    o.synchronize = ctx->synchronize ? 1 : 0;
  }
  if (grad_a && grad_b) {
    detail::desc<T, ndim_a> GA(*grad_a);
    detail::desc<T, ndim_b> GB(*grad_b);
    detail::raise(t1m_contract_vjp(&A.t, labels_a.c_str(), &B.t, labels_b.c_str(), &G.t,
                                   labels_c.c_str(), &GA.t, &GB.t, &o));
  } else if (grad_a) {
    detail::desc<T, ndim_a> GA(*grad_a);
    detail::raise(t1m_contract_vjp(&A.t, labels_a.c_str(), &B.t, labels_b.c_str(), &G.t,
                                   labels_c.c_str(), &GA.t, nullptr, &o));
  } else if (grad_b) {
    detail::desc<T, ndim_b> GB(*grad_b);
    detail::raise(t1m_contract_vjp(&A.t, labels_a.c_str(), &B.t, labels_b.c_str(), &G.t,
                                   labels_c.c_str(), nullptr, &GB.t, &o));
  }
}

}  // namespace t1m
