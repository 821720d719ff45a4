// Drop-in check of include/t1m/t1m.h: the reference's example
// (example/example.cpp:6-31) and its GEMM-like contraction test
// (test/test_contraction.cpp:33-165) written against this header, with host
// memory as the reference's callers use it, plus permuted-label checks
// against a naive loop.  Prints one line per case; exit status 0 = all pass.
#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "t1m/t1m.h"

namespace {

int g_fail = 0;

template <typename T>
T rnd(std::mt19937_64& g) {
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  if constexpr (t1m::TensorScalarCompound<T>)
    return T(typename T::value_type(u(g)), typename T::value_type(u(g)));
  else
    return T(u(g));
}

template <typename T>
double absval(const T& x) { return std::abs(x); }

// Naive C += A.B over labels for col-major tensors, accumulated in complex<double>.
template <typename T>
void naive(const std::vector<T>& A, const std::vector<size_t>& da, const std::string& la,
           const std::vector<T>& B, const std::vector<size_t>& db, const std::string& lb,
           std::vector<std::complex<double>>& C, const std::vector<size_t>& dc,
           const std::string& lc) {
  std::string labels = lc;
  std::vector<size_t> ext(dc);
  for (size_t i = 0; i < la.size(); ++i)
    if (lc.find(la[i]) == std::string::npos) { labels += la[i]; ext.push_back(da[i]); }
  auto stride_of = [](const std::vector<size_t>& d, const std::string& l, char ch) -> size_t {
    size_t s = 1;
    for (size_t i = 0; i < l.size(); ++i) {
      if (l[i] == ch) return s;
      s *= d[i];
    }
    return 0;
  };
  size_t total = 1;
  for (size_t e : ext) total *= e;
  for (size_t flat = 0; flat < total; ++flat) {
    size_t r = flat, oa = 0, ob = 0, oc = 0;
    for (size_t t = 0; t < labels.size(); ++t) {
      const size_t x = r % ext[t];
      r /= ext[t];
      oa += x * stride_of(da, la, labels[t]);
      ob += x * stride_of(db, lb, labels[t]);
      oc += x * stride_of(dc, lc, labels[t]);
    }
    C[oc] += std::complex<double>(A[oa]) * std::complex<double>(B[ob]);
  }
}

template <typename T, size_t na, size_t nb, size_t nc>
void check_case(const char* name, const std::string& la, const std::string& lb,
                const std::string& lc, const std::array<size_t, na>& dims_a,
                const std::array<size_t, nb>& dims_b, const std::array<size_t, nc>& dims_c,
                double tol) {
  std::mt19937_64 g(1234 + la.size() * 7 + lb.size());
  auto numel = [](const auto& d) { size_t n = 1; for (size_t x : d) n *= x; return n; };
  std::vector<T> A(numel(dims_a)), B(numel(dims_b)), C(numel(dims_c));
  for (auto& x : A) x = rnd<T>(g);
  for (auto& x : B) x = rnd<T>(g);
  for (auto& x : C) x = rnd<T>(g);
  std::vector<std::complex<double>> R(C.size());
  for (size_t i = 0; i < C.size(); ++i) R[i] = std::complex<double>(C[i]);
  naive(A, {dims_a.begin(), dims_a.end()}, la, B, {dims_b.begin(), dims_b.end()}, lb, R,
        {dims_c.begin(), dims_c.end()}, lc);
  t1m::tensor<T, na> a(dims_a, A.data());
  t1m::tensor<T, nb> b(dims_b, B.data());
  t1m::tensor<T, nc> c(dims_c, C.data());
  if constexpr (t1m::TensorScalarCompound<T>)
    t1m::contract(a, la, b, lb, c, lc);
  else
    t1m::contract(T(1), a, la, b, lb, T(0), c, lc);
  double num = 0, den = 0;
  for (size_t i = 0; i < C.size(); ++i) {
    num += std::norm(std::complex<double>(C[i]) - R[i]);
    den += std::norm(R[i]);
  }
  const double rel = std::sqrt(num / (den > 0 ? den : 1));
  const bool ok = rel <= tol;
  g_fail += ok ? 0 : 1;
  std::printf("%-4s %-34s rel_err=%.3e tol=%.1e\n", ok ? "ok" : "FAIL", name, rel, tol);
}

// test_contraction.cpp:33-112: GEMM-ordered labels, d = 2..max_d, estimator
// ||C - A*B|| / (eps * d * ||A|| * ||B||) <= 1 (norms over all entries here).
template <typename T, size_t na, size_t nb, size_t nc>
void gemm_like(const std::string& la, const std::string& lb, const std::string& lc,
               int em, int en, int ek, size_t max_d) {
  using U = t1m::underlying_scalar_t<T>;
  const double eps = std::numeric_limits<U>::epsilon();
  double worst = 0;
  for (size_t d = 2; d <= max_d; ++d) {
    std::array<size_t, na> da; da.fill(d);
    std::array<size_t, nb> db; db.fill(d);
    std::array<size_t, nc> dc; dc.fill(d);
    const size_t m = size_t(std::pow(d, em)), n = size_t(std::pow(d, en)),
                 k = size_t(std::pow(d, ek));
    std::mt19937_64 g(d);
    std::vector<T> A(m * k), B(k * n), C(m * n, T(0));
    for (auto& x : A) x = rnd<T>(g);
    for (auto& x : B) x = rnd<T>(g);
    t1m::tensor<T, na> a(da, A.data());
    t1m::tensor<T, nb> b(db, B.data());
    t1m::tensor<T, nc> c(dc, C.data());
    if constexpr (t1m::TensorScalarCompound<T>)
      t1m::contract(a, la, b, lb, c, lc);
    else
      t1m::contract(T(1), a, la, b, lb, T(0), c, lc);
    double diff = 0, na2 = 0, nb2 = 0;
    for (size_t j = 0; j < n; ++j)
      for (size_t i = 0; i < m; ++i) {
        std::complex<double> s = 0;
        for (size_t p = 0; p < k; ++p)
          s += std::complex<double>(A[i + p * m]) * std::complex<double>(B[p + j * k]);
        diff += std::norm(std::complex<double>(C[i + j * m]) - s);
      }
    for (auto& x : A) na2 += std::norm(std::complex<double>(x));
    for (auto& x : B) nb2 += std::norm(std::complex<double>(x));
    const double est = std::sqrt(diff) / (eps * d * std::sqrt(na2) * std::sqrt(nb2));
    worst = est > worst ? est : worst;
  }
  const bool ok = worst <= 1.0;
  g_fail += ok ? 0 : 1;
  std::printf("%-4s gemm_like<%s> %s,%s->%s max_estimator=%.3e\n", ok ? "ok" : "FAIL",
              t1m::TensorScalarCompound<T> ? (sizeof(U) == 4 ? "c64" : "c128")
                                           : (sizeof(U) == 4 ? "f32" : "f64"),
              la.c_str(), lb.c_str(), lc.c_str(), worst);
}

template <typename T>
void suite(size_t max_d) {
  gemm_like<T, 2, 2, 2>("ab", "bc", "ac", 1, 1, 1, max_d);
  gemm_like<T, 3, 3, 2>("abc", "bce", "ae", 1, 1, 2, max_d);
  gemm_like<T, 3, 2, 3>("abc", "cd", "abd", 2, 1, 1, max_d);
  gemm_like<T, 3, 3, 4>("abc", "cde", "abde", 2, 2, 1, max_d);
  gemm_like<T, 4, 3, 3>("abcd", "cde", "abe", 2, 1, 2, max_d);
  gemm_like<T, 4, 4, 4>("abcd", "cdef", "abef", 2, 2, 2, max_d);
}

// contract_vjp (this header's addition): grad_a += grad_c . conj(B) and
// grad_b += conj(A) . grad_c for 'asb,btc->astc', against the naive loop.
template <typename T>
void vjp_case(double tol) {
  std::mt19937_64 g(77);
  const std::array<size_t, 3> da{5, 4, 6}, db{6, 4, 3};
  const std::array<size_t, 4> dc{5, 4, 4, 3};
  std::vector<T> A(120), B(72), G(240), GA(120), GB(72);
  for (auto& x : A) x = rnd<T>(g);
  for (auto& x : B) x = rnd<T>(g);
  for (auto& x : G) x = rnd<T>(g);
  for (auto& x : GA) x = rnd<T>(g);
  for (auto& x : GB) x = rnd<T>(g);
  std::vector<T> cA(A.size()), cB(B.size());
  for (size_t i = 0; i < A.size(); ++i) cA[i] = std::conj(A[i]);
  for (size_t i = 0; i < B.size(); ++i) cB[i] = std::conj(B[i]);
  std::vector<std::complex<double>> RA(GA.begin(), GA.end()), RB(GB.begin(), GB.end());
  naive(G, {dc.begin(), dc.end()}, "astc", cB, {db.begin(), db.end()}, "btc", RA,
        {da.begin(), da.end()}, "asb");
  naive(cA, {da.begin(), da.end()}, "asb", G, {dc.begin(), dc.end()}, "astc", RB,
        {db.begin(), db.end()}, "btc");
  t1m::tensor<T, 3> a(da, A.data()), b(db, B.data()), ga(da, GA.data()), gb(db, GB.data());
  t1m::tensor<T, 4> gc(dc, G.data());
  t1m::contract_vjp(a, "asb", b, "btc", gc, "astc", &ga, &gb);
  auto rel = [](const std::vector<T>& x, const std::vector<std::complex<double>>& r) {
    double num = 0, den = 0;
    for (size_t i = 0; i < x.size(); ++i) {
      num += std::norm(std::complex<double>(x[i]) - r[i]);
      den += std::norm(r[i]);
    }
    return std::sqrt(num / den);
  };
  const double ea = rel(GA, RA), eb = rel(GB, RB);
  const bool ok = ea <= tol && eb <= tol;
  g_fail += ok ? 0 : 1;
  std::printf("%-4s contract_vjp asb,btc->astc %s rel_err=%.3e,%.3e tol=%.1e\n",
              ok ? "ok" : "FAIL", sizeof(typename T::value_type) == 4 ? "c64" : "c128", ea, eb,
              tol);
}

}  // namespace

int main() {
  // README example / example.cpp:6-31 (with initialised data).
  check_case<std::complex<float>, 3, 2, 3>("readme abc,bd->acd c64", "abc", "bd", "acd",
                                           {10, 10, 10}, {10, 10}, {10, 10, 10}, 1e-5);
  check_case<std::complex<float>, 4, 4, 4>("abcd,dbef->acef c64", "abcd", "dbef", "acef",
                                           {6, 5, 7, 4}, {4, 5, 3, 6}, {6, 7, 3, 6}, 1e-5);
  check_case<std::complex<double>, 3, 3, 4>("asb,btc->astc c128", "asb", "btc", "astc",
                                            {9, 4, 11}, {11, 4, 7}, {9, 4, 4, 7}, 1e-12);
  check_case<float, 4, 4, 4>("aebf,fdec->abcd f32", "aebf", "fdec", "abcd",
                             {5, 6, 4, 3}, {3, 7, 6, 2}, {5, 4, 2, 7}, 1e-5);
  check_case<double, 5, 2, 5>("efbad,cf->abcde f64", "efbad", "cf", "abcde",
                              {4, 3, 5, 2, 3}, {6, 3}, {2, 5, 6, 3, 4}, 1e-12);
  vjp_case<std::complex<float>>(1e-5);
  vjp_case<std::complex<double>>(1e-12);
  suite<float>(20);   // d = 2..20 as test_contraction.cpp:49,56
  suite<double>(20);
  suite<std::complex<float>>(20);
  suite<std::complex<double>>(20);
  // Errors the reference leaves undefined are reported.
  try {
    std::vector<std::complex<float>> x(8);
    t1m::tensor<std::complex<float>, 2> a({2, 4}, x.data()), b({3, 2}, x.data()),
        c({2, 2}, x.data());
    t1m::contract(a, "ab", b, "bc", c, "ac");
    std::printf("FAIL extent mismatch not reported\n");
    ++g_fail;
  } catch (const std::invalid_argument& e) {
    std::printf("ok   extent mismatch -> std::invalid_argument (%s)\n", e.what());
  }
  std::printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "PASSED", g_fail);
  return g_fail ? 1 : 0;
}
