// Harness utilities of src/util.rs:55-118: the randomized residual check,
// the flop-rate formula and a wall-clock timer.
#pragma once
#include <chrono>
#include <random>

#include "momms/composables.hpp"
#include "momms/matrix.hpp"

namespace momms {

// ||C w - beta*C0 w - alpha*A (B w)||^2 for a random w (src/util.rs:55-85,
// extended by the C0 term because inputs here carry a random C0). Host views.
template <class T>
T test_c_eq_a_b(Matrix<T>& a, Matrix<T>& b, Matrix<T>& c, Matrix<T>* c0, T alpha, T beta,
                std::mt19937_64& rng) {
  TripleLoop ref;
  const size_t m = c.height(), n = b.width(), k = a.width();
  Matrix<T> w(n, 1), bw(k, 1), abw(m, 1), cw(m, 1), c0w(m, 1);
  w.fill_rand(rng);
  bw.fill_zero();
  abw.fill_zero();
  cw.fill_zero();
  c0w.fill_zero();
  auto one = ThreadInfo::single_thread();
  ref.run(b, w, bw, one);
  ref.run(a, bw, abw, one);
  ref.run(c, w, cw, one);
  cw.axpy(-alpha, abw);
  if (c0) {
    ref.run(*c0, w, c0w, one);
    cw.axpy(-beta, c0w);
  }
  return cw.frosqr();
}

// 2mnk / t in GFLOP/s (src/util.rs:94-97).
inline double gflops(size_t m, size_t n, size_t k, double seconds) {
  return 2.0 * (double)m * (double)n * (double)k / seconds / 1e9;
}

inline double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace momms
