// Per-element distance terms in the reference's exact arithmetic.
//
// Each term is one summand of the left fold the reference computes; the
// caller accumulates `s = s + term` in f64 with separate roundings (the
// library is compiled with -ffp-contract=off so no FMA is formed).
#pragma once

#include <hip/hip_runtime.h>

namespace sc {

enum MetricId : int { kEuclid = 0, kManhattan = 1, kMinkowski = 2 };

template <typename T, int M>
struct Term;

// euclidian.rs:57-62, T = f64: r = a - b; (r * r) as f64
template <>
struct Term<double, kEuclid> {
  static __device__ __forceinline__ double f(double a, double b, double) {
    double r = a - b;
    return r * r;
  }
};
// euclidian.rs:57-62, T = f32: r and r * r rounded in f32, then widened
template <>
struct Term<float, kEuclid> {
  static __device__ __forceinline__ double f(float a, float b, double) {
    float r = a - b;
    float r2 = r * r;
    return static_cast<double>(r2);
  }
};
// manhattan.rs:56-59: (a - b).to_f64().abs()
template <>
struct Term<double, kManhattan> {
  static __device__ __forceinline__ double f(double a, double b, double) { return fabs(a - b); }
};
template <>
struct Term<float, kManhattan> {
  static __device__ __forceinline__ double f(float a, float b, double) {
    float r = a - b;
    return fabs(static_cast<double>(r));
  }
};
// minkowski.rs:61-64: (a - b).to_f64().abs().powf(p)
template <>
struct Term<double, kMinkowski> {
  static __device__ __forceinline__ double f(double a, double b, double p) {
    return pow(fabs(a - b), p);
  }
};
template <>
struct Term<float, kMinkowski> {
  static __device__ __forceinline__ double f(float a, float b, double p) {
    float r = a - b;
    return pow(fabs(static_cast<double>(r)), p);
  }
};

// Distance from the folded sum: euclidian.rs:71 sqrt, manhattan identity,
// minkowski.rs:66 powf(1/p).  The selection kernels filter on the sum, which
// is valid because each finaliser is monotone non-decreasing.
template <int M>
__device__ __forceinline__ double finalize(double s, double p) {
  if constexpr (M == kEuclid) {
    return __builtin_sqrt(s);
  } else if constexpr (M == kManhattan) {
    return s;
  } else {
    return pow(s, 1.0 / p);
  }
}

}  // namespace sc
