// Field types for the device kernels: real f64 and complex-symmetric c64.
// Complex-symmetric means no conjugation anywhere; |z| is the modulus and
// 1/z = conj(z)/|z|^2 (src/scalar.rs:208-236, num-complex Complex::inv).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace rsl {

struct cplx {
  double re, im;
};

__host__ __device__ inline cplx operator+(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cplx operator-(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cplx operator*(cplx a, cplx b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

template <class T>
struct F;

template <>
struct F<double> {
  __host__ __device__ static double zero() { return 0.0; }
  __host__ __device__ static double one() { return 1.0; }
  __host__ __device__ static double from_real(double x) { return x; }
  __host__ __device__ static double mag(double x) { return fabs(x); }
  __host__ __device__ static double magsq(double x) { return x * x; }
  __host__ __device__ static double real(double x) { return x; }
  __host__ __device__ static double recip(double x) { return 1.0 / x; }
  __host__ __device__ static bool is_zero(double x) { return x == 0.0; }
  __host__ __device__ static double scale(double x, double s) { return x * s; }
};

template <>
struct F<cplx> {
  __host__ __device__ static cplx zero() { return {0.0, 0.0}; }
  __host__ __device__ static cplx one() { return {1.0, 0.0}; }
  __host__ __device__ static cplx from_real(double x) { return {x, 0.0}; }
  __host__ __device__ static double mag(cplx x) { return hypot(x.re, x.im); }
  __host__ __device__ static double magsq(cplx x) { return x.re * x.re + x.im * x.im; }
  __host__ __device__ static double real(cplx x) { return x.re; }
  __host__ __device__ static cplx recip(cplx x) {
    double ns = x.re * x.re + x.im * x.im;
    return {x.re / ns, -x.im / ns};
  }
  __host__ __device__ static bool is_zero(cplx x) { return x.re == 0.0 && x.im == 0.0; }
  __host__ __device__ static cplx scale(cplx x, double s) { return {x.re * s, x.im * s}; }
};

}  // namespace rsl
