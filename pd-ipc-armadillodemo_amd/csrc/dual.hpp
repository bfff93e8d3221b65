// Forward-mode dual numbers for device code: a value and N tangents.  Used by
// the adjoint to differentiate the narrow-phase functions (primitive
// distances, ACCD, barrier projection) exactly as they are computed: the same
// templated code runs on double in the step and on Dual<N> in the adjoint, so
// the derivative follows every branch the step took.  Comparisons look at the
// value only.
#pragma once
#include <hip/hip_runtime.h>

namespace pdipc {

template <int N>
struct Dual {
  double v;
  double d[N];
  __host__ __device__ Dual() : v(0.0) {
#pragma unroll
    for (int k = 0; k < N; k++) d[k] = 0.0;
  }
  __host__ __device__ Dual(double a) : v(a) {
#pragma unroll
    for (int k = 0; k < N; k++) d[k] = 0.0;
  }
  __host__ __device__ static Dual seed(double a, int k) {
    Dual r(a);
    r.d[k] = 1.0;
    return r;
  }
  __host__ __device__ Dual& operator+=(const Dual& o) {
    v += o.v;
#pragma unroll
    for (int k = 0; k < N; k++) d[k] += o.d[k];
    return *this;
  }
  __host__ __device__ Dual& operator-=(const Dual& o) {
    v -= o.v;
#pragma unroll
    for (int k = 0; k < N; k++) d[k] -= o.d[k];
    return *this;
  }
  __host__ __device__ Dual& operator*=(const Dual& o) { return *this = *this * o; }
  __host__ __device__ Dual& operator/=(const Dual& o) { return *this = *this / o; }
  friend __host__ __device__ Dual operator+(Dual a, const Dual& b) { return a += b; }
  friend __host__ __device__ Dual operator-(Dual a, const Dual& b) { return a -= b; }
  friend __host__ __device__ Dual operator-(const Dual& a) {
    Dual r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < N; k++) r.d[k] = -a.d[k];
    return r;
  }
  friend __host__ __device__ Dual operator*(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
  }
  friend __host__ __device__ Dual operator/(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v / b.v;
#pragma unroll
    for (int k = 0; k < N; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
  }
  friend __host__ __device__ bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
  friend __host__ __device__ bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
  friend __host__ __device__ bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
  friend __host__ __device__ bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
  friend __host__ __device__ bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
  friend __host__ __device__ bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }
  friend __host__ __device__ Dual sqrt(const Dual& a) {
    Dual r;
    r.v = ::sqrt(a.v);
#pragma unroll
    for (int k = 0; k < N; k++) r.d[k] = r.v > 0.0 ? a.d[k] / (2.0 * r.v) : 0.0;
    return r;
  }
  friend __host__ __device__ Dual log(const Dual& a) {
    Dual r;
    r.v = ::log(a.v);
#pragma unroll
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] / a.v;
    return r;
  }
};

__host__ __device__ inline double value_of(double a) { return a; }
template <int N>
__host__ __device__ inline double value_of(const Dual<N>& a) { return a.v; }

}  // namespace pdipc
