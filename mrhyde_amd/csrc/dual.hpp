// dual.hpp — point-level forward-mode dual numbers (value + N derivative
// slots) for the assembly kernels: the point physics is evaluated once in
// dual arithmetic and its slots give the derivatives the element Jacobian
// needs through the chain rule (Sacado-Fad semantics restricted to the few
// quantities the physics actually depends on at a point).
#pragma once

#include <hip/hip_runtime.h>

namespace mhx {

template <int N>
struct QDual {
  double v;
  double d[N];
};
template <int N>
__device__ __forceinline__ QDual<N> qd_const(double x) {
  QDual<N> r;
  r.v = x;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = 0.0;
  return r;
}
template <int N>
__device__ __forceinline__ QDual<N> qd_seed(double x, int slot) {
  QDual<N> r = qd_const<N>(x);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (i == slot) ? 1.0 : 0.0;
  return r;
}
template <int N>
__device__ __forceinline__ QDual<N> operator*(const QDual<N>& a, const QDual<N>& b) {
  QDual<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ QDual<N> operator*(double a, const QDual<N>& b) {
  QDual<N> r;
  r.v = a * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a * b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ QDual<N> operator+(double a, const QDual<N>& b) {
  QDual<N> r = b;
  r.v = a + b.v;
  return r;
}
template <int N>
__device__ __forceinline__ QDual<N> operator+(const QDual<N>& a, const QDual<N>& b) {
  QDual<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}

}  // namespace mhx
