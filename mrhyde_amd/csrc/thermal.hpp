// thermal.hpp — pieces shared by the thermal assembly kernels: compile-time
// 1-D Gauss-Legendre / equispaced-Lagrange tables, the launch shape of the
// element-per-group kernel and the thermal point physics.
#pragma once

#include <cmath>

#include "dual.hpp"
#include "mhx_internal.hpp"

namespace mhx {

constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }
constexpr int pow2ceil(int x) {
  int r = 1;
  while (r < x) r *= 2;
  return r;
}

// ---------------------------------------------------------------------------
// 1-D reference tables as compile-time constants per (order, points): Gauss-
// Legendre points (descending) and weights, equispaced Lagrange values and
// slopes, and the linear geometry map.  Uniform-index reads become scalar
// loads of .rodata that the compiler can re-issue instead of keeping the whole
// table live in registers.
// ---------------------------------------------------------------------------
struct Tab1 {
  double X[kMaxQ1], W[kMaxQ1];
  double B[kMaxP1][kMaxQ1], D[kMaxP1][kMaxQ1];
  double N[2][kMaxQ1], M[2][kMaxQ1];
};

constexpr double cx_cos(double x) {  // Taylor series, |x| <= pi
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 40; ++k) {
    term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}

constexpr void cx_lagrange(int p, double xi, double* L, double* dL) {
  double nodes[kMaxP1] = {};
  for (int k = 0; k <= p; ++k) nodes[k] = -1.0 + 2.0 * k / p;
  for (int k = 0; k <= p; ++k) {
    double v = 1.0, d = 0.0;
    for (int m = 0; m <= p; ++m) {
      if (m == k) continue;
      const double den = nodes[k] - nodes[m];
      double prod = 1.0 / den;
      for (int l = 0; l <= p; ++l) {
        if (l == k || l == m) continue;
        prod *= (xi - nodes[l]) / (nodes[k] - nodes[l]);
      }
      d += prod;
      v *= (xi - nodes[m]) / den;
    }
    L[k] = v;
    dL[k] = d;
  }
}

constexpr Tab1 make_tab1(int p, int n) {
  Tab1 t{};
  for (int i = 0; i < n; ++i) {
    double z = cx_cos(3.14159265358979323846 * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 60; ++it) {
      double p0 = 1.0, p1 = 0.0;
      for (int k = 1; k <= n; ++k) {
        const double p2 = p1;
        p1 = p0;
        p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
      }
      dp = n * (z * p0 - p1) / (z * z - 1.0);
      z -= p0 / dp;
    }
    double p0 = 1.0, p1 = 0.0;
    for (int k = 1; k <= n; ++k) {
      const double p2 = p1;
      p1 = p0;
      p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
    }
    dp = n * (z * p0 - p1) / (z * z - 1.0);
    t.X[i] = (n % 2 == 1 && i == n / 2) ? 0.0 : z;
    t.W[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
  for (int q = 0; q < n; ++q) {
    double L[kMaxP1] = {}, dL[kMaxP1] = {}, L1[kMaxP1] = {}, dL1[kMaxP1] = {};
    cx_lagrange(p, t.X[q], L, dL);
    cx_lagrange(1, t.X[q], L1, dL1);
    for (int i = 0; i <= p; ++i) {
      t.B[i][q] = L[i];
      t.D[i][q] = dL[i];
    }
    for (int i = 0; i < 2; ++i) {
      t.N[i][q] = L1[i];
      t.M[i][q] = dL1[i];
    }
  }
  return t;
}

template <int P, int NQ1>
struct T1 {
  static constexpr Tab1 v = make_tab1(P, NQ1);
};

template <int DIM, int P, int NQ1>
struct Cfg {
  static constexpr int P1 = P + 1;
  static constexpr int NB = ipow(P1, DIM);
  static constexpr int NQ = ipow(NQ1, DIM);
  static constexpr int NV = 1 << DIM;
  static constexpr int NC = DIM + 1;
  static constexpr int G = pow2ceil(NB + 1) < 4 ? 4 : pow2ceil(NB + 1);
  static constexpr int BLOCK = 256;
  static constexpr int GPB = BLOCK / G;
  static_assert(G <= 64, "element needs more than one wavefront");
  static_assert(BLOCK <= 256 && BLOCK % 64 == 0, "k_thermal is built with __launch_bounds__(256)");
};

// ---------------------------------------------------------------------------
// thermal physics at one point (thermal.cpp:125-163, steady, no advection):
//   flux_c = k(T) dT/dx_c,   src = -f(x),   k = k0 (1 + a T^2)
// written once for a generic scalar type (the reference instantiates the
// module for ScalarT and AD types, thermal.cpp:453-467).
// ---------------------------------------------------------------------------
struct ThermalPhysics {
  static constexpr bool kSrcDependsOnU = false;
  template <int DIM, class T>
  __device__ static void volume(double k0, double a2, double f, const T& u, const T* g, T* flux, T& src) {
    T diff = k0 * (1.0 + a2 * (u * u));
#pragma unroll
    for (int c = 0; c < DIM; ++c) flux[c] = diff * g[c];
    src = qd_const<DIM + 1>(-f);
  }
};

__device__ __forceinline__ double thermal_source(const ThermalCoef& cf, long e, const double* x, int dim) {
  double f = cf.src ? cf.src[e] : 0.0;
  for (int m = 0; m < cf.n_modes; ++m) {
    double t = cf.amp[m];
    for (int a = 0; a < dim; ++a) {
      double w = cf.freq[m][a];
      if (w != 0.0) t = t * sin(w * x[a]);
    }
    f += t;
  }
  return f;
}

}  // namespace mhx
