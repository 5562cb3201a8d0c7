// expr.hpp — compiled coefficient expressions: the bytecode of the
// reference's FunctionManager trees and its evaluator in dual arithmetic.
//
// A deck's "Functions" strings ("thermal diffusion", "thermal source", ...)
// are parsed once, at context creation (expr.hip), into register bytecode:
// every branch of the reference's tree is a left fold over its dependencies
// (functionManager_evaluate.hpp:56-225), so one instruction per dependency
// suffices: r[dst] = op(r[a], r[b]).  Constant branches are folded on the
// host, as the reference evaluates them once at set-up
// (functionManager_create.hpp:522-534).  The evaluator runs inside the
// assembly kernels' point stage, one program per coefficient, in the
// kernel's point-level dual numbers; the same code, compiled for the host,
// serves mhx_expr_eval.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "dual.hpp"

namespace mhx {

constexpr int kExprRegs = 8;       // registers of the evaluator
constexpr int kExprMaxCode = 512;  // instructions over all programs of a context
constexpr int kExprMaxConst = 256;

// coefficients the thermal module reads (thermal.cpp:55-61): their programs
// and reference defaults
enum ExprFn { XF_DIFF = 0, XF_SRC, XF_CP, XF_RHO, XF_BX, XF_BY, XF_BZ, XF_COUNT };

// operand sources of XOP_VAR
enum ExprVar : uint8_t {
  XV_T = 0,   // "T"            (AD)
  XV_GX,      // "grad(T)[x]"   (AD)
  XV_GY,      // "grad(T)[y]"   (AD)
  XV_GZ,      // "grad(T)[z]"   (AD)
  XV_TT,      // "T_t"          (AD)
  XV_X,       // "x" (point coordinate)
  XV_Y,       // "y"
  XV_Z,       // "z"
  XV_ZERO,    // "u", "v", "w": mesh displacement, zero without mesh motion
  XV_TIME,    // "t"
  XV_K0,      // element data "k0"
  XV_SRCE,    // element data "src"
  XV_COUNT
};

enum ExprOp : uint8_t {
  XOP_CONST = 0,  // r[dst] = consts[a | b << 8]
  XOP_VAR,        // r[dst] = variable a
  XOP_COPY,       // r[dst] = r[a]
  XOP_ADD,        // r[a] + r[b]
  XOP_SUB,        // r[a] + (-r[b])   ("minus": data += -tdata)
  XOP_MUL,
  XOP_DIV,
  XOP_POW,        // pow(r[a], r[b]), both may carry derivatives
  XOP_POWS,       // r[a]^r[b], r[b] a scalar: 2..4 within 1e-12 -> repeated products
  XOP_LT, XOP_LTE, XOP_GT, XOP_GTE,
  XOP_MAX, XOP_MIN, XOP_MEAN,
  XOP_SIN, XOP_COS, XOP_TAN, XOP_EXP, XOP_LOG, XOP_ABS, XOP_SQRT, XOP_SINH, XOP_COSH,  // of r[b]
  XOP_COUNT
};

struct ExprIns {
  uint8_t op, dst, a, b;
};

// the compiled programs of a context, as the kernels see them
struct ExprSet {
  const ExprIns* code;      // all programs, back to back
  const double* consts;
  int off[XF_COUNT];        // first instruction of each coefficient's program
  int len[XF_COUNT];        // 0: not given, the value is dflt
  double dflt[XF_COUNT];
  double time;              // "t"
  int active;               // any coefficient given
};

// values of the variables at one point
template <int N>
struct ExprPoint {
  QDual<N> T, Tt, g[3];
  double x[3];
  double time, k0, src;
};

// ---------------------------------------------------------------------------
// dual arithmetic of the evaluator (Sacado Fad rules: product, quotient and
// chain rules on the value and every derivative slot)
// ---------------------------------------------------------------------------
template <int N>
__host__ __device__ inline QDual<N> xq_const(double v) {
  QDual<N> r;
  r.v = v;
  for (int i = 0; i < N; ++i) r.d[i] = 0.0;
  return r;
}
template <int N>
__host__ __device__ inline QDual<N> xq_add(const QDual<N>& a, const QDual<N>& b) {
  QDual<N> r;
  r.v = a.v + b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N>
__host__ __device__ inline QDual<N> xq_sub(const QDual<N>& a, const QDual<N>& b) {
  QDual<N> r;
  r.v = a.v + (-b.v);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + (-b.d[i]);
  return r;
}
template <int N>
__host__ __device__ inline QDual<N> xq_mul(const QDual<N>& a, const QDual<N>& b) {
  QDual<N> r;
  r.v = a.v * b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N>
__host__ __device__ inline QDual<N> xq_div(const QDual<N>& a, const QDual<N>& b) {
  QDual<N> r;
  r.v = a.v / b.v;
  const double b2 = b.v * b.v;
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / b2;
  return r;
}
// f(a) with f' = fp: value fv, slots a.d * fp
template <int N>
__host__ __device__ inline QDual<N> xq_chain(const QDual<N>& a, double fv, double fp) {
  QDual<N> r;
  r.v = fv;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * fp;
  return r;
}
// pow with derivative slots: d pow(a, b) = pow(a, b) (b da / a + log(a) db);
// no derivative where a = 0 (the rule divides by a)
template <int N>
__host__ __device__ inline QDual<N> xq_pow(const QDual<N>& a, const QDual<N>& b) {
  QDual<N> r;
  r.v = pow(a.v, b.v);
  if (a.v == 0.0) {
    for (int i = 0; i < N; ++i) r.d[i] = 0.0;
    return r;
  }
  const double la = log(a.v);
  for (int i = 0; i < N; ++i) {
    double s = 0.0;
    if (a.d[i] != 0.0) s = b.v * a.d[i] / a.v;
    if (b.d[i] != 0.0) s = s + la * b.d[i];
    r.d[i] = r.v * s;
  }
  return r;
}
template <int N>
__host__ __device__ inline QDual<N> xq_bool(bool c) {
  return xq_const<N>(c ? 1.0 : 0.0);
}

// run program [off, off + len) at one point; register 0 holds the result
template <int N>
__host__ __device__ inline QDual<N> expr_run(const ExprIns* code, const double* consts, int off, int len,
                                             const ExprPoint<N>& p) {
  QDual<N> r[kExprRegs];
  for (int k = off; k < off + len; ++k) {
    const ExprIns in = code[k];
    QDual<N> v;
    switch (in.op) {
      case XOP_CONST:
        v = xq_const<N>(consts[in.a | (in.b << 8)]);
        break;
      case XOP_VAR:
        switch (in.a) {
          case XV_T: v = p.T; break;
          case XV_GX: v = p.g[0]; break;
          case XV_GY: v = p.g[1]; break;
          case XV_GZ: v = p.g[2]; break;
          case XV_TT: v = p.Tt; break;
          case XV_X: v = xq_const<N>(p.x[0]); break;
          case XV_Y: v = xq_const<N>(p.x[1]); break;
          case XV_Z: v = xq_const<N>(p.x[2]); break;
          case XV_TIME: v = xq_const<N>(p.time); break;
          case XV_K0: v = xq_const<N>(p.k0); break;
          case XV_SRCE: v = xq_const<N>(p.src); break;
          default: v = xq_const<N>(0.0); break;
        }
        break;
      case XOP_COPY: v = r[in.a]; break;
      case XOP_ADD: v = xq_add(r[in.a], r[in.b]); break;
      case XOP_SUB: v = xq_sub(r[in.a], r[in.b]); break;
      case XOP_MUL: v = xq_mul(r[in.a], r[in.b]); break;
      case XOP_DIV: v = xq_div(r[in.a], r[in.b]); break;
      case XOP_POW: v = xq_pow(r[in.a], r[in.b]); break;
      case XOP_POWS: {
        // scalar exponent (functionManager_evaluate.hpp:952-995): an exponent
        // within 1e-12 of 2, 3 or 4 is applied as repeated products
        const QDual<N> a = r[in.a];
        const double e = r[in.b].v, er = round(e);
        if (er >= 2.0 && er <= 4.0 && fabs(e - er) < 1.0e-12) {
          v = a;
          for (int j = 1; j < (int)er; ++j) v = xq_mul(v, a);
        } else {
          v = xq_pow(a, xq_const<N>(e));
        }
        break;
      }
      case XOP_LT: v = xq_bool<N>(r[in.a].v < r[in.b].v); break;
      case XOP_LTE: v = xq_bool<N>(r[in.a].v <= r[in.b].v); break;
      case XOP_GT: v = xq_bool<N>(r[in.a].v > r[in.b].v); break;
      case XOP_GTE: v = xq_bool<N>(r[in.a].v >= r[in.b].v); break;
      case XOP_MAX: v = r[in.b].v > r[in.a].v ? r[in.b] : r[in.a]; break;
      case XOP_MIN: v = r[in.b].v < r[in.a].v ? r[in.b] : r[in.a]; break;
      case XOP_MEAN: {
        const QDual<N>& a = r[in.a];
        const QDual<N>& b = r[in.b];
        v.v = 0.5 * a.v + 0.5 * b.v;
        for (int i = 0; i < N; ++i) v.d[i] = 0.5 * a.d[i] + 0.5 * b.d[i];
        break;
      }
      default: {
        const QDual<N>& a = r[in.b];
        const double x = a.v;
        switch (in.op) {
          case XOP_SIN: v = xq_chain(a, sin(x), cos(x)); break;
          case XOP_COS: v = xq_chain(a, cos(x), -sin(x)); break;
          case XOP_TAN: {
            const double t = tan(x);
            v = xq_chain(a, t, 1.0 + t * t);
            break;
          }
          case XOP_EXP: {
            const double e = exp(x);
            v = xq_chain(a, e, e);
            break;
          }
          case XOP_LOG: {
            v.v = log(x);
            for (int i = 0; i < N; ++i) v.d[i] = a.d[i] / x;
            break;
          }
          case XOP_ABS:
            if (x < 0.0) {
              v.v = -x;
              for (int i = 0; i < N; ++i) v.d[i] = -a.d[i];
            } else {
              v = a;
            }
            break;
          case XOP_SQRT:
            if (x <= 0.0) {
              v = xq_const<N>(0.0);
            } else {
              v.v = sqrt(x);
              const double den = 2.0 * v.v;
              for (int i = 0; i < N; ++i) v.d[i] = a.d[i] / den;
            }
            break;
          case XOP_SINH: v = xq_chain(a, sinh(x), cosh(x)); break;
          case XOP_COSH: v = xq_chain(a, cosh(x), sinh(x)); break;
          default: v = xq_const<N>(0.0); break;
        }
        break;
      }
    }
    r[in.dst] = v;
  }
  return r[0];
}

template <int N>
__host__ __device__ inline QDual<N> expr_eval(const ExprSet& xs, int fn, const ExprPoint<N>& p) {
  if (xs.len[fn] == 0) return xq_const<N>(xs.dflt[fn]);
  return expr_run<N>(xs.code, xs.consts, xs.off[fn], xs.len[fn], p);
}

}  // namespace mhx
