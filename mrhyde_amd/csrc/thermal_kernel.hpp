// thermal_kernel.hpp — the fused thermal J+R assembly kernel k_thermal and
// its helpers (sum-factorised test-function application, the x-segment LDS
// image), shared by the translation units that instantiate it: assemble.hip
// (closed-form coefficients, every schedule) and assemble_expr.hip (compiled
// coefficient functions, element colouring).  The header comment of
// assemble.hip describes the kernel.
#pragma once

#include <cmath>

#include "dual.hpp"
#include "expr.hpp"
#include "lattice.hpp"
#include "mhx_internal.hpp"
#include "thermal.hpp"

namespace mhx {

// ---------------------------------------------------------------------------
// sum-factorised application of all test functions:
//   R[i] = sum_q sum_c f_c(q) ghat_i[c](q),  ghat = (d/dxi_0 .. d/dxi_{d-1}, value)
// f is produced point by point by the callable `fq(qx, qy, qz, f)`.  The
// outermost point axis is streamed: each of its planes is contracted over the
// inner axes and folded into R at once, so only one plane of partial sums is
// live.
// ---------------------------------------------------------------------------
template <int DIM, int P, int NQ1, class FQ>
__device__ __forceinline__ void test_apply(const Tab1& t, FQ&& fq, double* R) {
  constexpr int P1 = P + 1;
  constexpr int NB = DIM == 2 ? P1 * P1 : P1 * P1 * P1;
#pragma unroll
  for (int i = 0; i < NB; ++i) R[i] = 0.0;
  if constexpr (DIM == 2) {
#pragma unroll
    for (int qy = 0; qy < NQ1; ++qy) {
      double A0[P1], A1[P1];
#pragma unroll
      for (int i = 0; i < P1; ++i) A0[i] = A1[i] = 0.0;
#pragma unroll
      for (int qx = 0; qx < NQ1; ++qx) {
        double f[3];
        fq(qx, qy, 0, f);
#pragma unroll
        for (int ix = 0; ix < P1; ++ix) {
          A0[ix] = fma(t.D[ix][qx], f[0], fma(t.B[ix][qx], f[2], A0[ix]));
          A1[ix] = fma(t.B[ix][qx], f[1], A1[ix]);
        }
      }
#pragma unroll
      for (int iy = 0; iy < P1; ++iy)
#pragma unroll
        for (int ix = 0; ix < P1; ++ix)
          R[ix + P1 * iy] = fma(t.B[iy][qy], A0[ix], fma(t.D[iy][qy], A1[ix], R[ix + P1 * iy]));
    }
  } else {
#pragma unroll
    for (int qz = 0; qz < NQ1; ++qz) {
      // contract x over the plane qz: A0 (x-derivative and value parts), A1 (y), A2 (z)
      double A0[P1][NQ1], A1[P1][NQ1], A2[P1][NQ1];
#pragma unroll
      for (int i = 0; i < P1; ++i)
#pragma unroll
        for (int b = 0; b < NQ1; ++b) A0[i][b] = A1[i][b] = A2[i][b] = 0.0;
#pragma unroll
      for (int qy = 0; qy < NQ1; ++qy)
#pragma unroll
        for (int qx = 0; qx < NQ1; ++qx) {
          double f[4];
          fq(qx, qy, qz, f);
#pragma unroll
          for (int ix = 0; ix < P1; ++ix) {
            A0[ix][qy] = fma(t.D[ix][qx], f[0], fma(t.B[ix][qx], f[3], A0[ix][qy]));
            A1[ix][qy] = fma(t.B[ix][qx], f[1], A1[ix][qy]);
            A2[ix][qy] = fma(t.B[ix][qx], f[2], A2[ix][qy]);
          }
        }
      // contract y, then fold the plane into R through the z tables
#pragma unroll
      for (int iy = 0; iy < P1; ++iy)
#pragma unroll
        for (int ix = 0; ix < P1; ++ix) {
          double s = 0.0, z = 0.0;
#pragma unroll
          for (int qy = 0; qy < NQ1; ++qy) {
            s = fma(t.B[iy][qy], A0[ix][qy], fma(t.D[iy][qy], A1[ix][qy], s));
            z = fma(t.B[iy][qy], A2[ix][qy], z);
          }
#pragma unroll
          for (int iz = 0; iz < P1; ++iz) {
            double& r = R[ix + P1 * (iy + P1 * iz)];
            r = fma(t.B[iz][qz], s, fma(t.D[iz][qz], z, r));
          }
        }
    }
  }
}

// LDS image of a segment's Jacobian rows (x-segment scatter).  The segment's
// lattice nodes form x-planes 0..P*S; a plane holds RP rows (the pencil's
// y/z nodes), and each row's share of columns from this segment is CP (y,z)
// column blocks of W x-columns: W = 2P+1 on vertex planes (every P-th plane,
// shared by two elements), P+1 on interior planes.  The order within a row,
// (z', y', x'), is the CRS order, so each (row, z') block of the image is
// one contiguous run of the CRS row when it is written out.
template <int DIM, int P, int S>
struct SegCfg {
  static constexpr int P1 = P + 1;
  static constexpr int RP = ipow(P1, DIM - 1);
  static constexpr int CP = RP;
  static constexpr int WV = 2 * P + 1, WI = P + 1;
  static constexpr int SV = RP * CP * WV, SI = RP * CP * WI;
  static constexpr int GS = SV + (P - 1) * SI;  // a vertex plane and the interior planes after it
  static constexpr int NPL = P * S + 1;
  static constexpr int kSize = S * GS + SV;
  static constexpr int kRows = NPL * RP;
  __device__ static constexpr int plane_off(int X) {
    return (X / P) * GS + (X % P == 0 ? 0 : SV + (X % P - 1) * SI);
  }
};

// lattice range of the columns of a row along one axis (node c of n)
__device__ __forceinline__ void row_range(int c, int p, int n, int& lo, int& len) {
  if (c % p == 0) {
    lo = c - p < 0 ? 0 : c - p;
    const int hi = c + p > n - 1 ? n - 1 : c + p;
    len = hi - lo + 1;
  } else {
    lo = c - c % p;
    len = p + 1;
  }
}

// TR: transient (workset.cpp:601-680): per node the evaluated solution is
// alpha_u U + beta_u and its time derivative alpha_t U + beta_t, U the stage
// unknown (alpha_u = 1, beta_u = 0 for a single-stage tableau); the source
// row gains rho cp dT/dt (thermal.cpp:129) and every derivative slot the
// factor of its seed
// EX: the coefficients come from compiled expressions (a.xp) instead of the
// closed forms (element colouring only)
template <int DIM, int P, int NQ1, bool JAC, int MODE, bool TR = false, bool EX = false>
__global__ __launch_bounds__(256) void k_thermal(AsmArgs a) {
  using C = Cfg<DIM, P, NQ1>;
  constexpr int NB = C::NB, NQ = C::NQ, NV = C::NV, NC = C::NC, G = C::G, GPB = C::GPB, P1 = C::P1;
  constexpr bool BLOCKED = MODE == kBlockColouring;
  constexpr bool SEG = MODE == kSegmentMerge;
  using SG = SegCfg<DIM, P, GPB>;
  __shared__ double s_seg[SEG && JAC ? SG::kSize : 1];
  __shared__ double s_sres[SEG ? SG::kRows : 1];
  __shared__ int s_srp[SEG ? SG::kRows : 1];
  // rows of Dhat that can be nonzero: the flux rows, plus the source row when
  // the source depends on the solution (always when transient)
  constexpr int NR = (ThermalPhysics::kSrcDependsOnU || TR || EX) ? NC : DIM;
  __shared__ double s_u[GPB][NB];
  __shared__ double s_h[TR ? GPB : 1][TR ? NB : 1];
  __shared__ double s_X[GPB][NV * DIM];
  // Dhat = w T D T^T per point
  constexpr int NW = JAC ? NQ * NR * NC : 1;
  __shared__ double s_W[GPB][NW];
  __shared__ double s_F[GPB][NQ * NC];
  __shared__ int s_rp[GPB][NB];

  const int g = threadIdx.x / G, lane = threadIdx.x % G;
  // element colouring runs one element per group and no sub-phase loop (a
  // loop-carried kernel keeps more values live across iterations)
  constexpr int kMaxSub = BLOCKED ? (1 << DIM) : 1;
  for (int sub = 0; sub < kMaxSub; ++sub) {
  if (sub) __syncthreads();  // previous sub-phase's LDS reads and global stores are complete
  int ec[3];
  int sx = 0, s_act = 0;
  const bool active = SEG ? element_of_seg<DIM>(a, g, ec, sx, s_act) : element_of<DIM>(a, g, sub, ec);
  const long e = ec[0] + (long)a.m.nel[0] * (ec[1] + (long)a.m.nel[1] * ec[2]);
  // first-touch flags per axis: is this element (segment, along x in segment
  // scheduling) the first scheduled one among the (at most two) sharing its
  // low / high vertex plane
  bool first_lo[3] = {true, true, true}, first_hi[3] = {true, true, true};
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    if (SEG && d == 0) {
      first_lo[0] = sx == 0 || (sx & 1) < ((sx - 1) & 1);
      first_hi[0] = sx == a.nseg - 1 || (sx & 1) < ((sx + 1) & 1);
      continue;
    }
    const int k = axis_key(a, d, ec[d]);
    first_lo[d] = ec[d] == 0 || k < axis_key(a, d, ec[d] - 1);
    first_hi[d] = ec[d] == a.m.nel[d] - 1 || k < axis_key(a, d, ec[d] + 1);
  }

  // 1. gather --------------------------------------------------------------
  if (active) {
    for (int j = lane; j < NB; j += G) {
      const long nd = node_of<DIM, P>(a.m, ec, j);
      if constexpr (TR) {
        const double un = a.u[nd];
        s_u[g][j] = a.hist_u ? fma(a.alpha_u, un, a.hist_u[nd]) : un;
        s_h[g][j] = fma(a.alpha_t, un, a.hist[nd]);
      } else {
        s_u[g][j] = a.u[nd];
      }
    }
    for (int t = lane; t < NV * DIM; t += G) {
      int v = t / DIM, d = t % DIM;
      long vid = (long)(ec[0] + (v & 1)) +
                 (long)a.m.nvx[0] * ((long)(ec[1] + ((v >> 1) & 1)) + (long)a.m.nvx[1] * (long)(ec[2] + ((v >> 2) & 1)));
      s_X[g][t] = a.vx[vid * 3 + d];
    }
  }
  __syncthreads();

  // 2. point stage -----------------------------------------------------------
  constexpr const Tab1& t = T1<P, NQ1>::v;
  if (active) {
    const double k0 = a.coef.k0 ? a.coef.k0[e] : 1.0;
    for (int q = lane; q < NQ; q += G) {
      const int qi[3] = {q % NQ1, (q / NQ1) % NQ1, DIM == 3 ? q / (NQ1 * NQ1) : 0};
      // reference value / gradient of u (and the history term's value)
      double uq = 0.0, gu[DIM], hq = 0.0;
#pragma unroll
      for (int c = 0; c < DIM; ++c) gu[c] = 0.0;
      // tensor-product evaluation: the outer basis axis runs as a loop (its
      // factor read from the table), the inner axes are unrolled
      double bq[DIM - 1][P1], dq[DIM - 1][P1];
#pragma unroll
      for (int d = 0; d < DIM - 1; ++d)
#pragma unroll
        for (int i = 0; i < P1; ++i) {
          bq[d][i] = t.B[i][qi[d]];
          dq[d][i] = t.D[i][qi[d]];
        }
#pragma unroll 1
      for (int jo = 0; jo < P1; ++jo) {
        const double bo = t.B[jo][qi[DIM - 1]], dox = t.D[jo][qi[DIM - 1]];
        constexpr int NI = DIM == 2 ? P1 : P1 * P1;  // inner-plane basis functions
#pragma unroll
        for (int ji = 0; ji < NI; ++ji) {
          const int jx = ji % P1, jy = ji / P1;
          const double uj = s_u[g][ji + NI * jo];
          if constexpr (DIM == 2) {
            uq = fma(uj, bq[0][jx] * bo, uq);
            if constexpr (TR) hq = fma(s_h[g][ji + NI * jo], bq[0][jx] * bo, hq);
            gu[0] = fma(uj, dq[0][jx] * bo, gu[0]);
            gu[1] = fma(uj, bq[0][jx] * dox, gu[1]);
          } else {
            const double bxy = bq[0][jx] * bq[1][jy];
            uq = fma(uj, bxy * bo, uq);
            if constexpr (TR) hq = fma(s_h[g][ji + NI * jo], bxy * bo, hq);
            gu[0] = fma(uj, dq[0][jx] * bq[1][jy] * bo, gu[0]);
            gu[1] = fma(uj, bq[0][jx] * dq[1][jy] * bo, gu[1]);
            gu[2] = fma(uj, bxy * dox, gu[2]);
          }
        }
      }
      // geometry: x(xi) = sum_v N_v X_v,  J_ab = dx_a/dxi_b
      double J[DIM][DIM], x[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int s = 0; s < DIM; ++s) J[r][s] = 0.0;
      double nq[DIM][2], mq[DIM][2];
#pragma unroll
      for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          nq[d][i] = t.N[i][qi[d]];
          mq[d][i] = t.M[i][qi[d]];
        }
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const int vl[3] = {v & 1, (v >> 1) & 1, (v >> 2) & 1};
        double nv = 1.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) nv *= nq[d][vl[d]];
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
          const double X = s_X[g][v * DIM + r];
          x[r] = fma(nv, X, x[r]);
#pragma unroll
          for (int s = 0; s < DIM; ++s) {
            double dn = 1.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) dn *= (d == s) ? mq[d][vl[d]] : nq[d][vl[d]];
            J[r][s] = fma(X, dn, J[r][s]);
          }
        }
      }
      double Ji[DIM][DIM], det;
      if constexpr (DIM == 2) {
        det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        const double id = 1.0 / det;
        Ji[0][0] = J[1][1] * id;
        Ji[0][1] = -J[0][1] * id;
        Ji[1][0] = -J[1][0] * id;
        Ji[1][1] = J[0][0] * id;
      } else {
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        const double id = 1.0 / det;
        Ji[0][0] = c00 * id;
        Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
        Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
        Ji[1][0] = c01 * id;
        Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
        Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
        Ji[2][0] = c02 * id;
        Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
        Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
      }
      double w = fabs(det);
#pragma unroll
      for (int d = 0; d < DIM; ++d) w *= t.W[qi[d]];
      // physical gradient
      double gp[DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < DIM; ++b) s = fma(Ji[b][c], gu[b], s);
        gp[c] = s;
      }
      // physics in dual arithmetic, slots: 0..DIM-1 = d/d(grad u)_c, DIM = d/du
      QDual<NC> U = qd_seed<NC>(uq, DIM);
      QDual<NC> Gd[DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) Gd[c] = qd_seed<NC>(gp[c], c);
      if constexpr (TR) {
        // d(alpha_u U + beta_u)/dU
        U.d[DIM] = a.alpha_u;
#pragma unroll
        for (int c = 0; c < DIM; ++c) Gd[c].d[c] = a.alpha_u;
      }
      QDual<NC> flux[DIM], src;
      if constexpr (EX) {
        // thermal.cpp:80-92, 125-131: diff, source, cp, rho, advection from
        // the compiled functions; src = rho cp dT/dt - source + b . grad T
        ExprPoint<NC> pt;
        pt.T = U;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          pt.g[c] = c < DIM ? Gd[c] : xq_const<NC>(0.0);
          pt.x[c] = x[c];
        }
        pt.Tt = xq_const<NC>(0.0);
        if constexpr (TR) {
          pt.Tt.v = hq;
          pt.Tt.d[DIM] = a.alpha_t;
        }
        pt.time = a.xp.time;
        pt.k0 = k0;
        pt.src = a.coef.src ? a.coef.src[e] : 0.0;
        const QDual<NC> diff = expr_eval<NC>(a.xp, XF_DIFF, pt);
#pragma unroll
        for (int c = 0; c < DIM; ++c) flux[c] = xq_mul(diff, Gd[c]);
        const QDual<NC> fsrc = expr_eval<NC>(a.xp, XF_SRC, pt);
        if constexpr (TR) {
          const QDual<NC> rc = xq_mul(expr_eval<NC>(a.xp, XF_RHO, pt), expr_eval<NC>(a.xp, XF_CP, pt));
          src = xq_sub(xq_mul(rc, pt.Tt), fsrc);
        } else {
          src = xq_sub(xq_const<NC>(0.0), fsrc);
        }
        if (a.xp.len[XF_BX] + a.xp.len[XF_BY] + a.xp.len[XF_BZ] > 0) {
          QDual<NC> adv = xq_mul(expr_eval<NC>(a.xp, XF_BX, pt), Gd[0]);
          if constexpr (DIM > 1) adv = xq_add(adv, xq_mul(expr_eval<NC>(a.xp, XF_BY, pt), Gd[1]));
          if constexpr (DIM > 2) adv = xq_add(adv, xq_mul(expr_eval<NC>(a.xp, XF_BZ, pt), Gd[2]));
          src = xq_add(src, adv);
        }
      } else {
        const double f = thermal_source(a.coef, e, x, DIM);
        ThermalPhysics::volume<DIM>(k0, a.coef.diff_u2, f, U, Gd, flux, src);
        if constexpr (TR) {
          // + rho cp dT/dt, dT/dt interpolated from alpha_t U + beta_t: value and d/dU slot
          src.v += a.rho_cp * hq;
          src.d[DIM] += a.rho_cp * a.alpha_t;
        }
      }
      // Fhat = w T (flux, src);  Dhat = w T D T^T,  T = diag(J^-1, 1)
      double* F = &s_F[g][q * NC];
#pragma unroll
      for (int r = 0; r < DIM; ++r) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < DIM; ++c) s = fma(Ji[r][c], flux[c].v, s);
        F[r] = w * s;
      }
      F[DIM] = w * src.v;
      if constexpr (JAC) {
        // TD[r][c'] = sum_c T[r][c] D[c][c']  (only the NR rows the physics can make nonzero)
        double TD[NR][NC];
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) {
#pragma unroll
          for (int r = 0; r < DIM; ++r) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < DIM; ++c) s = fma(Ji[r][c], flux[c].d[cc], s);
            TD[r][cc] = s;
          }
          if constexpr (NR > DIM) TD[NR - 1][cc] = src.d[cc];
        }
        double* Dh = &s_W[g][q * NR * NC];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
#pragma unroll
          for (int s2 = 0; s2 < DIM; ++s2) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < DIM; ++c) s = fma(TD[r][c], Ji[s2][c], s);
            Dh[r * NC + s2] = w * s;
          }
          Dh[r * NC + DIM] = w * TD[r][DIM];
        }
      }
    }
  }
  __syncthreads();

  // 3. column stage ------------------------------------------------------------
  double R[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) R[i] = 0.0;
  const double* Dg = s_W[g];
  const double* Fg = s_F[g];
  if constexpr (JAC) {
    if (active && lane < NB) {
      const int jl[3] = {lane % P1, (lane / P1) % P1, DIM == 3 ? lane / (P1 * P1) : 0};
      // the lane's 1-D factors along the two inner axes; the outer-axis factor
      // is read per plane
      double bj[2][NQ1], dj[2][NQ1];
#pragma unroll
      for (int d = 0; d < 2; ++d)
#pragma unroll
        for (int q = 0; q < NQ1; ++q) {
          bj[d][q] = d < DIM - 1 ? t.B[jl[d]][q] : 0.0;
          dj[d][q] = d < DIM - 1 ? t.D[jl[d]][q] : 0.0;
        }
      test_apply<DIM, P, NQ1>(
          t,
          [&](int qx, int qy, int qz, double* f) {
            // outer axis (y in 2-D, z in 3-D) index and factor
            const int qo = DIM == 2 ? qy : qz;
            const double bo = t.B[jl[DIM - 1]][qo], dox = t.D[jl[DIM - 1]][qo];
            double gh[NC];
            if constexpr (DIM == 2) {
              gh[0] = dj[0][qx] * bo;
              gh[1] = bj[0][qx] * dox;
              gh[2] = bj[0][qx] * bo;
            } else {
              gh[0] = dj[0][qx] * bj[1][qy] * bo;
              gh[1] = bj[0][qx] * dj[1][qy] * bo;
              gh[2] = bj[0][qx] * bj[1][qy] * dox;
              gh[3] = bj[0][qx] * bj[1][qy] * bo;
            }
            const int q = DIM == 2 ? qx + NQ1 * qy : qx + NQ1 * (qy + NQ1 * qz);
            const double* Dh = Dg + q * NR * NC;
#pragma unroll
            for (int r = 0; r < NC; ++r) {
              if (r >= NR) {
                f[r] = 0.0;
                continue;
              }
              double s = 0.0;
#pragma unroll
              for (int c = 0; c < NC; ++c) s = fma(Dh[r * NC + c], gh[c], s);
              f[r] = s;
            }
          },
          R);
    } else if (active && lane == NB && a.res) {
      test_apply<DIM, P, NQ1>(
          t,
          [&](int qx, int qy, int qz, double* f) {
            const int q = DIM == 2 ? qx + NQ1 * qy : qx + NQ1 * (qy + NQ1 * qz);
#pragma unroll
            for (int r = 0; r < NC; ++r) f[r] = Fg[q * NC + r];
          },
          R);
    }
  } else {
    if (active && lane < NB) {
      const int il[3] = {lane % P1, (lane / P1) % P1, DIM == 3 ? lane / (P1 * P1) : 0};
      double s = 0.0;
      for (int q = 0; q < NQ; ++q) {
        const int qi[3] = {q % NQ1, (q / NQ1) % NQ1, DIM == 3 ? q / (NQ1 * NQ1) : 0};
        double b = 1.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) b *= t.B[il[d]][qi[d]];
        s = fma(Fg[q * NC + DIM], b, s);
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          double gg = 1.0;
#pragma unroll
          for (int d = 0; d < DIM; ++d) gg *= (d == c) ? t.D[il[d]][qi[d]] : t.B[il[d]][qi[d]];
          s = fma(Fg[q * NC + c], gg, s);
        }
      }
      R[0] = s;
    }
  }
  __syncthreads();  // all reads of s_u done before it is reused

  if constexpr (SEG) {
    // 4'. segment scatter: the segment's element columns are summed into an
    // LDS image of its rows (groups of even, then odd index, so the two
    // elements sharing a vertex plane never write at once), then the
    // workgroup writes the image out row run by row run.  An entry is shared
    // with another work unit only on the segment's end planes (x) or the
    // pencil's faces (y, z); the first scheduled unit stores it, later ones
    // add into it, as in element colouring.
    const int X0g = P * sx * GPB;  // global x-node of plane 0
    const int Xend = P * s_act;    // last plane holding this segment's entries
    const int yb = P * ec[1], zb = P * ec[2];
    if constexpr (JAC)
      for (int k = threadIdx.x; k < SG::kSize; k += C::BLOCK) s_seg[k] = 0.0;
    for (int k = threadIdx.x; k < SG::kRows; k += C::BLOCK) {
      s_sres[k] = 0.0;
      const int X = k / SG::RP, r = k % SG::RP;
      const int Xg = X0g + X;
      if (X <= Xend && Xg < a.m.nn[0]) {
        const long row = Xg + (long)a.m.nn[0] * ((yb + r % P1) + (long)a.m.nn[1] * (zb + (DIM == 3 ? r / P1 : 0)));
        s_srp[k] = a.fixed[row] ? -1 - a.rowptr[row] : a.rowptr[row];
      }
    }
    __syncthreads();
    for (int ph = 0; ph < 2; ++ph) {
      if (active && (g & 1) == ph) {
        if constexpr (JAC) {
          if (lane < NB) {
            const int jl[3] = {lane % P1, (lane / P1) % P1, DIM == 3 ? lane / (P1 * P1) : 0};
#pragma unroll
            for (int i = 0; i < NB; ++i) {
              const int ix = i % P1, iy = (i / P1) % P1, iz = DIM == 3 ? i / (P1 * P1) : 0;
              const bool vtx = ix == 0 || ix == P;
              const int W = vtx ? SG::WV : SG::WI;
              const int xo = vtx ? jl[0] - ix + P : jl[0];
              const int idx = SG::plane_off(P * g + ix) + ((iy + P1 * iz) * SG::CP + (jl[2] * P1 + jl[1])) * W + xo;
              s_seg[idx] += R[i];
            }
          } else if (lane == NB && a.res) {
#pragma unroll
            for (int i = 0; i < NB; ++i) {
              const int ix = i % P1, iy = (i / P1) % P1, iz = DIM == 3 ? i / (P1 * P1) : 0;
              s_sres[(P * g + ix) * SG::RP + iy + P1 * iz] += R[i];
            }
          }
        } else {
          if (lane < NB) {
            const int ix = lane % P1, iy = (lane / P1) % P1, iz = DIM == 3 ? lane / (P1 * P1) : 0;
            s_sres[(P * g + ix) * SG::RP + iy + P1 * iz] += R[0];
          }
        }
      }
      __syncthreads();
    }
    // residual rows
    if (a.res) {
      for (int k = threadIdx.x; k < SG::kRows; k += C::BLOCK) {
        const int X = k / SG::RP, r = k % SG::RP;
        const int Xg = X0g + X;
        if (X > Xend || Xg >= a.m.nn[0]) continue;
        const int iy = r % P1, iz = DIM == 3 ? r / P1 : 0;
        bool sh = true;
        if (X == 0) sh = sh && first_lo[0];
        if (X == Xend) sh = sh && first_hi[0];
        if (iy == 0) sh = sh && first_lo[1];
        if (iy == P) sh = sh && first_hi[1];
        if (DIM == 3 && iz == 0) sh = sh && first_lo[2];
        if (DIM == 3 && iz == P) sh = sh && first_hi[2];
        const long row = Xg + (long)a.m.nn[0] * ((yb + iy) + (long)a.m.nn[1] * (zb + iz));
        const bool fx = s_srp[k] < 0;
        const double v = -s_sres[k];
        if (a.first_touch && sh) a.res[row] = fx ? 0.0 : v;
        else if (!fx) a.res[row] += v;
      }
    }
    // Jacobian rows, CH entries per thread per pass: slots and flags, then
    // the loads of entries that earlier units already hold, then the stores
    if constexpr (JAC) {
      if (a.jac) {
        constexpr int CH = 8;
        for (int base = threadIdx.x; base < SG::kSize; base += C::BLOCK * CH) {
          int sl[CH], mode[CH];
          double val[CH], old[CH];
#pragma unroll
          for (int c = 0; c < CH; ++c) {
            mode[c] = 0;
            sl[c] = 0;
            val[c] = 0.0;
            const int eix = base + c * C::BLOCK;
            if (eix >= SG::kSize) continue;
            // decode (plane, row, column block, x offset); every divisor is a
            // compile-time constant (vertex and interior planes separately)
            const int grp = eix / SG::GS, rem = eix % SG::GS;
            int X, r, cb, xo;
            if (rem < SG::SV) {
              X = P * grp;
              r = rem / (SG::CP * SG::WV);
              cb = (rem / SG::WV) % SG::CP;
              xo = rem % SG::WV;
            } else {
              const int r2 = rem - SG::SV;
              X = P * grp + 1 + r2 / SG::SI;
              const int rr = r2 % SG::SI;
              r = rr / (SG::CP * SG::WI);
              cb = (rr / SG::WI) % SG::CP;
              xo = rr % SG::WI;
            }
            const int Xg = X0g + X;
            if (X > Xend || Xg >= a.m.nn[0]) continue;
            const int xcol = (X % P == 0) ? Xg - P + xo : Xg - X % P + xo;
            if (xcol < 0 || xcol >= a.m.nn[0]) continue;
            if ((X == 0 && xcol < Xg) || (X == Xend && xcol > Xg)) continue;  // a neighbouring segment's entries
            const int iy = r % P1, iz = DIM == 3 ? r / P1 : 0;
            const int jy = cb % P1, jz = DIM == 3 ? cb / P1 : 0;
            int lox, lx, loy, ly, loz = 0, lz = 1;
            row_range(Xg, P, a.m.nn[0], lox, lx);
            row_range(yb + iy, P, a.m.nn[1], loy, ly);
            if (DIM == 3) row_range(zb + iz, P, a.m.nn[2], loz, lz);
            int rp = s_srp[X * SG::RP + r];
            const bool fx = rp < 0;
            if (fx) rp = -1 - rp;
            sl[c] = rp + (((zb + jz - loz) * ly + (yb + jy - loy)) * lx + (xcol - lox));
            bool sh = true;
            if (X == 0 && xcol == Xg) sh = sh && first_lo[0];
            if (X == Xend && xcol == Xg) sh = sh && first_hi[0];
            if (iy == jy && iy == 0) sh = sh && first_lo[1];
            if (iy == jy && iy == P) sh = sh && first_hi[1];
            if (DIM == 3 && iz == jz && iz == 0) sh = sh && first_lo[2];
            if (DIM == 3 && iz == jz && iz == P) sh = sh && first_hi[2];
            const bool first = a.first_touch && sh;
            if (fx) {
              if (first) {
                mode[c] = 1;
                val[c] = (xcol == Xg && jy == iy && jz == iz) ? 1.0 : 0.0;
              }
            } else {
              mode[c] = first ? 1 : 2;
              val[c] = s_seg[eix];
            }
          }
#pragma unroll
          for (int c = 0; c < CH; ++c) old[c] = mode[c] == 2 ? a.jac[sl[c]] : 0.0;
#pragma unroll
          for (int c = 0; c < CH; ++c)
            if (mode[c]) a.jac[sl[c]] = val[c] + old[c];
        }
      }
    }
  } else {
  // 4. scatter -------------------------------------------------------------------
  // Lane i fetches row i's CRS offset once (negated minus one for a
  // strong-Dirichlet row); the value lane hands the residual to the row lanes
  // through s_u.
  if constexpr (JAC) {
    if (active && lane == NB && a.res) {
#pragma unroll
      for (int i = 0; i < NB; ++i) s_u[g][i] = R[i];
    }
  } else {
    if (active && lane < NB) s_u[g][lane] = R[0];
  }
  if (active && lane < NB) {
    const long row = node_of<DIM, P>(a.m, ec, lane);
    s_rp[g][lane] = a.fixed[row] ? -1 - a.rowptr[row] : a.rowptr[row];
  }
  __syncthreads();
  // An entry (row node r, column node c) is shared by the elements holding
  // both nodes: along each axis one element, or two when r and c sit on the
  // same vertex plane.  With a.first_touch the first scheduled of them writes
  // the entry (0 / 1 on strong-Dirichlet rows) and later ones add into it, so
  // no zero-fill pass and no separate dofConstraints pass are needed.
  // Otherwise every contribution is added and Dirichlet rows are skipped
  // (the caller zeroes; assemblyManager_scatter.hpp:208,253 semantics).
  auto shared_first = [&](const int* il, const int* jl) {
    bool f = true;
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      if (il[d] == jl[d] && (il[d] == 0 || il[d] == P)) f = f && (il[d] == 0 ? first_lo[d] : first_hi[d]);
    return f;
  };
  if (active && lane < NB && a.res) {
    const long row = node_of<DIM, P>(a.m, ec, lane);
    const int il[3] = {lane % P1, (lane / P1) % P1, DIM == 3 ? lane / (P1 * P1) : 0};
    const bool fx = s_rp[g][lane] < 0;
    const double v = -s_u[g][lane];
    if (a.first_touch && shared_first(il, il)) a.res[row] = fx ? 0.0 : v;
    else if (!fx) a.res[row] += v;
  }
  if constexpr (JAC) {
    if (active && lane < NB && a.jac) {
      // lane j holds column j; rows are unrolled so the per-axis slot logic
      // folds at compile time, and for a fixed row the group's lanes write the
      // row's runs of consecutive slots.  Per chunk of rows: slots and flags
      // first, the loads that are needed (later touches of non-Dirichlet
      // rows) issued together, then the stores; the next chunk's loads are
      // issued before this chunk's stores (double buffer).
      const int jl[3] = {lane % P1, (lane / P1) % P1, DIM == 3 ? lane / (P1 * P1) : 0};
      constexpr int CH = NB < 9 ? NB : 9;
      constexpr int NCH = (NB + CH - 1) / CH;
      // slots and flags of chunk c, and the loads of the entries it must add to
      auto prep = [&](int i0, int* sl, uint32_t& first_mask, uint32_t& fixed_mask, double* old) {
        first_mask = 0;
        fixed_mask = 0;
#pragma unroll
        for (int k = 0; k < CH; ++k) {
          const int i = i0 + k;
          if (i >= NB) break;
          const int il[3] = {i % P1, (i / P1) % P1, DIM == 3 ? i / (P1 * P1) : 0};
          int px, lx, py, ly, pz = 0, lz = 1;
          axis_slot<P>(il[0], jl[0], ec[0], a.m.nn[0], px, lx);
          axis_slot<P>(il[1], jl[1], ec[1], a.m.nn[1], py, ly);
          if (DIM == 3) axis_slot<P>(il[2], jl[2], ec[2], a.m.nn[2], pz, lz);
          int rp = s_rp[g][i];
          const bool fx = rp < 0;
          if (fx) rp = -1 - rp;
          sl[k] = rp + (pz * ly + py) * lx + px;
          const bool first = a.first_touch && shared_first(il, jl);
          if (first) first_mask |= 1u << k;
          if (fx) fixed_mask |= 1u << k;
          old[k] = (!first && !fx) ? a.jac[sl[k]] : 0.0;
        }
      };
      // double buffer: the loads of chunk c+1 are issued before the stores of
      // chunk c, so their latency overlaps
      int slA[CH], slB[CH];
      uint32_t fmA, xmA, fmB, xmB;
      double oldA[CH], oldB[CH];
      prep(0, slA, fmA, xmA, oldA);
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        int* sl = (c & 1) ? slB : slA;
        double* old = (c & 1) ? oldB : oldA;
        const uint32_t first_mask = (c & 1) ? fmB : fmA, fixed_mask = (c & 1) ? xmB : xmA;
        if (c + 1 < NCH) {
          if (c & 1) prep((c + 1) * CH, slA, fmA, xmA, oldA);
          else prep((c + 1) * CH, slB, fmB, xmB, oldB);
        }
#pragma unroll
        for (int k = 0; k < CH; ++k) {
          const int i = c * CH + k;
          if (i >= NB) break;
          const bool first = (first_mask >> k) & 1u, fx = (fixed_mask >> k) & 1u;
          // streaming stores: a later colour re-reads an entry long after
          // this launch, so keeping the line in L2 buys nothing
          if (fx) {
            if (first) __builtin_nontemporal_store((i == lane) ? 1.0 : 0.0, &a.jac[sl[k]]);
          } else {
            __builtin_nontemporal_store(old[k] + R[i], &a.jac[sl[k]]);
          }
        }
      }
    }
  }
  }  // element / block colouring scatter
  }  // sub-phase loop
}

// number of workgroups of colour `col` under element colouring (one group of
// G lanes per element, GPB groups per workgroup); sets a.color, a.cn, a.ncolor
template <int DIM, int P, int NQ1>
inline long element_colour_grid(const Ctx& c, AsmArgs& a, int col) {
  using C = Cfg<DIM, P, NQ1>;
  a.color = col;
  long cnt = 1;
  for (int d = 0; d < 3; ++d) {
    const int cd = (col >> d) & 1;
    int n = d < DIM ? (c.mesh.nel[d] - cd + 1) / 2 : 1;
    if (d >= DIM && cd) n = 0;
    a.cn[d] = n;
    cnt *= n;
  }
  a.ncolor = cnt;
  return (cnt + C::GPB - 1) / C::GPB;
}

}  // namespace mhx
