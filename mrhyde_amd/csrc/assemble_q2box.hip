// assemble_q2box.hip — nonlinear thermal residual/Jacobian assembly of Q2
// hexahedra with 3x3x3 Gauss points on meshes whose elements are axis-aligned
// boxes (the structured box without node perturbation), CDNA4 (gfx950).
//
// Replaces the same reference steps as k_thermal, for that mesh class:
//   AssemblyManager::performGather          assemblyManager_gather.hpp:183-234
//   Workset::computeSolnSteadySeeded        workset.cpp:865-905
//   Workset::evaluateSolutionField          workset.cpp:979-1115
//   thermal<EvalT>::volumeResidual          thermal.cpp:71-165
//   AssemblyManager::scatter                assemblyManager_scatter.hpp:164-278
//
// Work unit.  One 64-lane wavefront (one workgroup) owns an x-segment of one
// element pencil (fixed ey, ez) and walks it two elements per step: lanes
// 32h .. 32h+31 take element ex = xa + h of the step starting at xa.
//
// Element matrix by pair sum factorisation.  On a box element J = diag(J_c)
// and the thermal flux is isotropic (dflux_c / d(grad u)_c' = 0 for c != c'),
// so with reference derivatives
//   A_ij = sum_q sum_c [ alpha_c(q) dphi_i/dxi_c dphi_j/dxi_c
//                      + beta_c(q)  dphi_i/dxi_c phi_j ]
//   alpha_c = w|J| dflux_c/d(grad u)_c / J_c^2,   beta_c = w|J| dflux_c/du / J_c
// (the point physics in forward-mode dual arithmetic gives both slots).  Each
// term is a product over the axes of 1-D pair tables
//   M_t[(i_d, j_d)][q_d] = l_a(i_d, q_d) l_b(j_d, q_d),  l_0 = value, l_1 = slope
// with types t = (a, b): 0 = (0,0), 1 = (1,0), 3 = (1,1), so
//   A[(i1,j1),(i2,j2),(i3,j3)] = sum_q1 M_t1 sum_q2 M_t2 sum_q3 M_t3 field(q)
// follows from three contractions (q3, then q2, then q1) of the six point
// fields; terms are merged as soon as their remaining types agree:
//   after q3: groups (t1,t2) = (3,0) {alpha_x}, (1,0) {beta_x}, (0,3) {alpha_y},
//             (0,1) {beta_y}, (0,0) {alpha_z (t3 = 3), beta_z (t3 = 1)}
//   after q2: groups t1 = 3, 1, 0
// 11.7k FMA per element (27 sum-factorised columns: ~33k).  The products
// after q2 (3 q1 slices x 3 groups x 81) overwrite the point fields in LDS,
// and the q1 contraction runs per row task into register accumulators.
//
// Scatter.  Each lane's task in the last stage is one row (i1, i2, i3) and the
// x-run of three columns (j1 = 0..2) at (j2, j3).  Half 0 takes the rows in
// the mirrored x order of half 1 (i1 = 2 - m against m), so the lane holding
// element A's rows on the x-plane it shares with element B holds B's rows on
// that plane in the other half; the shared column is summed across the
// halves (shuffle) and the lane pair writes the row's x-run of five once.  The
// plane shared with the next step is carried in half 1's registers.  Entries
// shared with a neighbouring pencil (row and column on the same y or z vertex
// plane) or segment (the segment's end planes) are written by the first
// scheduled unit and added by later ones: 8 launches, one per colour of
// (segment parity, ey parity, ez parity), as in element colouring.  Strong
// Dirichlet rows are written as e_d by their first toucher
// (assemblyManager_scatter.hpp:208,253; constraints.hpp:125-141).
// Global residual = -F (scatter.hpp:227); Jacobian = +dF/du (:261).
//
// Memory order.  A later touch loads, adds and stores.  Vector-memory
// counters retire in issue order, so a load issued after a store waits for
// that store: the scatter issues all nine rounds' loads of a step before its
// first store, and the wavefront waits for memory about once per step.  The
// runs of one step are disjoint entries, and no entry is written in one step
// and read in the next (the carry), so the early loads see final values.
// The next step's gather is prefetched before the scatter.  Per-workgroup
// LDS (19 KB) and registers (<= 256) give 2 wavefronts per SIMD.

#include <cmath>
#include <type_traits>

#include "lattice.hpp"
#include "mhx_internal.hpp"
#include "thermal.hpp"

namespace mhx {

namespace {

constexpr const Tab1& kT = T1<2, 3>::v;

__device__ __forceinline__ double sel3(int i, double a0, double a1, double a2) {
  return i == 0 ? a0 : (i == 1 ? a1 : a2);
}

// LDS copy of the 1-D tables for reads at lane-dependent indices: value and
// slope of basis i at point q at [i * 3 + q] and [9 + i * 3 + q], points at
// [18 + q], weights at [21 + q].  (Selecting from the constants per lane
// would be hoisted out of the step loop and hold ~70 registers for the whole
// kernel; LDS reads behind the loop's barriers are re-issued instead.)
constexpr int kTabN = 24;

// Ordering of LDS traffic between the lanes of the kernel's single-wavefront
// workgroups: a workgroup-scope release / acquire pair restricted to LDS
// around the barrier.  Fences at wavefront scope did not keep one lane's LDS
// store ahead of another lane's load in every build; __syncthreads() orders
// global memory too, i.e. waits for every outstanding store of the scatter at
// each of the step's stages.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The first-touch rule of every entry: the first scheduled unit among those
// sharing the entry stores it (0/1 on a strong-Dirichlet row), later ones
// load, add and store; accumulate mode adds everywhere except on Dirichlet
// rows.  The first scheduled unit runs in an earlier launch than every other,
// and within one launch an entry has one writer, so the order of additions is
// the launch order and the sums are reproducible.

// column range of Q2 row coordinate r along an axis of n nodes: first column
// lo and count len of the row's lattice box
struct Range {
  int lo, len;
};
__device__ __forceinline__ Range q2_axis(int r, int n) {
  if (r & 1) return {r - 1, 3};
  const int lo = r - 2 < 0 ? 0 : r - 2;
  const int hi = r + 2 > n - 1 ? n - 1 : r + 2;
  return {lo, hi - lo + 1};
}

}  // namespace

// MODES: the source has sine modes (their double-precision sin sizes the
// register file of the whole kernel, so the modeless source has its own build)
template <bool JAC, bool MODES>
__global__ __launch_bounds__(64) void k_q2box(AsmArgs args) {
  // the arguments the kernel reads, as scalars: member accesses of the
  // by-value argument inside the lambdas below made the compiler copy the
  // whole structure to per-lane scratch memory and read it back from there
  const double* __restrict__ ga_u = args.u;
  const double* __restrict__ ga_vx = args.vx;
  const int32_t* __restrict__ ga_rowptr = args.rowptr;
  const uint8_t* __restrict__ ga_fixed = args.fixed;
  const double* __restrict__ ga_k0 = args.coef.k0;
  const double* __restrict__ ga_src = args.coef.src;
  double* const ga_res = args.res;
  double* const ga_jac = args.jac;
  const double ga_diff_u2 = args.coef.diff_u2;
  const int ga_n_modes = args.coef.n_modes;
  const int ga_first_touch = args.first_touch;
  const int nel0 = args.m.nel[0], nel1 = args.m.nel[1], nel2 = args.m.nel[2];
  const int nvx0 = args.m.nvx[0], nvx1 = args.m.nvx[1];
  const int ga_color = args.color, ga_cn0 = args.cn[0], ga_cn1 = args.cn[1], ga_seg = args.gpb, ga_nseg = args.nseg;
  __shared__ double s_u[45];         // nodal values: x 0..4, y 0..2, z 0..2 of the step
  __shared__ double s_v[4][3];       // corners: A low, A high, B low, B high
  // per half: the point fields (alpha_x, alpha_y, alpha_z, beta_x, beta_y,
  // beta_z at [0, 162); residual fields w|J| flux_c / J_c, w|J| src at
  // [162, 270)), later overwritten by the products after q2:
  // [q1][group][i2 * 3 + j2][i3 * 3 + j3]
  __shared__ double s_a[2][729];
  __shared__ double s_s1[2][5][81];  // after q3: [group][q1 + 3 q2][i3 * 3 + j3]
  __shared__ double s_t[kTabN];      // 1-D tables (see kTabN)
  __shared__ int s_rp[45];           // CRS offsets of the step's rows (-1 - offset: strong Dirichlet row)
  __shared__ double s_ek[4];         // k0 of A, B; element source of A, B
  // sine-mode sources: factor sin(w_md x_d(q_d)) (1 where w_md = 0) per
  // element, mode, axis and 1-D point
  __shared__ double s_sin[MODES ? 2 : 1][MODES ? MHX_MAX_MODES : 1][3][3];
  // the sine modes (amplitude, frequencies), copied from the kernel argument
  // with compile-time indices: indexing the by-value argument with a run-time
  // mode number changed the code generated for the rest of the kernel (wrong
  // Jacobians from the second step of a segment on, even with no modes)
  __shared__ double s_mode[MODES ? MHX_MAX_MODES : 1][4];

  const int lane = threadIdx.x, h = lane >> 5, l = lane & 31;
  if (lane < 9) {
    s_t[lane] = sel3(lane / 3, kT.B[0][lane % 3], kT.B[1][lane % 3], kT.B[2][lane % 3]);
    s_t[9 + lane] = sel3(lane / 3, kT.D[0][lane % 3], kT.D[1][lane % 3], kT.D[2][lane % 3]);
  } else if (lane < 12) {
    s_t[18 + lane - 9] = sel3(lane - 9, kT.X[0], kT.X[1], kT.X[2]);
    s_t[21 + lane - 9] = sel3(lane - 9, kT.W[0], kT.W[1], kT.W[2]);
  }
  if constexpr (MODES) {
#pragma unroll
    for (int m = 0; m < MHX_MAX_MODES; ++m)
      if (lane == m) {
        s_mode[m][0] = args.coef.amp[m];
#pragma unroll
        for (int d = 0; d < 3; ++d) s_mode[m][1 + d] = args.coef.freq[m][d];
      }
  }
  auto Bq = [&](int i, int q) __attribute__((always_inline)) { return s_t[i * 3 + q]; };
  auto Dq = [&](int i, int q) __attribute__((always_inline)) { return s_t[9 + i * 3 + q]; };
  auto Xq = [&](int q) __attribute__((always_inline)) { return s_t[18 + q]; };
  auto Wq = [&](int q) __attribute__((always_inline)) { return s_t[21 + q]; };
  int t = blockIdx.x;
  const int sx = 2 * (t % ga_cn0) + (ga_color & 1);
  t /= ga_cn0;
  const int ey = 2 * (t % ga_cn1) + ((ga_color >> 1) & 1);
  t /= ga_cn1;
  const int ez = 2 * t + ((ga_color >> 2) & 1);
  const int L = ga_seg;
  const int x0 = sx * L;
  const int x1 = min(x0 + L, nel0);
  const int nnx = (2 * nel0 + 1), nny = (2 * nel1 + 1), nnzn = (2 * nel2 + 1);
  // first scheduled unit along each axis among the (at most two) sharing a
  // vertex plane: even units run first (colour bit 0)
  const bool flx = sx == 0 || (sx & 1) == 0, fhx = sx == ga_nseg - 1 || (sx & 1) == 0;
  const bool fly = ey == 0 || (ey & 1) == 0, fhy = ey == nel1 - 1 || (ey & 1) == 0;
  const bool flz = ez == 0 || (ez & 1) == 0, fhz = ez == nel2 - 1 || (ez & 1) == 0;
  const int Y0 = 2 * ey, Z0 = 2 * ez;

  // (y, z) column ranges of the pencil's rows (i2, i3 = 0, 1, 2) and the
  // first-writer flags of their shared (y, z) entries
  const Range ry0 = q2_axis(Y0, nny), ry1 = q2_axis(Y0 + 1, nny), ry2 = q2_axis(Y0 + 2, nny);
  const Range rz0 = q2_axis(Z0, nnzn), rz1 = q2_axis(Z0 + 1, nnzn), rz2 = q2_axis(Z0 + 2, nnzn);

  // gather of the step starting at xs into registers: lane < 45 the value
  // and the CRS offset of node (2 xs + lane % 5, Y0 + lane / 5 % 3, Z0 + lane
  // / 15), lanes 48..59 the corner coordinates.  The next step's loads are
  // issued before this step's scatter stores, so their latency overlaps.
  // Every lane issues its loads on every step, from an address that is
  // always valid (roles without data of their own read a neighbour's), so
  // the prefetched registers are defined on all lanes: no value lives across
  // the step under a lane-divergent definition.
  double g_val = 0.0;
  int g_rp = 0;
  auto fetch = [&](int xs) __attribute__((always_inline)) {
    const bool hb = xs + 1 < x1;
    // nodal value and CRS offset (lanes < 45; others repeat node 0)
    const int li = lane < 45 ? lane : 0;
    int X = li % 5;
    if (!hb && X > 2) X = 2;
    const long node = (2 * xs + X) + (long)nnx * ((Y0 + (li / 5) % 3) + (long)nny * (Z0 + li / 15));
    const int rp = ga_rowptr[node];
    const bool fx = ga_fixed[node] != 0;
    // corner coordinate (lanes 48..59), element coefficient (lanes 60..63)
    const int lv = lane >= 48 && lane < 60 ? lane - 48 : 0;
    int v = lv / 3;
    if (!hb && v >= 2) v -= 2;
    const int hi = v & 1;
    const long vid = (long)(xs + (v >> 1) + hi) +
                     (long)nvx0 * ((long)(ey + hi) + (long)nvx1 * (long)(ez + hi));
    const long ee = (xs + (hb ? (lane & 1) : 0)) + (long)nel0 * (ey + (long)nel1 * ez);
    const double* cf = lane < 62 ? ga_k0 : ga_src;
    const double dflt = lane < 62 ? 1.0 : 0.0;
    const double* ptr = lane < 48 ? ga_u + node : (lane < 60 ? ga_vx + (vid * 3 + lv % 3) : (cf ? cf + ee : ga_u));
    const double val = *ptr;
    g_val = (lane >= 60 && !cf) ? dflt : val;
    g_rp = fx ? -1 - rp : rp;
  };
  fetch(x0);

  // the step-to-step carry of B's last-plane rows (their Jacobian runs of the
  // three last rounds, their residuals), held by the half-1 lanes in
  // registers and read by the half-0 lanes of the next step through
  // cross-half lane exchanges
  double carry_j[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  double carry_r = 0.0;
  for (int xa = x0; xa < x1; xa += 2) {
    const bool hasB = xa + 1 < x1;
    const bool act = h == 0 || hasB;
    const int ex = xa + h;
    const long e = ex + (long)nel0 * (ey + (long)nel1 * ez);
    const bool first_step = xa == x0, last_step = xa + 2 >= x1;

    // 1. gather ---------------------------------------------------------------
    wave_sync();  // the previous step's LDS reads are done
    if (lane < 45) {
      s_u[lane] = g_val;
      s_rp[lane] = g_rp;
    } else if (lane >= 48 && lane < 60) {
      s_v[(lane - 48) / 3][(lane - 48) % 3] = g_val;
    } else if (lane >= 60) {
      s_ek[lane - 60] = g_val;
    }
    wave_sync();
    // the LDS stores above read the registers the next gather loads into:
    // they complete before those loads are issued
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    if (xa + 2 < x1) fetch(xa + 2);

    // read-backs.  The scatters add to what is stored (later touches of rows
    // that are not strong Dirichlet rows).  The runs of one step are entries
    // no other run of the wavefront touches (shared-plane sums go through
    // lane exchanges and the step-to-step carry), so a batch of loads can go
    // before the stores; a load issued after a store waits for that store,
    // so each batch is issued as early as its compute allows.
    //
    // residual: lane (l < 27) = row (i1, i2, i3) of its half's element
    const int il[3] = {l % 3, (l / 3) % 3, l / 9};
    bool r_write = false, r_first = true, r_fixed = false;
    long r_row = 0;
    double r_old = 0.0;
    if (ga_res && act && l < 27) {
      const int i1 = il[0];
      const int X = 2 * ex + i1, Y = Y0 + il[1], Z = Z0 + il[2];
      bool xsh = false, xfirst = true;
      r_write = true;
      if (h == 0 && i1 == 2) {
        if (!hasB && x1 < nel0) xsh = true, xfirst = fhx;
      } else if (h == 0 && i1 == 0) {
        if (first_step && x0 > 0) xsh = true, xfirst = flx;
      } else if (h == 1 && i1 == 0) {
        r_write = false;  // summed into A's row
      } else if (h == 1 && i1 == 2) {
        r_write = last_step;  // otherwise carried to the next step
        if (last_step && x1 < nel0) xsh = true, xfirst = fhx;
      }
      r_row = X + (long)nnx * (Y + (long)nny * Z);
      const bool ysh = il[1] != 1, zsh = il[2] != 1;
      r_first = ga_first_touch && (!xsh || xfirst) && (!ysh || (il[1] == 0 ? fly : fhy)) &&
                (!zsh || (il[2] == 0 ? flz : fhz));
      r_fixed = s_rp[(X - 2 * xa) + 5 * il[1] + 15 * il[2]] < 0;
      if (r_write && !r_fixed && !r_first) r_old = ga_res[r_row];
    }
    // 2. point stage (lane = cubature point) ------------------------------------
    double Jc[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) Jc[d] = 0.5 * (s_v[2 * h + 1][d] - s_v[2 * h][d]);
    const double detJ = fabs(Jc[0] * Jc[1] * Jc[2]);
    // 1 / J_c once per element: the point stage multiplies (one division
    // per axis instead of four per axis and point)
    double iJ[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) iJ[d] = 1.0 / Jc[d];
    if constexpr (MODES) {
      // the sine factors of the source, in a stage of their own: the
      // double-precision sin needs many registers, and here few others are live
      const int nf = ga_n_modes * 9;
      for (int t = l; t < nf; t += 32) {
        const int m = t / 9, d = (t / 3) % 3, q = t % 3;
        const double w = s_mode[m][1 + d];
        const double xd = fma(Xq(q) + 1.0, Jc[d], s_v[2 * h][d]);
        s_sin[h][m][d][q] = w != 0.0 ? sin(w * xd) : 1.0;
      }
      wave_sync();
    }
    if (act && l < 27) {
      const int qv[3] = {l % 3, (l / 3) % 3, l / 9};
      double b[3][3], dd[3][3];  // [axis][j] = l_0(j, q_axis), l_1(j, q_axis)
#pragma unroll
      for (int ax = 0; ax < 3; ++ax)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          b[ax][j] = Bq(j, qv[ax]);
          dd[ax][j] = Dq(j, qv[ax]);
        }
      // u and its reference gradient by sum factorisation over the nodes
      double U = 0.0, G[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int j3 = 0; j3 < 3; ++j3) {
        double cu = 0.0, cx = 0.0, cy = 0.0;
#pragma unroll
        for (int j2 = 0; j2 < 3; ++j2) {
          double au = 0.0, axx = 0.0;
#pragma unroll
          for (int j1 = 0; j1 < 3; ++j1) {
            const double uj = s_u[2 * h + j1 + 5 * j2 + 15 * j3];
            au = fma(b[0][j1], uj, au);
            axx = fma(dd[0][j1], uj, axx);
          }
          cu = fma(b[1][j2], au, cu);
          cx = fma(b[1][j2], axx, cx);
          cy = fma(dd[1][j2], au, cy);
        }
        U = fma(b[2][j3], cu, U);
        G[0] = fma(b[2][j3], cx, G[0]);
        G[1] = fma(b[2][j3], cy, G[1]);
        G[2] = fma(dd[2][j3], cu, G[2]);
      }
      const double w = detJ * Wq(qv[0]) * Wq(qv[1]) * Wq(qv[2]);
      const double k0 = s_ek[h];
      // thermal source: element value plus the sine modes (thermal_source in
      // thermal.hpp, the same products in the same order)
      double f = s_ek[2 + h];
      if constexpr (MODES) {
        for (int m = 0; m < ga_n_modes; ++m) {
          double tm = s_mode[m][0];
#pragma unroll
          for (int d = 0; d < 3; ++d)
            if (s_mode[m][1 + d] != 0.0) tm = tm * s_sin[h][m][d][qv[d]];
          f += tm;
        }
      }
      // physics in dual arithmetic, slots 0..2 = d/d(grad u)_c, 3 = d/du
      QDual<4> Ud = qd_seed<4>(U, 3);
      QDual<4> Gd[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) Gd[c] = qd_seed<4>(G[c] * iJ[c], c);
      QDual<4> flux[3], src;
      ThermalPhysics::volume<3>(k0, ga_diff_u2, f, Ud, Gd, flux, src);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        s_a[h][162 + 27 * c + l] = w * flux[c].v * iJ[c];
        if constexpr (JAC) {
          s_a[h][27 * c + l] = w * flux[c].d[c] * (iJ[c] * iJ[c]);
          s_a[h][27 * (3 + c) + l] = w * flux[c].d[3] * iJ[c];
        }
      }
      s_a[h][162 + 81 + l] = w * src.v;
    }
    wave_sync();

    // 3. residual (lane = test function i) and the first contraction ------------
    double R = 0.0;
    if (act && l < 27 && ga_res) {
      double bi[3][3], di[3][3];  // [axis][q]
#pragma unroll
      for (int ax = 0; ax < 3; ++ax)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          bi[ax][q] = Bq(il[ax], q);
          di[ax][q] = Dq(il[ax], q);
        }
#pragma unroll
      for (int q3 = 0; q3 < 3; ++q3) {
        double tv = 0.0, tz = 0.0;
#pragma unroll
        for (int q2 = 0; q2 < 3; ++q2) {
          double svx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
          for (int q1 = 0; q1 < 3; ++q1) {
            const int q = q1 + 3 * q2 + 9 * q3;
            const double* fr = &s_a[h][162];
            svx = fma(bi[0][q1], fr[81 + q], fma(di[0][q1], fr[q], svx));
            sy = fma(bi[0][q1], fr[27 + q], sy);
            sz = fma(bi[0][q1], fr[54 + q], sz);
          }
          tv = fma(bi[1][q2], svx, fma(di[1][q2], sy, tv));
          tz = fma(bi[1][q2], sz, tz);
        }
        R = fma(bi[2][q3], tv, fma(di[2][q3], tz, R));
      }
    }
    if constexpr (JAC) {
      // q3 contraction: task (q1 + 3 q2, i3), outputs j3 = 0..2 of the five groups
      if (act && l < 27) {
        const int qq = l % 9, i3 = l / 9;
        double vax[3], vbx[3], vay[3], vby[3], vaz[3], vbz[3];
#pragma unroll
        for (int q3 = 0; q3 < 3; ++q3) {
          const double bb = Bq(i3, q3), db = Dq(i3, q3);
          const int q = qq + 9 * q3;
          vax[q3] = bb * s_a[h][q];
          vay[q3] = bb * s_a[h][27 + q];
          vaz[q3] = db * s_a[h][54 + q];
          vbx[q3] = bb * s_a[h][81 + q];
          vby[q3] = bb * s_a[h][108 + q];
          vbz[q3] = db * s_a[h][135 + q];
        }
#pragma unroll
        for (int j3 = 0; j3 < 3; ++j3) {
          double g30 = 0.0, g10 = 0.0, g03 = 0.0, g01 = 0.0, g00 = 0.0;
#pragma unroll
          for (int q3 = 0; q3 < 3; ++q3) {
            const double bj = kT.B[j3][q3], dj = kT.D[j3][q3];
            g30 = fma(bj, vax[q3], g30);
            g10 = fma(bj, vbx[q3], g10);
            g03 = fma(bj, vay[q3], g03);
            g01 = fma(bj, vby[q3], g01);
            g00 = fma(dj, vaz[q3], fma(bj, vbz[q3], g00));
          }
          const int p3 = 3 * i3 + j3;
          s_s1[h][0][qq * 9 + p3] = g30;
          s_s1[h][1][qq * 9 + p3] = g10;
          s_s1[h][2][qq * 9 + p3] = g03;
          s_s1[h][3][qq * 9 + p3] = g01;
          s_s1[h][4][qq * 9 + p3] = g00;
        }
      }
    }

    // residual scatter: A's x-plane 2 + B's x-plane 0 summed; A's x-plane 0
    // takes the carry of the previous step's B; B's x-plane 2 is carried
    if (ga_res) {
      const int i1 = il[0];
      // lanes of the partner rows in the other half (same i2, i3)
      const int srcA2 = 32 + l - 2;  // half 0, i1 == 2: B's row with i1 == 0
      const int srcA0 = 32 + l + 2;  // half 0, i1 == 0: the carried row of B with i1 == 2
      const double fromB = __shfl(R, (h == 0 && l < 27 && i1 == 2) ? srcA2 : lane, 64);
      const double cprev = __shfl(carry_r, (h == 0 && l < 27 && i1 == 0) ? srcA0 : lane, 64);
      carry_r = (h == 1 && l < 27 && i1 == 2 && act) ? R : carry_r;
      if (act && l < 27) {
        double v = R;
        if (h == 0 && i1 == 2 && hasB) v = R + fromB;
        else if (h == 0 && i1 == 0 && !first_step) v = cprev + R;
        // strong Dirichlet row: 0, written by its first toucher
        if (r_write && (!r_fixed || r_first)) ga_res[r_row] = r_fixed ? 0.0 : (r_first ? -v : r_old - v);
      }
    }
    if constexpr (!JAC) continue;

    // 4. q2 contraction for the three q1 slices ---------------------------------
    wave_sync();  // s_s1 complete; the point and residual fields are read
    if (act && l < 27) {
      // task (p3, i2), outputs j2 = 0..2 of the three groups
      const int p3 = l % 9, i2 = l / 9;
      const double bb[3] = {Bq(i2, 0), Bq(i2, 1), Bq(i2, 2)}, db[3] = {Dq(i2, 0), Dq(i2, 1), Dq(i2, 2)};
#pragma unroll 1
      for (int q1 = 0; q1 < 3; ++q1) {
        double v30[3], v10[3], v03[3], v01[3];
#pragma unroll
        for (int q2 = 0; q2 < 3; ++q2) {
          const int o = (q1 + 3 * q2) * 9 + p3;
          v30[q2] = bb[q2] * s_s1[h][0][o];
          v10[q2] = bb[q2] * s_s1[h][1][o];
          v03[q2] = db[q2] * s_s1[h][2][o];
          v01[q2] = fma(db[q2], s_s1[h][3][o], bb[q2] * s_s1[h][4][o]);
        }
        double* s2 = &s_a[h][q1 * 243];
#pragma unroll
        for (int j2 = 0; j2 < 3; ++j2) {
          double c3 = 0.0, c1 = 0.0, c0 = 0.0;
#pragma unroll
          for (int q2 = 0; q2 < 3; ++q2) {
            const double bj = kT.B[j2][q2], dj = kT.D[j2][q2];
            c3 = fma(bj, v30[q2], c3);
            c1 = fma(bj, v10[q2], c1);
            c0 = fma(dj, v03[q2], fma(bj, v01[q2], c0));
          }
          const int o = (3 * i2 + j2) * 9 + p3;
          s2[o] = c3;
          s2[81 + o] = c1;
          s2[162 + o] = c0;
        }
      }
    }
    wave_sync();

    // 5. q1 contraction and Jacobian scatter ---------------------------------
    // Jacobian read-backs, all issued before the first store.  Round r: tasks pp = l + 32 (r % 3) = j2 + 3
    // j3 + 9 i2 + 27 i3 < 81, rows i1 = (h ? m : 2 - m), m = r / 3; each lane
    // writes n consecutive columns of its row from column ox, at g_off[r];
    // g_meta[r] = n | ox << 3 | fixed << 6 | go << 7 | firstm << 8 (fixed:
    // strong Dirichlet row; go: the lane writes in this round; firstm bit k:
    // the lane writes column k first)
    double old[9][5];
    unsigned g_off[9], g_meta[9];
    if constexpr (JAC) {
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      const int m = r / 3, pp = l + 32 * (r % 3);
      const bool task = act && pp < 81;
      const int i1 = h ? m : 2 - m;
      const int j2 = pp % 3, j3 = (pp / 3) % 3, i2 = (pp / 9) % 3, i3 = pp / 27;
      int n = 3, ox = 0;
      bool xs_lo = false, xs_hi = false;  // first / last column shared with another segment
      bool go = task && ga_jac != nullptr;
      const int X = 2 * ex + i1;
      if (m == 0) {
        if (h == 0) {
          if (!hasB) xs_hi = x1 < nel0;
        } else {
          n = 2, ox = 3;
        }
      } else if (m == 2) {
        if (h == 0) {
          if (!first_step) {
            n = 5;
          } else {
            ox = X >= 2 ? 2 : 0;
            xs_lo = x0 > 0;
          }
        } else {
          go = go && last_step;  // otherwise carried to the next step
          xs_hi = x1 < nel0;
        }
      }
      const int rpl = s_rp[(X - 2 * xa) + 5 * i2 + 15 * i3];
      const bool fixed = rpl < 0;
      const int rp = fixed ? -1 - rpl : rpl;
      const int lx = (X & 1) ? 3 : ((X == 0 || X == 2 * nel0) ? 3 : 5);
      // (scalar selects: selecting a Range structure by a lane value would
      // place the structures in scratch memory)
      const int ylo = i2 == 0 ? ry0.lo : (i2 == 1 ? ry1.lo : ry2.lo);
      const int ylen = i2 == 0 ? ry0.len : (i2 == 1 ? ry1.len : ry2.len);
      const int zlo = i3 == 0 ? rz0.lo : (i3 == 1 ? rz1.lo : rz2.lo);
      const unsigned off = (unsigned)(rp + ((Z0 + j3 - zlo) * ylen + (Y0 + j2 - ylo)) * lx + ox);
      const bool ysh = i2 == j2 && i2 != 1, zsh = i3 == j3 && i3 != 1;
      const bool yzf = (!ysh || (i2 == 0 ? fly : fhy)) && (!zsh || (i3 == 0 ? flz : fhz));
      unsigned firstm = 0;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        bool f = yzf;
        if (k == 0 && xs_lo) f = f && flx;
        if (k == n - 1 && xs_hi) f = f && fhx;
        if (ga_first_touch && f) firstm |= 1u << k;
      }
      g_off[r] = off;
      g_meta[r] = (unsigned)n | (unsigned)ox << 3 | (unsigned)fixed << 6 | (unsigned)go << 7 | firstm << 8;
      // read-backs: later touches of entries of non-Dirichlet rows (a run
      // has at most 3 columns in rounds r < 6, 5 after)
#pragma unroll
      for (int k = 0; k < 5; ++k)
        old[r][k] = (k < (r < 6 ? 3 : 5) && go && k < n && !fixed && !((firstm >> k) & 1)) ? ga_jac[off + k] : 0.0;
    }
    }


#pragma unroll
      for (int r = 0; r < 9; ++r) {
        const int m = r / 3, pp = l + 32 * (r % 3);
        const bool task = act && pp < 81;
        const int i1 = h ? m : 2 - m;
        const int j2 = pp % 3, j3 = (pp / 3) % 3, i2 = (pp / 9) % 3, i3 = pp / 27;
        double acc[3] = {0.0, 0.0, 0.0};
        if (task) {
          const int o = (3 * i2 + j2) * 9 + 3 * i3 + j3;
#pragma unroll
          for (int q1 = 0; q1 < 3; ++q1) {
            const double* s2 = &s_a[h][q1 * 243];
            const double di = Dq(i1, q1), bi = Bq(i1, q1);
            const double v = di * s2[o];
            const double wv = fma(bi, s2[162 + o], di * s2[81 + o]);
#pragma unroll
            for (int j1 = 0; j1 < 3; ++j1) acc[j1] = fma(kT.D[j1][q1], v, fma(kT.B[j1][q1], wv, acc[j1]));
          }
        }
        // the column A shares with B on their common plane (m == 0) comes
        // from the other half; so does the previous step's carry (m == 2),
        // which half 1 then replaces with its own rows
        double fromB0 = 0.0, cin[3] = {0.0, 0.0, 0.0};
        if (m == 0) fromB0 = __shfl_xor(acc[0], 32, 64);
        if (m == 2) {
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            cin[k] = __shfl_xor(carry_j[r - 6][k], 32, 64);
            carry_j[r - 6][k] = (h == 1 && task) ? acc[k] : carry_j[r - 6][k];
          }
        }
        const unsigned off = g_off[r], meta = g_meta[r];
        if (!((meta >> 7) & 1)) continue;
        const int n = meta & 7, ox = (meta >> 3) & 7;
        const bool fixed = (meta >> 6) & 1;
        const unsigned firstm = meta >> 8;
        // the run's values
        double val[5] = {acc[0], acc[1], acc[2], 0.0, 0.0};
        if (m == 0) {
          if (h == 0) {
            if (hasB) val[2] = acc[2] + fromB0;
          } else {
            val[0] = acc[1], val[1] = acc[2];
          }
        } else if (m == 2 && h == 0 && !first_step) {
          val[0] = cin[0], val[1] = cin[1], val[2] = cin[2] + acc[0], val[3] = acc[1], val[4] = acc[2];
        }
        const int X = 2 * ex + i1;
        const int lox = (X & 1) ? X - 1 : (X - 2 < 0 ? 0 : X - 2);
        if (__any(fixed)) {
#pragma unroll
          for (int k = 0; k < (r < 6 ? 3 : 5); ++k) {
            const bool first = (firstm >> k) & 1;
            if (k >= n || (fixed && !first)) continue;
            // strong Dirichlet row: e_d, written by its first toucher
            const double v = fixed ? ((i2 == j2 && i3 == j3 && lox + ox + k == X) ? 1.0 : 0.0)
                                   : (first ? val[k] : old[r][k] + val[k]);
            ga_jac[off + k] = v;
          }
        } else {
          // no strong-Dirichlet row in the round (the common case): a first
          // touch read nothing (old = +0) and val is never -0 (sums started
          // from +0), so old + val is val exactly
#pragma unroll
          for (int k = 0; k < (r < 6 ? 3 : 5); ++k)
            if (k < n) ga_jac[off + k] = old[r][k] + val[k];
        }
      }
  }
}

// ---------------------------------------------------------------------------
// host launch
// ---------------------------------------------------------------------------
bool q2box_applicable(const Ctx& c) {
  return c.mesh.dim == 3 && c.mesh.p == 2 && c.mesh.nq1 == 3 && c.mesh.nf == 1 && c.box_mesh &&
         c.desc.physics == MHX_PHYS_THERMAL && !c.xh.active;
}

void launch_q2box(Ctx& c, AsmArgs& a, bool jac) {
  const int L = c.q2_segment;
  a.gpb = L;
  a.nseg = (c.mesh.nel[0] + L - 1) / L;
  for (int col = 0; col < 8; ++col) {
    a.color = col;
    const int cxb = col & 1, cyb = (col >> 1) & 1, czb = (col >> 2) & 1;
    a.cn[0] = (a.nseg - cxb + 1) / 2;
    a.cn[1] = (c.mesh.nel[1] - cyb + 1) / 2;
    a.cn[2] = (c.mesh.nel[2] - czb + 1) / 2;
    const long blocks = (long)a.cn[0] * a.cn[1] * a.cn[2];
    if (blocks == 0) continue;
    ScopedPhase ph(c.prof, jac ? "k_q2box<jacres>" : "k_q2box<res>", c.stream);
    const dim3 grid((unsigned)blocks), block(64);
    if (a.coef.n_modes > 0) {
      if (jac) hipLaunchKernelGGL((k_q2box<true, true>), grid, block, 0, c.stream, a);
      else hipLaunchKernelGGL((k_q2box<false, true>), grid, block, 0, c.stream, a);
    } else {
      if (jac) hipLaunchKernelGGL((k_q2box<true, false>), grid, block, 0, c.stream, a);
      else hipLaunchKernelGGL((k_q2box<false, false>), grid, block, 0, c.stream, a);
    }
    MHX_HIP(hipGetLastError());
  }
}

}  // namespace mhx
