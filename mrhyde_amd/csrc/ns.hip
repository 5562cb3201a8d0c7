// ns.hip — fused residual/Jacobian assembly for the steady incompressible
// Navier-Stokes module, Q1-Q1 (equal-order velocity/pressure), SUPG and PSPG,
// on structured quad/hex meshes, CDNA4 (gfx950).
//
// Replaces, as one fused kernel per element colour:
//   AssemblyManager::performGather          assemblyManager_gather.hpp:183-234
//   Workset::computeSolnSteadySeeded        workset.cpp:865-905
//   Workset::evaluateSolutionField          workset.cpp:979-1115
//   Workset::getElementSize                 workset.cpp:2699-2712
//   navierstokes<EvalT>::volumeResidual     navierstokes.cpp:252-494 (2-D), :494-~850 (3-D)
//   navierstokes<EvalT>::computeTau         navierstokes.cpp:1053-1081
//   AssemblyManager::scatter                assemblyManager_scatter.hpp:164-278
//
// Unknowns per node: ux, uy, [uz], pr (NF = DIM+1), DOF = node * NF + field.
// An element has NE = 2^DIM * NF unknowns; element-local index j = node*NF + f.
//
// Work mapping.  A workgroup holds EPW = 64 / 2^DIM elements and one
// wavefront per field (3-D: 8 elements, 4 wavefronts).
//   point stage   (wavefront 0, lane = (element, point)): geometry, the fields
//                 and their physical gradients, tau and d tau / d u, the
//                 strong momentum residual and the point residual in
//                 reference form, once per point, to LDS;
//   column stage  (wavefront F, lane = (element, node n)): residual row
//                 (n, F) from the point residuals, and Jacobian column (n, F)
//                 as the tangent of the point residual for the seed of basis
//                 n in field F.  The field is uniform in the wavefront, so
//                 each wavefront runs the tangent specialised for its field;
//                 it is the derivative part of the reference's element-level
//                 dual (Sacado SFad, one slot per DOF, workset.cpp:884) for
//                 that slot, written out term by term;
//   scatter       the columns are exchanged through LDS so that a lane
//                 writes the NF contiguous entries of a row and column node.
// Elements of one colour share no node, so the scatter needs no atomics; as
// in the thermal kernel the first scheduled element touching a CRS entry
// stores it (Dirichlet rows: e_d, residual 0) and later ones add into it.
// Global residual = -F (scatter.hpp:227); Jacobian = +dF/du (:261).

#include <cmath>

#include "lattice.hpp"
#include "mhx_internal.hpp"

namespace mhx {

struct NsCoef {
  double rho, mu;
  double force[3];
  int supg, pspg;
};

template <int DIM>
struct NsCfg {
  static constexpr int NV = 1 << DIM;  // nodes (= vertices) per element
  static constexpr int NQ = 1 << DIM;  // 2^DIM Gauss points
  static constexpr int NF = DIM + 1;   // unknowns per node
  static constexpr int NE = NV * NF;   // element unknowns
  static constexpr int EPW = 64 / NV;  // elements per workgroup: a lane per (element, node) in each wave
  static constexpr int NT = 64 * NF;   // one wave per field
  static constexpr int NFK = NF * (DIM + 1);  // (field, component) pairs: value and reference gradient
  // per point in LDS
  static constexpr int OJ = 0;                 // J^-1
  static constexpr int OW = DIM * DIM;         // w |J|
  static constexpr int OF = OW + 1;            // fields: physical gradient [f*(DIM+1) + c], value [.. + DIM]
  static constexpr int OT = OF + NFK;          // tau
  static constexpr int OTC = OT + 1;           // d tau / d u_a
  static constexpr int OS = OTC + DIM;         // strong momentum residual
  static constexpr int OH = OS + DIM;          // point residual Fhat[f][k] (reference form)
  static constexpr int NQS = OH + NFK;
  static_assert(NV * EPW == 64, "a wave holds whole elements");
};

// Q1 on [-1,1]: node bit b -> (1 + s xi)/2 with s = 2b - 1; Gauss points
// +-1/sqrt(3) in descending order (point bit 0 -> +1/sqrt(3)), weights 1.
__device__ __forceinline__ constexpr double q1_xi(int qbit) { return qbit ? -0.57735026918962576451 : 0.57735026918962576451; }
__device__ __forceinline__ constexpr double q1_val(int nbit, int qbit) {
  return 0.5 * (1.0 + (nbit ? 1.0 : -1.0) * q1_xi(qbit));
}
__device__ __forceinline__ constexpr double q1_der(int nbit) { return nbit ? 0.5 : -0.5; }

// Reference value and gradient of node basis n at point q: (dphi/dxi_0..DIM-1, phi).
template <int DIM>
__device__ __forceinline__ void q1_basis(int n, int q, double* t) {
  double val = 1.0;
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    double g = 1.0;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      const int nb = (n >> a) & 1, qb = (q >> a) & 1;
      g *= (a == c) ? q1_der(nb) : q1_val(nb, qb);
    }
    t[c] = g;
    val *= q1_val((n >> c) & 1, (q >> c) & 1);
  }
  t[DIM] = val;
}

// The same values as a table in constant memory, [point][node][(d/dxi.., value)]:
// read with a wavefront-uniform point index (scalar loads) in the column stage.
template <int DIM>
struct Q1Tab {
  double t[1 << DIM][1 << DIM][DIM + 1];
};
template <int DIM>
constexpr Q1Tab<DIM> make_q1tab() {
  Q1Tab<DIM> r{};
  for (int q = 0; q < (1 << DIM); ++q)
    for (int n = 0; n < (1 << DIM); ++n) {
      double val = 1.0;
      for (int c = 0; c < DIM; ++c) {
        double g = 1.0;
        for (int a = 0; a < DIM; ++a) {
          const int nb = (n >> a) & 1, qb = (q >> a) & 1;
          g *= (a == c) ? q1_der(nb) : q1_val(nb, qb);
        }
        r.t[q][n][c] = g;
        val *= q1_val((n >> c) & 1, (q >> c) & 1);
      }
      r.t[q][n][DIM] = val;
    }
  return r;
}
__constant__ Q1Tab<2> c_q1tab2 = make_q1tab<2>();
__constant__ Q1Tab<3> c_q1tab3 = make_q1tab<3>();
template <int DIM>
__device__ __forceinline__ const Q1Tab<DIM>& q1tab();
template <>
__device__ __forceinline__ const Q1Tab<2>& q1tab<2>() { return c_q1tab2; }
template <>
__device__ __forceinline__ const Q1Tab<3>& q1tab<3>() { return c_q1tab3; }

// Tangent of the point residual for a seed in field F (compile time) with
// value phi and physical gradient gj: dFh[f][k], the derivative of the
// reference-form point residual Fhat[f][k] (navierstokes.cpp:494-850,
// computeTau :1053-1081), from the point values o[] (LDS).  Written as the
// forward-mode derivative of the value expressions, term by term, with the
// terms that vanish for field F left out at compile time.
template <int DIM, int F>
__device__ __forceinline__ void ns_tangent(const double* o, const NsCoef& nc, double w, double phi,
                                           const double* gj, double dFh[DIM + 1][DIM + 1]) {
  using C = NsCfg<DIM>;
  constexpr bool VEL = F < DIM;
  constexpr int NFK1 = DIM + 1;
  const double* fld = o + C::OF;  // u_f at fld[f * NFK1 + DIM], d_c u_f at fld[f * NFK1 + c]
  const double tau = o[C::OT];
  // seed: du_F = phi, dgrad u_F = gj
  double dconv[DIM], dsr[DIM];
  double ugj = 0.0;
  if constexpr (VEL) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) ugj = fma(fld[c * NFK1 + DIM], gj[c], ugj);
  }
#pragma unroll
  for (int m = 0; m < DIM; ++m) {
    // conv_m = sum_c u_c d_c u_m
    double d = 0.0;
    if constexpr (VEL) {
      d = phi * fld[m * NFK1 + (VEL ? F : 0)];
      if (m == F) d += ugj;
    }
    dconv[m] = d;
    // stabres_m = rho conv_m + d_m p - rho f_m
    dsr[m] = nc.rho * d;
    if constexpr (!VEL) dsr[m] = gj[m];
  }
  const double dtau = VEL ? o[C::OTC + (VEL ? F : 0)] * phi : 0.0;
#pragma unroll
  for (int m = 0; m < DIM; ++m) {
    const double sr = o[C::OS + m];
    const double dts = dtau * sr + tau * dsr[m];  // d(tau stabres_m)
    const double ts = tau * sr;
    double dfl[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      double d = 0.0;
      if constexpr (VEL) {
        if (m == F) d = nc.mu * gj[c];  // mu d_c u_m
      } else {
        if (c == m) d = -phi;  // - p delta_mc
      }
      if (nc.supg) {
        // tau stabres_m u_c
        double sd = dts * fld[c * NFK1 + DIM];
        if constexpr (VEL) {
          if (c == F) sd = fma(ts, phi, sd);
        }
        d += sd;
      }
      dfl[c] = d * w;
    }
    dFh[m][DIM] = dconv[m] * (nc.rho * w);
#pragma unroll
    for (int r = 0; r < DIM; ++r) {
      double s = o[C::OJ + r * DIM] * dfl[0];
#pragma unroll
      for (int c = 1; c < DIM; ++c) s = fma(o[C::OJ + r * DIM + c], dfl[c], s);
      dFh[m][r] = s;
    }
  }
  // continuity: div u, PSPG tau stabres_c / rho
  dFh[DIM][DIM] = VEL ? gj[VEL ? F : 0] * w : 0.0;
  double dfl[DIM];
  const double kw = w / nc.rho;
#pragma unroll
  for (int c = 0; c < DIM; ++c) dfl[c] = nc.pspg ? (dsr[c] * (tau * kw) + o[C::OS + c] * (dtau * kw)) : 0.0;
#pragma unroll
  for (int r = 0; r < DIM; ++r) {
    double s = o[C::OJ + r * DIM] * dfl[0];
#pragma unroll
    for (int c = 1; c < DIM; ++c) s = fma(o[C::OJ + r * DIM + c], dfl[c], s);
    dFh[DIM][r] = s;
  }
}

// One workgroup: EPW elements, one wavefront per field.  Point stage (wave
// 0, lane = (element, point)): geometry, the fields and their gradients, tau
// and its derivative by the velocity, and the point residual in reference
// form, once per point, to LDS.  Column stage (wave F = field F, lane =
// (element, node n)): residual row (n, F) from the point residuals, and
// Jacobian column (n, F) from the tangent of the point residual for the seed
// of basis n in field F; the field is uniform in the wavefront, so each
// wavefront runs the tangent specialised for its field.
// Unit of the segment schedule: x-segment sx of EPW elements of the element
// row (ey, ez), one colour per (sx, ey, ez) parity; launch order is colour
// order, so along each axis the even unit of two sharing a vertex plane runs
// first.
struct NsSeg {
  int sx, ey, ez, x0, n;         // segment, pencil, first element, active elements
  bool flx, fhx, fly, fhy, flz, fhz;  // first toucher of the shared low / high plane
};
template <int DIM>
__device__ __forceinline__ NsSeg ns_seg(const AsmArgs& a, int S) {
  NsSeg u;
  int t = blockIdx.x;
  u.sx = 2 * (t % a.cn[0]) + (a.color & 1);
  t /= a.cn[0];
  u.ey = 2 * (t % a.cn[1]) + ((a.color >> 1) & 1);
  t /= a.cn[1];
  u.ez = DIM == 3 ? 2 * t + ((a.color >> 2) & 1) : 0;
  u.x0 = u.sx * S;
  u.n = min(S, a.m.nel[0] - u.x0);
  u.flx = u.sx == 0 || (u.sx & 1) == 0;
  u.fhx = u.sx == a.nseg - 1 || (u.sx & 1) == 0;
  u.fly = u.ey == 0 || (u.ey & 1) == 0;
  u.fhy = u.ey == a.m.nel[1] - 1 || (u.ey & 1) == 0;
  u.flz = DIM < 3 || u.ez == 0 || (u.ez & 1) == 0;
  u.fhz = DIM < 3 || u.ez == a.m.nel[2] - 1 || (u.ez & 1) == 0;
  return u;
}

template <int DIM, bool SEG>
__device__ __forceinline__ bool ns_element(const AsmArgs& a, const NsSeg& u, int e, int* ec) {
  if constexpr (SEG) {
    ec[0] = u.x0 + (e < u.n ? e : 0);
    ec[1] = u.ey;
    ec[2] = u.ez;
    return e < u.n;
  } else {
    return element_of<DIM>(a, e, 0, ec);
  }
}

// Scatter of an x-segment: element el (lanes of every wavefront, column (jn,
// F) per lane) puts its rows into LDS, even elements first (stores), then odd
// ones (adding where the entry lies on a plane shared with an even
// neighbour), so an entry on a plane shared inside the segment is the sum of
// exactly two contributions in a fixed order; each row's x-runs are then
// read back (if an earlier unit wrote them) and written once.  One round per
// row field (the LDS buffer is the point data's); the per-entry addressing
// and first-touch flags, which do not depend on the row field, are formed
// once.  An entry is shared with another unit only on the segment's end
// planes (x) or on a (y, z) vertex plane of the pencil; the first scheduled
// unit among the sharers stores it.
// Read-backs of one round: the CRS address of each of the lane's entries in
// the rows of field fr, and the value an earlier unit stored there (0 where
// this unit is the first toucher, on Dirichlet rows and for unused slots).
template <int NIT, int NF>
__device__ __forceinline__ void ns_seg_readback(const AsmArgs& a, const int* s_rps, const int* it_r,
                                                const int* it_off, const unsigned* it_f, int fr, long* addr,
                                                double* old, bool* fixd) {
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    int rp = s_rps[it_r[k] * NF + fr];
    fixd[k] = rp < 0;
    if (fixd[k]) rp = -1 - rp;
    addr[k] = (long)rp + it_off[k];
    const unsigned f = it_f[k];
    old[k] = ((f & 1u) && !fixd[k] && !(f & 2u)) ? a.jac[addr[k]] : 0.0;
  }
}

template <int DIM, bool JAC>
__device__ __forceinline__ void ns_seg_scatter(const AsmArgs& a, const NsSeg& u, double* buf, const int* s_rps,
                                               bool active, int el, int jn, int wv, int t, double rres,
                                               const double* R) {
  using C = NsCfg<DIM>;
  constexpr int NF = C::NF, NV = C::NV, EPW = C::EPW, NT = C::NT;
  constexpr int YZ = 1 << (DIM - 1), RPS = (EPW + 1) * YZ, CPR = YZ * 3 * NF;
  constexpr int NJ = JAC ? RPS * CPR : 0;  // Jacobian entries of a round
  constexpr int NIT = JAC ? (NJ + NT - 1) / NT : 1;
  static_assert(NJ + RPS <= EPW * C::NQ * C::NQS, "a round fits over the point data");
  double* jb = buf;       // [RPS][CPR]: row (x plane, yz line), column ((yz line) 3 + x offset) NF + field
  double* rb = buf + NJ;  // [RPS]: residual rows of the round's field
  const int jx = jn & 1, jyz = jn >> 1;
  const int X0 = u.x0, X1 = u.x0 + u.n;
  const int nn0 = a.m.nn[0], nn1 = a.m.nn[1];
  // this lane's entries q = t + k NT: row r, offset in the row's CRS box,
  // flags 1 write, 2 first, 4 diagonal node pair, field of the column << 4
  int it_r[NIT], it_off[NIT];
  unsigned it_f[NIT];
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    it_r[k] = 0;
    it_off[k] = 0;
    it_f[k] = 0;
    if constexpr (JAC) {
      const int q = t + k * NT;
      const int r = q / CPR, c = q - (q / CPR) * CPR;
      const int xp = r / YZ, yz = r % YZ, cyz = c / (3 * NF), cx = (c / NF) % 3, fc = c % NF;
      const int X = X0 + xp, Xc = X + cx - 1;
      const bool valid = q < NJ && xp <= u.n && Xc >= X0 && Xc <= X1;
      const int jy = yz & 1, jz = yz >> 1, cy = cyz & 1, cz = cyz >> 1;
      const int Y = u.ey + jy, Z = u.ez + jz, Yc = u.ey + cy, Zc = u.ez + cz;
      const int lox = X > 0 ? X - 1 : 0, loy = Y > 0 ? Y - 1 : 0, loz = Z > 0 ? Z - 1 : 0;
      const int lx = (X + 1 < nn0 ? X + 1 : nn0 - 1) - lox + 1;
      const int ly = (Y + 1 < nn1 ? Y + 1 : nn1 - 1) - loy + 1;
      const int pz = DIM == 3 ? Zc - loz : 0;
      const bool xsh = X == Xc && ((X == X0 && X0 > 0) || (X == X1 && X1 < a.m.nel[0]));
      const bool xr = !xsh || (X == X0 ? u.flx : u.fhx);
      const bool yr = Y != Yc || (jy == 0 ? u.fly : u.fhy);
      const bool zr = DIM < 3 || Z != Zc || (jz == 0 ? u.flz : u.fhz);
      const bool first = a.first_touch && xr && yr && zr;
      const bool dnode = X == Xc && Y == Yc && Z == Zc;
      it_r[k] = valid ? r : 0;
      it_off[k] = ((pz * ly + (Yc - loy)) * lx + (Xc - lox)) * NF + fc;
      it_f[k] = (valid ? 1u : 0u) | (first ? 2u : 0u) | (dnode ? 4u : 0u) | ((unsigned)fc << 4);
    }
  }
  // residual row t (< RPS) of each round: node and first-touch flag
  long res_node = 0;
  bool res_go = false, res_first = false;
  if (t < RPS) {
    const int xp = t / YZ, yz = t % YZ, jy = yz & 1, jz = yz >> 1;
    res_go = a.res != nullptr && xp <= u.n;
    res_node = ((long)(u.ez + jz) * nn1 + (u.ey + jy)) * nn0 + (X0 + xp);
    const bool xr = xp == 0 ? u.flx : (xp == u.n ? u.fhx : true);
    res_first = a.first_touch && xr && (jy == 0 ? u.fly : u.fhy) && (DIM < 3 || (jz == 0 ? u.flz : u.fhz));
  }
  // read-backs of round 0 (later touches of entries of non-Dirichlet rows)
  long addr[NIT], addr_n[NIT];
  double old[NIT], old_n[NIT];
  bool fixd[NIT], fix_n[NIT];
  if (JAC && a.jac) ns_seg_readback<NIT, NF>(a, s_rps, it_r, it_off, it_f, 0, addr, old, fixd);
  // odd element: its planes shared with the even neighbours (low: el - 1;
  // high: el + 1 when it exists)
  const bool odd = el & 1;
  const bool sh_lo = odd, sh_hi = odd && el + 1 < u.n;
  __syncthreads();  // every wavefront is done with the point data
#pragma unroll
  for (int fr = 0; fr < NF; ++fr) {
#pragma unroll
    for (int par = 0; par < 2; ++par) {
      if (active && (el & 1) == par) {
        if (wv == fr) {
          double* p = &rb[(el + jx) * YZ + jyz];
          const bool add = jx == 0 ? sh_lo : sh_hi;
          *p = add ? *p + rres : rres;
        }
        if constexpr (JAC) {
#pragma unroll
          for (int in = 0; in < NV; ++in) {
            const int ix = in & 1;
            double* p = &jb[((el + ix) * YZ + (in >> 1)) * CPR + (jyz * 3 + (jx - ix + 1)) * NF + wv];
            const bool add = ix == jx && (ix == 0 ? sh_lo : sh_hi);
            *p = add ? *p + R[in * NF + fr] : R[in * NF + fr];
          }
        }
      }
      __syncthreads();
    }
    if (res_go) {
      const long row = res_node * NF + fr;
      const bool fx = s_rps[t * NF + fr] < 0;
      const double v = -rb[t];
      if (res_first) a.res[row] = fx ? 0.0 : v;
      else if (!fx) a.res[row] += v;
    }
    if constexpr (JAC) {
      if (a.jac) {
        // the next round's read-backs go out before this round's stores
        // (disjoint rows; loads issued ahead of stores do not wait for them)
        if (fr + 1 < NF) ns_seg_readback<NIT, NF>(a, s_rps, it_r, it_off, it_f, fr + 1, addr_n, old_n, fix_n);
        bool anyfix = false;
#pragma unroll
        for (int k = 0; k < NIT; ++k) anyfix = anyfix || ((it_f[k] & 1u) && fixd[k]);
        if (__any(anyfix)) {
#pragma unroll
          for (int k = 0; k < NIT; ++k) {
            const unsigned f = it_f[k];
            if (!(f & 1u)) continue;
            const bool first = f & 2u;
            if (fixd[k]) {
              if (first) a.jac[addr[k]] = ((f & 4u) && (int)(f >> 4) == fr) ? 1.0 : 0.0;
            } else {
              const double v = jb[t + k * NT];
              a.jac[addr[k]] = first ? v : old[k] + v;
            }
          }
        } else {
          // no strong-Dirichlet row among the wavefront's entries (the
          // common case): a first touch read nothing (old = +0), so old + v
          // is v (up to the sign of a zero)
#pragma unroll
          for (int k = 0; k < NIT; ++k)
            if (it_f[k] & 1u) a.jac[addr[k]] = old[k] + jb[t + k * NT];
        }
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
          addr[k] = addr_n[k];
          old[k] = old_n[k];
          fixd[k] = fix_n[k];
        }
      }
    }
    __syncthreads();  // the next round overwrites the buffer
  }
}

// STAB: SUPG and PSPG both on, known at compile time (the stabilisation
// flags of the local copy become constants, so neither term is selected per
// lane at run time)
template <int DIM, bool JAC, bool SEG, bool STAB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_ns(AsmArgs a, NsCoef nc_arg) {
  NsCoef nc = nc_arg;
  if constexpr (STAB) nc.supg = 1, nc.pspg = 1;
  using C = NsCfg<DIM>;
  constexpr int NV = C::NV, NQ = C::NQ, NF = C::NF, NE = C::NE, EPW = C::EPW, NQS = C::NQS;
  __shared__ double s_u[EPW][NE];
  __shared__ double s_X[EPW][NV * DIM];
  __shared__ double s_q[EPW][NQ][NQS];
  __shared__ int s_rp[EPW][NE];
  constexpr int YZ = 1 << (DIM - 1);      // (y, z) node lines of an element row
  constexpr int RPS = (EPW + 1) * YZ;     // segment: node rows per field
  __shared__ int s_rps[SEG ? RPS * NF : 1];  // segment: CRS offset of row (x plane, yz line, field)

  const int t = threadIdx.x;
  const int wv = __builtin_amdgcn_readfirstlane(t / 64), lane = t % 64;
  const int el = lane / NV;  // element of this lane in the workgroup (column and point stages)
  const NsSeg u = SEG ? ns_seg<DIM>(a, EPW) : NsSeg{};

  // 1. gather: element unknowns (node-major, field-minor), CRS offsets, vertices
  for (int k = t; k < EPW * NE; k += C::NT) {
    const int e = k / NE, j = k % NE;
    int ec[3];
    const bool act = ns_element<DIM, SEG>(a, u, e, ec);
    if (act) {
      const int n = j / NF, f = j % NF;
      const long dof = node_of<DIM, 1>(a.m, ec, n) * NF + f;
      s_u[e][j] = a.u[dof];
      s_rp[e][j] = a.fixed[dof] ? -1 - a.rowptr[dof] : a.rowptr[dof];
      if constexpr (SEG) s_rps[((e + (n & 1)) * YZ + (n >> 1)) * NF + f] = s_rp[e][j];  // shared planes: same value
    }
  }
  for (int k = t; k < EPW * NV * DIM; k += C::NT) {
    const int e = k / (NV * DIM), r = k % (NV * DIM);
    int ec[3];
    if (ns_element<DIM, SEG>(a, u, e, ec)) {
      const int v = r / DIM, d = r % DIM;
      const long vid = node_of<DIM, 1>(a.m, ec, v);  // Q1: nodes are the vertices
      s_X[e][r] = a.vx[vid * 3 + d];
    }
  }
  __syncthreads();

  // 2. point stage, part 1: geometry and fields (wave 0, lane = (element, point))
  int ecp[3];
  const bool pact = wv == 0 && ns_element<DIM, SEG>(a, u, el, ecp);
  const int q = lane % NQ;
  if (pact) {
    double J[DIM][DIM];
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
      for (int s2 = 0; s2 < DIM; ++s2) J[r][s2] = 0.0;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double tv[DIM + 1];
      q1_basis<DIM>(v, q, tv);
#pragma unroll
      for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int s2 = 0; s2 < DIM; ++s2) J[r][s2] = fma(s_X[el][v * DIM + r], tv[s2], J[r][s2]);
    }
    double Ji[DIM][DIM], det;
    if constexpr (DIM == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      const double id = 1.0 / det;
      Ji[0][0] = J[1][1] * id; Ji[0][1] = -J[0][1] * id;
      Ji[1][0] = -J[1][0] * id; Ji[1][1] = J[0][0] * id;
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
    double* o = s_q[el][q];
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
      for (int s2 = 0; s2 < DIM; ++s2) o[C::OJ + r * DIM + s2] = Ji[r][s2];
    o[C::OW] = fabs(det);  // Gauss weights are 1
    double fv[NF][DIM + 1];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int c = 0; c <= DIM; ++c) fv[f][c] = 0.0;
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      double tv[DIM + 1];
      q1_basis<DIM>(n, q, tv);
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const double un = s_u[el][n * NF + f];
#pragma unroll
        for (int c = 0; c <= DIM; ++c) fv[f][c] = fma(un, tv[c], fv[f][c]);
      }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      double* of = o + C::OF + f * (DIM + 1);
      of[DIM] = fv[f][DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        double s2 = 0.0;
#pragma unroll
        for (int b = 0; b < DIM; ++b) s2 = fma(Ji[b][c], fv[f][b], s2);
        of[c] = s2;
      }
    }
  }
  __syncthreads();

  // 2'. point stage, part 2: tau (Workset::getElementSize h), convection,
  // strong momentum residual and the point residual values
  if (pact) {
    double* o = s_q[el][q];
    double vol = 0.0;
#pragma unroll
    for (int qq = 0; qq < NQ; ++qq) vol += s_q[el][qq][C::OW];
    const double h = DIM == 2 ? sqrt(vol) : cbrt(vol);
    const double w = o[C::OW];
    const double* fld = o + C::OF;
    double U[NF], GU[NF][DIM];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      U[f] = fld[f * (DIM + 1) + DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) GU[f][c] = fld[f * (DIM + 1) + c];
    }
    // computeTau, steady (C3 = 0): tau = (t1 + (C2/h |u|)^2)^(-1/2), with
    // |u|^2 used as |u| below 1e-12; its derivative by u_a
    double tau = 0.0, tc[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) tc[c] = 0.0;
    if (nc.supg || nc.pspg) {
      const double C1 = 4.0, C2 = 2.0;
      double nv = U[0] * U[0];
#pragma unroll
      for (int c = 1; c < DIM; ++c) nv = nv + U[c] * U[c];
      double dn = 2.0;  // d nvel / d(|u|^2) times 2 u_a
      if (nv > 1e-12) {
        const double r = sqrt(nv);
        dn = 2.0 * (0.5 / r);
        nv = r;
      }
      const double t1 = (C1 * nc.mu / h / h) * (C1 * nc.mu / h / h);
      const double c2 = (C2 / h) * nv;
      const double tt = c2 * c2 + t1;
      const double sq = sqrt(tt);
      tau = 1.0 / sq;
      // d tau = -(d tt (0.5/sq)) tau^2,  d tt = 2 c2 (C2/h) d nvel
      const double k = -(0.5 / sq) * tau * tau * (2.0 * c2) * (C2 / h) * dn;
#pragma unroll
      for (int c = 0; c < DIM; ++c) tc[c] = k * U[c];
    }
    o[C::OT] = tau;
#pragma unroll
    for (int c = 0; c < DIM; ++c) o[C::OTC + c] = tc[c];
    double conv[DIM], sr[DIM];
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
      double cv = U[0] * GU[m][0];
#pragma unroll
      for (int c = 1; c < DIM; ++c) cv = cv + U[c] * GU[m][c];
      conv[m] = cv;
      sr[m] = (nc.rho * cv + GU[DIM][m]) - nc.rho * nc.force[m];
      o[C::OS + m] = sr[m];
    }
    const double pr = U[DIM];
    double Fh[NF][DIM + 1];
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
      double fl[DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        double v = nc.mu * GU[m][c];
        if (c == m) v = v - pr;
        if (nc.supg) v = v + tau * sr[m] * U[c];
        fl[c] = v * w;
      }
      Fh[m][DIM] = (conv[m] - nc.force[m]) * (nc.rho * w);
#pragma unroll
      for (int r = 0; r < DIM; ++r) {
        double s2 = o[C::OJ + r * DIM] * fl[0];
#pragma unroll
        for (int c = 1; c < DIM; ++c) s2 = s2 + o[C::OJ + r * DIM + c] * fl[c];
        Fh[m][r] = s2;
      }
    }
    {
      double divu = GU[0][0];
#pragma unroll
      for (int c = 1; c < DIM; ++c) divu = divu + GU[c][c];
      Fh[DIM][DIM] = divu * w;
      double fl[DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) fl[c] = nc.pspg ? sr[c] * (tau * (w / nc.rho)) : 0.0;
#pragma unroll
      for (int r = 0; r < DIM; ++r) {
        double s2 = o[C::OJ + r * DIM] * fl[0];
#pragma unroll
        for (int c = 1; c < DIM; ++c) s2 = s2 + o[C::OJ + r * DIM + c] * fl[c];
        Fh[DIM][r] = s2;
      }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int k = 0; k <= DIM; ++k) o[C::OH + f * (DIM + 1) + k] = Fh[f][k];
  }
  __syncthreads();

  // 3. column stage: wave F, lane (element el, node jn): residual row and
  // Jacobian column (jn, F)
  int ec[3];
  const bool active = wv < NF && ns_element<DIM, SEG>(a, u, el, ec);
  const int F = wv, jn = lane % NV, jl = jn * NF + F;
  double R[JAC ? NE : 1];
#pragma unroll
  for (int i = 0; i < (JAC ? NE : 1); ++i) R[i] = 0.0;
  double rres = 0.0;
  if (active) {
    const Q1Tab<DIM>& tab = q1tab<DIM>();
    // a runtime point loop (uniform index): unrolled, the points' LDS reads
    // are hoisted together and the registers run out
#pragma unroll 1
    for (int qq = 0; qq < NQ; ++qq) {
      const double* o = s_q[el][qq];
      // the lane's basis at the point (reference gradient, value)
      double tjn[DIM + 1];
#pragma unroll
      for (int k = 0; k <= DIM; ++k) tjn[k] = tab.t[qq][jn][k];
      // residual row (jn, F)
#pragma unroll
      for (int k = 0; k <= DIM; ++k) rres = fma(o[C::OH + F * (DIM + 1) + k], tjn[k], rres);
      if constexpr (JAC) {
        double gj[DIM];  // physical gradient J^-T grad^
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          double s2 = 0.0;
#pragma unroll
          for (int b = 0; b < DIM; ++b) s2 = fma(o[C::OJ + b * DIM + c], tjn[b], s2);
          gj[c] = s2;
        }
        double dFh[DIM + 1][DIM + 1];
        const double w = o[C::OW];
        if (F == 0) ns_tangent<DIM, 0>(o, nc, w, tjn[DIM], gj, dFh);
        else if (F == 1) ns_tangent<DIM, 1>(o, nc, w, tjn[DIM], gj, dFh);
        else if (DIM == 3 && F == 2) ns_tangent<DIM, (DIM == 3 ? 2 : 0)>(o, nc, w, tjn[DIM], gj, dFh);
        else ns_tangent<DIM, DIM>(o, nc, w, tjn[DIM], gj, dFh);
        // column: against every test function (uniform table reads)
#pragma unroll
        for (int in = 0; in < NV; ++in) {
#pragma unroll
          for (int f = 0; f < NF; ++f) {
            double s2 = R[in * NF + f];
#pragma unroll
            for (int k = 0; k <= DIM; ++k) s2 = fma(dFh[f][k], tab.t[qq][in][k], s2);
            R[in * NF + f] = s2;
          }
        }
      }
    }
  }

  if constexpr (SEG) {
    ns_seg_scatter<DIM, JAC>(a, u, &s_q[0][0][0], s_rps, active, el, jn, wv, t, rres, R);
    return;
  }

  // 4. scatter: first scheduled element touching an entry stores it
  bool first_lo[3] = {true, true, true}, first_hi[3] = {true, true, true};
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const int k = axis_key(a, d, ec[d]);
    first_lo[d] = ec[d] == 0 || k < axis_key(a, d, ec[d] - 1);
    first_hi[d] = ec[d] == a.m.nel[d] - 1 || k < axis_key(a, d, ec[d] + 1);
  }
  auto shared_first = [&](int ni, int nj) {
    bool f = true;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int il = (ni >> d) & 1, jl2 = (nj >> d) & 1;
      if (il == jl2) f = f && (il == 0 ? first_lo[d] : first_hi[d]);
    }
    return f;
  };
  if (active && a.res) {
    const long row = node_of<DIM, 1>(a.m, ec, jn) * NF + F;
    const bool fx = s_rp[el][jl] < 0;
    if (a.first_touch && shared_first(jn, jn)) a.res[row] = fx ? 0.0 : -rres;
    else if (!fx) a.res[row] += -rres;
  }
  if constexpr (JAC) {
    // Column (jn, F) of every element sits in the lanes of wave F.  It is
    // exchanged through LDS (over the point data, no longer read), the rows
    // of two nodes per round, so that a lane then reads back and writes the
    // NF contiguous entries of one row and one column node, back to back.
    // Plain stores: the L2 merges a lane's NF stores into whole lines
    // (non-temporal stores of the same pattern went out as partial writes,
    // 2-4x the bytes).
    constexpr int RC = 2 * NF;  // rows per round
    static_assert(RC * EPW * NV * NF <= EPW * NQ * NQS, "the exchange fits over the point data");
    double* xb = &s_q[0][0][0];
    const int jlc[3] = {jn & 1, (jn >> 1) & 1, (jn >> 2) & 1};
    __syncthreads();  // every wavefront is done with the point data
#pragma unroll
    for (int i0 = 0; i0 < NE; i0 += RC) {
      if (active) {
#pragma unroll
        for (int r = 0; r < RC; ++r) xb[((r * EPW + el) * NV + jn) * NF + F] = R[i0 + r];
      }
      __syncthreads();
      if (active && a.jac) {
        // rows (in, F) for the round's two nodes in; lane: column node jn
        long base[2];
        bool first[2], fx[2];
        double old[2][NF];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int in = i0 / NF + h, i = in * NF + F;
          const int il[3] = {in & 1, (in >> 1) & 1, (in >> 2) & 1};
          int px, lx, py, ly, pz = 0, lz = 1;
          axis_slot<1>(il[0], jlc[0], ec[0], a.m.nn[0], px, lx);
          axis_slot<1>(il[1], jlc[1], ec[1], a.m.nn[1], py, ly);
          if (DIM == 3) axis_slot<1>(il[2], jlc[2], ec[2], a.m.nn[2], pz, lz);
          int rp = s_rp[el][i];
          fx[h] = rp < 0;
          if (fx[h]) rp = -1 - rp;
          base[h] = rp + (long)((pz * ly + py) * lx + px) * NF;
          first[h] = a.first_touch && shared_first(in, jn);
#pragma unroll
          for (int f = 0; f < NF; ++f) old[h][f] = (!first[h] && !fx[h]) ? a.jac[base[h] + f] : 0.0;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int in = i0 / NF + h, r = h * NF + F;
          const double* x = xb + ((r * EPW + el) * NV + jn) * NF;
#pragma unroll
          for (int f = 0; f < NF; ++f) {
            if (fx[h]) {
              if (first[h]) a.jac[base[h] + f] = (in == jn && f == F) ? 1.0 : 0.0;
            } else {
              a.jac[base[h] + f] = old[h][f] + x[f];
            }
          }
        }
      }
      __syncthreads();
    }
  }
}

// Element colouring (one launch per element colour, EPW elements of the
// colour per workgroup) or, with seg, x-segments of EPW consecutive elements
// per workgroup, one launch per (segment, ey, ez) parity.
template <int DIM>
static void launch_ns(Ctx& c, AsmArgs& a, const NsCoef& nc, bool jac, bool seg) {
  using C = NsCfg<DIM>;
  const int ncol = 1 << DIM;
  a.gpb = C::EPW;
  a.blocked = 0;
  a.nsub = 1;
  a.nseg = (c.mesh.nel[0] + C::EPW - 1) / C::EPW;
  for (int col = 0; col < ncol; ++col) {
    a.color = col;
    long cnt = 1;
    for (int d = 0; d < 3; ++d) {
      const int cd = (col >> d) & 1;
      const int units = d == 0 && seg ? a.nseg : (d < DIM ? c.mesh.nel[d] : 1);
      int n = d < DIM ? (units - cd + 1) / 2 : 1;
      if (d >= DIM && cd) n = 0;
      a.cn[d] = n;
      cnt *= n;
    }
    a.ncolor = cnt;
    if (cnt == 0) continue;
    const long blocks = seg ? cnt : (cnt + C::EPW - 1) / C::EPW;
    ScopedPhase ph(c.prof, jac ? "k_ns<jacres>" : "k_ns<res>", c.stream);
    const dim3 grid((unsigned)blocks), block(C::NT);
    auto go = [&](auto ST) {
      constexpr bool S = decltype(ST)::value;
      if (seg) {
        if (jac) hipLaunchKernelGGL((k_ns<DIM, true, true, S>), grid, block, 0, c.stream, a, nc);
        else hipLaunchKernelGGL((k_ns<DIM, false, true, S>), grid, block, 0, c.stream, a, nc);
      } else {
        if (jac) hipLaunchKernelGGL((k_ns<DIM, true, false, S>), grid, block, 0, c.stream, a, nc);
        else hipLaunchKernelGGL((k_ns<DIM, false, false, S>), grid, block, 0, c.stream, a, nc);
      }
    };
    if (nc.supg && nc.pspg) go(std::true_type{});
    else go(std::false_type{});
    MHX_HIP(hipGetLastError());
  }
}

bool ns_supported(int dim, int p, int nq1) { return (dim == 2 || dim == 3) && p == 1 && nq1 == 2; }

__global__ void k_dbc_diag_ns(const int32_t* __restrict__ list, long n, const int32_t* __restrict__ rowptr,
                              const int32_t* __restrict__ colidx, double* __restrict__ jac) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int row = list[i];
  int lo = rowptr[row], hi = rowptr[row + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (colidx[mid] < row) lo = mid + 1;
    else hi = mid;
  }
  jac[lo] = 1.0;
}

void assemble_ns(Ctx& c, const double* u, double* res, double* jac, bool want_res, bool want_jac, bool accumulate) {
  MHX_REQUIRE(ns_supported(c.mesh.dim, c.mesh.p, c.mesh.nq1), MHX_ERR_UNSUPPORTED,
              "navier stokes kernel built for Q1 with 2-point Gauss rules only");
  AsmArgs a{};
  a.m = c.mesh;
  a.vx = c.vx.p;
  a.u = u;
  a.rowptr = c.rowptr.p;
  a.fixed = c.fixed.p;
  a.res = want_res ? res : nullptr;
  a.jac = want_jac ? jac : nullptr;
  a.first_touch = accumulate ? 0 : 1;
  NsCoef nc{};
  nc.rho = c.desc.ns_rho;
  nc.mu = c.desc.ns_mu;
  for (int d = 0; d < 3; ++d) nc.force[d] = c.desc.ns_force[d];
  nc.supg = c.desc.ns_supg;
  nc.pspg = c.desc.ns_pspg;
  {
    ScopedPhase ph(c.prof, want_jac ? "AssemblyManager::assembleJacRes" : "AssemblyManager::assembleRes", c.stream);
    const bool seg = c.scatter_mode == kSegmentMerge;
    if (c.mesh.dim == 2) launch_ns<2>(c, a, nc, want_jac, seg);
    else launch_ns<3>(c, a, nc, want_jac, seg);
  }
  // traction sides (navierstokes.cpp:853-1000): the load vector, after the
  // volume terms (it adds into rows the volume launches wrote)
  if (want_res) neumann_residual(c, res);
  if (accumulate && want_jac && !c.h_dbc_list.empty()) {
    ScopedPhase ph(c.prof, "AssemblyManager::dofConstraints", c.stream);
    const long n = (long)c.h_dbc_list.size();
    hipLaunchKernelGGL(k_dbc_diag_ns, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, c.dbc_list.p, n,
                       c.rowptr.p, c.colidx.p, jac);
    MHX_HIP(hipGetLastError());
  }
}

}  // namespace mhx
