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
// Work mapping.  A group of G = next power of two >= NE lanes owns one element
// (3-D: 32 lanes, two elements per wavefront).  Lane j carries derivative
// slot j of the reference's element-level dual (Sacado SFad, one slot per
// DOF, workset.cpp:884): at every cubature point it evaluates the physics in
// a value + one-derivative dual whose seed is basis function j of its field,
// so the derivative part of the point residual is column j of the element
// Jacobian and the value part is the residual.  Geometry and the field values
// at the points are computed once per element into LDS (lanes 0..NQ-1).
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
  static constexpr int G = NE <= 16 ? 16 : 32;
  static constexpr int GPB = 256 / G;
  static constexpr int NQS = DIM * DIM + 1 + NF * (DIM + 1);  // per point: J^-1, w|J|, (u_f, grad u_f)
  static_assert(NE <= G, "element unknowns exceed the lane group");
};

// value + one derivative
struct D1 {
  double v, d;
};
__device__ __forceinline__ D1 operator+(D1 a, D1 b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ D1 operator-(D1 a, D1 b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ D1 operator*(D1 a, D1 b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ D1 operator*(double s, D1 a) { return {s * a.v, s * a.d}; }
__device__ __forceinline__ D1 operator*(D1 a, double s) { return {s * a.v, s * a.d}; }
__device__ __forceinline__ D1 operator-(D1 a, double s) { return {a.v - s, a.d}; }
__device__ __forceinline__ D1 d1_sqrt(D1 a) {
  const double r = sqrt(a.v);
  return {r, a.d * (0.5 / r)};
}
__device__ __forceinline__ D1 d1_inv(D1 a) {
  const double r = 1.0 / a.v;
  return {r, -a.d * r * r};
}

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

// navierstokes<EvalT>::computeTau, steady (C3 = 0)
__device__ __forceinline__ D1 ns_tau(const D1* vel, int dim, double mu, double h) {
  const double C1 = 4.0, C2 = 2.0;
  D1 nvel = vel[0] * vel[0];
  for (int a = 1; a < dim; ++a) nvel = nvel + vel[a] * vel[a];
  if (nvel.v > 1e-12) nvel = d1_sqrt(nvel);
  const double t1 = (C1 * mu / h / h) * (C1 * mu / h / h);
  const D1 c2 = (C2 / h) * nvel;
  D1 t = c2 * c2;
  t.v += t1;
  return d1_inv(d1_sqrt(t));
}

template <int DIM, bool JAC>
__global__ __launch_bounds__(256) void k_ns(AsmArgs a, NsCoef nc) {
  using C = NsCfg<DIM>;
  constexpr int NV = C::NV, NQ = C::NQ, NF = C::NF, NE = C::NE, G = C::G, GPB = C::GPB, NQS = C::NQS;
  __shared__ double s_u[GPB][NE];
  __shared__ double s_X[GPB][NV * DIM];
  __shared__ double s_q[GPB][NQ][NQS];
  __shared__ int s_rp[GPB][NE];

  const int g = threadIdx.x / G, lane = threadIdx.x % G;
  int ec[3];
  const bool active = element_of<DIM>(a, g, 0, ec);
  bool first_lo[3] = {true, true, true}, first_hi[3] = {true, true, true};
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const int k = axis_key(a, d, ec[d]);
    first_lo[d] = ec[d] == 0 || k < axis_key(a, d, ec[d] - 1);
    first_hi[d] = ec[d] == a.m.nel[d] - 1 || k < axis_key(a, d, ec[d] + 1);
  }

  // 1. gather: element unknowns (node-major, field-minor) and vertices
  if (active) {
    if (lane < NE) {
      const int n = lane / NF, f = lane % NF;
      const long dof = node_of<DIM, 1>(a.m, ec, n) * NF + f;
      s_u[g][lane] = a.u[dof];
      s_rp[g][lane] = a.fixed[dof] ? -1 - a.rowptr[dof] : a.rowptr[dof];
    }
    for (int t = lane; t < NV * DIM; t += G) {
      const int v = t / DIM, d = t % DIM;
      const long vid = node_of<DIM, 1>(a.m, ec, v);  // Q1: nodes are the vertices
      s_X[g][t] = a.vx[vid * 3 + d];
    }
  }
  __syncthreads();

  // 2. per point (lane q): J, J^-1, w|J|, field values and physical gradients
  if (active && lane < NQ) {
    const int q = lane;
    double J[DIM][DIM];
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
      for (int s = 0; s < DIM; ++s) J[r][s] = 0.0;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double t[DIM + 1];
      q1_basis<DIM>(v, q, t);
#pragma unroll
      for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int s = 0; s < DIM; ++s) J[r][s] = fma(s_X[g][v * DIM + r], t[s], J[r][s]);
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
    double* o = s_q[g][q];
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
      for (int s = 0; s < DIM; ++s) o[r * DIM + s] = Ji[r][s];
    o[DIM * DIM] = fabs(det);  // Gauss weights are 1
    // fields: value and reference gradient, then physical gradient J^-T grad^
    double fv[NF][DIM + 1];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int c = 0; c <= DIM; ++c) fv[f][c] = 0.0;
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      double t[DIM + 1];
      q1_basis<DIM>(n, q, t);
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const double un = s_u[g][n * NF + f];
#pragma unroll
        for (int c = 0; c <= DIM; ++c) fv[f][c] = fma(un, t[c], fv[f][c]);
      }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      double* of = o + DIM * DIM + 1 + f * (DIM + 1);
      of[DIM] = fv[f][DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < DIM; ++b) s = fma(Ji[b][c], fv[f][b], s);
        of[c] = s;
      }
    }
  }
  __syncthreads();

  // 3. lane j: residual row j and Jacobian column j
  const bool col = active && lane < NE;
  const int jn = lane / NF, jb = lane % NF;
  double R[JAC ? NE : 1];
#pragma unroll
  for (int i = 0; i < (JAC ? NE : 1); ++i) R[i] = 0.0;
  double rres = 0.0;
  if (col) {
    double vol = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) vol += s_q[g][q][DIM * DIM];
    const double h = DIM == 2 ? sqrt(vol) : cbrt(vol);  // Workset::getElementSize
    const bool stab = nc.supg || nc.pspg;
    // a runtime point loop: unrolled, every point's temporaries stay live
#pragma unroll 1
    for (int q = 0; q < NQ; ++q) {
      const double* o = s_q[g][q];
      double Ji[DIM][DIM];
#pragma unroll
      for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int s = 0; s < DIM; ++s) Ji[r][s] = o[r * DIM + s];
      const double w = o[DIM * DIM];
      // seed: basis function jn of field jb (reference and physical gradient)
      double tj[DIM + 1];
      q1_basis<DIM>(jn, q, tj);
      double gj[DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < DIM; ++b) s = fma(Ji[b][c], tj[b], s);
        gj[c] = s;
      }
      D1 U[NF], GU[NF][DIM];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const double* of = o + DIM * DIM + 1 + f * (DIM + 1);
        const bool seeded = JAC && f == jb;
        U[f] = {of[DIM], seeded ? tj[DIM] : 0.0};
#pragma unroll
        for (int c = 0; c < DIM; ++c) GU[f][c] = {of[c], seeded ? gj[c] : 0.0};
      }
      const D1 pr = U[DIM];
      const D1 tau = stab ? ns_tau(U, DIM, nc.mu, h) : D1{0.0, 0.0};
      D1 conv[DIM], stabres[DIM];
#pragma unroll
      for (int m = 0; m < DIM; ++m) {
        conv[m] = U[0] * GU[m][0];
#pragma unroll
        for (int c = 1; c < DIM; ++c) conv[m] = conv[m] + U[c] * GU[m][c];
        stabres[m] = (nc.rho * conv[m] + GU[DIM][m]) - nc.rho * nc.force[m];
      }
      // point residual in reference form: Fhat_f = (J^-1 flux_f, src_f)
      D1 Fh[NF][DIM + 1];
#pragma unroll
      for (int m = 0; m < DIM; ++m) {
        D1 fl[DIM];
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          fl[c] = nc.mu * GU[m][c];
          if (c == m) fl[c] = fl[c] - pr;
          if (nc.supg) fl[c] = fl[c] + tau * stabres[m] * U[c];
          fl[c] = fl[c] * w;
        }
        Fh[m][DIM] = (conv[m] - nc.force[m]) * (nc.rho * w);
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
          D1 s = Ji[r][0] * fl[0];
#pragma unroll
          for (int c = 1; c < DIM; ++c) s = s + Ji[r][c] * fl[c];
          Fh[m][r] = s;
        }
      }
      {
        D1 divu = GU[0][0];
#pragma unroll
        for (int c = 1; c < DIM; ++c) divu = divu + GU[c][c];
        Fh[DIM][DIM] = divu * w;
        D1 fl[DIM];
#pragma unroll
        for (int c = 0; c < DIM; ++c) fl[c] = nc.pspg ? stabres[c] * (tau * (w / nc.rho)) : D1{0.0, 0.0};
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
          D1 s = Ji[r][0] * fl[0];
#pragma unroll
          for (int c = 1; c < DIM; ++c) s = s + Ji[r][c] * fl[c];
          Fh[DIM][r] = s;
        }
      }
      // residual row (jn, jb): value part against test function jn
#pragma unroll
      for (int f = 0; f < NF; ++f)
        if (f == jb)
#pragma unroll
          for (int k = 0; k <= DIM; ++k) rres = fma(Fh[f][k].v, tj[k], rres);
      // Jacobian column j: derivative part against every test function
      if constexpr (JAC) {
#pragma unroll
        for (int in = 0; in < NV; ++in) {
          double ti[DIM + 1];
          q1_basis<DIM>(in, q, ti);
#pragma unroll
          for (int f = 0; f < NF; ++f) {
            double s = R[in * NF + f];
#pragma unroll
            for (int k = 0; k <= DIM; ++k) s = fma(Fh[f][k].d, ti[k], s);
            R[in * NF + f] = s;
          }
        }
      }
    }
  }

  // 4. scatter: first scheduled element touching an entry stores it
  auto shared_first = [&](int ni, int nj) {
    bool f = true;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int il = (ni >> d) & 1, jl = (nj >> d) & 1;
      if (il == jl) f = f && (il == 0 ? first_lo[d] : first_hi[d]);
    }
    return f;
  };
  if (col && a.res) {
    const long row = node_of<DIM, 1>(a.m, ec, jn) * NF + jb;
    const bool fx = s_rp[g][lane] < 0;
    if (a.first_touch && shared_first(jn, jn)) a.res[row] = fx ? 0.0 : -rres;
    else if (!fx) a.res[row] += -rres;
  }
  if constexpr (JAC) {
    if (col && a.jac) {
      const int jl[3] = {jn & 1, (jn >> 1) & 1, (jn >> 2) & 1};
      constexpr int CH = 8;
#pragma unroll
      for (int i0 = 0; i0 < NE; i0 += CH) {
        int sl[CH];
        uint32_t first_mask = 0, fixed_mask = 0;
        double old[CH];
#pragma unroll
        for (int k = 0; k < CH; ++k) {
          const int i = i0 + k;
          if (i >= NE) break;
          const int in = i / NF;
          const int il[3] = {in & 1, (in >> 1) & 1, (in >> 2) & 1};
          int px, lx, py, ly, pz = 0, lz = 1;
          axis_slot<1>(il[0], jl[0], ec[0], a.m.nn[0], px, lx);
          axis_slot<1>(il[1], jl[1], ec[1], a.m.nn[1], py, ly);
          if (DIM == 3) axis_slot<1>(il[2], jl[2], ec[2], a.m.nn[2], pz, lz);
          int rp = s_rp[g][i];
          const bool fx = rp < 0;
          if (fx) rp = -1 - rp;
          sl[k] = rp + ((pz * ly + py) * lx + px) * NF + jb;
          const bool first = a.first_touch && shared_first(in, jn);
          if (first) first_mask |= 1u << k;
          if (fx) fixed_mask |= 1u << k;
          old[k] = (!first && !fx) ? a.jac[sl[k]] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < CH; ++k) {
          const int i = i0 + k;
          if (i >= NE) break;
          const bool first = (first_mask >> k) & 1u, fx = (fixed_mask >> k) & 1u;
          // streaming stores (a later colour re-reads the entry long after this launch)
          if (fx) {
            if (first) __builtin_nontemporal_store((i == lane) ? 1.0 : 0.0, &a.jac[sl[k]]);
          } else {
            __builtin_nontemporal_store(old[k] + R[i], &a.jac[sl[k]]);
          }
        }
      }
    }
  }
}

template <int DIM>
static void launch_ns(Ctx& c, AsmArgs& a, const NsCoef& nc, bool jac) {
  using C = NsCfg<DIM>;
  const int ncol = 1 << DIM;
  a.gpb = C::GPB;
  a.blocked = 0;
  a.nsub = 1;
  for (int col = 0; col < ncol; ++col) {
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
    if (cnt == 0) continue;
    const long blocks = (cnt + C::GPB - 1) / C::GPB;
    ScopedPhase ph(c.prof, jac ? "k_ns<jacres>" : "k_ns<res>", c.stream);
    if (jac)
      hipLaunchKernelGGL((k_ns<DIM, true>), dim3((unsigned)blocks), dim3(256), 0, c.stream, a, nc);
    else
      hipLaunchKernelGGL((k_ns<DIM, false>), dim3((unsigned)blocks), dim3(256), 0, c.stream, a, nc);
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
    if (c.mesh.dim == 2) launch_ns<2>(c, a, nc, want_jac);
    else launch_ns<3>(c, a, nc, want_jac);
  }
  if (accumulate && want_jac && !c.h_dbc_list.empty()) {
    ScopedPhase ph(c.prof, "AssemblyManager::dofConstraints", c.stream);
    const long n = (long)c.h_dbc_list.size();
    hipLaunchKernelGGL(k_dbc_diag_ns, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, c.dbc_list.p, n,
                       c.rowptr.p, c.colidx.p, jac);
    MHX_HIP(hipGetLastError());
  }
}

}  // namespace mhx
