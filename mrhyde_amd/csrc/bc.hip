// bc.hip — boundary residual terms on Neumann sides: the thermal flux and
// the Navier-Stokes traction.
//
// Replaces thermal<EvalT>::boundaryResidual for "Neumann" sides
// (src/physics/thermal.cpp:210-222), navierstokes<EvalT>::boundaryResidual
// for "Neumann" (traction) sides of the velocity components
// (src/physics/navierstokes.cpp:853-1000), and the boundary-group half of
// AssemblyManager::assembleJacRes (assemblyManager_jacres.hpp:389-401):
//   thermal:        res_glob[row] += sum over side points of g w_side phi_row, g = q(x) . n;
//   Navier-Stokes:  res_glob[node, c] += sum over side points of q_c(x) w_side phi_node,
//                   c = velocity component (the reference's "Neumann u<c> <side>").
// The data do not depend on u, so the Jacobian is unchanged.
//
// One launch per Neumann side, one thread per lattice node of that side: the
// thread sums the contributions of the (at most 2^(dim-1)) boundary faces
// holding its node, so there are no atomics and the summation order is fixed.
// Side geometry comes from the element's Q1 map restricted to the face (its
// vertices, tangent vectors, surface measure and outward unit normal).

#include <cmath>

#include "dual.hpp"
#include "lattice.hpp"
#include "mhx_internal.hpp"

namespace mhx {

struct NeumannArgs {
  BoxMesh m;
  Tables1D t;
  const double* vx;
  const uint8_t* fixed;
  double* res;
  int nf;    // unknowns per node: 1 thermal (flux q . n), dim + 1 Navier-Stokes (traction q)
  int side;  // 2*axis + (high side)
  int n_terms;
  double amp[MHX_MAX_NQ_MODES][3], freq[MHX_MAX_NQ_MODES][3], phase[MHX_MAX_NQ_MODES][3];
};

template <int DIM>
__global__ __launch_bounds__(256) void k_neumann(NeumannArgs a) {
  constexpr int NT = DIM - 1;  // tangential axes
  const int ax = a.side / 2;
  const bool hi = a.side & 1;
  int ta[2] = {ax == 0 ? 1 : 0, ax == 2 ? 1 : 2};
  long n_side = 1;
  for (int t = 0; t < NT; ++t) n_side *= a.m.nn[ta[t]];
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n_side) return;
  const int p = a.m.p, nq1 = a.m.nq1;
  int c[3];
  c[ax] = hi ? a.m.nn[ax] - 1 : 0;
  c[ta[0]] = (int)(tid % a.m.nn[ta[0]]);
  if (NT == 2) c[ta[1]] = (int)(tid / a.m.nn[ta[0]]);
  if (DIM == 2) c[2] = 0;
  const long node = c[0] + (long)a.m.nn[0] * (c[1] + (long)a.m.nn[1] * c[2]);
  const int ncomp = a.nf == 1 ? 1 : DIM;  // residual rows of the node this side loads
  bool any = false;
  for (int k = 0; k < ncomp; ++k) any = any || !a.fixed[node * a.nf + k];
  if (!any) return;
  // candidate elements along each tangential axis
  int e0[2], ne_t[2];
  for (int t = 0; t < NT; ++t) {
    const int cb = c[ta[t]], nel = a.m.nel[ta[t]];
    if (cb % p == 0) {
      e0[t] = cb / p - 1 < 0 ? 0 : cb / p - 1;
      const int e1 = cb / p > nel - 1 ? nel - 1 : cb / p;
      ne_t[t] = e1 - e0[t] + 1;
    } else {
      e0[t] = cb / p;
      ne_t[t] = 1;
    }
  }
  double sum[3] = {0.0, 0.0, 0.0};
  const int nfc = NT == 1 ? ne_t[0] : ne_t[0] * ne_t[1];
  for (int f = 0; f < nfc; ++f) {
    int ec[3];
    ec[ax] = hi ? a.m.nel[ax] - 1 : 0;
    ec[ta[0]] = e0[0] + f % ne_t[0];
    if (NT == 2) ec[ta[1]] = e0[1] + f / ne_t[0];
    if (DIM == 2) ec[2] = 0;
    int lt[2];
    for (int t = 0; t < NT; ++t) lt[t] = c[ta[t]] - p * ec[ta[t]];
    // the face's vertices (bit ax fixed to the side), in face-local order
    double X[4][3];
    for (int v = 0; v < (1 << NT); ++v) {
      int vc[3] = {ec[0], ec[1], ec[2]};
      vc[ax] += hi ? 1 : 0;
      vc[ta[0]] += v & 1;
      if (NT == 2) vc[ta[1]] += (v >> 1) & 1;
      const long vid = vc[0] + (long)a.m.nvx[0] * (vc[1] + (long)a.m.nvx[1] * vc[2]);
      for (int d = 0; d < 3; ++d) X[v][d] = a.vx[vid * 3 + d];
    }
    const int nqs = NT == 1 ? nq1 : nq1 * nq1;
    for (int qs = 0; qs < nqs; ++qs) {
      const int qq[2] = {qs % nq1, qs / nq1};
      double phi = 1.0, w = 1.0;
      for (int t = 0; t < NT; ++t) {
        phi *= a.t.B[lt[t]][qq[t]];
        w *= a.t.W[qq[t]];
      }
      double x[3] = {0, 0, 0}, tg[2][3] = {{0, 0, 0}, {0, 0, 0}};
      for (int v = 0; v < (1 << NT); ++v) {
        const int vb[2] = {v & 1, (v >> 1) & 1};
        double N = 1.0, dN[2] = {1.0, 1.0};
        for (int t = 0; t < NT; ++t) {
          N *= a.t.N[vb[t]][qq[t]];
          for (int u = 0; u < NT; ++u) dN[u] *= (u == t) ? a.t.M[vb[t]][qq[t]] : a.t.N[vb[t]][qq[t]];
        }
        for (int d = 0; d < 3; ++d) {
          x[d] = fma(N, X[v][d], x[d]);
          for (int t = 0; t < NT; ++t) tg[t][d] = fma(dN[t], X[v][d], tg[t][d]);
        }
      }
      double nrm[3] = {0, 0, 0}, area;
      if (DIM == 2) {
        area = sqrt(tg[0][0] * tg[0][0] + tg[0][1] * tg[0][1]);
        nrm[0] = tg[0][1] / area;
        nrm[1] = -tg[0][0] / area;
      } else {
        nrm[0] = tg[0][1] * tg[1][2] - tg[0][2] * tg[1][1];
        nrm[1] = tg[0][2] * tg[1][0] - tg[0][0] * tg[1][2];
        nrm[2] = tg[0][0] * tg[1][1] - tg[0][1] * tg[1][0];
        area = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        for (int d = 0; d < 3; ++d) nrm[d] /= area;
      }
      if (nrm[ax] * (hi ? 1.0 : -1.0) < 0.0)
        for (int d = 0; d < 3; ++d) nrm[d] = -nrm[d];
      double g[3] = {0.0, 0.0, 0.0};
      for (int m = 0; m < a.n_terms; ++m) {
        double prod = 1.0;
        for (int d = 0; d < DIM; ++d) prod *= sin(a.freq[m][d] * x[d] + a.phase[m][d]);
        if (a.nf == 1) {
          for (int d = 0; d < DIM; ++d) g[0] += a.amp[m][d] * prod * nrm[d];
        } else {
          for (int d = 0; d < DIM; ++d) g[d] += a.amp[m][d] * prod;
        }
      }
      for (int k = 0; k < ncomp; ++k) sum[k] = fma(g[k] * w * area, phi, sum[k]);
    }
  }
  for (int k = 0; k < ncomp; ++k)
    if (!a.fixed[node * a.nf + k]) a.res[node * a.nf + k] += sum[k];
}

void neumann_residual(Ctx& c, double* res) {
  const mhx_problem_desc& d = c.desc;
  if (!res || !d.neumann_sides || d.n_nq_modes <= 0) return;
  ScopedPhase ph(c.prof, c.mesh.nf == 1 ? "thermal::boundaryResidual" : "navierstokes::boundaryResidual",
                 c.stream);
  NeumannArgs a{};
  a.m = c.mesh;
  a.t = c.tab;
  a.vx = c.vx.p;
  a.fixed = c.fixed.p;
  a.res = res;
  a.nf = c.mesh.nf;
  a.n_terms = d.n_nq_modes;
  for (int m = 0; m < d.n_nq_modes; ++m)
    for (int k = 0; k < 3; ++k) {
      a.amp[m][k] = d.nq_amp[m][k];
      a.freq[m][k] = d.nq_freq[m][k];
      a.phase[m][k] = d.nq_phase[m][k];
    }
  for (int s = 0; s < 2 * c.mesh.dim; ++s) {
    if (!(d.neumann_sides & (1u << s))) continue;
    a.side = s;
    const int ax = s / 2;
    long n_side = 1;
    for (int b = 0; b < c.mesh.dim; ++b)
      if (b != ax) n_side *= c.mesh.nn[b];
    const dim3 grid((unsigned)((n_side + 255) / 256));
    if (c.mesh.dim == 2) hipLaunchKernelGGL(k_neumann<2>, grid, dim3(256), 0, c.stream, a);
    else hipLaunchKernelGGL(k_neumann<3>, grid, dim3(256), 0, c.stream, a);
    MHX_HIP(hipGetLastError());
  }
}

// ---------------------------------------------------------------------------
// Weak Dirichlet (Nitsche) sides: thermal<EvalT>::boundaryResidual, "weak
// Dirichlet" (src/physics/thermal.cpp:226-281):
//   res(e,dof) += epen/h k (T - g) w phi - k (grad T . n) w phi - sf k (T - g) w (grad phi . n)
// with epen = 10, h = (sum of the side weights)^(1/(dim-1))
// (Workset::getSideElementSize, workset.cpp:2719-2733), sf = "form_param".
// T and grad T depend on every element DOF, so the side also adds an element
// matrix to J.  One wavefront per boundary element side; one launch per
// (side, colour), the colour being the parity of the element's tangential
// indices, so the elements of one launch share no DOF and the read-modify-
// write scatter needs no atomics.  Runs after the volume kernel (which has
// already written every entry).
//   lanes = side points: geometry, then T, grad T and k(T) in dual arithmetic
//     with d+1 slots (d/d grad T, d/dT), giving per point the coefficients
//     F = (b, a) of (grad phi, phi) in the residual integrand and their
//     derivatives D (the side analogue of the volume kernel's Dhat);
//   lanes = element DOFs j: column j of the side matrix, sum_q G_i^T D_q G_j,
//     and residual row j, sum_q F_q . G_j.
// ---------------------------------------------------------------------------
struct NitscheArgs {
  BoxMesh m;
  Tables1D t;
  double Ls[2][kMaxP1], Ds[2][kMaxP1];  // Lagrange values / slopes at xi = -1 and xi = +1
  const double* vx;
  const double* u;
  const double* k0;  // [ne] or nullptr (== 1)
  double diff_u2;
  const int32_t* rowptr;
  const uint8_t* fixed;
  double* res;
  double* jac;  // or nullptr (residual only)
  int side;     // 2*axis + (high side)
  int colour;   // parity bits of the tangential element indices
  int cn[2];    // elements of this colour per tangential axis
  double sf;
  double wd_const, wd_lin[3];
  int n_terms;
  double wd_amp[MHX_MAX_NQ_MODES], wd_freq[MHX_MAX_NQ_MODES][3], wd_phase[MHX_MAX_NQ_MODES][3];
};

template <int DIM, int P>
__global__ __launch_bounds__(64) void k_nitsche(NitscheArgs a) {
  constexpr int P1 = P + 1, NT = DIM - 1, NV = 1 << DIM, NC = DIM + 1;
  constexpr int NB = DIM == 2 ? P1 * P1 : P1 * P1 * P1;
  constexpr int NQS = NT == 1 ? kMaxQ1 : kMaxQ1 * kMaxQ1;
  static_assert(NB <= 64, "one wavefront per element side");
  __shared__ double s_u[NB];
  __shared__ double s_X[NV][3];
  __shared__ double s_w[NQS];
  __shared__ double s_Ji[NQS][DIM][DIM];
  __shared__ double s_F[NQS][NC];
  __shared__ double s_D[NQS][NC][NC];
  const int lane = threadIdx.x;
  const int ax = a.side / 2;
  const int hi = a.side & 1;
  const int ta[2] = {ax == 0 ? 1 : 0, ax == 2 ? 1 : 2};
  int ec[3] = {0, 0, 0};
  ec[ax] = hi ? a.m.nel[ax] - 1 : 0;
  {
    const int b = blockIdx.x;
    ec[ta[0]] = 2 * (b % a.cn[0]) + (a.colour & 1);
    if (NT == 2) ec[ta[1]] = 2 * (b / a.cn[0]) + ((a.colour >> 1) & 1);
  }
  const long e = ec[0] + (long)a.m.nel[0] * (ec[1] + (long)a.m.nel[1] * ec[2]);
  if (lane < NB) s_u[lane] = a.u[node_of<DIM, P>(a.m, ec, lane)];
  if (lane < NV * 3) {
    const int v = lane / 3, d = lane % 3;
    const long vid = (long)(ec[0] + (v & 1)) +
                     (long)a.m.nvx[0] * ((long)(ec[1] + ((v >> 1) & 1)) + (long)a.m.nvx[1] * (long)(ec[2] + ((v >> 2) & 1)));
    s_X[v][d] = a.vx[vid * 3 + d];
  }
  __syncthreads();
  const int nq1 = a.m.nq1;
  const int nqs = NT == 1 ? nq1 : nq1 * nq1;
  // 1-D factors of side point q along every axis: basis values/slopes (order
  // P) and the linear geometry map's values/slopes
  auto factors = [&](int q, double (*L)[P1], double (*dL)[P1], double (*N)[2], double (*M)[2], double& wt) {
    const int qq[2] = {q % nq1, q / nq1};
    wt = 1.0;
    for (int b = 0; b < DIM; ++b) {
      if (b == ax) {
        for (int k = 0; k < P1; ++k) {
          L[b][k] = a.Ls[hi][k];
          dL[b][k] = a.Ds[hi][k];
        }
        N[b][0] = hi ? 0.0 : 1.0;
        N[b][1] = hi ? 1.0 : 0.0;
        M[b][0] = -0.5;
        M[b][1] = 0.5;
      } else {
        const int t = b == ta[0] ? 0 : 1;
        const int qb = qq[t];
        for (int k = 0; k < P1; ++k) {
          L[b][k] = a.t.B[k][qb];
          dL[b][k] = a.t.D[k][qb];
        }
        for (int k = 0; k < 2; ++k) {
          N[b][k] = a.t.N[k][qb];
          M[b][k] = a.t.M[k][qb];
        }
        wt *= a.t.W[qb];
      }
    }
  };
  double nrm[3] = {0, 0, 0}, xq[3] = {0, 0, 0};
  if (lane < nqs) {
    double L[DIM][P1], dL[DIM][P1], N[DIM][2], M[DIM][2], wt;
    factors(lane, L, dL, N, M, wt);
    double J[DIM][DIM];
    for (int r = 0; r < DIM; ++r)
      for (int s = 0; s < DIM; ++s) J[r][s] = 0.0;
    for (int v = 0; v < NV; ++v) {
      const int vl[3] = {v & 1, (v >> 1) & 1, (v >> 2) & 1};
      double nv = 1.0;
      for (int b = 0; b < DIM; ++b) nv *= N[b][vl[b]];
      for (int r = 0; r < DIM; ++r) {
        xq[r] = fma(nv, s_X[v][r], xq[r]);
        for (int s = 0; s < DIM; ++s) {
          double dn = 1.0;
          for (int b = 0; b < DIM; ++b) dn *= (b == s) ? M[b][vl[b]] : N[b][vl[b]];
          J[r][s] = fma(s_X[v][r], dn, J[r][s]);
        }
      }
    }
    // tangents dx/dxi_t are the columns ta[t] of J; area and outward normal
    double area;
    if (DIM == 2) {
      const double t0 = J[0][ta[0]], t1 = J[1][ta[0]];
      area = sqrt(t0 * t0 + t1 * t1);
      nrm[0] = t1 / area;
      nrm[1] = -t0 / area;
    } else {
      double tg[2][3];
      for (int t = 0; t < 2; ++t)
        for (int r = 0; r < 3; ++r) tg[t][r] = J[r][ta[t]];
      nrm[0] = tg[0][1] * tg[1][2] - tg[0][2] * tg[1][1];
      nrm[1] = tg[0][2] * tg[1][0] - tg[0][0] * tg[1][2];
      nrm[2] = tg[0][0] * tg[1][1] - tg[0][1] * tg[1][0];
      area = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
      for (int r = 0; r < 3; ++r) nrm[r] /= area;
    }
    if (nrm[ax] * (hi ? 1.0 : -1.0) < 0.0)
      for (int r = 0; r < 3; ++r) nrm[r] = -nrm[r];
    s_w[lane] = wt * area;
    double Ji[DIM][DIM];
    if constexpr (DIM == 2) {
      const double id = 1.0 / (J[0][0] * J[1][1] - J[0][1] * J[1][0]);
      Ji[0][0] = J[1][1] * id;
      Ji[0][1] = -J[0][1] * id;
      Ji[1][0] = -J[1][0] * id;
      Ji[1][1] = J[0][0] * id;
    } else {
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      const double id = 1.0 / (J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02);
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
    for (int r = 0; r < DIM; ++r)
      for (int s = 0; s < DIM; ++s) s_Ji[lane][r][s] = Ji[r][s];
  }
  __syncthreads();
  double h = 0.0;
  for (int q = 0; q < nqs; ++q) h += s_w[q];
  if (DIM == 3) h = sqrt(h);
  if (lane < nqs) {
    double L[DIM][P1], dL[DIM][P1], N[DIM][2], M[DIM][2], wt;
    factors(lane, L, dL, N, M, wt);
    double T = 0.0, gr[DIM];
    for (int c = 0; c < DIM; ++c) gr[c] = 0.0;
    for (int j = 0; j < NB; ++j) {
      const int jl[3] = {j % P1, (j / P1) % P1, DIM == 3 ? j / (P1 * P1) : 0};
      double ph = 1.0;
      for (int b = 0; b < DIM; ++b) ph *= L[b][jl[b]];
      T = fma(s_u[j], ph, T);
      for (int c = 0; c < DIM; ++c) {
        double gb = 1.0;
        for (int b = 0; b < DIM; ++b) gb *= (b == c) ? dL[b][jl[b]] : L[b][jl[b]];
        gr[c] = fma(s_u[j], gb, gr[c]);
      }
    }
    // physical gradient, then the duals: slots 0..DIM-1 = d/d(grad T)_c, DIM = d/dT
    QDual<NC> Td = qd_seed<NC>(T, DIM), G[DIM];
    for (int c = 0; c < DIM; ++c) {
      double s = 0.0;
      for (int b = 0; b < DIM; ++b) s = fma(s_Ji[lane][b][c], gr[b], s);
      G[c] = qd_seed<NC>(s, c);
    }
    const double k0 = a.k0 ? a.k0[e] : 1.0;
    const QDual<NC> k = a.diff_u2 != 0.0 ? k0 * (1.0 + a.diff_u2 * (Td * Td)) : qd_const<NC>(k0);
    double g = a.wd_const;
    for (int b = 0; b < DIM; ++b) g = fma(a.wd_lin[b], xq[b], g);
    for (int m = 0; m < a.n_terms; ++m) {
      double prod = a.wd_amp[m];
      for (int b = 0; b < DIM; ++b) prod *= sin(a.wd_freq[m][b] * xq[b] + a.wd_phase[m][b]);
      g += prod;
    }
    const QDual<NC> Tm = (-g) + Td;
    QDual<NC> dTn = nrm[0] * G[0];
    for (int c = 1; c < DIM; ++c) dTn = dTn + nrm[c] * G[c];
    const double w = s_w[lane];
    const double epen = 10.0;
    // a: coefficient of phi; b_c: coefficient of (grad phi)_c
    const QDual<NC> kTm = k * Tm;
    const QDual<NC> aa = (w * epen / h) * kTm + (-w) * (k * dTn);
    for (int c = 0; c < DIM; ++c) {
      const QDual<NC> bc = (-w * a.sf * nrm[c]) * kTm;
      s_F[lane][c] = bc.v;
      for (int s = 0; s < NC; ++s) s_D[lane][c][s] = bc.d[s];
    }
    s_F[lane][DIM] = aa.v;
    for (int s = 0; s < NC; ++s) s_D[lane][DIM][s] = aa.d[s];
  }
  __syncthreads();
  if (lane >= NB) return;
  // column / residual row `lane`
  const int jl[3] = {lane % P1, (lane / P1) % P1, DIM == 3 ? lane / (P1 * P1) : 0};
  double R[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) R[i] = 0.0;
  double rres = 0.0;
  auto gvec = [&](double (*L)[P1], double (*dL)[P1], double (*Ji)[DIM], const int* il, double* Gv) {
    double gref[DIM];
    double ph = 1.0;
    for (int b = 0; b < DIM; ++b) ph *= L[b][il[b]];
    for (int c = 0; c < DIM; ++c) {
      double gb = 1.0;
      for (int b = 0; b < DIM; ++b) gb *= (b == c) ? dL[b][il[b]] : L[b][il[b]];
      gref[c] = gb;
    }
    for (int c = 0; c < DIM; ++c) {
      double s = 0.0;
      for (int b = 0; b < DIM; ++b) s = fma(Ji[b][c], gref[b], s);
      Gv[c] = s;
    }
    Gv[DIM] = ph;
  };
  for (int q = 0; q < nqs; ++q) {
    double L[DIM][P1], dL[DIM][P1], N[DIM][2], M[DIM][2], wt;
    factors(q, L, dL, N, M, wt);
    double Gj[NC];
    gvec(L, dL, s_Ji[q], jl, Gj);
    for (int r = 0; r < NC; ++r) rres = fma(s_F[q][r], Gj[r], rres);
    if (a.jac) {
      double f[NC];
      for (int r = 0; r < NC; ++r) {
        double s = 0.0;
        for (int c = 0; c < NC; ++c) s = fma(s_D[q][r][c], Gj[c], s);
        f[r] = s;
      }
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int il[3] = {i % P1, (i / P1) % P1, DIM == 3 ? i / (P1 * P1) : 0};
        double Gi[NC];
        gvec(L, dL, s_Ji[q], il, Gi);
        double s = 0.0;
        for (int r = 0; r < NC; ++r) s = fma(f[r], Gi[r], s);
        R[i] += s;
      }
    }
  }
  // scatter (global residual = -F, J += dF/du; strong Dirichlet rows skipped)
  const long rowj = node_of<DIM, P>(a.m, ec, lane);
  if (a.res && !a.fixed[rowj]) a.res[rowj] -= rres;
  if (a.jac) {
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int il[3] = {i % P1, (i / P1) % P1, DIM == 3 ? i / (P1 * P1) : 0};
      const long row = node_of<DIM, P>(a.m, ec, i);
      if (a.fixed[row]) continue;
      int px, lx, py, ly, pz = 0, lz = 1;
      axis_slot<P>(il[0], jl[0], ec[0], a.m.nn[0], px, lx);
      axis_slot<P>(il[1], jl[1], ec[1], a.m.nn[1], py, ly);
      if (DIM == 3) axis_slot<P>(il[2], jl[2], ec[2], a.m.nn[2], pz, lz);
      a.jac[a.rowptr[row] + (pz * ly + py) * lx + px] += R[i];
    }
  }
}

bool nitsche_supported(int dim, int p) { return (dim == 2 && p >= 1 && p <= 6) || (dim == 3 && p >= 1 && p <= 2); }

template <int DIM, int P>
static void launch_nitsche(Ctx& c, NitscheArgs& a) {
  const int ncol = 1 << (DIM - 1);
  const int ax = a.side / 2;
  const int ta[2] = {ax == 0 ? 1 : 0, ax == 2 ? 1 : 2};
  for (int col = 0; col < ncol; ++col) {
    a.colour = col;
    long cnt = 1;
    for (int t = 0; t < DIM - 1; ++t) {
      const int par = (col >> t) & 1;
      a.cn[t] = (c.mesh.nel[ta[t]] - par + 1) / 2;
      cnt *= a.cn[t];
    }
    if (DIM == 2) a.cn[1] = 1;
    if (cnt == 0) continue;
    hipLaunchKernelGGL((k_nitsche<DIM, P>), dim3((unsigned)cnt), dim3(64), 0, c.stream, a);
    MHX_HIP(hipGetLastError());
  }
}

void nitsche_boundary(Ctx& c, const double* u, double* res, double* jac) {
  const mhx_problem_desc& d = c.desc;
  if (!d.wdbc_sides || (!res && !jac)) return;
  MHX_REQUIRE(nitsche_supported(c.mesh.dim, c.mesh.p), MHX_ERR_UNSUPPORTED,
              "weak Dirichlet sides: no kernel built for this dimension and order");
  ScopedPhase ph(c.prof, c.mesh.nf == 1 ? "thermal::boundaryResidual" : "navierstokes::boundaryResidual",
                 c.stream);
  NitscheArgs a{};
  a.m = c.mesh;
  a.t = c.tab;
  for (int s = 0; s < 2; ++s) lagrange_equispaced(c.mesh.p, s ? 1.0 : -1.0, a.Ls[s], a.Ds[s]);
  a.vx = c.vx.p;
  a.u = u;
  a.k0 = c.k0.p;
  a.diff_u2 = d.diff_u2;
  a.rowptr = c.rowptr.p;
  a.fixed = c.fixed.p;
  a.res = res;
  a.jac = jac;
  a.sf = d.wdbc_form == 1 ? -1.0 : d.wdbc_form == 2 ? 0.0 : 1.0;
  a.wd_const = d.wd_const;
  for (int b = 0; b < 3; ++b) a.wd_lin[b] = d.wd_lin[b];
  a.n_terms = d.n_wd_modes;
  for (int m = 0; m < d.n_wd_modes; ++m) {
    a.wd_amp[m] = d.wd_amp[m];
    for (int b = 0; b < 3; ++b) {
      a.wd_freq[m][b] = d.wd_freq[m][b];
      a.wd_phase[m][b] = d.wd_phase[m][b];
    }
  }
  for (int s = 0; s < 2 * c.mesh.dim; ++s) {
    if (!(d.wdbc_sides & (1u << s))) continue;
    a.side = s;
    const int key = c.mesh.dim * 10 + c.mesh.p;
    switch (key) {
      case 21: launch_nitsche<2, 1>(c, a); break;
      case 22: launch_nitsche<2, 2>(c, a); break;
      case 23: launch_nitsche<2, 3>(c, a); break;
      case 24: launch_nitsche<2, 4>(c, a); break;
      case 25: launch_nitsche<2, 5>(c, a); break;
      case 26: launch_nitsche<2, 6>(c, a); break;
      case 31: launch_nitsche<3, 1>(c, a); break;
      default: launch_nitsche<3, 2>(c, a); break;
    }
  }
}

}  // namespace mhx
