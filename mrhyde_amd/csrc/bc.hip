// bc.hip — boundary residual terms of the thermal module on Neumann sides.
//
// Replaces thermal<EvalT>::boundaryResidual for "Neumann" sides
// (src/physics/thermal.cpp:210-222) and the boundary-group half of
// AssemblyManager::assembleJacRes (assemblyManager_jacres.hpp:389-401):
// res_glob[row] += sum over side points of g w_side phi_row, g = q(x) . n.
// g does not depend on u, so the Jacobian is unchanged.
//
// One launch per Neumann side, one thread per lattice node of that side: the
// thread sums the contributions of the (at most 2^(dim-1)) boundary faces
// holding its node, so there are no atomics and the summation order is fixed.
// Side geometry comes from the element's Q1 map restricted to the face (its
// vertices, tangent vectors, surface measure and outward unit normal).

#include <cmath>

#include "mhx_internal.hpp"

namespace mhx {

struct NeumannArgs {
  BoxMesh m;
  Tables1D t;
  const double* vx;
  const uint8_t* fixed;
  double* res;
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
  if (a.fixed[node]) return;
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
  double sum = 0.0;
  const int nf = NT == 1 ? ne_t[0] : ne_t[0] * ne_t[1];
  for (int f = 0; f < nf; ++f) {
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
      double g = 0.0;
      for (int m = 0; m < a.n_terms; ++m) {
        double prod = 1.0;
        for (int d = 0; d < DIM; ++d) prod *= sin(a.freq[m][d] * x[d] + a.phase[m][d]);
        for (int d = 0; d < DIM; ++d) g += a.amp[m][d] * prod * nrm[d];
      }
      sum = fma(g * w * area, phi, sum);
    }
  }
  a.res[node] += sum;
}

void neumann_residual(Ctx& c, double* res) {
  const mhx_problem_desc& d = c.desc;
  if (!res || !d.neumann_sides || d.n_nq_modes <= 0) return;
  ScopedPhase ph(c.prof, "thermal::boundaryResidual", c.stream);
  NeumannArgs a{};
  a.m = c.mesh;
  a.t = c.tab;
  a.vx = c.vx.p;
  a.fixed = c.fixed.p;
  a.res = res;
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

}  // namespace mhx
