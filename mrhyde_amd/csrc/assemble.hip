// assemble.hip — fused residual/Jacobian assembly for the thermal module on
// structured quad/hex meshes of any Lagrange order, CDNA4 (gfx950).
//
// Replaces, as one fused kernel per element colour:
//   AssemblyManager::performGather          assemblyManager_gather.hpp:183-234
//   Workset::computeSolnSteadySeeded        workset.cpp:865-905
//   Workset::evaluateSolutionField          workset.cpp:979-1115
//   FunctionManager::evaluate ("thermal diffusion", "thermal source")  functionManager_evaluate.hpp:14-233
//   thermal<EvalT>::volumeResidual          thermal.cpp:71-165
//   AssemblyManager::scatter                assemblyManager_scatter.hpp:164-278
//   AssemblyManager::dofConstraints         assemblyManager_constraints.hpp:241-270 (separate tiny kernel)
//
// Work mapping.  A group of G lanes (G = next power of two >= NB+1) owns one
// element; 64/G elements share a wavefront, 256/G a workgroup.
//   1. gather: lanes load the NB element DOF values and the 2^d vertex
//      coordinates into LDS.
//   2. point stage (lane = cubature point): u, grad u, the geometry Jacobian,
//      det J and J^-1 at the point; then the physics flux/source evaluated in
//      forward-mode dual arithmetic with d+1 derivative slots (d/d grad u,
//      d/du).  Sacado-Fad semantics: the element Jacobian column j follows by
//      the chain rule, dF/du_j = dF/d(grad u) . grad phi_j + dF/du phi_j.
//      The point writes  Dhat = w|J| T D T^T  and  Fhat = w|J| T (flux, src)
//      (T = diag(J^-1, 1)) to LDS, so the remaining stages work in reference
//      coordinates where the basis is a tensor product of 1-D tables.
//   3. column stage (lane = derivative slot): lane j < NB forms
//      f_q = Dhat_q (grad^ phi_j, phi_j)(q) and contracts it against every test
//      function with sum factorisation (three 1-D passes, O(p^{d+1} q) instead
//      of O(p^d q^d)), giving column j of the element Jacobian.  Lane NB does
//      the same with Fhat, giving the residual (the dual's value slot).
//   4. scatter: lane j adds column j into the CRS rows of the element's DOFs.
//      Elements of one colour share no DOF, so the read-modify-write needs no
//      atomics.  Slots come from lattice arithmetic plus rowptr; strong
//      Dirichlet rows are skipped (scatter.hpp:208,253).
// Global residual = -F (scatter.hpp:227); Jacobian = +dF/du (:261).

#include <cmath>

#include "thermal_kernel.hpp"

namespace mhx {

// strong Dirichlet rows: J(d,d) = 1 (setJacobianConstraints, constraints.hpp:125-141)
__global__ void k_dbc_diag(BoxMesh m, const int32_t* __restrict__ list, long n, const int32_t* __restrict__ rowptr,
                           const int32_t* __restrict__ colidx, double* __restrict__ jac) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int row = list[i];
  int lo = rowptr[row], hi = rowptr[row + 1];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (colidx[mid] < row) lo = mid + 1;
    else hi = mid;
  }
  jac[lo] = 1.0;
}

// ---------------------------------------------------------------------------
// host launch
// ---------------------------------------------------------------------------
// groups per axis for block colouring: split GPB over the axes, doubling the
// currently smallest one (x first)
static void block_shape(int dim, int gpb, int* s) {
  s[0] = s[1] = s[2] = 1;
  int prod = 1;
  while (prod < gpb) {
    int d = 0;
    for (int k = 1; k < dim; ++k)
      if (s[k] < s[d]) d = k;
    s[d] *= 2;
    prod *= 2;
  }
}

template <int DIM, int P, int NQ1>
static void launch_colors(Ctx& c, AsmArgs& a, bool jac) {
  using C = Cfg<DIM, P, NQ1>;
  const int ncol = 1 << DIM;
  // the transient variant is built for element colouring only, and so is
  // every order when pencil streaming is selected but does not apply
  const int mode = (a.hist || c.scatter_mode == kPencilStream) ? (int)kElementColouring : c.scatter_mode;
  // the segment schedule needs two elements per workgroup (even and odd
  // phases); every built order has at least four
  static_assert(C::GPB >= 2, "x-segments need two elements per workgroup");
  a.gpb = C::GPB;
  a.blocked = mode;
  a.nsub = mode == kBlockColouring ? ncol : 1;
  a.nseg = (c.mesh.nel[0] + C::GPB - 1) / C::GPB;
  block_shape(DIM, C::GPB, a.s);
  for (int d = 0; d < 3; ++d) a.bw[d] = 2 * a.s[d];
  for (int col = 0; col < ncol; ++col) {
    a.color = col;
    long cnt = 1;
    for (int d = 0; d < 3; ++d) {
      const int cd = (col >> d) & 1;
      // units along the axis: elements (element colouring), blocks (block
      // colouring) or, along x, segments (segment scheduling)
      int units = d < DIM ? c.mesh.nel[d] : 1;
      if (d < DIM && mode == kBlockColouring) units = (c.mesh.nel[d] + a.bw[d] - 1) / a.bw[d];
      if (d == 0 && mode == kSegmentMerge) units = a.nseg;
      int n = d < DIM ? (units - cd + 1) / 2 : 1;
      if (d >= DIM && cd) n = 0;
      a.cn[d] = n;
      cnt *= n;
    }
    a.ncolor = cnt;
    if (cnt == 0) continue;
    const long blocks = mode == kElementColouring ? (cnt + C::GPB - 1) / C::GPB : cnt;
    ScopedPhase ph(c.prof, jac ? "k_thermal<jacres>" : "k_thermal<res>", c.stream);
    const dim3 grid((unsigned)blocks), block(C::BLOCK);
    if (mode == kSegmentMerge) {
      if (jac) hipLaunchKernelGGL((k_thermal<DIM, P, NQ1, true, kSegmentMerge>), grid, block, 0, c.stream, a);
      else hipLaunchKernelGGL((k_thermal<DIM, P, NQ1, false, kSegmentMerge>), grid, block, 0, c.stream, a);
    } else if (mode == kBlockColouring) {
      if (jac) hipLaunchKernelGGL((k_thermal<DIM, P, NQ1, true, kBlockColouring>), grid, block, 0, c.stream, a);
      else hipLaunchKernelGGL((k_thermal<DIM, P, NQ1, false, kBlockColouring>), grid, block, 0, c.stream, a);
    } else if (a.hist) {
      if (jac) hipLaunchKernelGGL((k_thermal<DIM, P, NQ1, true, kElementColouring, true>), grid, block, 0, c.stream, a);
      else hipLaunchKernelGGL((k_thermal<DIM, P, NQ1, false, kElementColouring, true>), grid, block, 0, c.stream, a);
    } else {
      if (jac) hipLaunchKernelGGL((k_thermal<DIM, P, NQ1, true, kElementColouring>), grid, block, 0, c.stream, a);
      else hipLaunchKernelGGL((k_thermal<DIM, P, NQ1, false, kElementColouring>), grid, block, 0, c.stream, a);
    }
    MHX_HIP(hipGetLastError());
  }
}

using Launcher = void (*)(Ctx&, AsmArgs&, bool);
struct Entry {
  int dim, p, nq1;
  Launcher f;
};
static const Entry kTable[] = {
    {2, 1, 1, launch_colors<2, 1, 1>}, {2, 1, 2, launch_colors<2, 1, 2>}, {2, 1, 3, launch_colors<2, 1, 3>},
    {2, 2, 2, launch_colors<2, 2, 2>}, {2, 2, 3, launch_colors<2, 2, 3>}, {2, 2, 4, launch_colors<2, 2, 4>},
    {2, 3, 3, launch_colors<2, 3, 3>}, {2, 3, 4, launch_colors<2, 3, 4>}, {2, 4, 4, launch_colors<2, 4, 4>},
    {2, 4, 5, launch_colors<2, 4, 5>}, {2, 5, 6, launch_colors<2, 5, 6>}, {2, 6, 7, launch_colors<2, 6, 7>},
    {3, 1, 1, launch_colors<3, 1, 1>}, {3, 1, 2, launch_colors<3, 1, 2>}, {3, 1, 3, launch_colors<3, 1, 3>},
    {3, 2, 2, launch_colors<3, 2, 2>}, {3, 2, 3, launch_colors<3, 2, 3>}, {3, 2, 4, launch_colors<3, 2, 4>},
};

static Launcher find_launcher(int dim, int p, int nq1) {
  for (const Entry& en : kTable)
    if (en.dim == dim && en.p == p && en.nq1 == nq1) return en.f;
  return nullptr;
}

bool thermal_supported(int dim, int p, int nq1) { return find_launcher(dim, p, nq1) != nullptr; }

void assemble_thermal(Ctx& c, const double* u, double* res, double* jac, bool want_res, bool want_jac,
                      bool accumulate) {
  Launcher f = find_launcher(c.mesh.dim, c.mesh.p, c.mesh.nq1);
  MHX_REQUIRE(f, MHX_ERR_UNSUPPORTED, "no kernel built for this dim/order/quadrature");
  AsmArgs a{};
  a.m = c.mesh;
  a.coef.k0 = c.k0.p;
  a.coef.src = c.src.p;
  a.coef.diff_u2 = c.desc.diff_u2;
  a.coef.n_modes = c.desc.n_modes;
  for (int mth = 0; mth < MHX_MAX_MODES; ++mth) {
    a.coef.amp[mth] = c.desc.mode_amp[mth];
    for (int d = 0; d < 3; ++d) a.coef.freq[mth][d] = c.desc.mode_freq[mth][d];
  }
  a.vx = c.vx.p;
  a.u = u;
  a.rowptr = c.rowptr.p;
  a.fixed = c.fixed.p;
  a.res = want_res ? res : nullptr;
  a.jac = want_jac ? jac : nullptr;
  if (c.bdf_order > 0) {
    a.hist = c.hist.p;
    a.alpha_t = c.alpha_t;
    a.hist_u = c.hist_u.p;
    a.alpha_u = c.alpha_u;
    const double rho = c.desc.th_density > 0.0 ? c.desc.th_density : 1.0;
    const double cp = c.desc.th_specific_heat > 0.0 ? c.desc.th_specific_heat : 1.0;
    a.rho_cp = rho * cp;
  }
  if (c.xh.active) {
    a.xp.code = c.x_code.p;
    a.xp.consts = c.x_consts.p;
    for (int f = 0; f < XF_COUNT; ++f) {
      a.xp.off[f] = c.xh.off[f];
      a.xp.len[f] = c.xh.len[f];
      a.xp.dflt[f] = c.xh.dflt[f];
    }
    a.xp.time = c.desc.fn_time;
    a.xp.active = 1;
  }
  // Overwrite mode: the first scheduled element touching an entry writes it,
  // which also writes the Dirichlet rows (0, diagonal 1).  Accumulate mode:
  // contributions are added and the Dirichlet diagonal is set afterwards.
  a.first_touch = accumulate ? 0 : 1;
  {
    ScopedPhase ph(c.prof, want_jac ? "AssemblyManager::assembleJacRes" : "AssemblyManager::assembleRes", c.stream);
    if (a.xp.active) launch_thermal_expr(c, a, want_jac);
    else if (c.scatter_mode == kPencilStream && !a.hist && q2box_applicable(c)) launch_q2box(c, a, want_jac);
    else f(c, a, want_jac);
    // boundary groups, after the volume groups
    if (want_res) neumann_residual(c, res);
    nitsche_boundary(c, u, want_res ? res : nullptr, want_jac ? jac : nullptr);
  }
  if (accumulate && want_jac && !c.h_dbc_list.empty()) {
    ScopedPhase ph(c.prof, "AssemblyManager::dofConstraints", c.stream);
    long n = (long)c.h_dbc_list.size();
    hipLaunchKernelGGL(k_dbc_diag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, c.mesh, c.dbc_list.p, n,
                       c.rowptr.p, c.colidx.p, jac);
    MHX_HIP(hipGetLastError());
  }
}

}  // namespace mhx
