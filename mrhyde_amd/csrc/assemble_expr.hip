// assemble_expr.hip — thermal J+R assembly with the coefficients taken from
// compiled expressions (the reference's "Functions" block evaluated by
// FunctionManager, functionManager_evaluate.hpp:14-233, inside
// thermal<EvalT>::volumeResidual, thermal.cpp:71-165): the k_thermal kernel
// of thermal_kernel.hpp with EX = true, element colouring, steady and
// transient, for every order the closed-form path builds.  The bytecode is
// evaluated per cubature point in the kernel's dual numbers (expr.hpp).
#include "thermal_kernel.hpp"

namespace mhx {

namespace {

template <int DIM, int P, int NQ1>
void launch_expr(Ctx& c, AsmArgs& a, bool jac) {
  using C = Cfg<DIM, P, NQ1>;
  a.gpb = C::GPB;
  a.blocked = kElementColouring;
  a.nsub = 1;
  for (int col = 0; col < (1 << DIM); ++col) {
    const long blocks = element_colour_grid<DIM, P, NQ1>(c, a, col);
    if (blocks == 0) continue;
    ScopedPhase ph(c.prof, jac ? "k_thermal<jacres>" : "k_thermal<res>", c.stream);
    const dim3 grid((unsigned)blocks), block(C::BLOCK);
    if (a.hist) {
      if (jac) hipLaunchKernelGGL((k_thermal<DIM, P, NQ1, true, kElementColouring, true, true>), grid, block, 0, c.stream, a);
      else hipLaunchKernelGGL((k_thermal<DIM, P, NQ1, false, kElementColouring, true, true>), grid, block, 0, c.stream, a);
    } else {
      if (jac) hipLaunchKernelGGL((k_thermal<DIM, P, NQ1, true, kElementColouring, false, true>), grid, block, 0, c.stream, a);
      else hipLaunchKernelGGL((k_thermal<DIM, P, NQ1, false, kElementColouring, false, true>), grid, block, 0, c.stream, a);
    }
    MHX_HIP(hipGetLastError());
  }
}

using Launcher = void (*)(Ctx&, AsmArgs&, bool);
struct Entry {
  int dim, p, nq1;
  Launcher f;
};
const Entry kExprTable[] = {
    {2, 1, 1, launch_expr<2, 1, 1>}, {2, 1, 2, launch_expr<2, 1, 2>}, {2, 1, 3, launch_expr<2, 1, 3>},
    {2, 2, 2, launch_expr<2, 2, 2>}, {2, 2, 3, launch_expr<2, 2, 3>}, {2, 2, 4, launch_expr<2, 2, 4>},
    {2, 3, 3, launch_expr<2, 3, 3>}, {2, 3, 4, launch_expr<2, 3, 4>}, {2, 4, 4, launch_expr<2, 4, 4>},
    {2, 4, 5, launch_expr<2, 4, 5>}, {2, 5, 6, launch_expr<2, 5, 6>}, {2, 6, 7, launch_expr<2, 6, 7>},
    {3, 1, 1, launch_expr<3, 1, 1>}, {3, 1, 2, launch_expr<3, 1, 2>}, {3, 1, 3, launch_expr<3, 1, 3>},
    {3, 2, 2, launch_expr<3, 2, 2>}, {3, 2, 3, launch_expr<3, 2, 3>}, {3, 2, 4, launch_expr<3, 2, 4>},
};

}  // namespace

void launch_thermal_expr(Ctx& c, AsmArgs& a, bool jac) {
  for (const Entry& en : kExprTable)
    if (en.dim == c.mesh.dim && en.p == c.mesh.p && en.nq1 == c.mesh.nq1) {
      en.f(c, a, jac);
      return;
    }
  throw Error{MHX_ERR_UNSUPPORTED, "no kernel built for this dim/order/quadrature"};
}

}  // namespace mhx
