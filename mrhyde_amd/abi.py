"""Constants and ctypes structures of the C ABI (include/mhx.h).

Pure data description: importing this module does not load libmhx.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np


MHX_OK = 0
MHX_ERR_ARG = -1
MHX_ERR_DEVICE = -2
MHX_ERR_UNSUPPORTED = -3
MHX_ERR_NOT_CONVERGED = -4
MHX_ERR_NONFINITE = -5
MHX_ERR_NOMEM = -6
MHX_ERR_STATE = -7

MHX_PHYS_THERMAL = 1
MHX_PHYS_NAVIER_STOKES = 2

SIDE = {"left": 1, "right": 2, "bottom": 4, "top": 8, "back": 16, "front": 32, "all boundaries": 63}
MHX_SIDE_ALL = 63

MHX_RES, MHX_JAC, MHX_ACCUMULATE, MHX_DEVICE_PTRS = 1, 2, 4, 8
MHX_KRYLOV_CG, MHX_KRYLOV_GMRES = 1, 2
MHX_PREC_NONE, MHX_PREC_JACOBI, MHX_PREC_BLOCK_JACOBI, MHX_PREC_AMG, MHX_PREC_FP32 = 0, 1, 2, 3, 0x100
MHX_MAX_MODES = 8
MHX_MAX_NQ_MODES = 4
MHX_RNG_UNIFORM, MHX_RNG_LOGNORMAL, MHX_RNG_NORMAL = 0, 1, 2
MHX_BUF_U, MHX_BUF_RES, MHX_BUF_JAC, MHX_BUF_DU, MHX_BUF_AUX = 0, 1, 2, 3, 4

_dp = C.POINTER(C.c_double)


class ProblemDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("order", C.c_int32), ("quadrature", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double), ("ymax", C.c_double),
        ("zmin", C.c_double), ("zmax", C.c_double),
        ("perturb", C.c_double), ("perturb_seed", C.c_uint64),
        ("physics", C.c_int32), ("workset_size", C.c_int32),
        ("k0", _dp), ("diff_u2", C.c_double), ("src_elem", _dp),
        ("n_modes", C.c_int32),
        ("mode_amp", C.c_double * MHX_MAX_MODES),
        ("mode_freq", (C.c_double * 3) * MHX_MAX_MODES),
        ("th_density", C.c_double), ("th_specific_heat", C.c_double),
        ("neumann_sides", C.c_uint32), ("n_nq_modes", C.c_int32),
        ("nq_amp", (C.c_double * 3) * MHX_MAX_NQ_MODES), ("nq_freq", (C.c_double * 3) * MHX_MAX_NQ_MODES),
        ("nq_phase", (C.c_double * 3) * MHX_MAX_NQ_MODES),
        ("wdbc_sides", C.c_uint32), ("wdbc_form", C.c_int32), ("wd_const", C.c_double),
        ("wd_lin", C.c_double * 3), ("n_wd_modes", C.c_int32), ("wd_reserved", C.c_int32),
        ("wd_amp", C.c_double * MHX_MAX_NQ_MODES), ("wd_freq", (C.c_double * 3) * MHX_MAX_NQ_MODES),
        ("wd_phase", (C.c_double * 3) * MHX_MAX_NQ_MODES),
        ("dbc_sides", C.c_uint32), ("dbc_const", C.c_double), ("dbc_values", _dp),
        ("ns_rho", C.c_double), ("ns_mu", C.c_double), ("ns_supg", C.c_int32), ("ns_pspg", C.c_int32),
        ("ns_lid_sides", C.c_uint32), ("ns_lid_u", C.c_double * 3),
        ("ns_force", C.c_double * 3), ("ns_pin_pressure", C.c_int32),
        ("ns_comp_sides", C.c_uint32 * 3),
        ("n_functions", C.c_int32), ("fn_reserved", C.c_int32),
        ("function_names", C.POINTER(C.c_char_p)), ("function_exprs", C.POINTER(C.c_char_p)),
        ("fn_time", C.c_double),
    ]


class KrylovOpts(C.Structure):
    _fields_ = [
        ("solver", C.c_int32), ("precond", C.c_int32), ("max_iters", C.c_int32), ("restart", C.c_int32),
        ("tol", C.c_double), ("flags", C.c_int32), ("inner_iters", C.c_int32), ("inner_tol", C.c_double),
    ]


class NewtonOpts(C.Structure):
    _fields_ = [
        ("max_iters", C.c_int32), ("nl_tol", C.c_double), ("backtracking", C.c_int32), ("flags", C.c_int32),
        ("linear", KrylovOpts), ("relative_off", C.c_int32), ("use_abs_tol", C.c_int32), ("abs_tol", C.c_double),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("iters", C.c_int32), ("nl_iters", C.c_int32), ("rel_res", C.c_double), ("res_norm0", C.c_double),
        ("res_norm", C.c_double), ("refinements", C.c_int32), ("status", C.c_int32),
        ("t_assemble_ms", C.c_double), ("t_solve_ms", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerDesc(C.Structure):
    _fields_ = [
        ("hbm_budget", C.c_int64), ("max_active", C.c_int32), ("max_cached", C.c_int32),
        ("newton", NewtonOpts), ("reserved", C.c_int32 * 4),
    ]


MHX_BATCH_FP64, MHX_BATCH_FP32_REFINED = 0, 1


class BatchDesc(C.Structure):
    _fields_ = [
        ("n_inst", C.c_int32), ("n_axis", C.POINTER(C.c_int32)), ("k", _dp), ("g", _dp),
        ("tol", C.c_double), ("max_iters", C.c_int32), ("precision", C.c_int32), ("inner_tol", C.c_double),
        ("max_refine", C.c_int32), ("wg_size", C.c_int32), ("use_graph", C.c_int32), ("extent", _dp), ("src", _dp),
        ("reserved", C.c_int32 * 4),
    ]


def bdf_schedule(order: int, nsteps: int, startup_order=None, startup_steps=None):
    """BDF order of each time step as SolverManager sets it
    (solverManager_construct.hpp:145-194, solverManager_solvers.hpp:93-97):
    "transient startup BDF order" (default: the order itself) for the first
    "transient startup steps" (default: the order) steps, then the order."""
    so = order if startup_order is None else int(startup_order)
    ss = order if startup_steps is None else int(startup_steps)
    return [so if (order > 1 and step < ss) else order for step in range(nsteps)]


def butcher_tableau(name: str, A=None, b=None, c=None):
    """Butcher tableau (A, b, c) by the reference's "transient Butcher
    tableau" name (SolverManager::setButcherTableau,
    solverManager_setup.hpp:521-700): the diagonally implicit ones and CN;
    "custom" takes A, b, c.  The explicit tableaus (FWE, SSPRK-3,3, RK-4,4,
    leap-frog) run on the reference's separate explicit solver and are not
    built here."""
    r3 = np.sqrt(3.0)
    p = 0.4358665215
    tab = {
        "BWE": ([[1.0]], [1.0], [1.0]),
        "DIRK-1,1": ([[1.0]], [1.0], [1.0]),
        "CN": ([[0.0, 0.0], [0.5, 0.5]], [0.5, 0.5], [0.0, 1.0]),
        "DIRK-1,2": ([[0.5]], [1.0], [0.5]),
        "DIRK-2,2": ([[0.25, 0.0], [0.5, 0.25]], [0.5, 0.5], [0.25, 0.75]),
        "DIRK-2,3": ([[0.5 + r3 / 6.0, 0.0], [-r3 / 3.0, 0.5 + r3 / 6.0]], [0.5, 0.5],
                     [0.5 + r3 / 6.0, 0.5 - r3 / 6.0]),
        "DIRK-3,3": ([[p, 0.0, 0.0], [(1.0 - p) / 2.0, p, 0.0],
                      [-3.0 * p * p / 2.0 + 4.0 * p - 1.0 / 4.0, 3.0 * p * p / 2.0 - 5.0 * p + 5.0 / 4.0, p]],
                     [-3.0 * p * p / 2.0 + 4.0 * p - 1.0 / 4.0, 3.0 * p * p / 2.0 - 5.0 * p + 5.0 / 4.0, p],
                     [p, (1.0 + p) / 2.0, 1.0]),
    }
    if name == "custom":
        A, b, c = np.atleast_2d(np.asarray(A, float)), np.asarray(b, float).ravel(), np.asarray(c, float).ravel()
    elif name in tab:
        A, b, c = (np.asarray(v, float) for v in tab[name])
    else:
        raise ValueError(f"Butcher tableau {name!r} is not built (explicit tableaus run on the reference's "
                         "explicit solver)")
    if A.shape != (len(b), len(b)) or len(c) != len(b) or np.any(np.triu(A, 1) != 0.0):
        raise ValueError("Butcher tableau: A must be square, lower triangular, and match b and c")
    return np.ascontiguousarray(A), b, c


def subgrid_workload(n_inst: int = 4096, seed: int = 4096, sizes=(4, 6, 8, 12, 16), sigma: float = 1.0):
    """The batched-subgrid input set: per instance n ~ U(sizes), element
    conductivity exp(sigma z), z ~ N(0,1), and Dirichlet data U[-1,1] on the
    boundary nodes (interior entries are ignored), all from the counter-based
    generator with the given seed."""
    pick = rng_fill(seed, n_inst, MHX_RNG_UNIFORM, 0.0, float(len(sizes)))
    n_axis = np.array([sizes[min(int(p), len(sizes) - 1)] for p in pick], dtype=np.int32)
    ne = (n_axis.astype(np.int64)) ** 3
    nn = (n_axis.astype(np.int64) + 1) ** 3
    k = rng_fill(seed + 1, int(ne.sum()), MHX_RNG_LOGNORMAL, 0.0, sigma)
    g = rng_fill(seed + 2, int(nn.sum()), MHX_RNG_UNIFORM, -1.0, 1.0)
    return n_axis, k, g


def rng_fill(seed: int, n: int, kind: int = MHX_RNG_UNIFORM, a: float = 0.0, b: float = 1.0) -> np.ndarray:
    """The seeded counter-based generator of mhx_rng_fill, vectorised in numpy:
    u_i = top 53 bits of splitmix64(seed ^ (i*phi + c)) / 2^53; normals by
    Box-Muller from draws 2i and 2i+1."""
    M = np.uint64(0xFFFFFFFFFFFFFFFF)

    def mix(idx):
        with np.errstate(over="ignore"):
            z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) ^ (idx * np.uint64(0x9E3779B97F4A7C15) + np.uint64(0x632BE59BD9B4E019))
            z = z + np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    i = np.arange(int(n), dtype=np.uint64)
    u1 = mix(np.uint64(2) * i)
    if kind == MHX_RNG_UNIFORM:
        return a + (b - a) * u1
    u2 = mix(np.uint64(2) * i + np.uint64(1))
    z = np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * np.pi * u2)
    return a + b * z if kind == MHX_RNG_NORMAL else np.exp(a + b * z)


@dataclass
class Problem:
    """The deck subset this path reads (Mesh, Physics: thermal, Discretization,
    Dirichlet conditions).  `modes` lists (amplitude, (wx, wy, wz)) source terms
    f = sum amp * prod_d s(w_d x_d) with s(w x) = sin(w x) for w != 0, else 1."""
    dim: int = 2
    order: int = 1
    quadrature: int = 2
    n: Sequence[int] = (8, 8, 1)
    lo: Sequence[float] = (0.0, 0.0, 0.0)
    hi: Sequence[float] = (1.0, 1.0, 1.0)
    k0: Optional[np.ndarray] = None
    diff_u2: float = 0.0
    src_elem: Optional[np.ndarray] = None
    modes: Sequence = ()
    density: float = 1.0         # thermal "density" (scales dT/dt)
    specific_heat: float = 1.0   # thermal "specific heat"
    # Neumann sides and flux field q(x): terms (amp[3], freq[3], phase[3]),
    # q_c = sum amp_c prod_d sin(freq_d x_d + phase_d); g = q . n
    neumann_sides: int = 0
    neumann_flux: Sequence = ()
    # weak (Nitsche) Dirichlet sides: data g = const + lin . x + sum amp prod sin(freq x + phase),
    # terms (amp, freq[3], phase[3]); form 0 symmetric (form_param 1), 1 non-symmetric (-1), 2 incomplete (0)
    wdbc_sides: int = 0
    wdbc_form: int = 0
    wd_const: float = 0.0
    wd_lin: Sequence[float] = (0.0, 0.0, 0.0)
    wd_modes: Sequence = ()
    dbc_sides: int = MHX_SIDE_ALL
    dbc_const: float = 0.0
    dbc_values: Optional[np.ndarray] = None
    perturb: float = 0.0
    perturb_seed: int = 0
    workset_size: int = 100
    physics: int = MHX_PHYS_THERMAL
    # Navier-Stokes ("density", "viscosity", "useSUPG", "usePSPG", lid, body force)
    rho: float = 1.0
    mu: float = 1.0
    supg: bool = False
    pspg: bool = False
    lid_sides: int = 0
    lid_u: Sequence[float] = (0.0, 0.0, 0.0)
    force: Sequence[float] = (0.0, 0.0, 0.0)
    pin_pressure: bool = False
    comp_sides: Sequence[int] = (0, 0, 0)  # extra Dirichlet sides per velocity component
    # coefficient functions, the deck's "Functions" block: name -> expression
    # ("thermal diffusion", "thermal source", "specific heat", "density",
    # "advection x|y|z", or helper names they use); time = value of "t"
    functions: dict = field(default_factory=dict)
    time: float = 0.0
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n_field(self) -> int:
        return self.dim + 1 if self.physics == MHX_PHYS_NAVIER_STOKES else 1

    @property
    def n_elem(self) -> int:
        return int(np.prod(list(self.n)[: self.dim]))

    @property
    def n_node(self) -> int:
        return int(np.prod([self.order * int(x) + 1 for x in list(self.n)[: self.dim]]))

    @property
    def n_dof(self) -> int:
        return self.n_node * self.n_field

    def desc(self) -> ProblemDesc:
        d = ProblemDesc()
        d.dim, d.order, d.quadrature = self.dim, self.order, self.quadrature
        n = list(self.n) + [1] * (3 - len(self.n))
        d.nx, d.ny, d.nz = int(n[0]), int(n[1]), int(n[2]) if self.dim == 3 else 1
        lo = list(self.lo) + [0.0] * (3 - len(self.lo))
        hi = list(self.hi) + [1.0] * (3 - len(self.hi))
        d.xmin, d.ymin, d.zmin = map(float, lo[:3])
        d.xmax, d.ymax, d.zmax = map(float, hi[:3])
        d.perturb, d.perturb_seed = float(self.perturb), int(self.perturb_seed)
        d.physics, d.workset_size = self.physics, int(self.workset_size)
        self._keep = []

        def arr(a, count):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.size != count:
                raise ValueError(f"array of size {a.size}, expected {count}")
            self._keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_double))

        d.k0 = arr(self.k0, self.n_elem)
        d.src_elem = arr(self.src_elem, self.n_elem)
        d.diff_u2 = float(self.diff_u2)
        if len(self.modes) > MHX_MAX_MODES:
            raise ValueError("too many source modes")
        d.n_modes = len(self.modes)
        for m, (amp, freq) in enumerate(self.modes):
            d.mode_amp[m] = float(amp)
            fr = list(freq) + [0.0] * (3 - len(freq))
            for a in range(3):
                d.mode_freq[m][a] = float(fr[a])
        d.th_density, d.th_specific_heat = float(self.density), float(self.specific_heat)
        if len(self.neumann_flux) > MHX_MAX_NQ_MODES:
            raise ValueError("too many Neumann flux terms")
        d.neumann_sides, d.n_nq_modes = int(self.neumann_sides), len(self.neumann_flux)
        for m, (amp, freq, phase) in enumerate(self.neumann_flux):
            for a in range(3):
                d.nq_amp[m][a] = float((list(amp) + [0.0] * 3)[a])
                d.nq_freq[m][a] = float((list(freq) + [0.0] * 3)[a])
                d.nq_phase[m][a] = float((list(phase) + [0.0] * 3)[a])
        if len(self.wd_modes) > MHX_MAX_NQ_MODES:
            raise ValueError("too many weak Dirichlet data terms")
        d.wdbc_sides, d.wdbc_form, d.wd_const = int(self.wdbc_sides), int(self.wdbc_form), float(self.wd_const)
        d.n_wd_modes = len(self.wd_modes)
        for a in range(3):
            d.wd_lin[a] = float((list(self.wd_lin) + [0.0] * 3)[a])
        for m, (amp, freq, phase) in enumerate(self.wd_modes):
            d.wd_amp[m] = float(amp)
            for a in range(3):
                d.wd_freq[m][a] = float((list(freq) + [0.0] * 3)[a])
                d.wd_phase[m][a] = float((list(phase) + [0.0] * 3)[a])
        d.dbc_sides, d.dbc_const = int(self.dbc_sides), float(self.dbc_const)
        d.dbc_values = arr(self.dbc_values, self.n_dof)
        d.ns_rho, d.ns_mu = float(self.rho), float(self.mu)
        d.ns_supg, d.ns_pspg = int(bool(self.supg)), int(bool(self.pspg))
        d.ns_lid_sides = int(self.lid_sides)
        for a in range(3):
            d.ns_lid_u[a] = float((list(self.lid_u) + [0.0] * 3)[a])
            d.ns_force[a] = float((list(self.force) + [0.0] * 3)[a])
            d.ns_comp_sides[a] = int((list(self.comp_sides) + [0] * 3)[a])
        d.ns_pin_pressure = int(bool(self.pin_pressure))
        names = [str(k).encode() for k in self.functions]
        exprs = [str(v).encode() for v in self.functions.values()]
        d.n_functions = len(names)
        if names:
            na, ea = (C.c_char_p * len(names))(*names), (C.c_char_p * len(exprs))(*exprs)
            self._keep.extend([names, exprs, na, ea])
            d.function_names = C.cast(na, C.POINTER(C.c_char_p))
            d.function_exprs = C.cast(ea, C.POINTER(C.c_char_p))
        d.fn_time = float(self.time)
        return d

