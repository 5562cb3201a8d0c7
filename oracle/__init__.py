"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and used there only as the checker or the CPU
baseline.  The product (mrhyde_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from mrhyde_amd.abi import KrylovOpts, NewtonOpts, Problem, ProblemDesc, SolveStats

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB


def _load():
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    sig = {
        "orc_last_error": (C.c_char_p, []),
        "orc_create": (C.c_int, [C.POINTER(ProblemDesc), C.POINTER(vp)]),
        "orc_destroy": (None, [vp]),
        "orc_sizes": (C.c_int, [vp] + [C.POINTER(i64)] * 3 + [C.POINTER(i32)] * 3),
        "orc_pattern": (C.c_int, [vp, vp, vp]),
        "orc_dirichlet": (C.c_int, [vp, vp, vp]),
        "orc_set_dirichlet": (C.c_int, [vp, vp]),
        "orc_lids": (C.c_int, [vp, vp]),
        "orc_assemble": (C.c_int, [vp, vp, i32, vp, vp]),
        "orc_linear_solve": (C.c_int, [vp, vp, vp, vp, C.POINTER(KrylovOpts), C.POINTER(SolveStats)]),
        "orc_newton": (C.c_int, [vp, vp, C.POINTER(NewtonOpts), C.POINTER(SolveStats)]),
        "orc_l2_error": (C.c_int, [vp, vp, i32, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "orc_elem_basis": (C.c_int, [vp, i64, vp, vp, vp, vp]),
        "orc_gauss": (C.c_int, [i32, vp, vp]),
        "orc_dof_coords": (C.c_int, [vp, vp]),
        "orc_set_time_state": (C.c_int, [vp, C.c_double, i32, vp]),
        "orc_update": (C.c_int, [vp, C.POINTER(ProblemDesc)]),
        "orc_expr_eval": (C.c_int, [C.POINTER(ProblemDesc), C.c_char_p, i64, vp, vp, i32]),
        "orc_set_threads": (C.c_int, [vp, i32]),
        "orc_set_stage_state": (C.c_int, [vp, C.c_double, i32, vp, i32, vp, vp, i32, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    return lib


lib = _load()


def gauss(n: int):
    x = np.empty(n)
    w = np.empty(n)
    lib.orc_gauss(n, x.ctypes.data, w.ctypes.data)
    return x, w


def expr_eval(problem: Problem, name: str, pts: np.ndarray, one_element: bool = False) -> np.ndarray:
    """The oracle's FunctionManager::evaluate of function `name` (defined in
    problem.functions) at the rows of pts [n, 11] (T, T_t, grad(T)[x], [y],
    [z], x, y, z, t, k0, src); each row its own element, or all rows one
    element.  Returns [n, 6]: value and d/dT, d/dgrad(T)[x|y|z], d/dT_t."""
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 11)
    out = np.empty((pts.shape[0], 6))
    d = problem.desc()
    rc = lib.orc_expr_eval(C.byref(d), name.encode(), pts.shape[0], pts.ctypes.data, out.ctypes.data,
                           int(one_element))
    if rc != 0:
        raise RuntimeError(lib.orc_last_error().decode())
    return out


class Oracle:
    """Reference-algorithm restatement on one problem (same numbering as mhx)."""

    def __init__(self, problem: Problem):
        self.problem = problem
        self._desc = problem.desc()
        h = C.c_void_p()
        rc = lib.orc_create(C.byref(self._desc), C.byref(h))
        if rc != 0:
            raise RuntimeError(lib.orc_last_error().decode())
        self._h = h
        n, nnz, ne = C.c_int64(), C.c_int64(), C.c_int64()
        nb, nq, ad = C.c_int32(), C.c_int32(), C.c_int32()
        lib.orc_sizes(h, C.byref(n), C.byref(nnz), C.byref(ne), C.byref(nb), C.byref(nq), C.byref(ad))
        self.n_dof, self.nnz, self.n_elem = n.value, nnz.value, ne.value
        self.nb, self.nq, self.ad_size = nb.value, nq.value, ad.value

    def __del__(self):
        if getattr(self, "_h", None):
            lib.orc_destroy(self._h)
            self._h = None

    def crs_pattern(self):
        rp = np.empty(self.n_dof + 1, dtype=np.int32)
        ci = np.empty(self.nnz, dtype=np.int32)
        lib.orc_pattern(self._h, rp.ctypes.data, ci.ctypes.data)
        return rp, ci

    def dirichlet(self):
        m = np.empty(self.n_dof, dtype=np.int32)
        v = np.empty(self.n_dof)
        lib.orc_dirichlet(self._h, m.ctypes.data, v.ctypes.data)
        return m, v

    def setDirichlet(self, u):
        lib.orc_set_dirichlet(self._h, u.ctypes.data)
        return u

    def lids(self):
        L = np.empty((self.n_elem, self.nb * self.problem.n_field), dtype=np.int32)
        lib.orc_lids(self._h, L.ctypes.data)
        return L

    def assembleJacRes(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        res = np.zeros(self.n_dof)
        jac = np.zeros(self.nnz)
        lib.orc_assemble(self._h, u.ctypes.data, 1, res.ctypes.data, jac.ctypes.data)
        return res, jac

    def assembleRes(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        res = np.zeros(self.n_dof)
        lib.orc_assemble(self._h, u.ctypes.data, 0, res.ctypes.data, None)
        return res

    def linearSolver(self, jac, r, x=None, opts: KrylovOpts = None):
        x = np.zeros(self.n_dof) if x is None else np.ascontiguousarray(x, dtype=np.float64).copy()
        st = SolveStats()
        lib.orc_linear_solve(self._h, np.ascontiguousarray(jac).ctypes.data, np.ascontiguousarray(r).ctypes.data,
                             x.ctypes.data, C.byref(opts), C.byref(st))
        return x, st.as_dict()

    def nonlinearSolver(self, u, opts: NewtonOpts):
        u = np.ascontiguousarray(u, dtype=np.float64)
        st = SolveStats()
        lib.orc_newton(self._h, u.ctypes.data, C.byref(opts), C.byref(st))
        return u, st.as_dict()

    def set_time_state(self, dt: float, order: int, u_prev):
        """Backward-difference state for the following assemblies (u_prev:
        `order` vectors, most recent first)."""
        buf = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64) for v in u_prev[:order]])
                                   if order > 0 else np.zeros(1))
        rc = lib.orc_set_time_state(self._h, float(dt), int(order), buf.ctypes.data)
        if rc != 0:
            raise ValueError("orc_set_time_state: bad arguments")

    def set_threads(self, n: int) -> int:
        """OpenMP threads of the thermal assembly and of the Krylov vector
        work (1, the default, keeps the sequential order of every sum)."""
        return lib.orc_set_threads(self._h, int(n))

    def update(self, problem: Problem):
        """New source amplitudes / density / specific heat / Dirichlet constant."""
        self.problem = problem
        self._desc = problem.desc()
        lib.orc_update(self._h, C.byref(self._desc))

    def set_stage_state(self, dt: float, order: int, u_prev, A, b, stage: int, u_stages=()):
        """Stage `stage` of the Butcher tableau (A, b) (u_stages: the earlier
        stages' solutions)."""
        A = np.ascontiguousarray(A, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        prev = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64) for v in u_prev[:order]]))
        st = (np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64) for v in u_stages[:stage]]))
              if stage > 0 else np.zeros(1))
        rc = lib.orc_set_stage_state(self._h, float(dt), int(order), prev.ctypes.data, len(b), A.ctypes.data,
                                     b.ctypes.data, int(stage), st.ctypes.data)
        if rc != 0:
            raise ValueError("orc_set_stage_state: bad arguments")

    def transientSolver(self, u0, dt: float, nsteps: int, order: int = 1, startup_order=None, startup_steps=None,
                        t0: float = 0.0, at_time=None, opts: NewtonOpts = None, record=None, tableau: str = "BWE",
                        butcher=None):
        """SolverManager::transientSolver (solverManager_solvers.hpp:40-170):
        history vectors start as zeros and shift each step (u_prev[0] = u_n),
        the first `startup_steps` steps use `startup_order`; each step runs
        the stages of the Butcher tableau, every stage a Newton solve from u_n
        at the stage time t_n + c_s dt, and the step's solution is
        u_n + sum_s (U_s - u_n) (:125-151); at_time(t) -> Problem gives the
        coefficients at a stage time, and record(t, u) sees each step."""
        from mrhyde_amd.abi import MHX_KRYLOV_CG, MHX_PREC_JACOBI, bdf_schedule, butcher_tableau

        if opts is None:
            opts = NewtonOpts()
            opts.max_iters, opts.nl_tol, opts.backtracking = 10, 1e-12, 1
            opts.linear.solver, opts.linear.precond = MHX_KRYLOV_CG, MHX_PREC_JACOBI
            opts.linear.max_iters, opts.linear.tol = 10000, 1e-12
        A, b, c = butcher_tableau(tableau, *(butcher or (None, None, None)))
        one_stage = len(b) == 1 and A[0, 0] == 1.0 and b[0] == 1.0
        u = np.ascontiguousarray(u0, dtype=np.float64).copy()
        hist = [np.zeros_like(u) for _ in range(max(order, startup_order or order))]
        t = t0
        for k in bdf_schedule(order, nsteps, startup_order, startup_steps):
            hist = [u.copy()] + hist[:-1]
            un = u.copy()
            stages = []
            for s in range(len(b)):
                if at_time is not None:
                    self.update(at_time(t + c[s] * dt))
                if one_stage:
                    self.set_time_state(dt, k, hist)
                else:
                    self.set_stage_state(dt, k, hist, A, b, s, stages)
                us = self.setDirichlet(un.copy())
                us, _ = self.nonlinearSolver(us, opts)
                stages.append(us)
            u = stages[0] if one_stage else un.copy()
            if not one_stage:
                for us in stages:
                    u = u + us - un
            t = t + dt
            if record is not None:
                record(t, u)
        self.set_time_state(dt, 0, [])
        return u

    def computeError(self, u, modes):
        amp = np.array([m[0] for m in modes], dtype=np.float64)
        fr = np.zeros((len(modes), 3))
        for i, (_, f) in enumerate(modes):
            fr[i, : len(f)] = f
        l2, l2g = C.c_double(), C.c_double()
        lib.orc_l2_error(self._h, np.ascontiguousarray(u).ctypes.data, len(modes), amp.ctypes.data,
                         fr.ctypes.data, C.byref(l2), C.byref(l2g))
        return l2.value, l2g.value

    def elem_basis(self, e: int = 0):
        dim = self.problem.dim
        phi = np.empty((self.nb, self.nq))
        grad = np.empty((self.nb, self.nq, dim))
        wts = np.empty(self.nq)
        x = np.empty((self.nq, 3))
        lib.orc_elem_basis(self._h, e, phi.ctypes.data, grad.ctypes.data, wts.ctypes.data, x.ctypes.data)
        return phi, grad, wts, x

    def dof_coords(self):
        x = np.empty((self.n_dof, 3))
        lib.orc_dof_coords(self._h, x.ctypes.data)
        return x
