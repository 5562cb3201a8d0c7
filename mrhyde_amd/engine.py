"""Host-side mirror of the reference manager interface over the C ABI.

Method names follow the reference so call sites read the same:
  AssemblyManager::assembleJacRes / assembleRes
      (src/managers/assembly/assemblyManager.hpp:237-255, :330-345)
  LinearAlgebraInterface::linearSolver
      (src/interfaces/linear_algebra/linearAlgebraInterface.hpp:506)
  SolverManager::nonlinearSolver / setDirichlet
      (src/managers/solver/solverManager.hpp:139-145, solverManager_util.hpp:60-146)
  PostprocessManager::computeError
      (src/managers/postprocess/postprocessManager_error_estimation.hpp:173)

Arrays may be numpy (host) or torch CUDA tensors (device, zero copy).  All
work runs in libmhx.so on the GPU; a missing library raises at import.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import capi
from .abi import Problem, bdf_schedule, butcher_tableau, rng_fill  # noqa: F401
from .capi import lib


class MhxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mhx error {code}: {msg}")
        self.code = code


def rng(seed: int, n: int, kind: int = capi.MHX_RNG_UNIFORM, a: float = 0.0, b: float = 1.0) -> np.ndarray:
    """Seeded counter-based fields (uniform [a,b), lognormal exp(a + b z), normal a + b z)."""
    out = np.empty(int(n), dtype=np.float64)
    rc = lib.mhx_rng_fill(C.c_uint64(seed & (2**64 - 1)), int(n), int(kind), float(a), float(b),
                          out.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise MhxError(rc, "rng_fill")
    return out


def _ptr(a):
    """(pointer, is_device) for numpy arrays and torch tensors (contiguous, float64/int32)."""
    if a is None:
        return None, False
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data_as(C.c_void_p), False
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return C.c_void_p(a.data_ptr()), bool(a.is_cuda)
    raise TypeError(f"unsupported array type {type(a)}")


def function_eval(problem: Problem, name: str, pts) -> np.ndarray:
    """FunctionManager::evaluate(name, "point") on the host
    (functionManager_evaluate.hpp:14-50): function `name` of
    problem.functions, compiled as the assembly compiles it, at the rows of
    pts [n, 11] (T, T_t, grad(T)[x], [y], [z], x, y, z, t, k0, src).
    Returns [n, 6]: value and d/dT, d/dgrad(T)[x|y|z], d/dT_t."""
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 11)
    out = np.empty((pts.shape[0], 6))
    d = problem.desc()
    rc = lib.mhx_expr_eval(C.byref(d), name.encode(), pts.shape[0], pts.ctypes.data, out.ctypes.data)
    if rc != 0:
        raise MhxError(rc, lib.mhx_expr_last_error().decode())
    return out


def krylov_opts(solver: str = "gmres", tol: float = 1e-10, max_iters: int = 1000, restart: int = 50,
                precond: str = "jacobi", device: bool = False, inner_iters: int = 0,
                inner_tol: float = 0.0) -> capi.KrylovOpts:
    o = capi.KrylovOpts()
    o.solver = capi.MHX_KRYLOV_CG if solver.lower() in ("cg", "pcg", "pseudo block cg") else capi.MHX_KRYLOV_GMRES
    p = precond.lower()
    o.precond = {"none": 0, "jacobi": 1, "block jacobi": 2, "amg": 3, "muelu": 3}[p.replace("-fp32", "").strip()]
    if p.endswith("-fp32"):
        o.precond |= capi.MHX_PREC_FP32
    o.max_iters, o.restart, o.tol = int(max_iters), int(restart), float(tol)
    o.flags = capi.MHX_DEVICE_PTRS if device else 0
    o.inner_iters, o.inner_tol = int(inner_iters), float(inner_tol)
    return o


class Mhx:
    """One assembly/solve context (mesh, DOFs, colouring, CRS and device state)."""

    def __init__(self, problem: Problem):
        self.problem = problem
        self._desc = problem.desc()
        h = C.c_void_p()
        rc = lib.mhx_ctx_create(C.byref(self._desc), C.byref(h))
        if rc != 0:
            raise MhxError(rc, lib.mhx_last_error(None).decode())
        self._h = h
        n, nnz, ne, nc = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        lib.mhx_sizes(self._h, C.byref(n), C.byref(nnz), C.byref(ne), C.byref(nc))
        self.n_dof, self.nnz, self.n_elem, self.n_colors = n.value, nnz.value, ne.value, nc.value
        self._stream = None

    def _follow_torch_stream(self):
        """Order the engine's work after torch's current stream (device tensors)."""
        import torch

        s = torch.cuda.current_stream().cuda_stream
        if s != self._stream:
            self._check(lib.mhx_set_stream(self._h, C.c_void_p(s)), "set_stream")
            self._stream = s

    def set_stream(self, hip_stream: int):
        self._check(lib.mhx_set_stream(self._h, C.c_void_p(hip_stream)), "set_stream")
        self._stream = hip_stream

    def close(self):
        if getattr(self, "_h", None):
            lib.mhx_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc: int, what: str):
        if rc < 0:
            raise MhxError(rc, f"{what}: {lib.mhx_last_error(self._h).decode()}")
        return rc

    # ---- options ------------------------------------------------------------------
    def set_option(self, name: str, value: int):
        self._check(lib.mhx_set_option(self._h, name.encode(), int(value)), f"set_option({name})")

    def get_option(self, name: str) -> int:
        v = C.c_int64()
        self._check(lib.mhx_get_option(self._h, name.encode(), C.byref(v)), f"get_option({name})")
        return v.value

    # ---- device-resident state ------------------------------------------------
    def state_ptr(self, which: int):
        """Device address (int) and length of a context-owned buffer (MHX_BUF_*)."""
        p, n = C.c_void_p(), C.c_int64()
        self._check(lib.mhx_state_ptr(self._h, int(which), C.byref(p), C.byref(n)), "state_ptr")
        return p.value, n.value

    def state_upload(self, which: int, host: np.ndarray):
        host = np.ascontiguousarray(host, dtype=np.float64)
        self._check(lib.mhx_state_upload(self._h, int(which), host.ctypes.data, host.size), "state_upload")

    def state_download(self, which: int) -> np.ndarray:
        _, n = self.state_ptr(which)
        out = np.empty(n)
        self._check(lib.mhx_state_download(self._h, int(which), out.ctypes.data, n), "state_download")
        return out

    def assemble_resident(self, want_res: bool = True, want_jac: bool = True, accumulate: bool = False):
        """Assemble from the context's device u into its device res / J (no host traffic)."""
        pu, _ = self.state_ptr(capi.MHX_BUF_U)
        pr = self.state_ptr(capi.MHX_BUF_RES)[0] if want_res else None
        pj = self.state_ptr(capi.MHX_BUF_JAC)[0] if want_jac else None
        flags = capi.MHX_DEVICE_PTRS | (capi.MHX_RES if want_res else 0) | (capi.MHX_JAC if want_jac else 0)
        if accumulate:
            flags |= capi.MHX_ACCUMULATE
        self._check(lib.mhx_assemble(self._h, C.c_void_p(pu), flags, C.c_void_p(pr) if pr else None,
                                     C.c_void_p(pj) if pj else None), "assemble_resident")

    def solve_resident(self, opts: Optional[capi.KrylovOpts] = None, **kw):
        """Solve J du = res on the context's device buffers (J, res from
        assemble_resident; du is used as the initial guess)."""
        opts = opts or krylov_opts(**kw)
        opts.flags |= capi.MHX_DEVICE_PTRS
        pj, pr, px = (self.state_ptr(b)[0] for b in (capi.MHX_BUF_JAC, capi.MHX_BUF_RES, capi.MHX_BUF_DU))
        st = capi.SolveStats()
        rc = lib.mhx_linear_solve(self._h, C.c_void_p(pj), C.c_void_p(pr), C.c_void_p(px), C.byref(opts),
                                  C.byref(st))
        if rc < 0 and rc != capi.MHX_ERR_NOT_CONVERGED:
            self._check(rc, "solve_resident")
        return st.as_dict()

    # ---- setup products -------------------------------------------------------
    def crs_pattern(self):
        rp = np.empty(self.n_dof + 1, dtype=np.int32)
        ci = np.empty(self.nnz, dtype=np.int32)
        self._check(lib.mhx_crs_pattern(self._h, rp.ctypes.data, ci.ctypes.data), "crs_pattern")
        return rp, ci

    def dirichlet_mask(self):
        m = np.empty(self.n_dof, dtype=np.int32)
        self._check(lib.mhx_dirichlet_mask(self._h, m.ctypes.data), "dirichlet_mask")
        return m

    def dof_coords(self):
        x = np.empty((self.n_dof, 3), dtype=np.float64)
        self._check(lib.mhx_dof_coords(self._h, x.ctypes.data), "dof_coords")
        return x

    def setDirichlet(self, u):
        p, dev = _ptr(u)
        if dev:
            self._follow_torch_stream()
        self._check(lib.mhx_set_dirichlet(self._h, p, capi.MHX_DEVICE_PTRS if dev else 0), "setDirichlet")
        return u

    # ---- AssemblyManager ------------------------------------------------------
    def assemble(self, u, res=None, jac=None, accumulate: bool = False):
        flags = (capi.MHX_RES if res is not None else 0) | (capi.MHX_JAC if jac is not None else 0)
        if accumulate:
            flags |= capi.MHX_ACCUMULATE
        pu, du = _ptr(u)
        pr, dr = _ptr(res)
        pj, dj = _ptr(jac)
        devs = {d for d, a in ((du, u), (dr, res), (dj, jac)) if a is not None}
        if len(devs) != 1:
            raise ValueError("u/res/jac must all be host arrays or all device tensors")
        if devs.pop():
            flags |= capi.MHX_DEVICE_PTRS
            self._follow_torch_stream()
        self._check(lib.mhx_assemble(self._h, pu, flags, pr, pj), "assemble")
        return res, jac

    def assembleJacRes(self, u):
        res = np.zeros(self.n_dof)
        jac = np.zeros(self.nnz)
        self.assemble(np.ascontiguousarray(u, dtype=np.float64), res, jac)
        return res, jac

    def assembleRes(self, u):
        res = np.zeros(self.n_dof)
        self.assemble(np.ascontiguousarray(u, dtype=np.float64), res, None)
        return res

    # ---- LinearAlgebraInterface / SolverManager ---------------------------------
    def apply(self, jac, x, y=None):
        """y = J x (Tpetra CrsMatrix::apply)."""
        if y is None:
            y = np.zeros(self.n_dof) if isinstance(x, np.ndarray) else x.new_zeros(self.n_dof)
        pj, dj = _ptr(jac)
        px, dx = _ptr(x)
        py, dy = _ptr(y)
        if dx:
            self._follow_torch_stream()
        self._check(lib.mhx_spmv(self._h, pj, px, py, capi.MHX_DEVICE_PTRS if dx else 0), "apply")
        return y

    def linearSolver(self, jac, r, x=None, opts: Optional[capi.KrylovOpts] = None, **kw):
        opts = opts or krylov_opts(**kw)
        if x is None:
            x = np.zeros(self.n_dof)
        pj, dj = _ptr(jac)
        pr, dr = _ptr(r)
        px, dx = _ptr(x)
        if dr or dx:
            opts.flags |= capi.MHX_DEVICE_PTRS
            self._follow_torch_stream()
        st = capi.SolveStats()
        rc = lib.mhx_linear_solve(self._h, pj, pr, px, C.byref(opts), C.byref(st))
        if rc < 0 and rc != capi.MHX_ERR_NOT_CONVERGED:
            self._check(rc, "linearSolver")
        return x, st.as_dict()

    def buildAMGPreconditioner(self, jac=None, resident: bool = False) -> int:
        """Smoothed-aggregation hierarchy of J (None: the Jacobian of the last
        Newton step; resident=True: the context's device buffer MHX_BUF_JAC
        written by assemble_resident), kept in the context; returns the number
        of levels."""
        if resident:
            pj, dj = C.c_void_p(self.state_ptr(capi.MHX_BUF_JAC)[0]), True
        else:
            pj, dj = _ptr(jac)
        return self._check(lib.mhx_amg_setup(self._h, pj, capi.MHX_DEVICE_PTRS if dj else 0), "amg_setup")

    def amg_level(self, level: int, which: str = "A"):
        """One operator of the held hierarchy: which = "A", "P", "R" as scipy
        CSR matrices (level-0 "A" as its pattern only), or "Cinv" (dense
        inverse of the coarsest operator).  Returns (matrix, info) with info
        = {rows, cols, nnz, levels, lambda}."""
        import scipy.sparse as sp

        w = {"A": 0, "P": 1, "R": 2, "Cinv": 3}[which]
        info = (C.c_int64 * 4)()
        lam = C.c_double()
        self._check(lib.mhx_amg_level(self._h, level, w, info, C.byref(lam), None, None, None), "amg_level")
        rows, cols, nnz = info[0], info[1], info[2]
        meta = {"rows": rows, "cols": cols, "nnz": nnz, "levels": info[3], "lambda": lam.value}
        if w == 3:
            vals = np.empty(nnz)
            self._check(lib.mhx_amg_level(self._h, level, w, info, None, None, None, vals.ctypes.data), "amg_level")
            return vals.reshape(rows, cols), meta
        rp = np.empty(rows + 1, dtype=np.int32)
        ci = np.empty(nnz, dtype=np.int32)
        vals = np.ones(nnz)
        pv = None if (level == 0 and w == 0) else vals.ctypes.data
        self._check(lib.mhx_amg_level(self._h, level, w, info, None, rp.ctypes.data, ci.ctypes.data, pv),
                    "amg_level")
        return sp.csr_matrix((vals, ci, rp), shape=(rows, cols)), meta

    def nonlinearSolver(self, u, max_iters: int = 10, nl_tol: float = 1e-6, backtracking: bool = True,
                        linear: Optional[capi.KrylovOpts] = None, linear_max_iters: int = 1000,
                        use_relative_tol: bool = True, use_absolute_tol: bool = False, abs_tol: float = 0.0,
                        **kw):
        """Newton on u in place ("max nonlinear iters", "nonlinear TOL", "allow
        backtracking", "use relative TOL", "use absolute TOL", "absolute
        nonlinear TOL" with 0 meaning min(nl_tol, 1e-6)); `linear` / keyword
        arguments set the Krylov solve."""
        o = capi.NewtonOpts()
        o.max_iters, o.nl_tol, o.backtracking = int(max_iters), float(nl_tol), int(backtracking)
        o.relative_off, o.use_abs_tol, o.abs_tol = int(not use_relative_tol), int(use_absolute_tol), float(abs_tol)
        o.linear = linear or krylov_opts(max_iters=linear_max_iters, **kw)
        pu, du = _ptr(u)
        o.flags = capi.MHX_DEVICE_PTRS if du else 0
        o.linear.flags = o.flags
        if du:
            self._follow_torch_stream()
        st = capi.SolveStats()
        rc = lib.mhx_nonlinear_solve(self._h, pu, C.byref(o), C.byref(st))
        if rc < 0 and rc != capi.MHX_ERR_NOT_CONVERGED:
            self._check(rc, "nonlinearSolver")
        return u, st.as_dict()

    def writeToFile(self, jac=None, res=None, sol=None, jac_filename: str = "jacobian.mm",
                    res_filename: str = "residual.mm", sol_filename: str = "solution.mm"):
        """LinearAlgebraInterface::writeToFile: the given J (values in the
        context's pattern), residual and solution as MatrixMarket files;
        None skips a file (for J: pass "context" for the Newton Jacobian)."""
        if jac is not None:
            pj, dj = (None, False) if isinstance(jac, str) else _ptr(jac)
            self._check(lib.mhx_write_matrix_market(self._h, jac_filename.encode(), pj,
                                                    capi.MHX_DEVICE_PTRS if dj else 0), "write_matrix_market")
        for vec, name in ((res, res_filename), (sol, sol_filename)):
            if vec is not None:
                pv, dv = _ptr(vec)
                self._check(lib.mhx_write_vector_market(self._h, name.encode(), pv,
                                                        capi.MHX_DEVICE_PTRS if dv else 0), "write_vector_market")

    def readMatrixMarket(self, path: str) -> np.ndarray:
        """Values of a MatrixMarket matrix in the context's CRS pattern."""
        out = np.zeros(self.nnz)
        self._check(lib.mhx_read_matrix_market(self._h, path.encode(), out.ctypes.data, 0), "read_matrix_market")
        return out

    def readVectorMarket(self, path: str) -> np.ndarray:
        out = np.zeros(self.n_dof)
        self._check(lib.mhx_read_vector_market(self._h, path.encode(), out.ctypes.data, 0), "read_vector_market")
        return out

    def update(self, problem: Problem):
        """New coefficients on the same mesh (mhx_ctx_update): k0, sources,
        density / specific heat, Dirichlet values."""
        self.problem = problem
        self._desc = problem.desc()
        self._check(lib.mhx_ctx_update(self._h, C.byref(self._desc)), "update")

    def set_time_state(self, dt: float, order: int, u_prev=()):
        """Backward-difference state for the following assemblies and solves
        (u_prev: `order` vectors, most recent first; order 0 = steady)."""
        if order == 0:
            self._check(lib.mhx_set_time_state(self._h, 0.0, 0, None, 0), "set_time_state")
            return
        if all(isinstance(v, np.ndarray) for v in u_prev[:order]):
            buf = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64) for v in u_prev[:order]]))
            self._check(lib.mhx_set_time_state(self._h, float(dt), int(order), buf.ctypes.data, 0),
                        "set_time_state")
        else:
            import torch

            buf = torch.cat([v.reshape(-1) for v in u_prev[:order]]).contiguous()
            self._follow_torch_stream()
            self._check(lib.mhx_set_time_state(self._h, float(dt), int(order), C.c_void_p(buf.data_ptr()),
                                               capi.MHX_DEVICE_PTRS), "set_time_state")

    def set_stage_state(self, dt: float, order: int, u_prev, A, b, stage: int, u_stages=()):
        """Stage `stage` of the Butcher tableau (A, b) for the following
        assemblies and solves (mhx_set_stage_state): u_prev as in
        set_time_state, u_stages the earlier stages' solutions."""
        A = np.ascontiguousarray(A, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        prev = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64) for v in u_prev[:order]]))
        st = (np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64) for v in u_stages[:stage]]))
              if stage > 0 else None)
        self._check(lib.mhx_set_stage_state(self._h, float(dt), int(order), prev.ctypes.data, len(b), A.ctypes.data,
                                            b.ctypes.data, int(stage), st.ctypes.data if st is not None else None, 0),
                    "set_stage_state")

    def transientSolver(self, u0, dt: float, nsteps: int, order: int = 1, startup_order=None, startup_steps=None,
                        t0: float = 0.0, at_time=None, record=None, tableau: str = "BWE", butcher=None,
                        **newton_kw):
        """SolverManager::transientSolver (solverManager_solvers.hpp:40-170):
        history vectors start as zeros and shift each step (u_prev[0] = u_n);
        the first `startup_steps` steps use `startup_order` (bdf_schedule);
        each step runs the stages of the Butcher tableau `tableau` (or
        butcher = (A, b, c) for "custom"), every stage a Newton solve for its
        unknown from u_n at the stage time t_n + c_s dt, and the step's
        solution is u_n + sum_s (U_s - u_n) (:125-151); at_time(t) -> Problem
        gives the coefficients at a stage time; record(t, u) sees each step.
        Keyword arguments go to nonlinearSolver (defaults: CG + Jacobi to
        1e-12, Newton to 1e-12)."""
        A, b, c = butcher_tableau(tableau, *(butcher or (None, None, None)))
        kw = dict(max_iters=10, nl_tol=1e-12, solver="cg", tol=1e-12, linear_max_iters=10000)
        kw.update(newton_kw)
        u = np.ascontiguousarray(u0, dtype=np.float64).copy()
        hist = [np.zeros_like(u) for _ in range(max(order, startup_order or order))]
        t = t0
        one_stage = len(b) == 1 and A[0, 0] == 1.0 and b[0] == 1.0
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
                us, _ = self.nonlinearSolver(us, **kw)
                stages.append(us)
            u = un.copy()
            for us in stages:
                u = u + us - un if not one_stage else us
            t = t + dt
            if record is not None:
                record(t, u)
        self.set_time_state(dt, 0)
        return u

    def computeError(self, u, modes):
        amp = np.array([m[0] for m in modes], dtype=np.float64)
        fr = np.zeros((len(modes), 3))
        for i, (_, f) in enumerate(modes):
            fr[i, : len(f)] = f
        pu, du = _ptr(u)
        if du:
            self._follow_torch_stream()
        l2, l2g = C.c_double(), C.c_double()
        self._check(lib.mhx_l2_error(self._h, pu, len(modes), amp.ctypes.data, fr.ctypes.data,
                                     capi.MHX_DEVICE_PTRS if du else 0, C.byref(l2), C.byref(l2g)),
                    "computeError")
        return l2.value, l2g.value

    # ---- profiling --------------------------------------------------------------
    def profile(self, on: bool = True):
        lib.mhx_profile_enable(self._h, int(on))

    def profile_reset(self):
        lib.mhx_profile_reset(self._h)

    def profile_read(self) -> dict:
        cap = 64
        names = C.create_string_buffer(48 * cap)
        ms = np.zeros(cap)
        cnt = np.zeros(cap, dtype=np.int64)
        n = self._check(lib.mhx_profile_read(self._h, cap, names, ms.ctypes.data, cnt.ctypes.data), "profile_read")
        out = {}
        for i in range(min(n, cap)):
            nm = names.raw[48 * i: 48 * (i + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[i]), int(cnt[i]))
        return out

    def synchronize(self):
        self._check(lib.mhx_synchronize(self._h), "synchronize")


class SubgridBatch:
    """Batched subgrid solves (mhx_batch_*): many unit-cube Q1 thermal problems
    assembled and solved in one persistent launch.  Mirrors the role of
    SubGridDtN_Solver::nonlinearSolver looped over macro groups
    (src/subgrid/subgridDtN_solver.cpp:903-1043)."""

    def __init__(self, n_axis, k, g, tol: float = 1e-10, precision: str = "fp64", wg_size: int = 0,
                 use_graph: bool = True, max_iters: int = 10000, inner_tol: float = 1e-5, max_refine: int = 20,
                 extent=None, src=None):
        """extent: [n_inst, 3] box side lengths (default unit cubes); src:
        concatenated element sources like k (default none)."""
        self.n_axis = np.ascontiguousarray(n_axis, dtype=np.int32)
        self._k = np.ascontiguousarray(k, dtype=np.float64)
        self._g = np.ascontiguousarray(g, dtype=np.float64)
        d = capi.BatchDesc()
        d.n_inst = len(self.n_axis)
        d.n_axis = self.n_axis.ctypes.data_as(C.POINTER(C.c_int32))
        d.k = self._k.ctypes.data_as(C.POINTER(C.c_double))
        d.g = self._g.ctypes.data_as(C.POINTER(C.c_double))
        if extent is not None:
            self._ext = np.ascontiguousarray(extent, dtype=np.float64).reshape(len(self.n_axis), 3)
            d.extent = self._ext.ctypes.data_as(C.POINTER(C.c_double))
        if src is not None:
            self._src = np.ascontiguousarray(src, dtype=np.float64)
            if self._src.shape != self._k.shape:
                raise ValueError("src needs one value per element, like k")
            d.src = self._src.ctypes.data_as(C.POINTER(C.c_double))
        d.tol, d.max_iters = float(tol), int(max_iters)
        d.precision = capi.MHX_BATCH_FP32_REFINED if precision == "fp32-refined" else capi.MHX_BATCH_FP64
        d.inner_tol, d.max_refine = float(inner_tol), int(max_refine)
        d.wg_size, d.use_graph = int(wg_size), int(use_graph)
        h = C.c_void_p()
        rc = lib.mhx_batch_create(C.byref(d), C.byref(h))
        if rc != 0:
            raise MhxError(rc, "mhx_batch_create: " + lib.mhx_last_error(None).decode())
        self._h = h
        nn, wg, ms = C.c_int64(), C.c_int32(), C.c_double()
        lib.mhx_batch_info(h, C.byref(nn), C.byref(wg), C.byref(ms))
        self.total_nodes, self.wg_size = nn.value, wg.value
        self.offsets = np.concatenate([[0], np.cumsum((self.n_axis.astype(np.int64) + 1) ** 3)])

    def close(self):
        if getattr(self, "_h", None):
            lib.mhx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_wg(self, wg: int):
        if lib.mhx_batch_set_wg(self._h, int(wg)) != 0:
            raise MhxError(-1, "batch workgroup size must be 512 or 1024")
        self.wg_size = wg

    def solve(self) -> float:
        """One batched assemble+solve of every instance; returns device milliseconds."""
        rc = lib.mhx_batch_solve(self._h, 0)
        if rc != 0:
            raise MhxError(rc, "mhx_batch_solve")
        nn, wg, ms = C.c_int64(), C.c_int32(), C.c_double()
        lib.mhx_batch_info(self._h, C.byref(nn), C.byref(wg), C.byref(ms))
        return ms.value

    def results(self):
        n = len(self.n_axis)
        x = np.empty(self.total_nodes)
        it = np.empty(n, dtype=np.int32)
        rel = np.empty(n)
        refs = np.empty(n, dtype=np.int32)
        rc = lib.mhx_batch_results(self._h, x.ctypes.data, it.ctypes.data, rel.ctypes.data, refs.ctypes.data)
        if rc != 0:
            raise MhxError(rc, "mhx_batch_results")
        return x, it, rel, refs

    def instance(self, x, i):
        return x[self.offsets[i]:self.offsets[i + 1]]


class Server:
    """Request loop over mhx_server_*: independent steady solves of mixed mesh
    size, admitted by device-memory footprint, run concurrently (one stream
    per active request), with idle contexts of the same mesh reused."""

    def __init__(self, max_active: int = 8, max_cached: int = 32, hbm_budget: int = 0, solver: str = "cg",
                 precond: str = "jacobi", tol: float = 1e-10, max_iters: int = 20000, newton_iters: int = 2,
                 nl_tol: float = 1e-8, inner_tol: float = 0.0, restart: int = 50):
        d = capi.ServerDesc()
        d.hbm_budget, d.max_active, d.max_cached = int(hbm_budget), int(max_active), int(max_cached)
        d.newton.max_iters, d.newton.nl_tol, d.newton.backtracking = int(newton_iters), float(nl_tol), 1
        d.newton.linear = krylov_opts(solver=solver, tol=tol, max_iters=max_iters, restart=restart,
                                      precond=precond, inner_tol=inner_tol)
        h = C.c_void_p()
        rc = lib.mhx_server_create(C.byref(d), C.byref(h))
        if rc != 0:
            raise MhxError(rc, "mhx_server_create")
        self._h = h
        self._problems = {}

    def close(self):
        if getattr(self, "_h", None):
            lib.mhx_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, problem: Problem) -> int:
        desc = problem.desc()
        t = C.c_int64()
        rc = lib.mhx_server_submit(self._h, C.byref(desc), C.byref(t))
        if rc != 0:
            raise MhxError(rc, "mhx_server_submit")
        self._problems[t.value] = problem.n_dof
        return t.value

    def poll(self, ticket: int, wait: bool = True):
        """(x, stats, latency_ms) once done, None while pending (wait=False)."""
        x = np.empty(self._problems[ticket])
        st = capi.SolveStats()
        lat = C.c_double()
        rc = lib.mhx_server_poll(self._h, int(ticket), int(wait), x.ctypes.data, C.byref(st), C.byref(lat))
        if rc == 0:
            return None
        del self._problems[ticket]
        if rc < 0:
            raise MhxError(rc, "request failed: " + lib.mhx_server_last_error(self._h).decode())
        return x, st.as_dict(), lat.value

    def info(self) -> dict:
        out = np.zeros(8, dtype=np.int64)
        lib.mhx_server_info(self._h, out.ctypes.data)
        keys = ["done", "active", "queued", "bytes_held", "reused", "created", "idle", "budget"]
        return {k: int(v) for k, v in zip(keys, out)}

    @staticmethod
    def footprint(problem: Problem, opts: Optional[capi.KrylovOpts] = None):
        desc = problem.desc()
        n, z, b = C.c_int64(), C.c_int64(), C.c_int64()
        rc = lib.mhx_footprint(C.byref(desc), C.byref(opts) if opts is not None else None, C.byref(n), C.byref(z),
                               C.byref(b))
        if rc != 0:
            raise MhxError(rc, "mhx_footprint: " + lib.mhx_last_error(None).decode())
        return n.value, z.value, b.value
