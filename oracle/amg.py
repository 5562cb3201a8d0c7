"""CPU restatement of the lattice smoothed-aggregation AMG hierarchy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
hierarchy (mrhyde_amd/csrc/amg.hip); the product never imports it.

The reference builds its default preconditioner with
MueLu::CreateTpetraPreconditioner (linearAlgebraInterface_solvers.hpp:201-253:
"multigrid algorithm" sa, uncoupled aggregation, Chebyshev degree 2 with
eigenvalue ratio 7, coarse max size 500).  MueLu (Trilinos) is not in the
reference tree; this file restates the published smoothed-aggregation method
(Vanek, Mandel, Brezina, Computing 56 (1996)) with MueLu's defaults
(damping 4/3 over a 10-step power-iteration estimate of rho(D^-1 A),
R = P^T, Ifpack2's Chebyshev boost 1.1), on lattice-block aggregates of
(2r+1)^d nodes for a stencil of reach r.

Two routes are kept apart on purpose:
  * `build_level` follows the device kernels' index arithmetic (closed-form
    box row pointers, per-row slot boxes) and asserts every index against the
    size of the buffer it addresses;
  * the tests compare its P, R and A_c with scipy products of the same
    operators ((I - omega D^-1 A) T, P^T, R A P), which share no code with it.
Parity pinned by: T ns_c = ns (the tentative prolongator reproduces the
near-null space), the scipy Galerkin products, and PCG convergence with the
V-cycle below on SPD problems.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

MAX_LEVELS = 10
COARSE_MAX = 128  # the device coarsens to at most 128 rows before its dense coarsest solve
POWER_ITERS = 10
DAMPING = 4.0 / 3.0
CHEB_RATIO = 7.0
CHEB_BOOST = 1.1


class Axis:
    """Per-axis tables of one level (see amg.hip make_axis)."""

    def __init__(self, n: int, r: int, rc: int):
        self.n, self.b = n, 2 * r + 1
        self.na = max(1, n // self.b)
        b, na = self.b, self.na
        p = np.arange(n)
        self.agg = np.minimum(p // b, na - 1)
        self.plo = self.agg[np.maximum(0, p - r)]
        self.pext = np.minimum(self.plo + 1, na - 1) - self.plo + 1
        self.aplo = self.agg[np.maximum(0, p - 2 * r)]
        self.apext = np.minimum(self.aplo + 2, na - 1) - self.aplo + 1
        self.sp = np.concatenate([[0], np.cumsum(self.pext)])
        rlo = np.full(na, 1 << 30)
        rhi = np.full(na, -1)
        for q in range(n):
            for I in range(self.plo[q], self.plo[q] + self.pext[q]):
                rlo[I] = min(rlo[I], q)
                rhi[I] = max(rhi[I], q)
        assert (rhi >= 0).all()
        self.rlo, self.rext = rlo, rhi - rlo + 1
        self.sr = np.concatenate([[0], np.cumsum(self.rext)])
        I = np.arange(na)
        self.clo = np.maximum(0, I - rc)
        self.cext = np.minimum(na - 1, I + rc) - self.clo + 1
        self.sc = np.concatenate([[0], np.cumsum(self.cext)])


def box_rowptr(N, E, S, nf):
    """Closed-form row pointers of a separable box pattern (k_box_rowptr)."""
    nrows = N[0] * N[1] * N[2] * nf
    rp = np.empty(nrows + 1, dtype=np.int64)
    Tx, Ty = S[0][N[0]], S[1][N[1]]
    for row in range(nrows):
        node, f = divmod(row, nf)
        x, yz = node % N[0], node // N[0]
        y, z = yz % N[1], yz // N[1]
        base = nf * nf * (S[2][z] * Ty * Tx + E[2][z] * (S[1][y] * Tx + E[1][y] * S[0][x]))
        rp[row] = base + f * E[0][x] * E[1][y] * E[2][z] * nf
    rp[nrows] = nf * nf * S[0][N[0]] * S[1][N[1]] * S[2][N[2]]
    # the closed form equals the running sum of the row lengths
    lens = np.array([E[0][(r // nf) % N[0]] * E[1][(r // nf // N[0]) % N[1]] * E[2][r // nf // N[0] // N[1]] * nf
                     for r in range(nrows)], dtype=np.int64)
    assert np.array_equal(rp, np.concatenate([[0], np.cumsum(lens)]))
    return rp


def diag_inv(A: sp.csr_matrix):
    d = A.diagonal()
    out = np.zeros_like(d)
    nz = d != 0
    out[nz] = 1.0 / d[nz]
    return out


def power_estimate(A: sp.csr_matrix, dinv):
    """10 steps from v_i = 0.5 + frac(i * 0.618...), as the device does."""
    n = A.shape[0]
    x = np.arange(n, dtype=np.float64) * 0.6180339887498949
    v = 0.5 + (x - np.floor(x))
    s = 1.0 / np.sqrt(v @ v)
    lam = 1.0
    for _ in range(POWER_ITERS):
        w = s * dinv * (A @ v)
        lam = s * (v @ w)
        wn = np.sqrt(w @ w)
        if not (wn > 0) or not np.isfinite(wn):
            break
        v, s = w, 1.0 / wn
    return lam if (np.isfinite(lam) and lam > 0) else 1.0


def build_level(A: sp.csr_matrix, N, nf: int, reach: int, ns):
    """One coarsening step.  Returns dict(P, R, Ac, ns_c, Nc, rc, lam, omega,
    axes) or None when the lattice cannot coarsen further."""
    r = max(1, reach)
    b = 2 * r + 1
    rc = (3 * r - 1) // b + 1
    ax = [Axis(N[d], r, rc) for d in range(3)]
    Nc = [a.na for a in ax]
    n = A.shape[0]
    nc = Nc[0] * Nc[1] * Nc[2] * nf
    if nc >= n:
        return None
    A = A.tocsr()
    rp, ci, av = A.indptr, A.indices, A.data
    dinv = diag_inv(A)
    lam = power_estimate(A, dinv)
    omega = DAMPING / lam

    def xyz(node, dims):
        return node % dims[0], (node // dims[0]) % dims[1], node // (dims[0] * dims[1])

    # aggregate norms of the near-null space (k_amg_nrm)
    nrm = np.zeros(nc)
    for cr in range(nc):
        I, f = divmod(cr, nf)
        Ic = xyz(I, Nc)
        lo = [Ic[d] * ax[d].b for d in range(3)]
        hi = [ax[d].n - 1 if Ic[d] == ax[d].na - 1 else lo[d] + ax[d].b - 1 for d in range(3)]
        s = 0.0
        for z in range(lo[2], hi[2] + 1):
            for y in range(lo[1], hi[1] + 1):
                for x in range(lo[0], hi[0] + 1):
                    idx = (x + N[0] * (y + N[1] * z)) * nf + f
                    assert 0 <= idx < n
                    s += ns[idx] ** 2
        nrm[cr] = np.sqrt(s)

    def t_entry(j):
        q, g = divmod(j, nf)
        qc = xyz(q, N)
        a3 = [int(ax[d].agg[qc[d]]) for d in range(3)]
        if ns[j] == 0:
            return 0.0, a3, g
        I = a3[0] + Nc[0] * (a3[1] + Nc[1] * a3[2])
        assert 0 <= I * nf + g < nc
        nv = nrm[I * nf + g]
        return (ns[j] / nv if nv > 0 else 0.0), a3, g

    # P (k_amg_P)
    prp = box_rowptr(N, [a.pext for a in ax], [a.sp for a in ax], nf)
    pnnz = prp[-1]
    pci = np.full(pnnz, -1, dtype=np.int64)
    pv = np.zeros(pnnz)
    for i in range(n):
        p = i // nf
        pc = xyz(p, N)
        lo = [int(ax[d].plo[pc[d]]) for d in range(3)]
        ex = [int(ax[d].pext[pc[d]]) for d in range(3)]
        cnt = ex[0] * ex[1] * ex[2] * nf
        assert prp[i + 1] - prp[i] == cnt
        acc = np.zeros(cnt)

        def slot(a3, g):
            dx, dy, dz = a3[0] - lo[0], a3[1] - lo[1], a3[2] - lo[2]
            assert 0 <= dx < ex[0] and 0 <= dy < ex[1] and 0 <= dz < ex[2], "column outside the prolongator box"
            return ((dz * ex[1] + dy) * ex[0] + dx) * nf + g

        ti, a3, g = t_entry(i)
        if ti != 0:
            acc[slot(a3, g)] += ti
        sc = -omega * dinv[i]
        for k in range(rp[i], rp[i + 1]):
            tj, a3, g = t_entry(ci[k])
            if tj == 0:
                continue
            acc[slot(a3, g)] += sc * av[k] * tj
        for s in range(cnt):
            gg, cell = s % nf, s // nf
            dx, dy, dz = cell % ex[0], (cell // ex[0]) % ex[1], cell // (ex[0] * ex[1])
            J = (lo[0] + dx) + Nc[0] * ((lo[1] + dy) + Nc[1] * (lo[2] + dz))
            assert 0 <= prp[i] + s < pnnz
            pci[prp[i] + s] = J * nf + gg
            pv[prp[i] + s] = acc[s]
    assert (pci >= 0).all() and (pci < nc).all()
    P = sp.csr_matrix((pv, pci, prp), shape=(n, nc))

    # R = P^T (k_amg_R)
    rrp = box_rowptr(Nc, [a.rext for a in ax], [a.sr for a in ax], nf)
    rnnz = rrp[-1]
    rci = np.full(rnnz, -1, dtype=np.int64)
    rv = np.zeros(rnnz)
    for cr in range(nc):
        I, f = divmod(cr, nf)
        Ic = xyz(I, Nc)
        rl = [int(ax[d].rlo[Ic[d]]) for d in range(3)]
        re = [int(ax[d].rext[Ic[d]]) for d in range(3)]
        for e in range(rrp[cr + 1] - rrp[cr]):
            g, cell = e % nf, e // nf
            x, y, z = rl[0] + cell % re[0], rl[1] + (cell // re[0]) % re[1], rl[2] + cell // (re[0] * re[1])
            assert x < N[0] and y < N[1] and z < N[2]
            i = (x + N[0] * (y + N[1] * z)) * nf + g
            o = [Ic[0] - ax[0].plo[x], Ic[1] - ax[1].plo[y], Ic[2] - ax[2].plo[z]]
            pe = [ax[0].pext[x], ax[1].pext[y], ax[2].pext[z]]
            assert all(0 <= o[d] < pe[d] for d in range(3))
            idx = prp[i] + ((o[2] * pe[1] + o[1]) * pe[0] + o[0]) * nf + f
            assert prp[i] <= idx < prp[i + 1]
            rci[rrp[cr] + e] = i
            rv[rrp[cr] + e] = pv[idx]
    R = sp.csr_matrix((rv, rci, rrp), shape=(nc, n))

    # A P in 3x3x3 boxes (k_amg_AP)
    AP = np.zeros((n * nf, 27))
    for t in range(n * nf):
        i, g = divmod(t, nf)
        pc = xyz(i // nf, N)
        alo = [int(ax[d].aplo[pc[d]]) for d in range(3)]
        aex = [int(ax[d].apext[pc[d]]) for d in range(3)]
        for k in range(rp[i], rp[i + 1]):
            j = ci[k]
            qc = xyz(j // nf, N)
            pl = [int(ax[d].plo[qc[d]]) for d in range(3)]
            pe = [int(ax[d].pext[qc[d]]) for d in range(3)]
            for dz in range(pe[2]):
                for dy in range(pe[1]):
                    for dx in range(pe[0]):
                        s3 = (pl[0] + dx - alo[0], pl[1] + dy - alo[1], pl[2] + dz - alo[2])
                        assert all(0 <= s3[d] < aex[d] for d in range(3)), "A P outside its box"
                        idx = prp[j] + ((dz * pe[1] + dy) * pe[0] + dx) * nf + g
                        assert prp[j] <= idx < prp[j + 1]
                        AP[t, (s3[2] * 3 + s3[1]) * 3 + s3[0]] += av[k] * pv[idx]

    # A_c = R (A P) (k_amg_Ac)
    arp = box_rowptr(Nc, [a.cext for a in ax], [a.sc for a in ax], nf)
    annz = arp[-1]
    aci = np.full(annz, -1, dtype=np.int64)
    acv = np.zeros(annz)
    for cr in range(nc):
        I, f = divmod(cr, nf)
        Ic = xyz(I, Nc)
        cl = [int(ax[d].clo[Ic[d]]) for d in range(3)]
        ce = [int(ax[d].cext[Ic[d]]) for d in range(3)]
        cnt = arp[cr + 1] - arp[cr]
        assert cnt <= 512
        for e in range(cnt):
            g, cell = e % nf, e // nf
            J = (cl[0] + cell % ce[0], cl[1] + (cell // ce[0]) % ce[1], cl[2] + cell // (ce[0] * ce[1]))
            col = (J[0] + Nc[0] * (J[1] + Nc[1] * J[2])) * nf + g
            s = 0.0
            for k in range(rrp[cr], rrp[cr + 1]):
                i = rci[k]
                pc = xyz(i // nf, N)
                s3 = [J[d] - ax[d].aplo[pc[d]] for d in range(3)]
                if any(s3[d] < 0 or s3[d] >= ax[d].apext[pc[d]] for d in range(3)):
                    continue
                s += rv[k] * AP[i * nf + g, (s3[2] * 3 + s3[1]) * 3 + s3[0]]
            aci[arp[cr] + e] = col
            acv[arp[cr] + e] = 1.0 if (col == cr and nrm[cr] == 0) else s
    Ac = sp.csr_matrix((acv, aci, arp), shape=(nc, nc))
    T = tentative(N, nf, ax, ns, nrm)
    return dict(P=P, R=R, Ac=Ac, ns_c=nrm, Nc=Nc, rc=rc, lam=lam, omega=omega, T=T, dinv=dinv, nrm=nrm)


def tentative(N, nf, ax, ns, nrm):
    """T as a sparse matrix (for the scipy cross-checks)."""
    Nc = [a.na for a in ax]
    n = N[0] * N[1] * N[2] * nf
    rows, cols, vals = [], [], []
    for i in range(n):
        q, g = divmod(i, nf)
        x, y, z = q % N[0], (q // N[0]) % N[1], q // (N[0] * N[1])
        I = ax[0].agg[x] + Nc[0] * (ax[1].agg[y] + Nc[1] * ax[2].agg[z])
        if ns[i] != 0 and nrm[I * nf + g] > 0:
            rows.append(i)
            cols.append(I * nf + g)
            vals.append(ns[i] / nrm[I * nf + g])
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, Nc[0] * Nc[1] * Nc[2] * nf))


def hierarchy(A: sp.csr_matrix, N, nf: int, reach: int, fixed):
    """All levels down to at most COARSE_MAX rows (or MAX_LEVELS)."""
    ns = np.where(fixed, 0.0, 1.0)
    levels = [dict(A=A.tocsr(), N=list(N), reach=reach)]
    while levels[-1]["A"].shape[0] > COARSE_MAX and len(levels) < MAX_LEVELS:
        L = levels[-1]
        out = build_level(L["A"], L["N"], nf, L["reach"], ns)
        if out is None:
            break
        L.update(out)
        ns = out["ns_c"]
        levels.append(dict(A=out["Ac"], N=out["Nc"], reach=out["rc"]))
    levels[-1]["Cinv"] = np.linalg.inv(levels[-1]["A"].toarray())
    return levels


def cheb_coef(lam):
    lmax = CHEB_BOOST * lam
    lmin = lmax / CHEB_RATIO
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    sigma = theta / delta
    rho0 = 1.0 / sigma
    rho1 = 1.0 / (2.0 * sigma - rho0)
    return 1.0 / theta, rho1 * rho0, 2.0 * rho1 / delta


def chebyshev(A, dinv, lam, b, x=None):
    """Two Chebyshev steps on D^-1 A (Ifpack2's three-term recurrence with the
    coefficients of cheb_coef), from x, or from zero when x is None (the
    pre-smoother, whose first residual is b).  The error x* - x is multiplied
    by the degree-2 residual polynomial T_2((theta - t) / delta) / T_2(theta / delta)
    of t = D^-1 A."""
    it, a2, b2 = cheb_coef(lam)
    if x is None:
        d = it * dinv * b
        x = d.copy()
    else:
        d = it * dinv * (b - A @ x)
        x = x + d
    d = a2 * d + b2 * dinv * (b - A @ x)
    return x + d


BJ_OMEGA = 0.4
BJ_SWEEPS = 3


def block_inverses(A: sp.csr_matrix, nf: int):
    """Inverse of every node's nf x nf diagonal block (k_amg_block_inv); a
    singular block gives a zero block.  Returns an array [n_nodes, nf, nf]."""
    n = A.shape[0] // nf
    out = np.zeros((n, nf, nf))
    A = A.tocsr()
    for k in range(n):
        blk = A[k * nf:(k + 1) * nf, k * nf:(k + 1) * nf].toarray()
        try:
            out[k] = np.linalg.inv(blk)
        except np.linalg.LinAlgError:
            pass
    return out


def node_colours(N, reach):
    """Colour of every lattice node for the multicolour sweep: the node's
    coordinates modulo s = reach + 1, numbered x fastest over the axes that
    have more than one node (gs_sweep)."""
    s = max(1, reach) + 1
    nc = [s if N[d] > 1 else 1 for d in range(3)]
    idx = np.arange(N[0] * N[1] * N[2])
    x, y, z = idx % N[0], (idx // N[0]) % N[1], idx // (N[0] * N[1])
    return (x % s) + nc[0] * ((y % s) + nc[1] * (z % s)), nc[0] * nc[1] * nc[2]


def block_apply(binv, r, nf):
    return np.einsum("kfg,kg->kf", binv, r.reshape(-1, nf)).reshape(-1)


def gs_sweep(A, binv, nf, N, reach, b, x, forward):
    """One multicolour nodal block Gauss-Seidel sweep, in place on x: colour
    by colour, x_S += Binv_S (b - A x)_S for the rows S of the colour's nodes."""
    col, ncol = node_colours(N, reach)
    rows_of = np.repeat(col, nf)
    order = range(ncol) if forward else range(ncol - 1, -1, -1)
    for c in order:
        S = np.nonzero(rows_of == c)[0]
        if S.size == 0:
            continue
        r = b[S] - A[S] @ x
        nodes = S[::nf] // nf
        x[S] += np.einsum("kfg,kg->kf", binv[nodes], r.reshape(-1, nf)).reshape(-1)
    return x


def vcycle(levels, b, smoother="chebyshev", nf=1):
    """x = M^-1 b, exact coarsest solve; pre/post smoother "chebyshev"
    (degree 2 on D^-1 A), "block_jacobi" (BJ_SWEEPS sweeps damped by
    BJ_OMEGA on each side) or
    "block_gauss_seidel" (one multicolour sweep, forward before the coarse
    correction and backward after it, so the cycle is symmetric for
    symmetric A)."""
    def rec(l, b):
        L = levels[l]
        if l == len(levels) - 1:
            return L["Cinv"] @ b
        A, dinv = L["A"], L["dinv"]
        b = np.asarray(b, dtype=np.float64)
        if smoother == "chebyshev":
            x = chebyshev(A, dinv, L["lam"], b)
        else:
            if "binv" not in L:
                L["binv"] = block_inverses(A, nf)
            if smoother == "block_jacobi":
                x = BJ_OMEGA * block_apply(L["binv"], b, nf)
                for _ in range(BJ_SWEEPS - 1):
                    x = x + BJ_OMEGA * block_apply(L["binv"], b - A @ x, nf)
            else:
                x = gs_sweep(A, L["binv"], nf, L["N"], L["reach"], b, np.zeros_like(b), True)
        xc = rec(l + 1, L["R"] @ (b - A @ x))
        x = x + L["P"] @ xc
        if smoother == "chebyshev":
            return chebyshev(A, dinv, L["lam"], b, x)
        if smoother == "block_jacobi":
            for _ in range(BJ_SWEEPS):
                x = x + BJ_OMEGA * block_apply(L["binv"], b - A @ x, nf)
            return x
        return gs_sweep(A, L["binv"], nf, L["N"], L["reach"], b, x, False)
    return rec(0, b)
