"""Synthetic problems of the benchmark configurations, from fixed seeds (the
reference has no benchmark decks; these are the deck subsets the engine
reads, with their coefficient fields drawn by the engine's counter-based
generator so every process builds the same problem).

  poisson_2d(n)      steady linear Poisson on [0,1]^2, n x n Q1 quads, 2x2 Gauss,
                     k = 1, f = 2 pi^2 sin(pi x) sin(pi y), T = 0 on all sides
  thermal_q2(n)      nonlinear thermal on [0,1]^3, n^3 Q2 hexes, 3x3x3 Gauss,
                     k = k0_e (1 + T^2), k0_e = exp(0.5 z_e), f_e ~ U[0,1],
                     T = 0 on x = 0 and x = 1, zero flux elsewhere
  cavity(n)          steady Navier-Stokes, Q1-Q1 with SUPG and PSPG, lid-driven
                     cavity on [0,1]^3 with n^3 hexes, Re = 100 (viscosity 0.01,
                     lid u = (1, 0, 0) on y = 1, no-slip elsewhere, pressure
                     pinned at the origin); cavity_state gives the seeded
                     N(0, 1e-3) initial velocity
  serve_requests(k)  thermal solve requests: 2-D Q1 n^2 (n log-uniform on
                     [32, 256]) or 3-D Q1/Q2 n^3 (n log-uniform on [16, 64]),
                     k = exp(0.5 z_e), a 1-4 mode sine source, T = 0 on the
                     boundary
"""
from __future__ import annotations

import numpy as np

from .abi import MHX_PHYS_NAVIER_STOKES, MHX_RNG_LOGNORMAL, MHX_RNG_NORMAL, MHX_RNG_UNIFORM, Problem, rng_fill

PI = np.pi


def poisson_2d(n: int = 64) -> Problem:
    return Problem(dim=2, order=1, quadrature=2, n=(n, n), modes=[(2 * PI * PI, (PI, PI))])


def thermal_q2(n: int = 96) -> Problem:
    ne = n ** 3
    k0 = rng_fill(2024, ne, MHX_RNG_LOGNORMAL, 0.0, 0.5)
    f = rng_fill(2025, ne, MHX_RNG_UNIFORM, 0.0, 1.0)
    return Problem(dim=3, order=2, quadrature=4, n=(n, n, n), k0=k0, src_elem=f, diff_u2=1.0, dbc_sides=1 | 2)


def cavity(n: int = 64) -> Problem:
    return Problem(dim=3, order=1, quadrature=2, n=(n, n, n), physics=MHX_PHYS_NAVIER_STOKES, mu=0.01, supg=True,
                   pspg=True, dbc_sides=1 | 2 | 4 | 16 | 32, lid_sides=8, lid_u=(1.0, 0.0, 0.0), pin_pressure=True)


def cavity_state(n_dof: int, seed: int = 2027) -> np.ndarray:
    """Initial state before the Dirichlet values are set: N(0, 1e-3) noise,
    drawn by the library's generator (libmhx.so)."""
    from .engine import rng

    return 1e-3 * rng(seed, n_dof, MHX_RNG_NORMAL, 0.0, 1.0)


def serve_requests(count: int, seed: int):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        if rng.random() < 0.5:
            dim, order, n = 2, 1, int(round(np.exp(rng.uniform(np.log(32), np.log(256)))))
        else:
            dim, order, n = 3, int(rng.integers(1, 3)), int(round(np.exp(rng.uniform(np.log(16), np.log(64)))))
        ne = n ** dim
        modes = [(float(rng.uniform(0, 1)), [PI * int(rng.integers(1, 5)) for _ in range(dim)])
                 for _ in range(int(rng.integers(1, 5)))]
        out.append(Problem(dim=dim, order=order, quadrature=2 * order, n=(n,) * dim,
                           k0=np.exp(0.5 * rng.standard_normal(ne)), modes=modes))
    return out


def poisson_arrivals(count: int, rate: float, seed: int) -> np.ndarray:
    """Arrival times (s) of a Poisson process of the given rate."""
    return np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, count))
