"""Extract the reference's own known-answer data into small JSON fixtures.

Reads (from a MrHyDE checkout, default /root/reference):
  regression/discretization/HGRAD/mrhyde.gold     basis values/gradients at Gauss points
  regression/thermal/*/mrhyde.gold                L2 error norms of manufactured solutions
  regression/navierstokes/channel/mrhyde.gold     L2 error norms of the channel flow
and writes tests/golden/hgrad_kat.json, thermal_gold.json and ns_gold.json.
Only numbers (inputs and expected outputs) are stored, with each deck's
"Functions" strings (its coefficient expressions, input data); the deck
parameters of each case are transcribed next to its gold values.

Usage: python tests/golden/make_golden.py [/path/to/MrHyDE]
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
PI = "pi"


def parse_hgrad(path: Path) -> dict:
    out = {"quad": {"phi": {}, "grad": {}}, "hex": {"phi": {}, "grad": {}}}
    cur = None
    for line in path.read_text().splitlines():
        m = re.match(r"description = (\w+)", line)
        if m:
            desc = m.group(1)
            cur = ("hex" if "Hex" in desc else "quad", "phi" if desc.startswith("Projection") else "grad")
            continue
        m = re.match(r"dof (\d+), point (\d+) grad: \(([^)]*)\)", line)
        if m and cur:
            out[cur[0]]["grad"].setdefault(m.group(1), {})[m.group(2)] = [float(v) for v in m.group(3).split(",")]
            continue
        m = re.match(r"dof (\d+), point (\d+): (\S+)", line)
        if m and cur:
            out[cur[0]]["phi"].setdefault(m.group(1), {})[m.group(2)] = float(m.group(3))
    res = {}
    for top in ("quad", "hex"):
        nd = len(out[top]["phi"])
        res[top] = {
            "phi": [[out[top]["phi"][str(d)][str(q)] for q in range(nd)] for d in range(nd)],
            "grad": [[out[top]["grad"][str(d)][str(q)] for q in range(nd)] for d in range(nd)],
        }
    return res


def norms(path: Path) -> dict:
    vals = {}
    for line in path.read_text().splitlines():
        m = re.search(r"L2 norm of the error for (\S+) = (\S+)", line)
        if m and m.group(1) not in vals:
            vals[m.group(1)] = float(m.group(2))
        m = re.search(r"L2-face norm of the error for (\S+) = (\S+)", line)
        if m and ("face " + m.group(1)) not in vals:
            vals["face " + m.group(1)] = float(m.group(2))
    return vals


# deck transcriptions: (gold directory, mesh/discretisation/source/true solution)
# source and true solution as sums of amp * prod_d sin(w_d x_d) with w in units of pi
CASES = {
    "2D_verification": dict(dim=2, order=1, quadrature=2, n=[40, 40], workset=100,
                            source=[[8.0, [2, 2]]], true=[[1.0, [2, 2]]], nl_iters=2, nl_tol=1e-7),
    "2D_LinearSolvers": dict(dim=2, order=1, quadrature=2, n=[8, 8], workset=100,
                             source=[[8.0, [2, 2]]], true=[[1.0, [2, 2]]], nl_iters=2, nl_tol=1e-7),
    "3D_verification": dict(dim=3, order=1, quadrature=2, n=[10, 10, 10], workset=100,
                            source=[[12.0, [2, 2, 2]]], true=[[1.0, [2, 2, 2]]], nl_iters=2, nl_tol=1e-7),
    "2D_verification_highorder": dict(dim=2, order=4, quadrature=8, n=[10, 10], workset=10,
                                      source=[[8.0, [2, 2]]], true=[[1.0, [2, 2]]], nl_iters=2, nl_tol=1e-11),
    # Dirichlet on left/right, Neumann on bottom/top with the flux field
    # q = (0, 2 pi sin(2 pi x) cos(2 pi y)): terms [amp, freq, phase], all in units of pi
    "2D_mixed_bcs": dict(dim=2, order=1, quadrature=2, n=[40, 40], workset=100,
                         source=[[8.0, [2, 2]]], true=[[1.0, [2, 2]]], nl_iters=4, nl_tol=1e-7,
                         dbc_sides=3, neumann_sides=12, neumann_flux=[[[0.0, 2.0], [2.0, 2.0], [0.0, 0.5]]]),
}


# navier stokes: regression/navierstokes/channel (Q1-Q1, PSPG, body force 1 in x)
NS_CASES = {
    "navierstokes/channel": dict(dim=2, order=1, quadrature=2, n=[50, 10], lo=[0.0, 0.0], hi=[5.0, 1.0],
                                 vel_dbc_sides=12, pspg=True, supg=False, force=[1.0, 0.0],
                                 rho=1.0, mu=1.0),
}


# transient: regression/thermal/2D_verification_transient (BWE, BDF order 1,
# 40x40 Q1, dt = 1/20, T(t=0) = 0, T = 0 on the boundary); the source
# (8 pi^2 sin(2 pi t) + 2 pi cos(2 pi t)) sin(2 pi x) sin(2 pi y) is stored as
# the two time factors' amplitudes, the true solution sin(2 pi t) sin(2 pi x) sin(2 pi y)
TRANSIENT_CASES = {
    "2D_verification_transient": dict(dim=2, order=1, quadrature=2, n=[40, 40], workset=100, final_time=1.0,
                                      steps=20, bdf_order=1, tableau="BWE",
                                      source_sin_amp_pi2=8.0, source_cos_amp_pi=2.0, omega_pi=2.0,
                                      space_freq_pi=[2, 2]),
}


def transient_norms(path: Path) -> list:
    out = []
    for line in path.read_text().splitlines():
        m = re.search(r"L2 norm of the error for (\S+) = (\S+)\s+\(time = (\S+)\)", line)
        if m:
            out.append([float(m.group(3)), float(m.group(2))])
    return out


def deck_functions(deck_dir: Path) -> dict:
    """The deck's "Functions" block verbatim (name -> expression string), from
    input.yaml or, where a directory holds several decks, the first input*.yaml."""
    import yaml

    files = [deck_dir / "input.yaml"] if (deck_dir / "input.yaml").exists() else sorted(deck_dir.glob("input*.yaml"))
    if not files:
        return {}
    doc = yaml.safe_load(files[0].read_text())

    def find(o):
        if isinstance(o, dict):
            for k, v in o.items():
                if k == "Functions" and isinstance(v, dict):
                    return {str(a): str(b) for a, b in v.items() if not isinstance(b, dict)}
                r = find(v)
                if r is not None:
                    return r
        return None

    return find(doc) or {}


def main(ref: Path):
    kat = parse_hgrad(ref / "regression/discretization/HGRAD/mrhyde.gold")
    (HERE / "hgrad_kat.json").write_text(json.dumps(kat, indent=1))
    gold = {}
    for name, deck in CASES.items():
        g = norms(ref / "regression/thermal" / name / "mrhyde.gold")
        gold[name] = dict(deck, functions=deck_functions(ref / "regression/thermal" / name), gold=g)
    gold["_note"] = ("amplitudes of 'source' are in units of pi^2 (source = amp*pi^2*prod sin(w*pi*x)); "
                     "'true' amplitudes are plain; frequencies are in units of pi")
    (HERE / "thermal_gold.json").write_text(json.dumps(gold, indent=1))
    ns = {}
    for name, deck in NS_CASES.items():
        ns[name] = dict(deck, gold=norms(ref / "regression" / name / "mrhyde.gold"))
    ns["_note"] = ("velocity Dirichlet sides as a mask (left 1, right 2, bottom 4, top 8); true solution "
                   "ux = 0.5*y*(1-y), uy = pr = 0; density and viscosity at their defaults (1)")
    (HERE / "ns_gold.json").write_text(json.dumps(ns, indent=1))
    tr = {}
    for name, deck in TRANSIENT_CASES.items():
        tr[name] = dict(deck, functions=deck_functions(ref / "regression/thermal" / name),
                        gold=transient_norms(ref / "regression/thermal" / name / "mrhyde.gold"))
    tr["_note"] = "gold: [time, L2 norm of the error for T] per step"
    (HERE / "transient_gold.json").write_text(json.dumps(tr, indent=1))
    print("wrote", HERE / "hgrad_kat.json", HERE / "thermal_gold.json", HERE / "ns_gold.json",
          HERE / "transient_gold.json")


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference"))
