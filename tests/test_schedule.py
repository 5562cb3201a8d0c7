"""CPU replay of the assembly kernels' scatter schedules (k_thermal, k_ns).

The three schedules of `assembly.colouring` are restated here from the kernel's
index arithmetic (lattice.hpp `element_of`, `element_of_seg`, `axis_key`;
assemble.hip `launch_colors`, the first-touch flags and the x-segment LDS image)
and replayed on small lattices:

* every element is assembled exactly once;
* race freedom: two work units that may run at the same time never write the
  same CRS entry or residual row (element / segment colouring: one launch;
  block colouring: one launch, different workgroups, or one workgroup and one
  sub-phase);
* first touch: every entry of the CRS pattern has exactly one writer whose
  first-touch flag is set, and it runs strictly before every other writer of
  that entry, so overwrite mode needs no zero fill;
* the segment image: the LDS slot an element's (row, column) contribution is
  summed into decodes, in the write-out loop, to the same global entry, and
  the workgroup's strided loops visit every slot exactly once.

The reference adds every contribution with atomics and a column search
(assemblyManager_scatter.hpp:164-278); these are the invariants that let the
coloured, atomic-free form produce the same matrix.
"""
from __future__ import annotations

import itertools

import pytest

# ---------------------------------------------------------------------------
# kernel configuration (assemble.hip Cfg / SegCfg)
# ---------------------------------------------------------------------------


def pow2ceil(x):
    r = 1
    while r < x:
        r *= 2
    return r


def cfg(dim, p, block=256):
    nb = (p + 1) ** dim
    g = max(4, pow2ceil(nb + 1))
    return dict(NB=nb, G=g, GPB=block // g, BLOCK=block)


def seg_cfg(dim, p, s):
    p1 = p + 1
    rp = p1 ** (dim - 1)
    cp = rp
    wv, wi = 2 * p + 1, p + 1
    sv, si = rp * cp * wv, rp * cp * wi
    gs = sv + (p - 1) * si
    return dict(P1=p1, RP=rp, CP=cp, WV=wv, WI=wi, SV=sv, SI=si, GS=gs, NPL=p * s + 1, kSize=s * gs + sv,
                kRows=(p * s + 1) * rp)


def plane_off(sc, p, x):
    return (x // p) * sc["GS"] + (0 if x % p == 0 else sc["SV"] + (x % p - 1) * sc["SI"])


def block_shape(dim, gpb):
    s = [1, 1, 1]
    prod = 1
    while prod < gpb:
        d = 0
        for k in range(1, dim):
            if s[k] < s[d]:
                d = k
        s[d] *= 2
        prod *= 2
    return s


# ---------------------------------------------------------------------------
# launches: (launch index, workgroup, sub-phase, unit) per element, as the
# kernel enumerates them
# ---------------------------------------------------------------------------
ELEM, BLOCKED, SEG = 0, 1, 2


def schedule(mode, dim, nel, gpb):
    """Yield dicts: launch, wg, sub, elems (list of ec tuples), seg (sx, s_act) or None,
    plus the AsmArgs fields the key functions read."""
    nel = list(nel) + [1] * (3 - len(nel))
    s = block_shape(dim, gpb)
    bw = [2 * x for x in s]
    nseg = (nel[0] + gpb - 1) // gpb
    ncol = 1 << dim
    for col in range(ncol):
        cn = [0, 0, 0]
        cnt = 1
        for d in range(3):
            cd = (col >> d) & 1
            units = nel[d] if d < dim else 1
            if d < dim and mode == BLOCKED:
                units = (nel[d] + bw[d] - 1) // bw[d]
            if d == 0 and mode == SEG:
                units = nseg
            n = (units - cd + 1) // 2 if d < dim else 1
            if d >= dim and cd:
                n = 0
            cn[d] = n
            cnt *= n
        if cnt == 0:
            continue
        blocks = (cnt + gpb - 1) // gpb if mode == ELEM else cnt
        for b in range(blocks):
            if mode == ELEM:
                for g in range(gpb):
                    idx = b * gpb + g
                    if idx >= cnt:
                        continue
                    t = idx
                    e0 = 2 * (t % cn[0]) + (col & 1)
                    t //= cn[0]
                    e1 = 2 * (t % cn[1]) + ((col >> 1) & 1)
                    t //= cn[1]
                    e2 = 2 * t + ((col >> 2) & 1) if dim == 3 else 0
                    yield dict(launch=col, wg=b, sub=0, elems=[(e0, e1, e2)], seg=None, bw=bw, nseg=nseg)
            elif mode == BLOCKED:
                for sub in range(ncol):
                    for g in range(gpb):
                        gc = [g % s[0], (g // s[0]) % s[1], g // (s[0] * s[1])]
                        t = b
                        ec = [0, 0, 0]
                        ok = True
                        for d in range(dim):
                            bi = 2 * (t % cn[d]) + ((col >> d) & 1)
                            t //= cn[d]
                            ec[d] = bi * bw[d] + 2 * gc[d] + ((sub >> d) & 1)
                            ok = ok and ec[d] < nel[d]
                        if ok:
                            yield dict(launch=col, wg=b, sub=sub, elems=[tuple(ec)], seg=None, bw=bw, nseg=nseg)
            else:
                t = b
                sx = 2 * (t % cn[0]) + (col & 1)
                t //= cn[0]
                e1 = 2 * (t % cn[1]) + ((col >> 1) & 1)
                t //= cn[1]
                e2 = 2 * t + ((col >> 2) & 1) if dim == 3 else 0
                rest = nel[0] - sx * gpb
                s_act = min(rest, gpb)
                elems = [(sx * gpb + g, e1, e2) for g in range(gpb) if g < s_act]
                yield dict(launch=col, wg=b, sub=0, elems=elems, seg=(sx, s_act), bw=bw, nseg=nseg)


def axis_key(mode, bw, d, e):
    return (((e // bw[d]) & 1) * 2 if mode == BLOCKED else 0) + (e & 1)


def first_flags(mode, dim, nel, unit, ec):
    lo, hi = [True] * 3, [True] * 3
    for d in range(dim):
        if mode == SEG and d == 0:
            sx = unit["seg"][0]
            lo[0] = sx == 0 or (sx & 1) < ((sx - 1) & 1)
            hi[0] = sx == unit["nseg"] - 1 or (sx & 1) < ((sx + 1) & 1)
            continue
        k = axis_key(mode, unit["bw"], d, ec[d])
        lo[d] = ec[d] == 0 or k < axis_key(mode, unit["bw"], d, ec[d] - 1)
        hi[d] = ec[d] == nel[d] - 1 or k < axis_key(mode, unit["bw"], d, ec[d] + 1)
    return lo, hi


def local(i, p, dim):
    p1 = p + 1
    return (i % p1, (i // p1) % p1, i // (p1 * p1) if dim == 3 else 0)


def gnode(ec, loc, p):
    return tuple(p * ec[d] + loc[d] for d in range(3))


# ---------------------------------------------------------------------------
# writes of one work unit: {(row node, col node): first?}
# ---------------------------------------------------------------------------


def unit_writes_element(mode, dim, p, nel, unit):
    (ec,) = unit["elems"]
    lo, hi = first_flags(mode, dim, nel, unit, ec)
    nb = (p + 1) ** dim
    out = {}
    for i in range(nb):
        il = local(i, p, dim)
        for j in range(nb):
            jl = local(j, p, dim)
            f = True
            for d in range(dim):
                if il[d] == jl[d] and il[d] in (0, p):
                    f = f and (lo[d] if il[d] == 0 else hi[d])
            out[(gnode(ec, il, p), gnode(ec, jl, p))] = f
    return out


def unit_writes_segment(dim, p, nel, unit, nn):
    """The write-out loop of the x-segment scatter: decode every LDS slot
    back to (row, column) and apply the kernel's skip and first-touch rules."""
    sx, s_act = unit["seg"]
    gpb = unit["gpb"]
    sc = seg_cfg(dim, p, gpb)
    ec0 = unit["elems"][0]
    lo, hi = first_flags(SEG, dim, nel, unit, ec0)
    x0g = p * sx * gpb
    xend = p * s_act
    yb, zb = p * ec0[1], p * ec0[2]
    p1 = p + 1
    out = {}
    for eix in range(sc["kSize"]):
        grp, rem = divmod(eix, sc["GS"])
        if rem < sc["SV"]:
            X = p * grp
            r = rem // (sc["CP"] * sc["WV"])
            cb = (rem // sc["WV"]) % sc["CP"]
            xo = rem % sc["WV"]
        else:
            r2 = rem - sc["SV"]
            X = p * grp + 1 + r2 // sc["SI"]
            rr = r2 % sc["SI"]
            r = rr // (sc["CP"] * sc["WI"])
            cb = (rr // sc["WI"]) % sc["CP"]
            xo = rr % sc["WI"]
        xg = x0g + X
        if X > xend or xg >= nn[0]:
            continue
        xcol = xg - p + xo if X % p == 0 else xg - X % p + xo
        if xcol < 0 or xcol >= nn[0]:
            continue
        if (X == 0 and xcol < xg) or (X == xend and xcol > xg):
            continue
        iy, iz = r % p1, (r // p1 if dim == 3 else 0)
        jy, jz = cb % p1, (cb // p1 if dim == 3 else 0)
        sh = True
        if X == 0 and xcol == xg:
            sh = sh and lo[0]
        if X == xend and xcol == xg:
            sh = sh and hi[0]
        if iy == jy == 0:
            sh = sh and lo[1]
        if iy == jy == p:
            sh = sh and hi[1]
        if dim == 3 and iz == jz == 0:
            sh = sh and lo[2]
        if dim == 3 and iz == jz == p:
            sh = sh and hi[2]
        out[((xg, yb + iy, zb + iz), (xcol, yb + jy, zb + jz))] = (sh, eix)
    return out


def element_entries(ec, p, dim):
    nb = (p + 1) ** dim
    return {(gnode(ec, local(i, p, dim), p), gnode(ec, local(j, p, dim), p)) for i in range(nb) for j in range(nb)}


def pattern_entries(dim, p, nel):
    """The CRS pattern: union of element couplings (linearAlgebraInterface_construct.hpp:243-258)."""
    nel3 = list(nel) + [1] * (3 - len(nel))
    ents = set()
    for ec in itertools.product(*(range(n) for n in nel3)):
        ents |= element_entries(ec, p, dim)
    return ents


# ---------------------------------------------------------------------------
# checks
# ---------------------------------------------------------------------------
CASES = [
    # dim, p, nel
    (2, 1, (9, 7)),
    (2, 2, (11, 6)),
    (2, 3, (7, 5)),
    (3, 1, (9, 5, 4)),
    (3, 2, (9, 4, 3)),
    (3, 2, (17, 3, 3)),
]


def replay(mode, dim, p, nel, gpb):
    nel3 = list(nel) + [1] * (3 - len(nel))
    nn = [p * n + 1 for n in nel3]
    writes = {}  # entry -> list of (time, concurrency key, first)
    seen = {}
    for unit in schedule(mode, dim, nel, gpb):
        unit["gpb"] = gpb
        for ec in unit["elems"]:
            seen[ec] = seen.get(ec, 0) + 1
        if mode == SEG:
            w = {k: v[0] for k, v in unit_writes_segment(dim, p, nel3, unit, nn).items()}
            # every contribution of the segment's elements is in its image
            covered = set()
            for ec in unit["elems"]:
                covered |= element_entries(ec, p, dim)
            assert covered == set(w), "segment image misses or adds entries"
        else:
            w = unit_writes_element(mode, dim, p, nel3, unit)
        # time: launch, then sub-phase inside one workgroup (block colouring)
        for ent, f in w.items():
            writes.setdefault(ent, []).append(((unit["launch"], unit["sub"]), unit["wg"], f))
    return nel3, seen, writes


@pytest.mark.parametrize("mode", [ELEM, BLOCKED, SEG], ids=["element", "block", "segment"])
@pytest.mark.parametrize("gpb", [4, 8])
@pytest.mark.parametrize("dim,p,nel", CASES)
def test_schedule_race_free_and_first_touch(mode, gpb, dim, p, nel):
    nel3, seen, writes = replay(mode, dim, p, nel, gpb)
    # every element exactly once
    all_el = set(itertools.product(*(range(n) for n in nel3)))
    assert set(seen) == all_el and all(v == 1 for v in seen.values())
    # the union of writes is the CRS pattern
    assert set(writes) == pattern_entries(dim, p, nel)
    for ent, ws in writes.items():
        # race freedom: writers with the same launch run concurrently unless
        # they are sub-phases of one workgroup (barrier between sub-phases)
        for (t1, wg1, _), (t2, wg2, _) in itertools.combinations(ws, 2):
            if t1[0] == t2[0]:
                assert mode == BLOCKED and wg1 == wg2 and t1[1] != t2[1], f"race on {ent}: {ws}"
        firsts = [w for w in ws if w[2]]
        assert len(firsts) == 1, f"entry {ent}: {len(firsts)} first touchers in {ws}"
        t0 = firsts[0][0]
        assert all(w[0] > t0 for w in ws if not w[2]), f"entry {ent}: a later touch runs before the first"


def test_residual_rows_race_free():
    """Residual rows follow the diagonal entries' schedule: the same replay
    restricted to (row, row)."""
    for mode in (ELEM, BLOCKED, SEG):
        nel3, _, writes = replay(mode, 3, 2, (9, 4, 3), 8)
        diag = {e: w for e, w in writes.items() if e[0] == e[1]}
        nn = [2 * n + 1 for n in nel3]
        assert len(diag) == nn[0] * nn[1] * nn[2]
        for ws in diag.values():
            assert sum(1 for w in ws if w[2]) == 1


# ---------------------------------------------------------------------------
# the x-segment LDS image
# ---------------------------------------------------------------------------
BUILT = [(2, p) for p in range(1, 7)] + [(3, 1), (3, 2)]


@pytest.mark.parametrize("dim,p", BUILT)
def test_segment_image_slot_decodes_to_its_entry(dim, p):
    """The slot an element's contribution is summed into (scatter phase) and
    the (row, column) the write-out loop decodes from that slot agree, and
    no two contributions of different entries share a slot."""
    c = cfg(dim, p)
    gpb = c["GPB"]
    sc = seg_cfg(dim, p, gpb)
    p1 = p + 1
    nel = (gpb, 1, 1) if dim == 2 else (gpb, 1, 1)
    nn = [p * n + 1 for n in nel]
    unit = dict(launch=0, wg=0, sub=0, elems=[(g, 0, 0) for g in range(gpb)], seg=(0, gpb), bw=[2, 2, 2], nseg=1,
                gpb=gpb)
    decoded = {v[1]: k for k, v in unit_writes_segment(dim, p, list(nel), unit, nn).items()}
    slot_of = {}
    for g in range(gpb):
        for i in range(c["NB"]):
            ix, iy, iz = local(i, p, dim)
            for j in range(c["NB"]):
                jx, jy, jz = local(j, p, dim)
                vtx = ix == 0 or ix == p
                w = sc["WV"] if vtx else sc["WI"]
                xo = jx - ix + p if vtx else jx
                idx = plane_off(sc, p, p * g + ix) + ((iy + p1 * iz) * sc["CP"] + (jz * p1 + jy)) * w + xo
                assert 0 <= idx < sc["kSize"]
                ent = ((p * g + ix, iy, iz), (p * g + jx, jy, jz))
                assert slot_of.setdefault(ent, idx) == idx
                assert decoded[idx] == ent
    assert len(set(slot_of.values())) == len(slot_of)
    # residual rows: (plane, row-in-plane) slots inside kRows
    assert max((p * (gpb - 1) + p) * sc["RP"] + (p1 ** (dim - 1) - 1), 0) < sc["kRows"]


def strided_visits(n, block, stride, chunk=1):
    """How often each slot is visited by `for (b = tid; b < n; b += stride * chunk)
    for c < chunk: k = b + c * stride` over tid < block."""
    cnt = [0] * n
    for tid in range(block):
        b = tid
        while b < n:
            for c in range(chunk):
                k = b + c * stride
                if k < n:
                    cnt[k] += 1
            b += stride * chunk
    return cnt


@pytest.mark.parametrize("dim,p", BUILT)
@pytest.mark.parametrize("block", [128, 256])
def test_segment_loops_visit_every_slot_once(dim, p, block):
    """The zeroing, residual and write-out loops stride by the workgroup size.
    With the stride equal to the block every slot is visited once; a fixed
    256-lane stride in a 128-lane build leaves half the image unzeroed and
    unwritten (the disagreement a 128-lane build once showed)."""
    c = cfg(dim, p, block)
    if c["GPB"] < 2:
        pytest.skip("fewer than two elements per workgroup")
    sc = seg_cfg(dim, p, c["GPB"])
    for n, chunk in ((sc["kSize"], 1), (sc["kRows"], 1), (sc["kSize"], 8)):
        assert strided_visits(n, block, block, chunk) == [1] * n
    if block == 128:
        bad = strided_visits(sc["kSize"], 128, 256)
        assert 0 in bad
