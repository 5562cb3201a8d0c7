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


# ---------------------------------------------------------------------------
# pencil streaming (assemble_q2box.hip, assembly.colouring = 3, Q2 hexahedra):
# a 64-lane wavefront walks an x-segment of one element pencil two elements
# per step.  Restated from the kernel's round / lane arithmetic: every write
# is (entry, contributing elements, first-touch flag); the replay checks that
# each CRS entry receives the contribution of every element holding both of
# its nodes exactly once, has exactly one first writer that runs before all
# other writers, and that no two writers of an entry can run at once.
# ---------------------------------------------------------------------------
def pencil_units(nel, seg):
    nseg = (nel[0] + seg - 1) // seg
    for col in range(8):
        cx, cy, cz = col & 1, (col >> 1) & 1, (col >> 2) & 1
        cn = [(nseg - cx + 1) // 2, (nel[1] - cy + 1) // 2, (nel[2] - cz + 1) // 2]
        for t in range(cn[0] * cn[1] * cn[2]):
            tt = t
            sx = 2 * (tt % cn[0]) + cx
            tt //= cn[0]
            ey = 2 * (tt % cn[1]) + cy
            tt //= cn[1]
            ez = 2 * tt + cz
            yield col, t, sx, ey, ez, nseg


def pencil_writes(nel, seg):
    """[(launch, wg, entry, contributors, first)] for the Jacobian and
    [(launch, wg, row, contributors, first)] for the residual."""
    nn = [2 * n + 1 for n in nel]
    jw, rw = [], []
    for col, wg, sx, ey, ez, nseg in pencil_units(nel, seg):
        x0 = sx * seg
        x1 = min(x0 + seg, nel[0])
        flx, fhx = sx == 0 or sx % 2 == 0, sx == nseg - 1 or sx % 2 == 0
        fly, fhy = ey == 0 or ey % 2 == 0, ey == nel[1] - 1 or ey % 2 == 0
        flz, fhz = ez == 0 or ez % 2 == 0, ez == nel[2] - 1 or ez % 2 == 0
        Y0, Z0 = 2 * ey, 2 * ez
        carry = {}   # (r, l) -> contributions of B's last-plane run, by column
        rcarry = {}  # l -> residual contributions
        for xa in range(x0, x1, 2):
            hasB = xa + 1 < x1
            first_step, last_step = xa == x0, xa + 2 >= x1
            el = {0: (xa, ey, ez), 1: (xa + 1, ey, ez)}
            # residual (lane l < 27 of half h: test function i = l)
            R = {}
            for h in (0, 1):
                if h == 1 and not hasB:
                    continue
                for l in range(27):
                    R[(h, l)] = [el[h]]
            newr = {}
            for h in (0, 1):
                if h == 1 and not hasB:
                    continue
                ex = xa + h
                for l in range(27):
                    i1, i2, i3 = l % 3, (l // 3) % 3, l // 9
                    X = 2 * ex + i1
                    contrib = list(R[(h, l)])
                    write, xsh, xfirst = True, False, True
                    if h == 0 and i1 == 2:
                        if hasB:
                            contrib = contrib + R[(1, l - 2)]
                        elif x1 < nel[0]:
                            xsh, xfirst = True, fhx
                    elif h == 0 and i1 == 0:
                        if not first_step:
                            contrib = rcarry[l + 2] + contrib
                        elif x0 > 0:
                            xsh, xfirst = True, flx
                    elif h == 1 and i1 == 0:
                        write = False
                    elif h == 1 and i1 == 2:
                        newr[l] = contrib
                        write = last_step
                        if last_step and x1 < nel[0]:
                            xsh, xfirst = True, fhx
                    if write:
                        ysh, zsh = i2 != 1, i3 != 1
                        first = (not xsh or xfirst) and (not ysh or (fly if i2 == 0 else fhy)) and \
                                (not zsh or (flz if i3 == 0 else fhz))
                        rw.append((col, wg, (X, Y0 + i2, Z0 + i3), tuple(contrib), first))
            rcarry.update(newr)
            # Jacobian rounds
            for r in range(9):
                m = r // 3
                for h in (0, 1):
                    if h == 1 and not hasB:
                        continue
                    ex = xa + h
                    for l in range(32):
                        pp = l + 32 * (r % 3)
                        if pp >= 81:
                            continue
                        i1 = m if h else 2 - m
                        j2, j3, i2, i3 = pp % 3, (pp // 3) % 3, (pp // 9) % 3, pp // 27
                        X = 2 * ex + i1
                        # this lane's run: columns 2 ex + j1, j1 = 0..2, element ex
                        own = {2 * ex + j1: [el[h]] for j1 in range(3)}
                        cols = {}
                        xsh_first, xsh_last = False, False
                        if m == 1:
                            cols = own
                        elif m == 0:
                            if h == 0:
                                cols = dict(own)
                                if hasB:
                                    cols[2 * ex + 2] = own[2 * ex + 2] + [el[1]]
                                else:
                                    xsh_last = x1 < nel[0]
                            else:
                                cols = {2 * ex + 1: own[2 * ex + 1], 2 * ex + 2: own[2 * ex + 2]}
                        elif h == 0:
                            if not first_step:
                                c = carry[(r, l)]
                                cols = {2 * ex - 2: c[2 * ex - 2], 2 * ex - 1: c[2 * ex - 1],
                                        2 * ex: c[2 * ex] + own[2 * ex], 2 * ex + 1: own[2 * ex + 1],
                                        2 * ex + 2: own[2 * ex + 2]}
                            else:
                                cols = own
                                xsh_first = x0 > 0
                        else:
                            carry[(r, l)] = own
                            if last_step:
                                cols = own
                                xsh_last = x1 < nel[0]
                        if not cols:
                            continue
                        ysh = i2 == j2 and i2 != 1
                        zsh = i3 == j3 and i3 != 1
                        yz = (not ysh or (fly if i2 == 0 else fhy)) and (not zsh or (flz if i3 == 0 else fhz))
                        ks = sorted(cols)
                        for k, cxv in enumerate(ks):
                            first = yz
                            if xsh_first and k == 0:
                                first = first and flx
                            if xsh_last and k == len(ks) - 1:
                                first = first and fhx
                            ent = ((X, Y0 + i2, Z0 + i3), (cxv, Y0 + j2, Z0 + j3))
                            jw.append((col, wg, ent, tuple(cols[cxv]), first))
    return jw, rw


def q2_node_elems(node, nel):
    """Elements holding a Q2 lattice node."""
    out = []
    rng = []
    for d in range(3):
        c = node[d]
        cand = [c // 2 - 1, c // 2] if c % 2 == 0 else [c // 2]
        rng.append([e for e in cand if 0 <= e < nel[d]])
    for e in itertools.product(*rng):
        out.append(e)
    return out


@pytest.mark.parametrize("seg", [2, 4, 6, 12])
@pytest.mark.parametrize("nel", [(5, 4, 3), (7, 3, 2), (8, 1, 1), (1, 1, 1), (2, 5, 7), (13, 2, 3)])
def test_pencil_schedule(nel, seg):
    jw, rw = pencil_writes(nel, seg)
    for writes, is_res in ((jw, False), (rw, True)):
        by = {}
        for col, wg, ent, contrib, first in writes:
            by.setdefault(ent, []).append((col, wg, contrib, first))
        if is_res:
            expect = {tuple(n) for n in itertools.product(*(range(2 * n + 1) for n in nel))}
        else:
            expect = pattern_entries(3, 2, nel)
        assert set(by) == expect, "writes do not cover the pattern exactly"
        for ent, ws in by.items():
            # each element holding both nodes contributes exactly once
            if is_res:
                want = sorted(q2_node_elems(ent, nel))
            else:
                er, ec_ = set(q2_node_elems(ent[0], nel)), set(q2_node_elems(ent[1], nel))
                want = sorted(er & ec_)
            got = sorted(e for w in ws for e in w[2])
            assert got == want, f"{ent}: contributions {got} != {want}"
            # concurrency: one writer per entry per launch (across and within workgroups)
            launches = [w[0] for w in ws]
            assert len(launches) == len(set(launches)), f"{ent}: two writers in one launch {ws}"
            firsts = [w for w in ws if w[3]]
            assert len(firsts) == 1, f"{ent}: {len(firsts)} first writers {ws}"
            assert all(w[0] > firsts[0][0] for w in ws if not w[3]), f"{ent}: a later touch precedes the first"
