"""Pins of the multi-level FP64 oracle (oracle/amr_multilevel_oracle.py), CPU
only.

- With two levels it is the two-level oracle (oracle/amr_oracle.py) bit for
  bit, body included: the generalised passes lose nothing at level 1.
- A three-level hierarchy whose third level holds no tile is the two-level
  run bit for bit (the schedule only adds work where tiles exist).
- Three and four levels conserve the composite mass and momentum to float64
  rounding (explode and coalesce with orphans are exactly conservative at a
  pool parent as at L0); with orphans off the same check fails.
- A uniform flow stays uniform through every level.
- At a seed, reading a parent cell in the tile that owns it and reading it in
  the ring of the tile that holds the child block give the same bits (the
  ring was filled from the same parent positions), so the owner read changes
  nothing a static hierarchy sees.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "getting-air_amd"))

import amr_oracle as ao  # noqa: E402
import amr_multilevel_oracle as mo  # noqa: E402
from test_amr_oracle import smooth_field  # noqa: E402


def level_topology(L, coords):
    """Block maps of level L refining the listed block coordinates, slots in
    ascending block id."""
    ids = sorted(L.block_id(*c) for c in coords)
    bs = np.full(L.nblocks, -1, dtype=np.int32)
    sb = np.full(L.max_slots, -1, dtype=np.int32)
    for k, b in enumerate(ids):
        bs[b], sb[k] = k, b
    return bs, sb


def cube(lo, hi):
    r = range(lo, hi)
    return [(x, y, z) for z in r for y in r for x in r]


def hierarchy(desc, lv1, lv2=None, lv3=None):
    o = mo.MultiLevelOracle(desc)
    o.upload_topology(1, *level_topology(o.lv[1], lv1))
    if lv2 is not None:
        o.upload_topology(2, *level_topology(o.lv[2], lv2))
    if lv3 is not None:
        o.upload_topology(3, *level_topology(o.lv[3], lv3))
    return o


BASE = {"q": 19, "nx": 16, "ny": 16, "nz": 16, "rb": 4, "maxSlots": 16, "omega": 1.0 / 0.62}
# level 1 refines L0 cells [4, 12)^3; level 2 refines level-1 cells [12, 20)^3,
# whose ring's parent cells [11, 20] lie inside level 1's [8, 24)
LV1 = cube(1, 3)
LV2 = cube(3, 5)


@pytest.mark.parametrize("body", [False, True])
def test_two_levels_is_the_two_level_oracle(body):
    desc = dict(BASE, q=27)
    if body:
        desc["body"] = {"a": 1.6, "centre": [8.0, 8.0, 8.0]}
    a = ao.AmrOracle(desc)
    bs, sb = np.full(a.nblocks, -1, dtype=np.int32), np.full(a.max_slots, -1, dtype=np.int32)
    for k, b in enumerate(sorted(a.block_id(*c) for c in LV1)):
        bs[b], sb[k] = k, b
    a.upload_topology(bs, sb)
    m = hierarchy(dict(desc, levels=2), LV1)
    mac = smooth_field(16, 16, 16, 3)
    a.upload_macro(mac)
    m.upload_macro(mac)
    for _ in range(3):
        a.macro_step()
        m.macro_step()
    assert np.array_equal(a.f0, m.f0, equal_nan=True)
    assert np.array_equal(a.pool, m.lv[1].pool, equal_nan=True)
    assert np.array_equal(a.mac0, m.mac0, equal_nan=True)
    if body:
        assert np.array_equal(np.array(a.forces), np.array(m.forces))
        assert abs(np.array(m.forces)[-1, 0]) > 0


def test_empty_third_level_is_the_two_level_run():
    mac = smooth_field(16, 16, 16, 5)
    two = hierarchy(dict(BASE, levels=2), LV1)
    three = hierarchy(dict(BASE, levels=3), LV1, [])
    two.upload_macro(mac)
    three.upload_macro(mac)
    for _ in range(3):
        two.macro_step()
        three.macro_step()
    assert np.array_equal(two.f0, three.f0, equal_nan=True)
    assert np.array_equal(two.lv[1].pool, three.lv[1].pool, equal_nan=True)


@pytest.mark.parametrize("levels,linear", [(3, True), (3, False), (4, True)])
def test_deep_composite_mass_and_momentum_are_conserved(levels, linear):
    desc = dict(BASE, levels=levels, explodeLinear=int(linear), exactWeights=True)
    lv3 = None
    if levels == 4:
        # level 3 refines level-2 cells [28, 36)^3 inside level 2's [24, 40)
        desc.update(nx=16, maxSlots=16)
        lv3 = cube(7, 9)
    o = hierarchy(desc, LV1, LV2, lv3)
    o.upload_macro(smooth_field(16, 16, 16, 7))
    t0 = o.totals()
    for _ in range(3):
        o.macro_step()
    t1 = o.totals()
    # no NaN reached a solved cell
    assert np.isfinite(o.f0[:, ~o.lv[1].covered_mask()]).all()
    assert abs(t1[0] / t0[0] - 1) < 1e-13, (t0, t1)
    assert np.abs(t1[1:] - t0[1:]).max() < 1e-13 * t0[0], (t0, t1)


def test_deep_conservation_check_has_teeth_without_orphans():
    desc = dict(BASE, levels=3, orphans=0, exactWeights=True)
    o = hierarchy(desc, LV1, LV2)
    o.upload_macro(smooth_field(16, 16, 16, 7))
    t0 = o.totals()
    for _ in range(3):
        o.macro_step()
    assert abs(o.totals()[0] / t0[0] - 1) > 1e-9


def test_deep_uniform_flow_stays_uniform():
    desc = dict(BASE, levels=3, exactWeights=True)
    o = hierarchy(desc, LV1, LV2)
    u = (0.03, -0.02, 0.01)
    mac = np.tile(np.array([1.0, *u], dtype=np.float32), (16 ** 3, 1))
    o.upload_macro(mac)
    for _ in range(2):
        o.macro_step()
    want = ao.feq(o.T, np.float64(1.0), *(np.float64(np.float32(v)) for v in u))
    cov = o.lv[1].covered_mask()
    assert np.abs(o.f0[:, ~cov] - want[:, None]).max() < 1e-14
    for m in (1, 2):
        L = o.lv[m]
        lo, hi = mo.GHOST, mo.GHOST + L.rb2
        for _, s in L.tiles():
            inner = L.pool[:, s, lo:hi, lo:hi, lo:hi].reshape(o.q, -1)
            inner = inner[:, np.isfinite(inner[0])]
            assert np.abs(inner - want[:, None]).max() < 1e-14, m


def test_owner_read_equals_holding_ring_read_at_a_seed():
    desc = dict(BASE, levels=4, q=27)
    o = hierarchy(desc, LV1, LV2, cube(7, 9))
    o.upload_macro(smooth_field(16, 16, 16, 9))
    for m in (2, 3):
        before = o.lv[m].pool.copy()
        o.interp(m, owner=False)
        assert np.array_equal(before, o.lv[m].pool)
    # and the owner read does reach other tiles: level-2 block (4, 4, 4) sits
    # at the low edge of level-1 tile (2, 2, 2), and its stencil's lowest
    # parent cells (origin - 2) belong to level-1 tile (1, 1, 1)
    L1, L2 = o.lv[1], o.lv[2]
    assert (4 * L2.rb) // L1.rb2 == 2 and (4 * L2.rb - 2) // L1.rb2 == 1
    assert L1.block_slot[L1.block_id(1, 1, 1)] >= 0 and L2.block_slot[L2.block_id(4, 4, 4)] >= 0


# --- the manager at depth ---------------------------------------------------

def node_cascade(wants, nbs, levels):
    """d3-amr.mjs cascade21 on the listed want sets (block ids per level)."""
    script = """
import { cascade21 } from '%s';
const a = JSON.parse(process.argv[1]);
const sets = [null];
for (let m = 1; m < a.levels; m++) {
  const nb = a.nbs[m];
  sets.push(a.wants[m].map((id) => [id %% nb[0], Math.floor(id / nb[0]) %% nb[1], Math.floor(id / (nb[0] * nb[1]))].join(',')));
}
const r = cascade21(sets, (m) => a.nbs[m], { levels: a.levels });
const out = [null];
for (let m = 1; m < a.levels; m++) {
  const nb = a.nbs[m];
  out.push(Array.from(r.sets[m]).map((k) => { const b = k.split(',').map(Number); return (b[2] * nb[1] + b[1]) * nb[0] + b[0]; }).sort((x, y) => x - y));
}
process.stdout.write(JSON.stringify(out));
""" % os.path.join(ROOT, "getting-air_amd", "js", "d3-amr.mjs")
    arg = json.dumps({"levels": levels, "nbs": nbs, "wants": wants})
    out = subprocess.run(["node", "--input-type=module", "-e", script, arg], check=True, capture_output=True, text=True)
    return json.loads(out.stdout)


@pytest.mark.parametrize("levels", [3, 4])
def test_closure_is_d3_amr_cascade21(levels):
    """The oracle's two-pass closure (octets, then the 64-child parent
    gather) equals the drop-in module's cascade21 on random want sets at
    every level, periodic wrap included."""
    o = mo.MultiLevelOracle(dict(BASE, levels=levels))
    rng = np.random.default_rng(levels)
    wants = [None] + [(rng.random(o.lv[m].nblocks) < 0.02).astype(np.uint8) for m in range(1, levels)]
    # a want on the domain edge exercises the wrap
    wants[levels - 1][0] = 1
    closed = o.close(wants)
    nbs = [None] + [[o.lv[m].nbx, o.lv[m].nby, o.lv[m].nbz] for m in range(1, levels)]
    ref = node_cascade([None] + [np.nonzero(w)[0].tolist() for w in wants[1:]], nbs, levels)
    for m in range(1, levels):
        assert np.nonzero(closed[m])[0].tolist() == ref[m], m
        assert closed[m].sum() > wants[m].sum() or m == levels - 1


def test_manager_builds_refine_hierarchy_for_a_pinned_sphere():
    """From an empty pool the manager's first pass (geometry at the finest
    level, the closure, ascending-id slots) builds exactly refineHierarchy's
    maps (d3-amr.mjs:603-651, through the drop-in module), which is why a
    dynamic run of a pinned body can reproduce the static one."""
    from pylbm import scenarios
    n, c, r = 32, (16.3, 15.7, 16.1), 3.0
    maps = scenarios.sphere_hierarchy(n, n, n, 4, 3, c, r)
    slots = maps[1][1].size
    desc = dict(BASE, nx=n, ny=n, nz=n, maxSlots=slots, levels=3, manageEvery=1,
                body={"a": r, "centre": list(c)})
    o = mo.MultiLevelOracle(desc)
    o.upload_macro(np.tile(np.array([1.0, 0.0, 0.0, 0.0], np.float32), (n ** 3, 1)))
    o.manage()
    for m in (1, 2):
        assert np.array_equal(o.lv[m].block_slot, maps[m][0]), m
        assert np.array_equal(o.lv[m].slot_to_block, maps[m][1]), m
    assert o.exhausted == 0 and (o.lv[2].block_slot >= 0).sum() > 0


@pytest.mark.parametrize("kind,octaves", [("q", 1.0), ("vorticity", 1.0), ("q", 2.0)])
def test_level_wants_follow_the_ladder(kind, octaves):
    """Each level's own want is desiredLevelAt (common_refine.wgsl:54-79)
    computed in the log domain: desired = the largest k <= levels - 1 with
    log2(rate) >= log2(rate at refineAt) + octaves (k - 1), rate = sqrt(Q) or
    |curl u|, and level m wants a block when desired >= m; a held block is
    judged against keepAt's ladder."""
    refine, keep = 1e-4, 4e-5
    desc = dict(BASE, levels=4, manageEvery=1, criterion=kind, refineAt=refine, keepAt=keep,
                ladderOctaves=octaves)
    o = mo.MultiLevelOracle(desc)
    rng = np.random.default_rng(5)
    L1 = o.lv[1]
    # criterion values well inside a rung: base x r^(j + 0.5), j = -1 .. 3
    r = 2.0 ** (octaves * (2.0 if kind == "q" else 1.0))
    j = rng.integers(-1, 4, L1.nblocks)
    crit = (refine * r ** (j + 0.5)).astype(np.float32)
    # some blocks held at every level
    for m in range(1, 4):
        L = o.lv[m]
        held = rng.random(L.nblocks) < 0.3
        L.block_slot = np.where(held, 0, -1).astype(np.int32)
    w = o.level_wants(crit)

    def desired(c, base):
        e = 0.5 * np.log2(c.astype(np.float64)) if kind == "q" else np.log2(c.astype(np.float64))
        e0 = 0.5 * np.log2(base) if kind == "q" else np.log2(base)
        lv = np.zeros(c.shape, dtype=int)
        for k in range(1, 4):
            lv = np.where(e >= e0 + octaves * (k - 1), k, lv)
        return lv

    for m in range(1, 4):
        L = o.lv[m]
        b = np.arange(L.nblocks)
        bx, by, bz = L.block_xyz(b)
        c = crit[((bz >> (m - 1)) * L1.nby + (by >> (m - 1))) * L1.nbx + (bx >> (m - 1))]
        want = np.where(L.block_slot >= 0, desired(c, keep) >= m, desired(c, refine) >= m)
        assert np.array_equal(w[m].astype(bool), want), m
        assert 0 < w[m].sum() < L.nblocks


def ring_parents_present(o, m):
    """check_level_topology's rule: every ring cell of a level-m tile over a
    parent cell no level-m tile covers lies in a level m - 1 tile."""
    C, P = o.lv[m], o.lv[m - 1]
    for b, _ in C.tiles():
        bx, by, bz = C.block_xyz(b)
        n = C.fb
        f = np.arange(n)
        fz, fy, fx = (a.ravel() for a in np.meshgrid(f, f, f, indexing="ij"))
        g = [(o_ * C.rb2 + v - mo.GHOST) % (2 * N) for o_, v, N in ((bx, fx, C.nx), (by, fy, C.ny), (bz, fz, C.nz))]
        vx, vy, vz = (v // 2 for v in g)
        open_ = ~C.covered(vx, vy, vz)
        has = P.block_slot[P.block_id(vx // P.rb2, vy // P.rb2, vz // P.rb2)] >= 0
        if not has[open_].all():
            return False
    return True


def test_dynamic_hierarchy_keeps_its_invariants():
    """A field-driven three-level run under the manager (a quarter-octave
    ladder): after every pass each level-2 tile nests in a level-1 tile,
    level 2 holds whole octets, every ring parent exists and the solved cells
    stay finite, while level 2 is refined and then drained as the field
    decays."""
    desc = dict(BASE, levels=3, manageEvery=2, criterion="vorticity", refineAt=0.024, keepAt=0.022,
                ladderOctaves=0.25, maxSlots=64)
    o = mo.MultiLevelOracle(desc)
    o.upload_macro(smooth_field(16, 16, 16, 21))
    seen = set()
    changes = 0
    for _ in range(8):
        prev = [None] + [o.lv[m].block_slot.copy() for m in (1, 2)]
        o.macro_step()
        if o.last_fresh is not None:
            for m in (1, 2):
                changes += int(((prev[m] >= 0) != (o.lv[m].block_slot >= 0)).sum())
            L1, L2 = o.lv[1], o.lv[2]
            for b, _ in L2.tiles():
                bx, by, bz = L2.block_xyz(b)
                assert L1.block_slot[L1.block_id(bx // 2, by // 2, bz // 2)] >= 0
            w = (L2.block_slot >= 0).reshape(L2.nbz // 2, 2, L2.nby // 2, 2, L2.nbx // 2, 2)
            assert (w.all(axis=(1, 3, 5)) == w.any(axis=(1, 3, 5))).all()
            assert ring_parents_present(o, 2)
            seen.add(tuple(int((o.lv[m].block_slot >= 0).sum()) for m in (1, 2)))
        assert np.isfinite(o.f0[:, ~o.lv[1].covered_mask()]).all()
    assert o.exhausted == 0 and o.band == 0
    # level 2 was refined and later drained back as the field decayed
    assert changes > 0 and len(seen) > 2 and max(t[1] for t in seen) > 0 and min(t[1] for t in seen) == 0
