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
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

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
