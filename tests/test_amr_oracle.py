"""Pins of the FP64 block-pool oracle (oracle/amr_oracle.py), CPU only.

Each check is chosen so that a plausible mistake in a pass fails it: a wrong
ring index or periodic image breaks the dense equivalences, a dropped or
mis-signed orphan term breaks conservation (the check is shown to have teeth
by switching orphans off), a wrong child offset breaks the restriction /
interpolation identity, and the criterion meets closed forms.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import amr_oracle as ao  # noqa: E402
from oracle_runner import run as run_js  # noqa: E402


def f32(v):
    return float(np.float32(v))


def smooth_field(nx, ny, nz, seed, amp=0.03, modes=3):
    """Smooth periodic [rho, ux, uy, uz] (N, 4) float32 field."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    out = np.zeros((4, nz, ny, nx))
    out[0] = 1.0
    for c in range(4):
        for _ in range(modes):
            k = rng.integers(1, 3, size=3)
            ph = rng.uniform(0, 2 * np.pi, size=3)
            a = rng.uniform(-1, 1) * amp * (0.3 if c == 0 else 1.0)
            out[c] += a * (np.sin(2 * np.pi * k[0] * x / nx + ph[0]) * np.cos(2 * np.pi * k[1] * y / ny + ph[1])
                           * np.cos(2 * np.pi * k[2] * z / nz + ph[2]))
    return np.moveaxis(out, 0, -1).reshape(-1, 4).astype(np.float32)


def topology(o, blocks):
    bs = np.full(o.nblocks, -1, dtype=np.int32)
    sb = np.full(o.max_slots, -1, dtype=np.int32)
    for k, b in enumerate(sorted(blocks)):
        bs[b], sb[k] = k, b
    return bs, sb


def ids(o, coords):
    return [o.block_id(bx, by, bz) for bx, by, bz in coords]


def test_tables_match_lattice_header():
    for q in (19, 27):
        T = ao.Tables(q)
        assert abs(T.w.sum() - 1.0) < 1e-15
        e = np.stack([T.ex, T.ey, T.ez], 1)
        assert (e[T.opp] == -e).all()
        # second moment isotropy with exact weights
        Te = ao.Tables(q, True)
        m2 = np.einsum("i,ia,ib->ab", Te.w, e, e)
        assert np.allclose(m2, np.eye(3) / 3, atol=1e-15)


@pytest.mark.parametrize("trt", [False, True])
def test_no_tiles_is_the_dense_js_oracle(trt):
    n = 12
    desc = {"q": 19, "nx": n, "ny": n, "nz": n, "rb": 4, "maxSlots": 4, "omega": f32(1 / 0.62),
            "force": [f32(1e-5), 0.0, f32(-2e-6)]}
    if trt:
        desc["q"] = 27
        desc["omegaMinus"] = f32(1.3)
    mac = smooth_field(n, n, n, 3)
    o = ao.AmrOracle(desc)
    o.upload_macro(mac)
    for _ in range(3):
        o.macro_step()
    js = run_js({k: desc[k] for k in ("q", "nx", "ny", "nz", "omega", "force") if k in desc}
                | ({"omegaMinus": desc["omegaMinus"]} if trt else {}), mac, steps=3, want=("mac", "f"))
    assert np.abs(o.l0_f() - js["f"]).max() < 1e-14
    assert np.abs(o.mac() - js["mac"]).max() < 1e-14


def fine_global(o):
    """Tile interiors of an all-refined pool as one fine lattice [q, 2nz, 2ny, 2nx]."""
    lo, hi = ao.GHOST, ao.GHOST + o.rb2
    out = np.empty((o.q, 2 * o.nz, 2 * o.ny, 2 * o.nx))
    for b in range(o.nblocks):
        bx, by, bz = o.block_xyz(b)
        s = o.block_slot[b]
        out[:, bz * o.rb2:(bz + 1) * o.rb2, by * o.rb2:(by + 1) * o.rb2, bx * o.rb2:(bx + 1) * o.rb2] = \
            o.pool[:, s, lo:hi, lo:hi, lo:hi]
    return out


def test_all_refined_is_a_dense_fine_lattice():
    n, rb = 12, 4
    desc = {"q": 27, "nx": n, "ny": n, "nz": n, "rb": rb, "maxSlots": 27, "omega": f32(1 / 0.7),
            "force": [f32(2e-6), f32(1e-6), 0.0]}
    o = ao.AmrOracle(desc)
    o.upload_topology(*topology(o, range(o.nblocks)))
    o.upload_macro(smooth_field(n, n, n, 5))
    # start both from the same binary32 fine state
    o.pool = o.pool.astype(np.float32).astype(np.float64)
    f_start = fine_global(o).reshape(o.q, -1)
    for _ in range(2):
        o.macro_step()
    js = run_js({"q": 27, "nx": 2 * n, "ny": 2 * n, "nz": 2 * n, "omega": o.omega_f, "force": desc["force"]},
                f=f_start, steps=4, want=("f",))
    # the JS oracle rounds omega_f to binary32; compare at that size
    assert np.abs(fine_global(o).reshape(o.q, -1) - js["f"]).max() < 2e-9


@pytest.mark.parametrize("q", [19, 27])
def test_all_refined_body_is_the_dense_body_at_twice_the_resolution(q):
    """With every block refined the fine level is a dense lattice of 2n
    cells whose cell J sits at (J - 1/2) / 2 in L0 units, so a sphere of
    radius a at c on the fine level is the dense sphere of radius 2a at
    2c + 1/2 (the binary32 signed distance scales exactly by 2).  The pool
    state must follow the dense JS oracle's bounce-back and SOLID_EQ step,
    and each macro-step's force record must be the dense force of the same
    fine state times dx^2 = 1/4 (torque 1/8: the arm is in L0 cells).  A
    wrong source position, reflected direction, weight or arm fails this."""
    n, rb = 12, 4
    c, a = [f32(5.3), f32(6.1), f32(5.7)], f32(2.6)
    desc = {"q": q, "nx": n, "ny": n, "nz": n, "rb": rb, "maxSlots": 27, "omega": f32(1 / 0.62),
            "body": {"a": a, "centre": c}}
    o = ao.AmrOracle(desc)
    o.upload_topology(*topology(o, range(o.nblocks)))
    mac = smooth_field(n, n, n, 9)
    mac[:, 1] += 0.04
    o.upload_macro(mac)
    o.pool = o.pool.astype(np.float32).astype(np.float64)
    f_start = fine_global(o).reshape(o.q, -1)
    for _ in range(3):
        o.macro_step()
    cf = [2 * v + 0.5 for v in c]
    js = run_js({"q": q, "nx": 2 * n, "ny": 2 * n, "nz": 2 * n, "omega": o.omega_f, "solidEq": 1,
                 "body": {"a": 2 * a, "centre": cf}}, f=f_start, steps=6, want=("f", "force", "cls"))
    assert (js["cls"] & 1).sum() > 100  # a real body, not an empty set
    assert np.abs(fine_global(o).reshape(o.q, -1) - js["f"]).max() < 2e-9
    # the JS oracle rounds omega_f to binary32, so the states agree to 2e-9
    # after the first record and the forces to that relative size
    amr_f = np.array(o.forces)
    dense_f = js["force"][::2]
    assert np.abs(amr_f[0] - np.r_[0.25 * dense_f[0, :3], 0.125 * dense_f[0, 3:]]).max() < 1e-13
    assert np.abs(dense_f[:, 0]).min() > 1e-3
    np.testing.assert_allclose(amr_f[:, :3], 0.25 * dense_f[:, :3], rtol=0, atol=1e-8 * np.abs(dense_f).max())
    np.testing.assert_allclose(amr_f[:, 3:], 0.125 * dense_f[:, 3:], rtol=0, atol=1e-8 * np.abs(dense_f).max())


SEAMS = {
    # flat faces, convex edges and corners, a concave corner, a lone block
    # and a block that wraps the periodic boundary
    "mixed": [(1, 1, 1), (2, 1, 1), (1, 2, 1), (3, 3, 3), (0, 3, 1)],
    "bar": [(0, 1, 1), (1, 1, 1), (2, 1, 1), (3, 1, 1)],
}


@pytest.mark.parametrize("linear", [0, 1])
@pytest.mark.parametrize("shape", sorted(SEAMS))
def test_composite_mass_and_momentum_are_conserved(linear, shape):
    n, rb = 16, 4
    desc = {"q": 27, "nx": n, "ny": n, "nz": n, "rb": rb, "maxSlots": 8, "omega": 1 / 0.6,
            "explodeLinear": linear, "exactWeights": True}
    o = ao.AmrOracle(desc)
    o.upload_topology(*topology(o, ids(o, SEAMS[shape])))
    o.upload_macro(smooth_field(n, n, n, 11, amp=0.04))
    t0 = o.totals()
    for _ in range(3):
        o.macro_step()
        t = o.totals()
        assert abs(t[0] - t0[0]) < 1e-12 * t0[0]
        assert np.abs(t[1:] - t0[1:]).max() < 1e-12 * t0[0]


def test_conservation_check_has_teeth_without_orphans():
    n, rb = 16, 4
    desc = {"q": 27, "nx": n, "ny": n, "nz": n, "rb": rb, "maxSlots": 8, "omega": 1 / 0.6,
            "orphans": 0, "exactWeights": True}
    o = ao.AmrOracle(desc)
    o.upload_topology(*topology(o, ids(o, SEAMS["mixed"])))
    o.upload_macro(smooth_field(n, n, n, 11, amp=0.04))
    t0 = o.totals()
    for _ in range(3):
        o.macro_step()
    assert abs(o.totals()[0] - t0[0]) > 1e-8 * t0[0]


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("shape", sorted(SEAMS))
def test_interp_interface_with_reflux_is_conservative(q, shape):
    # common_d3_amr_flux.wgsl:8-11: with the flux correction the composite
    # mass and momentum are conserved exactly; without it they drift
    n, rb = 16, 4
    drift = {}
    for reflux in (0, 1):
        desc = {"q": q, "nx": n, "ny": n, "nz": n, "rb": rb, "maxSlots": 8, "omega": 1 / 0.6,
                "interface": "interp", "reflux": reflux, "exactWeights": True}
        o = ao.AmrOracle(desc)
        o.upload_topology(*topology(o, ids(o, SEAMS[shape])))
        o.upload_macro(smooth_field(n, n, n, 11, amp=0.04))
        t0 = o.totals()
        worst = 0.0
        for _ in range(3):
            o.macro_step()
            t = o.totals()
            worst = max(worst, abs(t[0] - t0[0]) / t0[0], np.abs(t[1:] - t0[1:]).max() / t0[0])
        drift[reflux] = worst
    assert drift[1] < 1e-14, drift
    assert drift[0] > 1e-7, drift


def test_interp_interface_all_refined_is_a_dense_fine_lattice():
    # with every block refined no ring cell is read (DIRECT_GHOST), so the
    # interp interface steps the same dense fine lattice as explode/coalesce
    n, rb = 12, 4
    desc = {"q": 27, "nx": n, "ny": n, "nz": n, "rb": rb, "maxSlots": 27, "omega": f32(1 / 0.7),
            "force": [f32(2e-6), f32(1e-6), 0.0], "interface": "interp", "reflux": 1}
    o = ao.AmrOracle(desc)
    o.upload_topology(*topology(o, range(o.nblocks)))
    o.upload_macro(smooth_field(n, n, n, 5))
    o.pool = o.pool.astype(np.float32).astype(np.float64)
    f_start = fine_global(o).reshape(o.q, -1)
    for _ in range(2):
        o.macro_step()
    js = run_js({"q": 27, "nx": 2 * n, "ny": 2 * n, "nz": 2 * n, "omega": o.omega_f, "force": desc["force"]},
                f=f_start, steps=4, want=("f",))
    assert np.abs(fine_global(o).reshape(o.q, -1) - js["f"]).max() < 2e-9


def test_uniform_flow_stays_uniform():
    n, rb = 16, 4
    u = (0.03, -0.02, 0.01)
    desc = {"q": 19, "nx": n, "ny": n, "nz": n, "rb": rb, "maxSlots": 8, "omega": 1 / 0.58, "exactWeights": True}
    o = ao.AmrOracle(desc)
    o.upload_topology(*topology(o, ids(o, SEAMS["mixed"])))
    mac = np.tile(np.array([1.0, *u], dtype=np.float32), (n ** 3, 1))
    o.upload_macro(mac)
    for _ in range(3):
        o.macro_step()
    eq = ao.feq(o.T, 1.0, *[float(np.float32(v)) for v in u])
    cov = o.covered_mask()
    l0 = o.f0[:, ~cov]
    assert np.abs(l0 - eq[:, None]).max() < 1e-15
    lo, hi = ao.GHOST, ao.GHOST + o.rb2
    for s in ao.active_list(o.slot_to_block):
        inner = o.pool[:, s, lo:hi, lo:hi, lo:hi].reshape(o.q, -1)
        assert np.abs(inner - eq[:, None]).max() < 1e-15


def test_restriction_after_interpolation_is_identity_on_linear_fields():
    n, rb = 16, 4
    # exact weights: the binary32 weights' second moment is off by ~1e-8,
    # which alone moves the recovered velocity by ~1e-10
    desc = {"q": 27, "nx": n, "ny": n, "nz": n, "rb": rb, "maxSlots": 2, "omega": 1 / 0.8, "exactWeights": True}
    o = ao.AmrOracle(desc)
    T = o.T
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    rho = 1.0 + 0.01 * x - 0.005 * y + 0.002 * z
    u = (0.02, -0.01, 0.015)
    rng = np.random.default_rng(4)
    # fneq with no zeroth or first moment, linear in space
    A = rng.normal(size=(T.q, 4)) * 1e-4
    E = np.stack([np.ones(T.q), T.ex, T.ey, T.ez]).astype(float)
    A -= E.T @ np.linalg.solve(E @ E.T, E @ A)
    fneq = A[:, 0, None, None, None] + A[:, 1, None, None, None] * x + A[:, 2, None, None, None] * y \
        + A[:, 3, None, None, None] * z
    o.f0 = ao.feq(T, rho, *u) + fneq
    o.mac0 = np.zeros((4, n, n, n))
    b = o.block_id(1, 2, 1)
    o.block_slot[b], o.slot_to_block[0] = 0, b
    before = o.f0.copy()
    o.interp([(b, 0)])
    o.f0[:] = np.nan
    o.average([(b, 0)])
    cells = (slice(4, 8), slice(8, 12), slice(4, 8))  # z, y, x of block (1, 2, 1)
    assert np.abs(o.f0[(slice(None),) + cells] - before[(slice(None),) + cells]).max() < 1e-15
    # the Dupuis-Chopard pair is an exact inverse
    assert abs(o.c2f * o.f2c - 1.0) < 1e-15


def test_criterion_closed_forms():
    n = 16
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    gamma = 2.0 ** -6
    mac = np.zeros((4, n, n, n))
    mac[0] = 1.0
    mac[1] = gamma * y  # pure shear: |w| = gamma (test-d3-criterion.js:59)
    for dt in (np.float32, np.float64):
        vm = ao.vorticity_magnitude(mac, dt)
        assert vm.dtype == dt
        assert (vm[:, 1:-1, :] == dt(gamma)).all()
    Om = 2.0 ** -7
    mac[1] = -Om * (y - 7.5)
    mac[2] = Om * (x - 7.5)
    vm = ao.vorticity_magnitude(mac, np.float32)  # solid rotation: |w| = 2 Omega
    assert (vm[:, 1:-1, 1:-1] == np.float32(2 * Om)).all()
    # block maximum over a dilated box sees a spike one cell outside
    vm = np.zeros((n, n, n), dtype=np.float32)
    vm[0, 0, 4] = 1.0
    c0 = ao.block_criterion(vm, 4, 0)
    c1 = ao.block_criterion(vm, 4, 1)
    assert c0[1] == 1.0 and c0[0] == 0.0
    assert c1[0] == 1.0 and c1[1] == 1.0 and c1[2] == 0.0


def test_q_criterion_closed_forms():
    """Q = 1/2 (|Omega|^2 - |S|^2) (d3-criterion.mjs qOfGrad; the closed forms
    of the reference's test-d3-criterion.js): pure shear is exactly 0 (the
    reason the criterion is Q and not |omega|), solid rotation at rate W gives
    W^2, pure strain -W^2; and |omega| differs from Q where it should (shear
    has |omega| = gamma > 0).  Q does not change when the gradient is
    transposed (Omega only changes sign), so the last case pins the
    off-diagonal terms instead: u = (a y, b x, 0) has Q = -a b."""
    n = 16
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    gamma, W = 2.0 ** -6, 2.0 ** -7
    mac = np.zeros((4, n, n, n))
    mac[0] = 1.0
    mac[1] = gamma * y
    for dt in (np.float32, np.float64):
        q = ao.q_criterion(mac, dt)
        assert q.dtype == dt and (q[:, 1:-1, :] == 0).all()
        assert (ao.vorticity_magnitude(mac, dt)[:, 1:-1, :] == dt(gamma)).all()
    mac[1] = -W * (y - 7.5)
    mac[2] = W * (x - 7.5)
    assert (ao.q_criterion(mac, np.float32)[:, 1:-1, 1:-1] == np.float32(W * W)).all()
    mac[1] = W * (x - 7.5)
    mac[2] = -W * (y - 7.5)
    assert (ao.q_criterion(mac, np.float32)[:, 1:-1, 1:-1] == np.float32(-W * W)).all()
    # u = (a y, b x, 0): Omega_xy = (b - a)/2, S_xy = (a + b)/2, so
    # Q = ((b - a)^2 - (a + b)^2) / 4 = -a b; with a = 2 gamma, b = -gamma:
    # Q = 2 gamma^2 (the diagonal terms alone would give 0).
    mac[1] = 2 * gamma * (y - 7.5)
    mac[2] = -gamma * (x - 7.5)
    assert np.allclose(ao.q_criterion(mac)[:, 1:-1, 1:-1], 2 * gamma * gamma, rtol=0, atol=1e-18)


def test_manager_ascending_ids_and_hysteresis():
    bs = np.array([-1, 0, -1, 2, -1, 1], dtype=np.int32)
    sb = np.array([1, 5, 3, -1], dtype=np.int32)
    crit = np.array([0.9, 0.35, 0.6, 0.1, 0.45, 0.5], dtype=np.float32)
    want = ao.decide(bs, crit, np.float32(0.5), np.float32(0.3))
    # held blocks judged against keep_at, free ones against refine_at
    assert want.tolist() == [1, 1, 1, 0, 0, 1]
    nbs, nsb, dying, fresh, exh = ao.manage_slots(bs, sb, want)
    assert dying == [(3, 2)]
    # free slots after coarsening: 2, 3 (ascending); new blocks 0, 2 (ascending)
    assert fresh == [(0, 2), (2, 3)] and exh == 0
    assert nbs.tolist() == [2, 0, 3, -1, -1, 1] and nsb.tolist() == [1, 5, 0, 2]
    assert ao.active_list(nsb).tolist() == [0, 1, 2, 3]
    # exhaustion: more wanted blocks than free slots
    nbs, nsb, dying, fresh, exh = ao.manage_slots(np.full(6, -1), np.full(2, -1), np.ones(6, dtype=np.uint8))
    assert fresh == [(0, 0), (1, 1)] and exh == 4


def test_js_manage_slots_matches_the_oracle():
    rng = np.random.default_rng(9)
    cases = []
    for _ in range(40):
        nb, ns = int(rng.integers(1, 60)), int(rng.integers(1, 30))
        bs = np.full(nb, -1, dtype=np.int32)
        sb = np.full(ns, -1, dtype=np.int32)
        for b in rng.permutation(nb)[: int(rng.integers(0, min(nb, ns) + 1))]:
            s = int(rng.choice(np.nonzero(sb < 0)[0]))
            bs[b], sb[s] = s, b
        want = (rng.random(nb) < 0.5).astype(np.uint8)
        cases.append((bs, sb, want))
    js = """
import { manageSlots, activeSlotList } from '%s';
const cases = JSON.parse(require('fs').readFileSync(0, 'utf8'));
""" % os.path.join(ROOT, "getting-air_amd", "js", "d3-amr.mjs")
    script = ("import { createRequire } from 'module';\nconst require = createRequire(import.meta.url);\n" + js +
              "const out = cases.map(([bs, sb, w]) => { const r = manageSlots(bs, sb, w);"
              " return [Array.from(r.blockSlot), Array.from(r.slotToBlock), r.exhausted,"
              " Array.from(activeSlotList(r.slotToBlock))]; });\n"
              "process.stdout.write(JSON.stringify(out));\n")
    path = os.path.join(ROOT, "tests", "js", ".manage_twin_probe.mjs")
    with open(path, "w") as fh:
        fh.write(script)
    try:
        payload = json.dumps([[c[0].tolist(), c[1].tolist(), c[2].tolist()] for c in cases])
        r = subprocess.run(["node", path], input=payload, capture_output=True, text=True, timeout=60)
    finally:
        os.remove(path)
    assert r.returncode == 0, r.stderr
    for (bs, sb, want), (jbs, jsb, jexh, jact) in zip(cases, json.loads(r.stdout)):
        nbs, nsb, _, _, exh = ao.manage_slots(bs, sb, want)
        assert nbs.tolist() == jbs and nsb.tolist() == jsb and exh == jexh
        act = ao.active_list(nsb)
        assert jact == [len(act)] + act.tolist()


def test_sphere_refinement_matches_the_js_module():
    """pylbm.scenarios.refine_near_sphere (the host side of ?refine=body)
    against the JS module's refineNearBody with the same sphere SDF, and the
    geometric fact the fine-level body relies on: every unrefined block lies
    more than the margin from the surface (checked by brute force on its
    box's cell centres)."""
    sys.path.insert(0, os.path.join(ROOT, "getting-air_amd"))
    from pylbm import scenarios as sc
    cases = [(192, 128, 128, 4, (48.0, 64.0, 64.0), 8.0, 2.0), (32, 32, 32, 4, (15.3, 16.7, 14.1), 3.7, 2.0),
             (48, 32, 40, 8, (20.5, 15.0, 19.25), 6.0, 1.0)]
    js = """
import { makePool, refineNearBody } from '%s';
const cases = JSON.parse(require('fs').readFileSync(0, 'utf8'));
""" % os.path.join(ROOT, "getting-air_amd", "js", "d3-amr.mjs")
    script = ("import { createRequire } from 'module';\nconst require = createRequire(import.meta.url);\n" + js +
              "const out = cases.map(([nx, ny, nz, rb, c, a, m]) => {"
              " const pool = makePool({ dims: [nx, ny, nz], rb: rb, maxSlots: nx * ny * nz / (rb * rb * rb) });"
              " const sdf = (p) => Math.hypot(p[0] - c[0], p[1] - c[1], p[2] - c[2]) - a;"
              " return Array.from(refineNearBody(pool, sdf, m).blockSlot); });\n"
              "process.stdout.write(JSON.stringify(out));\n")
    path = os.path.join(ROOT, "tests", "js", ".near_body_probe.mjs")
    with open(path, "w") as fh:
        fh.write(script)
    try:
        r = subprocess.run(["node", path], input=json.dumps(cases), capture_output=True, text=True, timeout=60)
    finally:
        os.remove(path)
    assert r.returncode == 0, r.stderr
    for (nx, ny, nz, rb, c, a, m), jbs in zip(cases, json.loads(r.stdout)):
        bs, sb = sc.refine_near_sphere(nx, ny, nz, rb, c, a, m)
        assert bs.tolist() == jbs and (sb >= 0).sum() > 0
        nbx, nby = nx // rb, ny // rb
        for b in np.nonzero(bs < 0)[0]:
            lo = np.array([b % nbx, (b // nbx) % nby, b // (nbx * nby)]) * rb
            g = np.stack(np.meshgrid(*[np.arange(v, v + rb) for v in lo], indexing="ij"), -1).reshape(-1, 3)
            assert (np.sqrt(((g - np.array(c)) ** 2).sum(1)) - a).min() > m
