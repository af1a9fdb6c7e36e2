"""Pins of the D2Q9 quadtree oracle (oracle/amr2d_oracle.py), CPU only.

As for the 3D pool: with no tile the macro-step is the dense FP64 JS
oracle's D2Q9 step; with every block refined the pool is a dense fine
lattice, the JS oracle run at twice the resolution and tau_f; explode,
coalesce and the orphans conserve the composite mass and momentum to float64
rounding on faces and corners (and fail to without the orphans); a uniform
flow stays uniform; restriction after interpolation is the identity on linear
fields; the criterion meets its closed forms.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import amr2d_oracle as a2  # noqa: E402
from oracle_runner import run as run_js  # noqa: E402


def f32(v):
    return float(np.float32(v))


def smooth_field_2d(nx, ny, seed, amp=0.03, modes=3):
    """Smooth periodic [rho, ux, uy, 0] (N, 4) float32 field."""
    rng = np.random.default_rng(seed)
    y, x = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    out = np.zeros((4, ny, nx))
    out[0] = 1.0
    for c in range(3):
        for _ in range(modes):
            k = rng.integers(1, 3, size=2)
            ph = rng.uniform(0, 2 * np.pi, size=2)
            a = rng.uniform(-1, 1) * amp * (0.3 if c == 0 else 1.0)
            out[c] += a * np.sin(2 * np.pi * k[0] * x / nx + ph[0]) * np.cos(2 * np.pi * k[1] * y / ny + ph[1])
    return np.moveaxis(out, 0, -1).reshape(-1, 4).astype(np.float32)


def topology(o, blocks):
    bs = np.full(o.nblocks, -1, dtype=np.int32)
    sb = np.full(o.max_slots, -1, dtype=np.int32)
    for k, b in enumerate(sorted(blocks)):
        bs[b], sb[k] = k, b
    return bs, sb


def ids(o, coords):
    return [o.block_id(bx, by) for bx, by in coords]


# faces, convex and concave corners, a lone block, a block across the
# periodic boundary; and a bar of blocks
SEAMS2 = {
    "mixed": [(1, 1), (2, 1), (1, 2), (4, 4), (0, 3), (5, 2)],
    "bar": [(0, 2), (1, 2), (2, 2), (3, 2), (4, 2), (5, 2)],
}


@pytest.mark.parametrize("trt", [False, True])
def test_no_tiles_is_the_dense_js_oracle(trt):
    n = 24
    desc = {"q": 9, "nx": n, "ny": n, "rb": 4, "maxSlots": 4, "omega": f32(1 / 0.62)}
    if trt:
        desc["omegaMinus"] = f32(1.3)
    mac = smooth_field_2d(n, n, 3)
    o = a2.Amr2DOracle(desc)
    o.upload_macro(mac)
    for _ in range(3):
        o.macro_step()
    js = run_js({k: desc[k] for k in ("q", "nx", "ny", "omega", "omegaMinus") if k in desc}, mac, steps=3,
                want=("mac", "f"))
    assert np.abs(o.f0.reshape(9, -1) - js["f"]).max() < 1e-14
    assert np.abs(o.mac()[:, :3] - js["mac"][:, :3]).max() < 1e-14


def test_all_refined_is_a_dense_fine_lattice():
    n, rb = 12, 4
    desc = {"q": 9, "nx": n, "ny": n, "rb": rb, "maxSlots": 9, "omega": f32(1 / 0.7)}
    o = a2.Amr2DOracle(desc)
    o.upload_topology(*topology(o, range(o.nblocks)))
    o.upload_macro(smooth_field_2d(n, n, 5))
    o.pool = o.pool.astype(np.float32).astype(np.float64)
    lo, hi = a2.GHOST, a2.GHOST + o.rb2

    def fine():
        out = np.empty((9, 2 * n, 2 * n))
        for b in range(o.nblocks):
            bx, by = o.block_xy(b)
            out[:, by * o.rb2:(by + 1) * o.rb2, bx * o.rb2:(bx + 1) * o.rb2] = o.pool[:, o.block_slot[b], lo:hi, lo:hi]
        return out.reshape(9, -1)

    f_start = fine()
    for _ in range(2):
        o.macro_step()
    js = run_js({"q": 9, "nx": 2 * n, "ny": 2 * n, "omega": o.omega_f}, f=f_start, steps=4, want=("f",))
    assert np.abs(fine() - js["f"]).max() < 2e-9


@pytest.mark.parametrize("linear", [0, 1])
@pytest.mark.parametrize("shape", sorted(SEAMS2))
def test_composite_mass_and_momentum_are_conserved(linear, shape):
    n = 24
    desc = {"q": 9, "nx": n, "ny": n, "rb": 4, "maxSlots": 8, "omega": 1 / 0.6, "explodeLinear": linear,
            "exactWeights": True}
    o = a2.Amr2DOracle(desc)
    o.upload_topology(*topology(o, ids(o, SEAMS2[shape])))
    o.upload_macro(smooth_field_2d(n, n, 11, amp=0.04))
    t0 = o.totals()
    for _ in range(3):
        o.macro_step()
        t = o.totals()
        assert abs(t[0] - t0[0]) < 1e-12 * t0[0]
        assert np.abs(t[1:] - t0[1:]).max() < 1e-12 * t0[0]


def test_conservation_check_has_teeth_without_orphans():
    n = 24
    desc = {"q": 9, "nx": n, "ny": n, "rb": 4, "maxSlots": 8, "omega": 1 / 0.6, "orphans": 0,
            "exactWeights": True}
    o = a2.Amr2DOracle(desc)
    o.upload_topology(*topology(o, ids(o, SEAMS2["mixed"])))
    o.upload_macro(smooth_field_2d(n, n, 11, amp=0.04))
    t0 = o.totals()
    for _ in range(3):
        o.macro_step()
    assert abs(o.totals()[0] - t0[0]) > 1e-8 * t0[0]


def test_uniform_flow_stays_uniform():
    n = 24
    u = (0.03, -0.02)
    desc = {"q": 9, "nx": n, "ny": n, "rb": 4, "maxSlots": 8, "omega": 1 / 0.58, "exactWeights": True}
    o = a2.Amr2DOracle(desc)
    o.upload_topology(*topology(o, ids(o, SEAMS2["mixed"])))
    o.upload_macro(np.tile(np.array([1.0, *u, 0.0], dtype=np.float32), (n * n, 1)))
    for _ in range(3):
        o.macro_step()
    eq = a2.feq(o.w, 1.0, *[float(np.float32(v)) for v in u])
    cov = o.covered_mask()
    assert np.abs(o.f0[:, ~cov] - eq[:, None]).max() < 1e-15
    lo, hi = a2.GHOST, a2.GHOST + o.rb2
    for _, s in o.tiles():
        assert np.abs(o.pool[:, s, lo:hi, lo:hi].reshape(9, -1) - eq[:, None]).max() < 1e-15


def test_restriction_after_interpolation_is_identity_on_linear_fields():
    n = 16
    desc = {"q": 9, "nx": n, "ny": n, "rb": 4, "maxSlots": 2, "omega": 1 / 0.8, "exactWeights": True}
    o = a2.Amr2DOracle(desc)
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    rho = 1.0 + 0.01 * x - 0.005 * y
    u = (0.02, -0.01)
    rng = np.random.default_rng(4)
    # fneq with no zeroth or first moment, linear in space
    A = rng.normal(size=(9, 3)) * 1e-4
    E = np.stack([np.ones(9), a2.EX, a2.EY]).astype(float)
    A -= E.T @ np.linalg.solve(E @ E.T, E @ A)
    fneq = A[:, 0, None, None] + A[:, 1, None, None] * x + A[:, 2, None, None] * y
    o.f0 = a2.feq(o.w, rho, *u) + fneq
    o.mac0 = np.zeros((3, n, n))
    b = o.block_id(1, 2)
    o.block_slot[b], o.slot_to_block[0] = 0, b
    before = o.f0.copy()
    o.interp([(b, 0)])
    o.f0[:] = np.nan
    o.average([(b, 0)])
    cells = (slice(8, 12), slice(4, 8))  # y, x of block (1, 2)
    assert np.abs(o.f0[(slice(None),) + cells] - before[(slice(None),) + cells]).max() < 1e-15
    assert abs(o.c2f * o.f2c - 1.0) < 1e-15


def test_criterion_closed_forms():
    n = 16
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    gamma = 2.0 ** -6
    mac = np.zeros((3, n, n))
    mac[0] = 1.0
    mac[1] = gamma * y  # pure shear: |w| = gamma
    vm = a2.vorticity_2d(mac)
    assert vm.dtype == np.float32 and (vm[1:-1, :] == np.float32(gamma)).all()
    Om = 2.0 ** -7
    mac[1] = -Om * (y - 7.5)
    mac[2] = Om * (x - 7.5)
    vm = a2.vorticity_2d(mac)  # solid rotation: |w| = 2 Omega
    assert (vm[1:-1, 1:-1] == np.float32(2 * Om)).all()
    vm = np.zeros((n, n), dtype=np.float32)
    vm[0, 4] = 1.0
    c0 = a2.block_criterion(vm, 4, 0)
    c1 = a2.block_criterion(vm, 4, 1)
    assert c0[1] == 1.0 and c0[0] == 0.0
    assert c1[0] == 1.0 and c1[1] == 1.0 and c1[2] == 0.0


def test_f32_weight_sum_drift_is_omega_times_the_excess():
    """The D2Q9 f32 weights sum to 1 + 7.45e-9: a collision moves a cell's
    mass by omega (sum feq - sum f) = omega x that excess (x rho), so a dense
    run drifts by omega x 7.45e-9 per step, and with exact weights not at
    all.  The GPU test of the composite totals follows this drift."""
    excess = a2.W_F32.sum() - 1.0
    assert 7e-9 < excess < 8e-9
    n = 24
    mac = np.tile(np.array([1.0, 0.02, -0.01, 0.0], np.float32), (n * n, 1))
    for exact, want in ((False, 1.0), (True, 0.0)):
        o = a2.Amr2DOracle({"q": 9, "nx": n, "ny": n, "rb": 4, "maxSlots": 1, "omega": 1 / 0.62,
                            "exactWeights": exact})
        o.upload_macro(mac)
        m0 = o.totals()[0]
        for _ in range(10):
            o.macro_step()
        per_step = (o.totals()[0] - m0) / m0 / 10
        if want:
            assert abs(per_step / (o.omega * excess) - 1.0) < 1e-3, per_step
        else:
            assert abs(per_step) < 1e-15
