"""GPU parity of the D2Q9 quadtree pool (lbm_amr_desc q = 9, nz = 1) against
the FP64 oracle (oracle/amr2d_oracle.py), and the reference's 2D Taylor-Green
AMR gate.

- No tile: the L0 step is the oracle's dense D2Q9 step.
- Static seams (faces, corners, a wrapped block, a bar), uniform and linear
  explosion, RB 4, 8 and 16: L0 moments within 1e-5, L0 and tile-interior
  populations within 1e-6 after several macro-steps.
- The |curl u| manager: the criterion is the oracle's binary32 restatement of
  the engine's own L0 field bit for bit, the block maps follow the oracle's
  ascending-id allocator through topology changes, and the fields follow.
- The composite mass and momentum (fine interiors weighted 1/4) follow the
  f32-weight oracle's: conserved but for the D2Q9 weights' sum of 1 +
  7.45e-9, which each collision adds, as the reference's f32 collision does.
- benchmarks/tgv.json amr-N64-L2: the Taylor-Green vortex at N 64, u0 0.04,
  tau 0.8, here with a central block of tiles refined (not idle): the L0
  velocity field against the exact decaying solution within L2rel 0.01 at
  0.5, 1.5 and 3 t_d, and the fitted decay rate within 1 %.
"""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "getting-air_amd"))

import amr2d_oracle as a2  # noqa: E402
import amr_oracle as ao  # noqa: E402
import pylbm  # noqa: E402
from pylbm import scenarios  # noqa: E402
from test_amr2d_oracle import SEAMS2, ids, smooth_field_2d, topology  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def base(n=24, rb=4, **kw):
    d = {"q": 9, "nx": n, "ny": n, "nz": 1, "rb": rb, "maxSlots": 16, "omega": 1.0 / 0.62}
    d.update(kw)
    return d


def make(desc, mac, blocks=()):
    amr = pylbm.Amr(desc)
    orc = a2.Amr2DOracle(desc)
    if blocks:
        bs, sb = topology(orc, blocks)
        amr.upload_topology(bs, sb)
        orc.upload_topology(bs, sb)
    amr.upload_macro(mac)
    orc.upload_macro(mac)
    return amr, orc


def compare(amr, orc, tol=1e-5, label=""):
    mac = amr.read_macro().reshape(-1, 4).astype(np.float64)
    ref = orc.mac()
    umax = np.abs(ref[:, 1:3]).max()
    e_rho = np.abs(mac[:, 0] - ref[:, 0]).max()
    e_u = np.abs(mac[:, 1:3] - ref[:, 1:3]).max() / umax
    f0, fp = amr.read_f()
    cov = orc.covered_mask().reshape(-1)
    e_f0 = np.abs(f0.reshape(9, -1)[:, ~cov] - orc.f0.reshape(9, -1)[:, ~cov]).max() if (~cov).any() else 0.0
    act = ao.active_list(orc.slot_to_block)
    lo, hi = a2.GHOST, a2.GHOST + orc.rb2
    gp = fp.reshape(9, amr.max_slots, amr.fb, amr.fb)[:, act, lo:hi, lo:hi]
    op = orc.pool[:, act, lo:hi, lo:hi]
    e_fp = np.abs(gp - op).max() if act.size else 0.0
    say(label, "rho %.2e  u(rel) %.2e  f_l0 %.2e  f_pool %.2e" % (e_rho, e_u, e_f0, e_fp))
    assert e_rho < tol and e_u < tol
    assert e_f0 < tol * 0.1 and e_fp < tol * 0.1


def test_no_tiles_matches_the_oracle():
    desc = base()
    say("case 2D no tiles")
    amr, orc = make(desc, smooth_field_2d(24, 24, 3))
    with amr:
        for k in range(4):
            amr.step(1)
            orc.macro_step()
        compare(amr, orc, label="no tiles")


@pytest.mark.parametrize("rb", [4, 8, 16])
@pytest.mark.parametrize("linear", [0, 1])
@pytest.mark.parametrize("shape", sorted(SEAMS2))
def test_static_seams_match_the_oracle(shape, linear, rb):
    n = 6 * rb  # six blocks per axis: the seam shapes use block coordinates 0-5
    desc = base(n=n, rb=rb, explodeLinear=linear)
    say("case 2D seams %s linear %d rb %d" % (shape, linear, rb))
    orc0 = a2.Amr2DOracle(desc)
    amr, orc = make(desc, smooth_field_2d(n, n, 7, amp=0.04), ids(orc0, SEAMS2[shape]))
    with amr:
        for k in range(4):
            amr.step(1)
            orc.macro_step()
            compare(amr, orc, label="%s step %d" % (shape, k + 1))


@pytest.mark.parametrize("rb", [4, 8])
def test_manager_is_bit_exact_and_fields_follow(rb):
    n = 12 * rb  # twelve blocks per axis
    desc = base(n=n, rb=rb, maxSlots=64, manageEvery=2, lead=0)
    mac = smooth_field_2d(n, n, 21, amp=0.05)
    vm = a2.vorticity_2d(np.moveaxis(mac.reshape(n, n, 4), -1, 0)[:3].astype(np.float64))
    c = a2.block_criterion(vm, rb, 0)
    desc.update(refineAt=float(np.quantile(c, 0.8)), keepAt=float(np.quantile(c, 0.7)))
    say("case 2D manager rb %d, thresholds" % rb, desc["refineAt"], desc["keepAt"])
    amr, orc = make(desc, mac)
    changes = 0
    with amr:
        prev_bs, prev_sb = amr.read_topology()
        for step in range(10):
            manage = step % 2 == 0
            if manage:
                m = amr.read_macro().reshape(n, n, 4)
                want_crit = a2.block_criterion(a2.vorticity_2d(np.moveaxis(m, -1, 0)[:3].astype(np.float64)), rb, 0)
            amr.step(1)
            if manage:
                crit, want = amr.read_criterion()
                assert (crit.view(np.uint32) == want_crit.view(np.uint32)).all(), step
                dec = ao.decide(prev_bs, crit, np.float32(desc["refineAt"]), np.float32(desc["keepAt"]))
                assert (want == dec).all(), step
                bs, sb = amr.read_topology()
                tbs, tsb, dying, fresh, exh = ao.manage_slots(prev_bs, prev_sb, dec)
                assert (bs == tbs).all() and (sb == tsb).all(), step
                assert exh == 0
                changes += len(dying) + len(fresh)
                prev_bs, prev_sb = bs, sb
                orc.macro_step(crit=crit)
            else:
                orc.macro_step()
            compare(amr, orc, label="manager step %d" % step)
        say("topology changes", changes, "active", amr.info()["active_tiles"])
    assert changes > 0


def test_composite_totals_follow_the_oracle():
    """Explode, coalesce and the orphans conserve the composite mass and
    momentum (the exact-weight oracle pins that to 1e-13).  The D2Q9 f32
    weights sum to 1 + 7.45e-9, so each collision adds about omega x 7.45e-9
    to a cell's mass, as the reference's f32 collision does: over 50
    macro-steps the engine's drift equals the f32-weight oracle's."""
    n = 48
    desc = base(n=n, rb=4)
    mac = smooth_field_2d(n, n, 9, amp=0.04)
    orc = a2.Amr2DOracle(desc)
    topo = topology(orc, ids(orc, SEAMS2["mixed"]))
    orc.upload_topology(*topo)
    orc.upload_macro(mac)
    o0 = orc.totals()
    for _ in range(50):
        orc.macro_step()
    o1 = orc.totals()
    with pylbm.Amr(desc) as amr:
        amr.upload_topology(*topo)
        amr.upload_macro(mac)
        t0 = amr.read_totals()
        amr.step(50)
        t1 = amr.read_totals()
    dm_gpu, dm_orc = (t1[0] - t0[0]) / t0[0], (o1[0] - o0[0]) / o0[0]
    say("2D mass drift over 50 macro-steps: engine %.3e, oracle %.3e" % (dm_gpu, dm_orc))
    assert dm_orc > 3e-7  # the weight-sum drift is there to follow
    assert abs(dm_gpu - dm_orc) < 1e-7
    assert np.abs((t1[1:3] - t0[1:3]) - (o1[1:3] - o0[1:3])).max() < 1e-6 * t0[0]


def test_taylor_green_with_a_refined_centre_matches_the_exact_decay():
    n, tau, u0, rb = 64, 0.8, 0.04, 8
    d, mac = scenarios.tgv2d(n=n, tau=tau, u0=u0)
    desc = base(n=n, rb=rb, omega=d["omega"], maxSlots=16)
    orc0 = a2.Amr2DOracle(desc)
    blocks = [(x, y) for y in range(2, 6) for x in range(2, 6)]  # the central 32 x 32 cells
    nu = (tau - 0.5) / 3.0
    k = 2 * math.pi / n
    td = 1.0 / (nu * 2 * k * k)
    marks = [int(round(c * td)) for c in (0.5, 1.5, 3.0)]
    say("case 2D TGV N64 tau 0.8 with 16 tiles, checkpoints", marks)
    errs, rms = [], []
    with pylbm.Amr(desc) as amr:
        amr.upload_topology(*topology(orc0, ids(orc0, blocks)))
        amr.upload_macro(mac)
        done = 0
        for t in marks:
            amr.step(t - done)
            done = t
            m = amr.read_macro().reshape(-1, 4).astype(np.float64)
            ax, ay = scenarios.tgv2d_analytic(n, tau, u0, t)
            err = math.sqrt(((m[:, 1] - ax) ** 2 + (m[:, 2] - ay) ** 2).sum() / (ax ** 2 + ay ** 2).sum())
            errs.append(err)
            rms.append(math.sqrt((m[:, 1] ** 2 + m[:, 2] ** 2).mean()))
    rate = -math.log(rms[2] / rms[1]) / (marks[2] - marks[1])
    rate_err = abs(rate * td / 1.0 - 1.0)
    say("2D TGV AMR: field L2rel %s, decay rate error %.2e" % (["%.3e" % e for e in errs], rate_err))
    assert max(errs) < 0.01
    assert rate_err < 0.01
