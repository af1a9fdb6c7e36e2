"""GPU parity of the two-level block pool (lbm_amr_* through the C ABI)
against the FP64 oracle (oracle/amr_oracle.py).

Float state: L0 rho/u everywhere and the tile interiors within 1e-5 of the
oracle after several macro-steps.  Integer state (criterion flags, block
maps, active list) bit-exact: the oracle's binary32 criterion is evaluated on
the engine's own macroscopic field, so both sides decide from the same input.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "getting-air_amd"))

import amr_oracle as ao  # noqa: E402
import pylbm  # noqa: E402
from pylbm import scenarios  # noqa: E402
from test_amr_oracle import SEAMS, ids, smooth_field, topology  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def pool_interiors(amr, f_pool):
    """[q, slots, 2rb, 2rb, 2rb] interior view of a read_f pool array."""
    fb, lo, hi = amr.fb, ao.GHOST, ao.GHOST + 2 * amr.rb
    p = f_pool.reshape(amr.q, amr.max_slots, fb, fb, fb)
    return p[:, :, lo:hi, lo:hi, lo:hi]


def compare(amr, orc, tol=1e-5, label=""):
    mac = amr.read_macro().reshape(-1, 4).astype(np.float64)
    ref = orc.mac()
    umax = np.abs(ref[:, 1:]).max()
    e_rho = np.abs(mac[:, 0] - ref[:, 0]).max()
    e_u = np.abs(mac[:, 1:] - ref[:, 1:]).max() / umax
    f0, fp = amr.read_f()
    cov = orc.covered_mask().reshape(-1)
    e_f0 = np.abs(f0.reshape(amr.q, -1)[:, ~cov] - orc.l0_f()[:, ~cov]).max() if (~cov).any() else 0.0
    act = ao.active_list(orc.slot_to_block)
    lo, hi = ao.GHOST, ao.GHOST + orc.rb2
    gp = pool_interiors(amr, fp)[:, act]
    op = orc.pool[:, act, lo:hi, lo:hi, lo:hi]
    e_fp = np.abs(gp - op).max() if act.size else 0.0
    say(label, "rho %.2e  u(rel) %.2e  f_l0 %.2e  f_pool %.2e" % (e_rho, e_u, e_f0, e_fp))
    assert e_rho < tol and e_u < tol
    assert e_f0 < tol * 0.1 and e_fp < tol * 0.1


def make(desc, mac, blocks=()):
    amr = pylbm.Amr(desc)
    orc = ao.AmrOracle(desc)
    if blocks:
        bs, sb = topology(orc, blocks)
        amr.upload_topology(bs, sb)
        orc.upload_topology(bs, sb)
    amr.upload_macro(mac)
    orc.upload_macro(mac)
    return amr, orc


@pytest.mark.parametrize("q", [19, 27])
def test_no_tiles_matches_the_oracle(q):
    n = 16
    desc = {"q": q, "nx": n, "ny": n, "nz": n, "rb": 4, "maxSlots": 4, "omega": 1 / 0.6,
            "force": [1e-5, 0.0, -2e-6]}
    say("no tiles q", q)
    amr, orc = make(desc, smooth_field(n, n, n, 3))
    with amr:
        amr.step(10)
        for _ in range(10):
            orc.macro_step()
        compare(amr, orc, label="no tiles")
        assert amr.read_active().size == 0


@pytest.mark.parametrize("shape", sorted(SEAMS))
@pytest.mark.parametrize("linear", [0, 1])
def test_static_seams_match_the_oracle(shape, linear):
    n = 16
    desc = {"q": 27, "nx": n, "ny": n, "nz": n, "rb": 4, "maxSlots": 8, "omega": 1 / 0.6,
            "explodeLinear": linear}
    say("static seams", shape, "linear", linear)
    orc0 = ao.AmrOracle(desc)
    amr, orc = make(desc, smooth_field(n, n, n, 11, amp=0.04), ids(orc0, SEAMS[shape]))
    with amr:
        t0 = amr.read_totals()
        compare(amr, orc, label="after seed")
        for k in range(3):
            amr.step(2)
            orc.macro_step()
            orc.macro_step()
            compare(amr, orc, label="macro-step %d" % (2 * k + 2))
        t = amr.read_totals()
        # binary32 populations: mass to ~1e-7 over a few steps
        assert abs(t[0] - t0[0]) < 1e-6 * t0[0]
        assert np.abs(t[1:] - t0[1:]).max() < 1e-6 * t0[0]
        bs, sb = amr.read_topology()
        assert (bs == orc.block_slot).all() and (sb == orc.slot_to_block).all()
        assert (amr.read_active() == ao.active_list(orc.slot_to_block)).all()


@pytest.mark.parametrize("q,reflux,rb", [(19, 0, 4), (19, 1, 4), (27, 0, 4), (27, 1, 4), (27, 0, 8), (27, 1, 8)])
def test_interp_interface_and_reflux_match_the_oracle(q, reflux, rb):
    n = 4 * rb  # four blocks per axis: the seam shape uses block coordinates 0-3
    desc = {"q": q, "nx": n, "ny": n, "nz": n, "rb": rb, "maxSlots": 8, "omega": 1 / 0.6,
            "interface": "interp", "reflux": reflux}
    say("interp interface q", q, "reflux", reflux, "rb", rb)
    orc0 = ao.AmrOracle(desc)
    amr, orc = make(desc, smooth_field(n, n, n, 11, amp=0.04), ids(orc0, SEAMS["mixed"]))
    with amr:
        t0 = amr.read_totals()
        for k in range(3):
            amr.step(2)
            orc.macro_step()
            orc.macro_step()
            compare(amr, orc, label="macro-step %d" % (2 * k + 2))
        t = amr.read_totals()
        drift = max(abs(t[0] - t0[0]), np.abs(t[1:] - t0[1:]).max()) / t0[0]
        ot = orc.totals()
        say("composite drift gpu %.2e" % drift)
        if reflux:
            # exact in the oracle; binary32 rounding on the GPU
            assert drift < 1e-6
        bs, sb = amr.read_topology()
        assert (bs == orc.block_slot).all() and (sb == orc.slot_to_block).all()
        assert ot.shape == (4,)


def test_all_refined_matches_the_oracle():
    n = 12
    desc = {"q": 19, "nx": n, "ny": n, "nz": n, "rb": 4, "maxSlots": 27, "omega": 1 / 0.7,
            "force": [2e-6, 1e-6, 0.0]}
    say("all refined")
    amr, orc = make(desc, smooth_field(n, n, n, 5), range(27))
    with amr:
        amr.step(4)
        for _ in range(4):
            orc.macro_step()
        compare(amr, orc, label="all refined")


def test_manager_is_bit_exact_and_fields_follow():
    # two decaying rings; thresholds chosen so that blocks refine after the
    # first pass and others coarsen later, reusing freed slots
    n, rb = 32, 4
    desc = {"q": 27, "nx": n, "ny": n, "nz": n, "rb": rb, "maxSlots": 64, "omega": 1 / 0.56,
            "manageEvery": 2, "lead": 0, "refineAt": 0.0014, "keepAt": 0.00138}
    _, mac, _ = scenarios.vortex_rings(n=n, radius=5.0, core=1.5, circulation=0.05, seed=3, jitter=2.0,
                                       tau=0.56, q=27, rings=2)
    say("manager")
    amr, orc = make(desc, mac)
    with amr:
        prev_bs, prev_sb = amr.read_topology()
        changes = 0
        for step in range(8):
            manage = step % 2 == 0
            if manage:
                m = amr.read_macro().reshape(n, n, n, 4)
                want_crit = ao.block_criterion(
                    ao.vorticity_magnitude(np.moveaxis(m, -1, 0), np.float32), rb, desc["lead"])
            amr.step(1)
            if manage:
                crit, want = amr.read_criterion()
                # bit-exact criterion and decisions from the same field
                assert (crit.view(np.uint32) == want_crit.view(np.uint32)).all(), step
                dec = ao.decide(prev_bs, crit, np.float32(desc["refineAt"]), np.float32(desc["keepAt"]))
                assert (want == dec).all(), step
                bs, sb = amr.read_topology()
                tbs, tsb, dying, fresh, exh = ao.manage_slots(prev_bs, prev_sb, dec)
                assert (bs == tbs).all() and (sb == tsb).all(), step
                assert (amr.read_active() == ao.active_list(tsb)).all(), step
                assert exh == 0 and amr.info()["exhausted"] == 0
                changes += len(dying) + len(fresh)
                prev_bs, prev_sb = bs, sb
                orc.macro_step(crit=crit)
            else:
                orc.macro_step()
            compare(amr, orc, label="step %d" % step)
        say("topology changes", changes, "active", amr.info()["active_tiles"])
        assert amr.info()["active_tiles"] > 0 and changes > 38


def test_q_criterion_is_bit_exact_and_drives_the_manager():
    """criterion 'q': the block maximum of Q = 1/2 (|Omega|^2 - |S|^2)
    (qAt, common_d3_criterion.wgsl:102-158) bit-exact against the oracle's
    binary32 restatement, over a box dilated by one cell, and the manager's
    decisions and maps from it as the oracle twin's."""
    n, rb = 32, 4
    desc = {"q": 19, "nx": n, "ny": n, "nz": n, "rb": rb, "maxSlots": 128, "omega": 1 / 0.56,
            "manageEvery": 2, "lead": 1, "refineAt": 4.0e-6, "keepAt": 3.0e-6, "criterion": "q"}
    _, mac, _ = scenarios.vortex_rings(n=n, radius=5.0, core=1.5, circulation=0.05, seed=3, jitter=2.0,
                                       tau=0.56, q=19, rings=2)
    say("Q criterion")
    amr, orc = make(desc, mac)
    with amr:
        prev_bs, prev_sb = amr.read_topology()
        refined = 0
        for step in range(6):
            manage = step % 2 == 0
            if manage:
                m = amr.read_macro().reshape(n, n, n, 4)
                want_crit = ao.block_criterion(ao.q_criterion(np.moveaxis(m, -1, 0), np.float32), rb, desc["lead"])
            amr.step(1)
            if manage:
                crit, want = amr.read_criterion()
                assert (crit.view(np.uint32) == want_crit.view(np.uint32)).all(), step
                dec = ao.decide(prev_bs, crit, np.float32(desc["refineAt"]), np.float32(desc["keepAt"]))
                assert (want == dec).all(), step
                bs, sb = amr.read_topology()
                tbs, tsb, _, _, exh = ao.manage_slots(prev_bs, prev_sb, dec)
                assert (bs == tbs).all() and (sb == tsb).all() and exh == 0, step
                refined = max(refined, int((bs >= 0).sum()))
                say("step", step, "Q max %.3e" % crit.max(), "refined", int((bs >= 0).sum()))
                prev_bs, prev_sb = bs, sb
                orc.macro_step(crit=crit)
            else:
                orc.macro_step()
            compare(amr, orc, label="step %d" % step)
        assert 0 < refined < (n // rb) ** 3


def test_pool_exhaustion_is_a_hard_failure():
    n = 16
    desc = {"q": 19, "nx": n, "ny": n, "nz": n, "rb": 4, "maxSlots": 2, "omega": 1 / 0.6,
            "manageEvery": 1, "refineAt": -1.0, "keepAt": -1.0}
    amr = pylbm.Amr(desc)
    with amr:
        amr.upload_macro(smooth_field(n, n, n, 2))
        with pytest.raises(pylbm.LbmError) as e:
            amr.step(1)
        assert e.value.code == -3 and "exhausted" in str(e.value)
        info = amr.info()
        assert info["exhausted"] == 64 - 2 and info["active_tiles"] == 2
        bs, sb = amr.read_topology()
        assert sb.tolist() == [0, 1] and (bs[2:] == -1).all()


def beltrami_mac(n, u0):
    """The reference's Beltrami field at t = 0 (d3-scenarios.mjs:160-175,
    seedMacro3): rho = 1 - |u|^2 / (2 cs^2)."""
    k = 2 * np.pi / n
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    u = np.stack([u0 * (np.sin(k * z) + np.cos(k * y)), u0 * (np.sin(k * x) + np.cos(k * z)),
                  u0 * (np.sin(k * y) + np.cos(k * x))], -1).reshape(-1, 3)
    mac = np.empty((n ** 3, 4), np.float32)
    mac[:, 0] = 1.0 - (u ** 2).sum(1) * 1.5
    mac[:, 1:] = u
    return mac.reshape(-1)


def beltrami_run(interface, blocks, checkpoints):
    """Beltrami flow (N 48, tau 0.8, u0 0.04, D3Q19, RB 4) with the given
    blocks refined, static; the L0 field sampled every 2nd cell against the
    analytic decay at each checkpoint (in decay times), as
    tools/lib/d3-metrics.js:109-160 runBeltramiCase scores it.  Returns the
    worst field L2rel and the decay-rate relative error fitted from the RMS
    speed of the last two samples."""
    n, tau, u0 = 48, 0.8, 0.04
    nu = (tau - 0.5) / 3.0
    kk = 2 * np.pi / n
    td = 1.0 / (nu * kk * kk)
    desc = {"q": 19, "nx": n, "ny": n, "nz": n, "rb": 4, "maxSlots": len(blocks), "omega": 1.0 / tau,
            "interface": interface}
    bs = np.full(12 ** 3, -1, np.int32)
    sb = np.full(len(blocks), -1, np.int32)
    for s, b in enumerate(sorted(blocks)):
        bs[b], sb[s] = s, b
    z, y, x = np.meshgrid(np.arange(0, n, 2), np.arange(0, n, 2), np.arange(0, n, 2), indexing="ij")
    shape = np.stack([np.sin(kk * z) + np.cos(kk * y), np.sin(kk * x) + np.cos(kk * z),
                      np.sin(kk * y) + np.cos(kk * x)], -1).reshape(-1, 3) * u0
    samples = []
    with pylbm.Amr(desc) as amr:
        amr.upload_topology(bs, sb)
        amr.upload_macro(beltrami_mac(n, u0))
        prev = 0
        for m in checkpoints:
            t = int(round(m * td))
            amr.step(t - prev)
            prev = t
            u = amr.read_macro().reshape(n, n, n, 4)[::2, ::2, ::2].reshape(-1, 4)[:, 1:].astype(np.float64)
            e = shape * np.exp(-nu * kk * kk * t)
            samples.append((t, np.sqrt(((u - e) ** 2).sum() / (e ** 2).sum()), np.sqrt((u ** 2).sum(1).mean())))
    (ta, _, ra), (tb, _, rb_) = samples[-2:]
    rate = -np.log(rb_ / ra) / (tb - ta)
    return max(s[1] for s in samples), (rate - 1.0 / td) * td, [s[1] for s in samples]


BOX = [(bz * 12 + by) * 12 + bx for bz in range(3, 9) for by in range(3, 9) for bx in range(3, 9)]


@pytest.mark.parametrize("interface,l2_tol,rate_tol,l2_ref,rate_ref,margin", [
    ("explode", 0.0035, 0.004, 2.484e-3, 1.826e-3, 0.10),
    ("interp", 0.012, 0.014, 9.012e-3, 1.024e-2, 0.02)])
def test_box_refined_beltrami_matches_the_reference_measurements(interface, l2_tol, rate_tol, l2_ref, rate_ref,
                                                                   margin):
    """benchmarks/d3.json amr_box_explode_cases[0] / amr_box_cases[0]: the
    centred 6x6x6 cube of 4-cell blocks refined (blocks whose centres lie in
    [N/4, 3N/4)), checkpoints 0.25 and 0.5 decay times.  Held to the
    reference's own gates (field_l2_tol, decay_rate_tol) and to its recorded
    measurements: interp reproduces 9.012e-3 / 1.024e-2 to 0.1 %.  Explode
    reads 2.475e-3 at 0.5 decay times, where the reference records 2.48e-3
    (benchmarks/d3.json amr_interface_note), and its decay rate to 1.3 %;
    its worst checkpoint is the 0.25 one at 2.654e-3, 6.8 % above the
    recorded worst, so the early transient off the seam is where the two
    differ and the margin is 10 %.  The no-interface control is
    test_all_refined_beltrami_matches_the_reference_control."""
    say("box-refined Beltrami", interface)
    l2, rate_err, per = beltrami_run(interface, BOX, (0.25, 0.5))
    say("%s: field L2rel %s worst %.4e (reference %.4e), decay-rate rel err %.4e (reference %.4e)"
        % (interface, ["%.4e" % v for v in per], l2, l2_ref, rate_err, rate_ref))
    assert l2 < l2_tol and abs(rate_err) < rate_tol
    assert abs(l2 / l2_ref - 1.0) < margin
    assert abs(rate_err / rate_ref - 1.0) < 0.1


def test_all_refined_beltrami_matches_the_reference_control():
    """benchmarks/d3.json amr_cases[0] (amr-all-RB4): every block refined, so
    there is no coarse/fine interface; checkpoints 0.5, 1.5 and 3 decay
    times, gates field_l2_tol 0.01 and decay_rate_tol 0.005.  The reference
    records 2.293e-3 worst field L2rel (the no-interface control of the box
    case, field_l2_control_all) and a decay-rate error of -1.222e-4."""
    say("all-refined Beltrami")
    l2, rate_err, per = beltrami_run("explode", list(range(12 ** 3)), (0.5, 1.5, 3.0))
    say("all: field L2rel %s worst %.4e (reference 2.293e-3), decay-rate rel err %.4e (reference -1.222e-4)"
        % (["%.4e" % v for v in per], l2, rate_err))
    assert l2 < 0.01 and abs(rate_err) < 0.005
    assert abs(l2 / 2.293e-3 - 1.0) < 0.02


def body_case(q, n=16):
    c, a = [8.3, 7.9, 8.1], 2.6
    bs, sb = scenarios.refine_near_sphere(n, n, n, 4, c, a, 2.0)
    desc = {"q": q, "nx": n, "ny": n, "nz": n, "rb": 4, "maxSlots": len(sb), "omega": 1 / 0.62,
            "body": {"a": a, "centre": c}}
    mac = smooth_field(n, n, n, 21, amp=0.02)
    mac[:, 1] += 0.04
    return desc, mac, bs, sb


@pytest.mark.parametrize("q", [19, 27])
def test_fine_level_body_matches_the_oracle(q):
    """A pinned sphere on the fine level of a partially refined grid (the
    blocks within 2 cells of it refined, so the seam has faces, edges and
    corners): pool and L0 populations and the per-macro-step force records
    against the oracle's bounce-back, SOLID_EQ and force pass."""
    desc, mac, bs, sb = body_case(q)
    say("fine-level body q", q, "tiles", int((sb >= 0).sum()))
    amr, orc = pylbm.Amr(desc), ao.AmrOracle(desc)
    with amr:
        amr.upload_topology(bs, sb)
        orc.upload_topology(bs, sb)
        amr.upload_macro(mac)
        orc.upload_macro(mac)
        for k in range(3):
            amr.step(1)
            orc.macro_step()
            compare(amr, orc, label="body macro-step %d" % (k + 1))
        f = amr.read_forces()
        ref = np.array(orc.forces)
        scale = np.abs(ref).max()
        say("forces", f.tolist(), "oracle", ref.tolist())
        assert f.shape == ref.shape and abs(ref[:, 0]).min() > 1e-3
        assert np.abs(f - ref).max() < 1e-5 * scale


def test_fine_level_body_with_the_manager_matches_the_oracle():
    """The pinned sphere with the manager running (the geometry term of
    common_d3_manage.wgsl blockWanted OR'ed with the |curl u| criterion,
    :443-462): every block within 2 L0 cells of the body is forced in at the
    first pass and kept whatever the field says, while the field refines and
    coarsens blocks in the wake.  Decisions, block maps and active list
    bit-exact against the oracle's twin; fields and force records follow."""
    n, rb = 24, 4
    c, a = [9.3, 11.9, 12.1], 2.6
    desc = {"q": 27, "nx": n, "ny": n, "nz": n, "rb": rb, "maxSlots": 96, "omega": 1 / 0.6,
            "manageEvery": 2, "lead": 0, "refineAt": 0.0016, "keepAt": 0.0014, "body": {"a": a, "centre": c}}
    _, mac, _ = scenarios.vortex_rings(n=n, radius=4.0, core=1.5, circulation=0.05, seed=5, jitter=2.0,
                                       tau=0.6, q=27, rings=2)
    mac = np.asarray(mac, np.float32).reshape(-1, 4).copy()
    mac[:, 1] += 0.03
    mac = mac.reshape(-1)
    geom = ao.geometry_mask(n // rb, n // rb, n // rb, rb, [float(v) for v in np.float32(c)], float(np.float32(a)),
                            2.0)
    say("AMR body with the manager: forced blocks", int((geom & 1).sum()))
    amr, orc = pylbm.Amr(desc), ao.AmrOracle(desc)
    with amr:
        amr.upload_macro(mac)
        orc.upload_macro(mac)
        prev_bs, prev_sb = amr.read_topology()
        changes, field_only = 0, 0
        for step in range(8):
            manage = step % 2 == 0
            amr.step(1)
            if manage:
                crit, want = amr.read_criterion()
                dec = ao.decide(prev_bs, crit, np.float32(desc["refineAt"]), np.float32(desc["keepAt"]), geom)
                assert (want == dec).all(), step
                assert (want[(geom & 1) != 0] == 1).all(), step
                bs, sb = amr.read_topology()
                tbs, tsb, dying, fresh, exh = ao.manage_slots(prev_bs, prev_sb, dec)
                assert (bs == tbs).all() and (sb == tsb).all(), step
                assert (amr.read_active() == ao.active_list(tsb)).all(), step
                assert exh == 0 and amr.info()["exhausted"] == 0
                changes += len(dying) + len(fresh)
                field_only = max(field_only, int(((bs >= 0) & ((geom & 1) == 0)).sum()))
                prev_bs, prev_sb = bs, sb
                orc.macro_step(crit=crit)
            else:
                orc.macro_step()
            compare(amr, orc, label="step %d" % step)
        f = amr.read_forces()
        ref = np.array(orc.forces)
        say("topology changes", changes, "field-refined blocks (max)", field_only, "forces", f[-1].tolist())
        assert changes > int((geom & 1).sum()) and field_only > 0
        assert f.shape == ref.shape and np.abs(f - ref).max() < 1e-5 * np.abs(ref).max()


def test_amr_body_refusals():
    desc, mac, bs, sb = body_case(19)
    say("AMR body refusals")
    with pytest.raises(pylbm.LbmError):
        pylbm.Amr(dict(desc, interface="interp"))
    with pylbm.Amr(desc) as amr:
        amr.upload_macro(mac)
        with pytest.raises(pylbm.LbmError):
            amr.step(1)  # no topology holds the body yet
        hole = bs.copy()
        b = (2 * 4 + 1) * 4 + 2  # the block holding the centre (8.3, 7.9, 8.1)
        hole_sb = sb.copy()
        hole_sb[hole[b]] = -1
        hole[b] = -1
        with pytest.raises(pylbm.LbmError):
            amr.upload_topology(hole, hole_sb)  # a block next to the body left coarse
        amr.upload_topology(bs, sb)
        amr.step(1)
        assert amr.read_forces().shape == (1, 6)
    with pylbm.Amr(dict(desc, sponge={"width": 6.0, "axes": [1, 0, 0], "u": [0.04, 0, 0]})) as amr:
        with pytest.raises(pylbm.LbmError):
            amr.upload_topology(bs, sb)  # refined blocks inside the sponge band


def test_sphere_amr_re20_matches_the_reference_measurement():
    """benchmarks/d3.json sphere-amr-Re20-D16: the Re 20, D 16 sphere of
    sphere-Re20-D16 (192 x 128 x 128, tau 0.62, sponge 8) with the blocks
    within 2 cells of it refined (RB 4) and the body on the fine level alone;
    after 30 D/U the reference measured AMR Cd 2.9416 against its dense
    2.9436, and gates the pair at 0.5 %."""
    import math
    desc, mac, (bs, sb), meta = scenarios.sphere_amr()
    steps = 30 * 320
    say("case sphere AMR Re20 D16, tiles", int((sb >= 0).sum()))
    with pylbm.Amr(desc) as amr:
        amr.upload_topology(bs, sb)
        amr.upload_macro(mac)
        amr.step(steps)
        f = amr.read_forces(1)
    cd = float(f[-1, 0] / (0.5 * 0.05 ** 2 * math.pi * 64.0))
    say("sphere AMR Re20 D16 Cd %.4f (reference AMR 2.9416, dense 2.9436)" % cd)
    assert abs(cd / 2.9416 - 1.0) < 0.005
    assert abs(cd / 2.9436 - 1.0) < 0.005


@pytest.mark.parametrize("dirs", [0, 1, 2, 27])
def test_rb8_seams_match_the_oracle_for_every_explode_split(dirs):
    """The block size config 3 runs (RB 8: 20^3 tiles, k_pool_step<27, 8> and
    k_explode<27, 8>) on seams with faces, convex edges and corners, a
    concave corner, a lone block and a periodic wrap.  The explode kernel
    gives each thread `dirs` of a ring cell's directions (0: the automatic
    split, here all 27 directions in separate threads); every split must give
    the oracle's populations."""
    n, rb = 32, 8
    desc = {"q": 27, "nx": n, "ny": n, "nz": n, "rb": rb, "maxSlots": 8, "omega": 1 / 0.6, "explodeDirs": dirs}
    say("rb 8 seams, explode dirs per thread", dirs)
    orc0 = ao.AmrOracle(desc)
    amr, orc = make(desc, smooth_field(n, n, n, 17, amp=0.04), ids(orc0, SEAMS["mixed"]))
    with amr:
        t0 = amr.read_totals()
        for k in range(2):
            amr.step(2)
            orc.macro_step()
            orc.macro_step()
            compare(amr, orc, label="rb 8 dirs %d macro-step %d" % (dirs, 2 * k + 2))
        t = amr.read_totals()
        assert abs(t[0] - t0[0]) < 1e-6 * t0[0]


def test_rb16_seams_match_the_oracle():
    """RB 16 (36^3-cell tiles): a face pair, an edge neighbour and a corner
    across the periodic wrap on a 48^3 root (three blocks per axis), linear
    explosion, four macro-steps against the oracle; composite mass
    conserved."""
    n, rb = 48, 16
    desc = {"q": 27, "nx": n, "ny": n, "nz": n, "rb": rb, "maxSlots": 8, "omega": 1 / 0.6}
    say("rb 16 seams")
    orc0 = ao.AmrOracle(desc)
    amr, orc = make(desc, smooth_field(n, n, n, 23, amp=0.04), ids(orc0, [(0, 0, 0), (1, 0, 0), (0, 1, 1), (2, 2, 2)]))
    with amr:
        t0 = amr.read_totals()
        for k in range(2):
            amr.step(2)
            orc.macro_step()
            orc.macro_step()
            compare(amr, orc, label="rb 16 macro-step %d" % (2 * k + 2))
        t = amr.read_totals()
        assert abs(t[0] - t0[0]) < 1e-6 * t0[0]


def test_rb8_manager_coarsens_and_refines_bit_exact_and_fields_follow():
    """RB 8 under the field-criterion manager: a planted layer of refined
    blocks at z = 0 partly coarsens (drain through k_average<27, 8>) while the
    vortex rings' blocks refine (interp fill) on the first pass; criterion,
    decisions, block maps and active list bit-exact, fields within 1e-5."""
    n, rb = 32, 8
    desc = {"q": 27, "nx": n, "ny": n, "nz": n, "rb": rb, "maxSlots": 40, "omega": 1 / 0.7,
            "manageEvery": 2, "lead": 0, "refineAt": 8.2e-4, "keepAt": 7.6e-4}
    _, mac, _ = scenarios.vortex_rings(n=n, radius=6.0, core=1.5, circulation=0.05, seed=3, jitter=2.0,
                                       tau=0.7, q=27, rings=2)
    say("rb 8 manager")
    orc0 = ao.AmrOracle(desc)
    amr, orc = make(desc, mac, [orc0.block_id(bx, by, 0) for bx in range(4) for by in range(4)])
    with amr:
        prev_bs, prev_sb = amr.read_topology()
        dying_total = fresh_total = 0
        for step in range(4):
            manage = step % 2 == 0
            if manage:
                m = amr.read_macro().reshape(n, n, n, 4)
                want_crit = ao.block_criterion(
                    ao.vorticity_magnitude(np.moveaxis(m, -1, 0), np.float32), rb, desc["lead"])
            amr.step(1)
            if manage:
                crit, want = amr.read_criterion()
                assert (crit.view(np.uint32) == want_crit.view(np.uint32)).all(), step
                dec = ao.decide(prev_bs, crit, np.float32(desc["refineAt"]), np.float32(desc["keepAt"]))
                assert (want == dec).all(), step
                bs, sb = amr.read_topology()
                tbs, tsb, dying, fresh, exh = ao.manage_slots(prev_bs, prev_sb, dec)
                assert (bs == tbs).all() and (sb == tsb).all(), step
                assert (amr.read_active() == ao.active_list(tsb)).all(), step
                assert exh == 0
                dying_total += len(dying)
                fresh_total += len(fresh)
                prev_bs, prev_sb = bs, sb
                orc.macro_step(crit=crit)
            else:
                orc.macro_step()
            compare(amr, orc, label="rb 8 manager step %d" % step)
        say("coarsened", dying_total, "refined", fresh_total)
        assert dying_total > 0 and fresh_total > 0
