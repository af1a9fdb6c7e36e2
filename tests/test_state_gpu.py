"""GPU tests of exact checkpoints (lbm_save_state / lbm_load_state) and of
serve requests that resume from one (lbm_serve_submit_state).

A run split at a checkpoint must give the uninterrupted run's macroscopic
field and force records bit for bit, in every storage format; resuming
through the f32 wire format instead (lbm_read_f / lbm_upload_f) does not,
which is what the native state is for.
"""
import os
import struct
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "getting-air_amd"))

import pylbm  # noqa: E402
from pylbm import scenarios as sc  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def sphere(storage):
    desc, mac, _ = sc.sphere(nx=64, ny=32, nz=32, radius=5.0, centre=(20.0, 16.0, 16.0), u0=0.05, re=100.0,
                             storage=storage)
    return dict(desc, autotune=0), mac


def straight(desc, mac, steps):
    with pylbm.Solve(desc) as s:
        s.upload_macro(mac)
        s.step(steps)
        return s.read_macro(), s.read_forces()


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_checkpoint_resume_is_bit_exact(storage):
    desc, mac = sphere(storage)
    say("checkpoint", storage)
    ref_mac, ref_f = straight(desc, mac, 70)
    with pylbm.Solve(desc) as b:
        b.upload_macro(mac)
        b.step(30)
        state = b.save_state()
    magic, version, q, nx, ny, nz, st, salt, step, payload = struct.unpack_from("<8sIiiiiiI4xqQ", state)
    assert magic == b"LBMSTATE" and version == 1 and (q, nx, ny, nz) == (19, 64, 32, 32)
    assert step == 30 and salt == 30 and payload == len(state) - 64
    with pylbm.Solve(desc) as c:
        c.load_state(state)
        c.step(40)
        mac2, f2 = c.read_macro(), c.read_forces()
    assert np.array_equal(mac2.view(np.uint32), ref_mac.view(np.uint32))
    assert ref_f.shape == (70, 6) and f2.shape == (40, 6)
    assert np.array_equal(f2, ref_f[30:])


def test_resume_through_the_wire_format_is_not_exact():
    desc, mac = sphere("f32")
    ref_mac, _ = straight(desc, mac, 70)
    with pylbm.Solve(desc) as b:
        b.upload_macro(mac)
        b.step(30)
        wire = b.read_f()
    with pylbm.Solve(desc) as c:
        c.upload_f(wire)
        c.step(40)
        mac2 = c.read_macro()
    n = int(np.count_nonzero(mac2.view(np.uint32) != ref_mac.view(np.uint32)))
    say("cells differing after a wire-format resume:", n)
    assert n > 0
    assert np.abs(mac2 - ref_mac).max() < 1e-5


def test_mismatched_state_is_refused():
    desc, mac = sphere("f16")
    with pylbm.Solve(desc) as b:
        b.upload_macro(mac)
        state = b.save_state()
    other = dict(desc, storage="bf16")
    with pylbm.Solve(other) as c:
        with pytest.raises(pylbm.LbmError) as e:
            c.load_state(state)
        assert "does not match" in str(e.value)
        with pytest.raises(pylbm.LbmError):
            c.load_state(b"LBMSTATE" + bytes(8))


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_serve_requests_hand_their_state_on(storage):
    desc, mac = sphere(storage)
    ref_mac, ref_f = straight(desc, mac, 100)
    with pylbm.Server(budget_bytes=1 << 30, chunk_steps=16) as sv:
        t1 = sv.submit(desc, mac, 60, keep_state=True)
        sv.wait(t1, 120)
        state = sv.read_state(t1)
        f1 = sv.read_forces(t1)
    # a second server (another process could do it) picks the solve up
    with pylbm.Server(budget_bytes=1 << 30, chunk_steps=24) as sv:
        t2 = sv.submit(desc, None, 40, state=state, keep_state=True)
        sv.wait(t2, 120)
        mac2, f2 = sv.read_macro(t2), sv.read_forces(t2)
        final = sv.read_state(t2)
    assert np.array_equal(mac2.view(np.uint32), ref_mac.view(np.uint32))
    assert np.array_equal(np.vstack([f1, f2]), ref_f)
    assert struct.unpack_from("<q", final, 40)[0] == 100
    # and a plain solve continues from a served state
    with pylbm.Solve(desc) as c:
        c.load_state(state)
        c.step(40)
        assert np.array_equal(c.read_macro().view(np.uint32), ref_mac.view(np.uint32))
