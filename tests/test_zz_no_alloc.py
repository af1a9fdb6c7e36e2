"""GPU test of lbm_totals_noalloc, the totals entry point that allocates
nothing (its partial sums live in a fixed device array of the library).

For a solve's current lattice it must give lbm_read_totals' four FP64 sums
bit for bit: the test takes the lattice out of an exact checkpoint
(plane-major, storage bits), lays it out in the device block layout that
include/lbm.h documents, copies it to device memory it allocates itself, and
calls the entry point through the C ABI on calls in a row of decreasing
size (the fixed array keeps the previous call's contents), in every
lattice and storage format, with grids that are not a multiple of the
64-cell block."""
import ctypes
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

H2D, D2H = 1, 2


def say(*a):
    print(*a, flush=True)


def hip():
    pylbm.load()  # liblbm.so brings the HIP runtime in; bind the same library
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


def device_copy(h, arr):
    p = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(p), max(arr.nbytes, 256)) == 0
    assert h.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D) == 0
    return p


BLOCK = 64  # cells per layout block (include/lbm.h, lbm_totals_noalloc)


def block_layout(plane, q, n):
    """[q][n] plane-major -> population i of cell c at ((c // B) q + i) B + c % B."""
    nb = (n + BLOCK - 1) // BLOCK
    padded = np.zeros((q, nb * BLOCK), dtype=plane.dtype)
    padded[:, :n] = plane
    return np.ascontiguousarray(padded.reshape(q, nb, BLOCK).transpose(1, 0, 2)).reshape(-1)


def case(kind, storage):
    if kind == "sphere19":
        desc, mac, _ = sc.sphere(nx=48, ny=24, nz=24, radius=4.0, centre=(14.0, 12.0, 12.0), re=80.0,
                                 storage=storage)
    elif kind == "field27":
        desc, mac = sc.random_field(24, 18, 9, q=27, seed=3)  # 3888 cells: a ragged last block
        desc = dict(desc, storage=storage)
    else:
        desc, mac = sc.random_field(40, 27, 1, q=9, seed=4)  # 1080 cells
        desc = dict(desc, storage=storage)
    return dict(desc, autotune=0), mac


def run_case(h, kind, storage, steps):
    desc, mac = case(kind, storage)
    with pylbm.Solve(desc) as s:
        s.upload_macro(mac)
        s.step(steps)
        s.sync()
        want = s.read_totals()
        state = s.save_state()
        q, n = desc["q"], s.ncells
    magic, _, q_, _, _, _, st, _, _, payload = struct.unpack_from("<8sIiiiiiI4xqQ", state)
    assert magic == b"LBMSTATE" and q_ == q and st == pylbm.STORE[storage]
    dt = np.float32 if storage == "f32" else np.uint16
    plane = np.frombuffer(state, dtype=dt, count=payload // np.dtype(dt).itemsize, offset=64).reshape(q, n)
    lat = device_copy(h, block_layout(plane, q, n))
    out = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(out), 32) == 0
    try:
        got = []
        for _ in range(2):  # twice on the same input: the array holds the last call's partials
            rc = pylbm.load().lbm_totals_noalloc(q, pylbm.STORE[storage], n, lat, out, None)
            assert rc == 0, pylbm.load().lbm_last_error()
            res = np.zeros(4, dtype=np.float64)
            assert h.hipMemcpy(res.ctypes.data, out, 32, D2H) == 0
            got.append(res)
    finally:
        h.hipFree(lat)
        h.hipFree(out)
    say(kind, storage, "cells", n, "read_totals", want, "noalloc", got[0])
    for g in got:
        assert np.array_equal(g.view(np.uint64), want.view(np.uint64)), (kind, storage, g, want)
    return n


def test_totals_noalloc_equals_read_totals_on_shrinking_inputs():
    h = hip()
    sizes = []
    # each call smaller than the one before
    for kind, storage, steps in (("sphere19", "f32", 30), ("field27", "f16", 12), ("field9", "bf16", 20),
                                 ("field9", "f32", 7)):
        sizes.append(run_case(h, kind, storage, steps))
    assert sizes[0] > sizes[1] > sizes[2] == sizes[3]


def test_totals_noalloc_refuses_sizes_beyond_its_limit():
    lib = pylbm.load()
    assert lib.lbm_totals_noalloc(19, 0, (1 << 31) + 1, ctypes.c_void_p(0x1000), ctypes.c_void_p(0x1000), None) == -1
    assert b"ncells" in lib.lbm_last_error()
