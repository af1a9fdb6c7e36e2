"""CPU checks of the C ABI boundary: liblbm.so loads, exports every function
include/lbm.h declares, and the ctypes descriptor mirrors the C struct
layout byte for byte.  No compute calls (no GPU here)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pylbm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbm.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(lbm_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = pylbm.load()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(pylbm.EXPORTS) == names


def test_abi_version_and_error_text_without_gpu():
    lib = pylbm.load()
    assert lib.lbm_abi_version() == pylbm.ABI_VERSION
    # an invalid descriptor is refused before any device call
    d = pylbm.make_desc({"q": 13, "nx": 4, "ny": 4, "nz": 4, "omega": 1.0})
    h = ctypes.c_void_p()
    rc = lib.lbm_solve_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == -1 and b"q must be" in lib.lbm_last_error()
    d = pylbm.make_desc({"q": 19, "nx": 4, "ny": 4, "nz": 4, "omega": 2.5})
    assert lib.lbm_solve_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"omega" in lib.lbm_last_error()


def test_descriptor_layout_matches_c():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "lbm.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(lbm_desc), offsetof(lbm_desc, body),
         offsetof(lbm_desc, solid_eq), offsetof(lbm_desc, autotune), sizeof(lbm_body),
         offsetof(lbm_desc, sponge_u));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        got = [int(v) for v in subprocess.check_output([exe]).split()]
    D = pylbm.Desc
    want = [ctypes.sizeof(D), D.body.offset, D.solid_eq.offset, D.autotune.offset,
            ctypes.sizeof(pylbm.Body), D.sponge_u.offset]
    assert got == want


def test_amr_descriptor_layout_matches_c():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "lbm.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(lbm_amr_desc), offsetof(lbm_amr_desc, omega),
         offsetof(lbm_amr_desc, force), offsetof(lbm_amr_desc, manage_every),
         offsetof(lbm_amr_desc, lead), offsetof(lbm_amr_desc, orphans), offsetof(lbm_amr_desc, reflux),
         offsetof(lbm_amr_desc, body), offsetof(lbm_amr_desc, explode_dirs), offsetof(lbm_amr_desc, criterion),
         offsetof(lbm_amr_desc, levels), offsetof(lbm_amr_desc, ladder_octaves));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        got = [int(v) for v in subprocess.check_output([exe]).split()]
    A = pylbm.AmrDesc
    want = [ctypes.sizeof(A), A.omega.offset, A.force.offset, A.manage_every.offset, A.lead.offset,
            A.orphans.offset, A.reflux.offset, A.body.offset, A.explode_dirs.offset, A.criterion.offset,
            A.levels.offset, A.ladder_octaves.offset]
    assert got == want


def test_amr_refusals_without_gpu():
    lib = pylbm.load()
    h = ctypes.c_void_p()
    base = {"q": 27, "nx": 32, "ny": 32, "nz": 32, "rb": 8, "maxSlots": 8, "omega": 1.0 / 0.6}
    # tau_c = 1 makes the non-equilibrium rescale singular (main-3d.js:430-442)
    d = pylbm.make_amr_desc(dict(base, omega=1.0))
    assert lib.lbm_amr_create(ctypes.byref(d), ctypes.byref(h)) == -4
    assert b"tau_c" in lib.lbm_last_error()
    d = pylbm.make_amr_desc(dict(base, nx=30))
    assert lib.lbm_amr_create(ctypes.byref(d), ctypes.byref(h)) == -1
    d = pylbm.make_amr_desc(dict(base, nx=16))  # two blocks on x
    assert lib.lbm_amr_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"fewer than 3 blocks" in lib.lbm_last_error()
    d = pylbm.make_amr_desc(dict(base, nx=36, ny=36, nz=36, rb=6))
    assert lib.lbm_amr_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"rb = 4, 8, 16" in lib.lbm_last_error()
    d = pylbm.make_amr_desc(dict(base, q=9))
    assert lib.lbm_amr_create(ctypes.byref(d), ctypes.byref(h)) == -1
    # the seam flux correction needs the interp interface (main-3d.js:1333-1336)
    d = pylbm.make_amr_desc(dict(base, reflux=1))
    assert lib.lbm_amr_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"interp" in lib.lbm_last_error()
    d = pylbm.make_amr_desc(dict(base, explodeDirs=28))
    assert lib.lbm_amr_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"explode_dirs" in lib.lbm_last_error()
    d = pylbm.make_amr_desc(base)
    d.criterion = 2
    assert lib.lbm_amr_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"criterion" in lib.lbm_last_error()
    d = pylbm.make_amr_desc(dict(base, nx=1032, rb=8))  # L0 coordinates are packed in 10 bits
    assert lib.lbm_amr_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"1024" in lib.lbm_last_error()
    # deeper hierarchies: 2 to 4 levels, explode only, parent grids within
    # the 10-bit packing, no parent level at tau = 1, a ladder of at most 16
    # octaves per level
    for lv in (1, 5):
        d = pylbm.make_amr_desc(dict(base, levels=lv))
        assert lib.lbm_amr_create(ctypes.byref(d), ctypes.byref(h)) == -1
        assert b"levels" in lib.lbm_last_error()
    d = pylbm.make_amr_desc(dict(base, levels=3, interface="interp"))
    assert lib.lbm_amr_create(ctypes.byref(d), ctypes.byref(h)) == -4
    d = pylbm.make_amr_desc(dict(base, levels=3, manageEvery=10, ladderOctaves=17))
    assert lib.lbm_amr_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"ladder_octaves" in lib.lbm_last_error()
    d = pylbm.make_amr_desc(dict(base, levels=3, manageEvery=10, ladderOctaves=float("nan")))
    assert lib.lbm_amr_create(ctypes.byref(d), ctypes.byref(h)) == -1
    d = pylbm.make_amr_desc(dict(base, levels=4, nx=264, ny=264, nz=264))
    assert lib.lbm_amr_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"1024" in lib.lbm_last_error()
    d = pylbm.make_amr_desc(dict(base, levels=3, omega=1.0 / 0.75))  # tau_1 = 1
    assert lib.lbm_amr_create(ctypes.byref(d), ctypes.byref(h)) == -4
    assert b"level 1" in lib.lbm_last_error()
    # the D2Q9 quadtree: one cell deep, two levels, explode interface, no body
    b2 = dict(base, q=9, nz=1)
    for kw, code, text in ((dict(nz=8), -1, b"nz = 1"), (dict(levels=3), -4, b"two levels"),
                           (dict(interface="interp"), -4, b"explode"),
                           (dict(body={"a": 2.0, "centre": [12.0, 12.0, 0.0]}), -4, b"no body")):
        d = pylbm.make_amr_desc(dict(b2, **kw))
        assert lib.lbm_amr_create(ctypes.byref(d), ctypes.byref(h)) == code, kw
        assert text in lib.lbm_last_error(), (kw, lib.lbm_last_error())


def test_serve_struct_layouts_match_c():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "lbm.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(lbm_serve_desc),
         offsetof(lbm_serve_desc, budget_bytes), offsetof(lbm_serve_desc, autotune),
         offsetof(lbm_serve_desc, persistent), sizeof(lbm_serve_result),
         offsetof(lbm_serve_result, t_submit), offsetof(lbm_serve_result, force), sizeof(lbm_serve_stats),
         offsetof(lbm_serve_stats, peak_resident), offsetof(lbm_serve_stats, p99_latency),
         offsetof(lbm_serve_stats, device_bytes), offsetof(lbm_serve_stats, graph_updates));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        got = [int(v) for v in subprocess.check_output([exe]).split()]
    D, R, S = pylbm.ServeDesc, pylbm.ServeResult, pylbm.ServeStats
    want = [ctypes.sizeof(D), D.budget_bytes.offset, D.autotune.offset, D.persistent.offset, ctypes.sizeof(R),
            R.t_submit.offset,
            R.force.offset, ctypes.sizeof(S), S.peak_resident.offset, S.p99_latency.offset, S.device_bytes.offset,
            S.graph_updates.offset]
    assert got == want


def test_serve_refusals_without_gpu():
    lib = pylbm.load()
    h = ctypes.c_void_p()
    d = pylbm.ServeDesc()
    d.abi_version = pylbm.ABI_VERSION
    d.budget_bytes = 1000.0
    assert lib.lbm_serve_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"budget" in lib.lbm_last_error()
    d.budget_bytes = float(1 << 30)
    d.persistent = 2
    assert lib.lbm_serve_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"persistent" in lib.lbm_last_error()


def test_totals_noalloc_refusals_without_gpu():
    """The no-allocation totals refuse bad arguments before any device call:
    sizes beyond LBM_NOALLOC_MAX_CELLS, empty lattices, unknown q or
    storage, null pointers."""
    lib = pylbm.load()
    text = open(HEADER).read()
    max_cells = int(re.search(r"#define LBM_NOALLOC_MAX_CELLS (\d+)LL", text).group(1))
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: every call below is refused first
    for q, st, n, lat, out, what in ((19, 0, max_cells + 1, fake, fake, b"ncells"), (19, 0, 0, fake, fake, b"ncells"),
                                     (13, 0, 64, fake, fake, b"q must be"), (19, 3, 64, fake, fake, b"storage"),
                                     (19, 0, 64, None, fake, b"null"), (27, 1, 64, fake, None, b"null")):
        assert lib.lbm_totals_noalloc(q, st, n, lat, out, None) == -1, (q, st, n)
        assert what in lib.lbm_last_error(), lib.lbm_last_error()


def test_state_header_layout():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "lbm.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(lbm_state_header), offsetof(lbm_state_header, version),
         offsetof(lbm_state_header, storage), offsetof(lbm_state_header, salt), offsetof(lbm_state_header, step),
         offsetof(lbm_state_header, payload_bytes));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        got = [int(v) for v in subprocess.check_output([exe]).split()]
    # the layout tests/test_state_gpu.py reads with struct "<8sIiiiiiI4xqQ"
    assert got == [64, 8, 28, 32, 40, 48]


def test_rigid_struct_layouts_match_c():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "lbm.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(lbm_rigid_body), offsetof(lbm_rigid_body, a),
         offsetof(lbm_rigid_body, q), offsetof(lbm_rigid_body, v_max), offsetof(lbm_rigid_body, planar),
         offsetof(lbm_rigid_body, window_anchor), sizeof(lbm_rigid_state), offsetof(lbm_rigid_state, mass));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        got = [int(v) for v in subprocess.check_output([exe]).split()]
    B, S = pylbm.RigidBody, pylbm.RigidState
    want = [ctypes.sizeof(B), B.a.offset, B.q.offset, B.v_max.offset, B.planar.offset, B.window_anchor.offset,
            ctypes.sizeof(S), S.mass.offset]
    assert got == want


def test_rigid_attach_refusals_without_gpu():
    lib = pylbm.load()
    assert lib.lbm_rigid_attach(None, None) == -1
    assert b"null" in lib.lbm_last_error()
    st = pylbm.RigidState()
    assert lib.lbm_rigid_read(None, ctypes.byref(st)) == -1
