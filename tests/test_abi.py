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
