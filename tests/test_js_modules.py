"""Runs the Node test suites under tests/js (oracle closed forms, drop-in JS
modules).  CPU only."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JS = os.path.join(ROOT, "tests", "js")
SUITES = sorted(f for f in os.listdir(JS) if f.startswith("test-") and f.endswith((".js", ".mjs")))


@pytest.mark.parametrize("suite", SUITES)
def test_node_suite(suite):
    r = subprocess.run(["node", os.path.join(JS, suite)], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
