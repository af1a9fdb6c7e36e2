"""GPU test of the JavaScript host path (lbm-host.mjs -> N-API addon ->
liblbm.so), checked against the FP64 oracle inside the same Node process."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.gpu
def test_js_host_over_napi_addon():
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js_gpu", "test-host-gpu.mjs")], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
