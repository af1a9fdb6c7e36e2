"""Runs the FP64 JS oracle (oracle/run.js) on a job and returns numpy arrays.

Test infrastructure: used by tests/, the smoke check and bench.py's
cpu_baseline leg only.
"""
import json
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN_JS = os.path.join(ROOT, "oracle", "run.js")


def run(desc, mac=None, f=None, steps=0, threads=1, want=("mac", "force"), timeout=900, rigid=None):
    """desc: oracle descriptor dict; mac: float32 [rho,ux,uy,uz]*N; rigid:
    rigid-body parameters (oracle/rigid_oracle.js makeBody) or None; returns
    {'mac': (N,4) f64, 'force': (steps,6) f64, 'f': (Q,N) f64, 'cls': (N,) u8,
     'body': (steps,19) f64 [x, q, v, L, w, d], 'summary': {...}}."""
    q = int(desc["q"])
    nz = 1 if q == 9 else int(desc["nz"])
    n = int(desc["nx"]) * int(desc["ny"]) * nz
    with tempfile.TemporaryDirectory(prefix="lbm_oracle_") as d:
        job = {"desc": desc, "steps": int(steps), "threads": int(threads), "out": {}}
        if rigid is not None:
            job["rigid"] = rigid
        if mac is not None:
            np.ascontiguousarray(mac, dtype=np.float32).tofile(os.path.join(d, "mac_in.f32"))
            job["macIn"] = "mac_in.f32"
        if f is not None:
            np.ascontiguousarray(f, dtype=np.float32).tofile(os.path.join(d, "f_in.f32"))
            job["fIn"] = "f_in.f32"
        for k in want:
            job["out"][k] = k + ".bin"
        with open(os.path.join(d, "job.json"), "w") as fh:
            json.dump(job, fh)
        r = subprocess.run(["node", RUN_JS, os.path.join(d, "job.json")], capture_output=True, text=True,
                           timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError("oracle failed: " + r.stderr[-2000:])
        out = {"summary": json.loads(r.stdout.strip().splitlines()[-1])}
        if "mac" in want:
            out["mac"] = np.fromfile(os.path.join(d, "mac.bin"), dtype=np.float64).reshape(n, 4)
        if "force" in want:
            out["force"] = np.fromfile(os.path.join(d, "force.bin"), dtype=np.float64).reshape(-1, 6)
        if "f" in want:
            out["f"] = np.fromfile(os.path.join(d, "f.bin"), dtype=np.float64).reshape(q, n)
        if "cls" in want:
            out["cls"] = np.fromfile(os.path.join(d, "cls.bin"), dtype=np.uint8)
        if "body" in want:
            out["body"] = np.fromfile(os.path.join(d, "body.bin"), dtype=np.float64).reshape(-1, 19)
        return out
