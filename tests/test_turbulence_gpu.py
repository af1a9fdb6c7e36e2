"""Config 4, decaying turbulence in a periodic D3Q19 box (E(k) ~ k^4
exp(-2 (k/k_p)^2), u_rms 0.04, tau 0.55, seeded random phases): fp32,
shifted fp16 and shifted bf16 storage, collision in fp32.

- At 64^3 (the spectrum's peak wavenumber scaled by 64/512) every format is
  scored against the FP64 oracle after 300 steps: fp32 u and rho within
  1e-5; the 16-bit formats within the bounds stated below (their rounding
  error, measured: f16 ~1e-4, bf16 ~1e-3 relative in u).
- At the full 512^3, 100 steps of each format with the properties the size
  allows: finite fields, FP64 mass drift of the fp32 run under 1e-6 relative
  (the periodic box conserves mass exactly up to fp32 rounding), the energy
  budget of decaying turbulence, and the 16-bit runs within stated bounds of
  the fp32 run.  Enstrophy is not monotone here: from random phases it first
  grows (vortex stretching builds the cascade), while the kinetic energy
  E = <|u|^2>/2 falls at the rate dE/dt = -2 nu Omega.  With Omega from
  central differences (which underestimate |curl u|^2) and the trapezoidal
  rule over the 100 steps, the FP64 oracle at 64^3 gives a ratio of measured
  to predicted energy loss of 1.04-1.11; the test asks for 0.9-1.25.
"""
import numpy as np
import pytest

import oracle_runner
import pylbm
from pylbm import scenarios as sc

pytestmark = pytest.mark.gpu

# u relative L2 error bounds against FP64 (64^3, 300 steps) and against the
# fp32 run (512^3, 100 steps); enstrophy relative bounds likewise
BOUNDS = {"f16": {"u": 2e-3, "ens": 1e-3}, "bf16": {"u": 2e-2, "ens": 1e-2}}


def u_err(m, ref):
    u = m.reshape(-1, 4)[:, 1:].astype(np.float64)
    r = ref.reshape(-1, 4)[:, 1:].astype(np.float64)
    return float(np.sqrt(((u - r) ** 2).sum() / (r ** 2).sum()))


def run(desc, mac, storage, steps):
    with pylbm.Solve(dict(desc, storage=storage)) as s:
        s.upload_macro(mac)
        t0 = s.read_totals()
        s.step(steps)
        s.sync()
        return s.read_macro(), t0, s.read_totals()


def test_turbulence_64_every_format_against_the_fp64_oracle():
    n, steps = 64, 300
    desc, mac = sc.turbulence(n=n, kpeak=8.0 * n / 512)
    print("case turbulence %d^3, %d steps" % (n, steps), flush=True)
    o = oracle_runner.run(desc, mac, steps=steps, threads=8, want=("mac",))
    ref = o["mac"].reshape(-1)
    ens_ref = sc.enstrophy(ref, n)
    rows = {}
    for st in ("f32", "f16", "bf16"):
        m, _, _ = run(desc, mac, st, steps)
        g = m.reshape(-1, 4).astype(np.float64)
        rho = float(np.linalg.norm(g[:, 0] - o["mac"][:, 0]) / np.linalg.norm(o["mac"][:, 0]))
        rows[st] = {"u": u_err(m, ref), "rho": rho, "ens": sc.enstrophy(m, n) / ens_ref - 1.0}
    print("vs FP64:", rows)
    assert rows["f32"]["u"] < 1e-5 and rows["f32"]["rho"] < 1e-5, rows["f32"]
    for st in ("f16", "bf16"):
        assert rows[st]["u"] < BOUNDS[st]["u"] and abs(rows[st]["ens"]) < BOUNDS[st]["ens"], (st, rows[st])


def kinetic_energy(m):
    u = m.reshape(-1, 4)[:, 1:].astype(np.float64)
    return 0.5 * float((u * u).sum(axis=1).mean())


def test_turbulence_512_full_size_properties():
    n, steps = 512, 100
    desc, mac = sc.turbulence(n=n)
    print("case turbulence %d^3 initial field ready" % n, flush=True)
    ens0 = sc.enstrophy(mac, n, np.float32)
    ref, t0, t1 = run(desc, mac, "f32", steps)
    assert np.isfinite(ref).all()
    drift = abs(t1[0] - t0[0]) / t0[0]
    ens_ref = sc.enstrophy(ref, n, np.float32)
    e0, e1 = kinetic_energy(mac), kinetic_energy(ref)
    nu = (1.0 / desc["omega"] - 0.5) / 3.0
    ratio = (e1 - e0) / (-2.0 * nu * 0.5 * (ens0 + ens_ref) * steps)
    print("f32: mass drift %.3e, enstrophy %.6e -> %.6e, energy %.6e -> %.6e, loss / (2 nu Omega t) %.4f" %
          (drift, ens0, ens_ref, e0, e1, ratio), flush=True)
    assert drift < 1e-6
    assert e1 < e0 and 0.9 < ratio < 1.25, ratio
    for st in ("f16", "bf16"):
        m, _, _ = run(desc, mac, st, steps)
        assert np.isfinite(m).all(), st
        eu = u_err(m, ref)
        ee = sc.enstrophy(m, n, np.float32) / ens_ref - 1.0
        print("%s vs f32: u %.3e, enstrophy %.3e" % (st, eu, ee), flush=True)
        assert eu < BOUNDS[st]["u"] and abs(ee) < BOUNDS[st]["ens"], (st, eu, ee)
        del m
