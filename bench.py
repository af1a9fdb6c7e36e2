"""Benchmark: D3Q19 dense flow past a pinned sphere on one MI355X.

Workload (the D3Q19 TRT sphere case): 256x128x128 cells, sphere radius 12 at
(64, 64, 64) with halfway bounce-back, freestream u = 0.05 at Re = 200, ALBC
sponge (width 12) on every face, FP64 drag/lift reduction every step.  One
"step" = force reduction + fused pull-stream/collide over the whole grid.
Inputs are resident in HBM before the timed region starts.

Prints one JSON line:
  value        whole-job MLUPS (fp32 storage), host clock around K steps
  roofline     the fused stream-collide kernel: algorithmic bytes per launch
               (2 * Q * 4 B * cells = 152 B/cell) over its mean launch time
               measured with HIP events on the solve's stream
  cpu_baseline the FP64 oracle (node oracle/run.js, worker threads) on a
               bounded number of steps of the same grid
  fp16_store / bf16_store  the same workload with shifted 16-bit storage
  towed_sphere_d3q19_256x128x128  the sphere as a moving rigid body towed at
               -u0 through fluid at rest under a moving window (its drag
               beside the pinned sphere's: the same flow in another frame)
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "getting-air_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

import pylbm  # noqa: E402
from pylbm import scenarios as sc  # noqa: E402

METRIC = "D3Q19 MLUPS & % HBM peak, fp32 vs fp16-store; solves/s + p99 latency served"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_d3q19_sphere_256x128x128.json")


def measure(storage, steps, warmup, kernel_reps, periodic=False, tile=None):
    desc, mac, meta = sc.sphere(storage=storage)
    if periodic:
        # the same grid with no body, sponge or walls: the step kernel's
        # periodic fast path, for comparison only
        desc = {k: v for k, v in desc.items() if k not in ("body", "sponge")}
    desc["autotune"] = 0 if tile else 1
    q = desc["q"]
    with pylbm.Solve(desc) as s:
        if tile:
            s.set_tile(*tile)
        s.upload_macro(mac)
        s.step(warmup)
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        t1 = time.perf_counter()
        forces = s.read_forces(steps)
        kernel_ms = s.time_step_kernel(kernel_reps)
        info = s.info()
        ncells = s.ncells
    secs = t1 - t0
    elem = 4 if storage == "f32" else 2
    bytes_per_launch = 2 * q * elem * ncells
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9
    # time-averaged drag over the last 1000 steps (the 16-bit formats carry
    # per-step dither noise of a few 0.1 %; its mean is unbiased)
    cd = float(forces[-1000:, 0].mean() / (0.5 * meta["u0"] ** 2 * meta["area"])) if len(forces) else None
    return {
        "mlups": ncells * steps / secs / 1e6,
        "ms_per_step": secs * 1e3 / steps,
        "kernel_ms": kernel_ms,
        "achieved_gbs": achieved,
        "bytes_per_launch": bytes_per_launch,
        "ncells": ncells,
        "tile": [info["block_x"], info["block_y"], info["flags"]],
        "cd_last1000": cd,
        "finite": bool(np.all(np.isfinite(forces))),
    }


def measure_towed(steps, warmup, kernel_reps, re=36.0):
    """The same sphere moving at -u0 through fluid at rest (towed: force
    measured, never applied), with a moving window along x that keeps it at
    x = 64 and carries the sponge with it: the Galilean partner of the pinned
    sphere in a +u0 freestream.  One step = the rigid body's force reduction
    and integrator, its reclassification, and the fused stream-collide.
    Re = 36 puts tau at 0.6, where the moving-body bounce-back coupling is
    sound; its force deficit grows as 1/(tau - 1/2) toward the headline
    case's tau = 0.518 (the reference records the same, benchmarks/d3.json
    tow_amr_note).  The pinned sphere at the same tau is run beside it."""
    desc, mac, meta = sc.sphere(storage="f32", re=re)
    pinned = dict(desc, autotune=1)
    with pylbm.Solve(pinned) as s:
        s.upload_macro(mac)
        s.step(warmup + steps)
        s.sync()
        fp = s.read_forces(steps)
    cd_pinned = float(fp[-1000:, 0].mean() / (0.5 * meta["u0"] ** 2 * meta["area"]))
    body = desc.pop("body")
    desc["sponge"] = dict(desc["sponge"], u=[0.0, 0.0, 0.0])
    desc["autotune"] = 1
    u0 = meta["u0"]
    m = np.asarray(mac, np.float32).reshape(-1, 4).copy()
    m[:, 0] = 1.0
    m[:, 1:] = 0.0
    rigid = {"shape": {"kind": 0, "a": body["a"]}, "x": list(body["centre"]), "v": [-u0, 0.0, 0.0],
             "noFluidForce": True, "window": [1, 0, 0], "anchor": [int(body["centre"][0]), 0, 0]}
    with pylbm.Solve(desc) as s:
        s.upload_macro(m.reshape(-1))
        s.attach_rigid(rigid)
        s.step(warmup)
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        t1 = time.perf_counter()
        forces = s.read_forces(steps)
        kernel_ms = s.time_step_kernel(kernel_reps)
        st = s.read_rigid()
        ncells = s.ncells
    secs = t1 - t0
    cd = float(forces[-1000:, 0].mean() / (0.5 * u0 ** 2 * meta["area"])) if len(forces) else None
    return {"mlups": ncells * steps / secs / 1e6, "ms_per_step": secs * 1e3 / steps, "step_kernel_ms": kernel_ms,
            "cd_last1000": cd, "cd_pinned_last1000": cd_pinned, "tau": meta["tau"], "re": re,
            "travel_cells": -st["d"][0], "finite": bool(np.all(np.isfinite(forces)))}


def measure_plate():
    """The reference's blunt-plate gate (benchmarks/d3.json plate_cases, C16):
    a 16 x 16 x 2 plate normal to the flow at Re 200 (tau 0.512), pinned in a
    -0.05 stream and half-towed (tow 0.025, stream -0.025), Cd averaged as its
    harness does (the last 30 % of 20 chord/U)."""
    out = {}
    for name, tow, stream in (("pinned", 0.0, -0.05), ("half_tow", 0.025, -0.025)):
        desc, mac, rigid, meta = sc.plate(tow=tow, stream=stream)
        settle = int(round(20 * meta["convective"]))
        warm = int(round(settle * 0.7))
        with pylbm.Solve(desc) as s:
            s.upload_macro(mac)
            s.attach_rigid(rigid)
            s.step(warm)
            s.step(settle - warm)
            s.sync()
            f = s.read_forces(settle - warm)
        out[name] = float((meta["sign"] * f[:, 0] / (0.5 * meta["u_rel"] ** 2 * meta["area"])).mean())
    # the reference's dense sphere pair at tau 0.6 (D 12, U 0.025, 16 D/U;
    # benchmarks/d3.json tow_amr_note): pinned 5.8957, towed 5.8481
    for name, tow, stream in (("sphere_pinned", 0.0, -0.025), ("sphere_towed", 0.025, 0.0)):
        desc, mac, rigid, meta = sc.fall(tow=tow, stream=stream)
        settle = int(round(16 * meta["convective"]))
        warm = int(round(settle * 0.7))
        with pylbm.Solve(desc) as s:
            s.upload_macro(mac)
            s.attach_rigid(rigid)
            s.step(settle)
            s.sync()
            f = s.read_forces(settle - warm)
        out[name] = float((meta["sign"] * f[:, 0] / (0.5 * meta["u_rel"] ** 2 * meta["area"])).mean())
    # the reference's recorded dense pinned sphere (sphere-Re20-D16): 2.9436
    desc, mac, _ = sc.sphere(nx=192, ny=128, nz=128, radius=8.0, centre=(48.0, 64.0, 64.0), u0=0.05, re=20.0,
                             trt=False, sponge_width=8)
    with pylbm.Solve(desc) as s:
        s.upload_macro(mac)
        s.step(30 * 320)
        s.sync()
        f = s.read_forces(1)
    out["sphere_re20"] = float(f[-1, 0] / (0.5 * 0.05 ** 2 * np.pi * 64.0))
    # the same sphere on the fine level of two (sphere-amr-Re20-D16): 2.9416
    adesc, amac, (bs, sb), _ = sc.sphere_amr()
    with pylbm.Amr(adesc) as a:
        a.upload_topology(bs, sb)
        a.upload_macro(amac)
        a.step(1)
        a.read_forces(1)  # waits for the stream
        t0 = time.perf_counter()
        a.step(30 * 320 - 1)
        f = a.read_forces(1)
        amr_ms = (time.perf_counter() - t0) * 1e3 / (30 * 320 - 1)
    out["sphere_amr"] = float(f[-1, 0] / (0.5 * 0.05 ** 2 * np.pi * 64.0))
    # the same sphere on the finest of three levels (sphere-amr-L3-Re20-D16,
    # refineHierarchy with margin 2): 2.9015
    adesc, amac, topo, _ = sc.sphere_amr(levels=3)
    with pylbm.Amr(adesc) as a:
        for m in (1, 2):
            a.upload_level_topology(m, *topo[m])
        a.upload_macro(amac)
        a.step(1)
        a.read_forces(1)
        t0 = time.perf_counter()
        a.step(30 * 320 - 1)
        f_static = a.read_forces(30 * 320)
        l3_ms = (time.perf_counter() - t0) * 1e3 / (30 * 320 - 1)
    out["sphere_amr_l3"] = float(f_static[-1, 0] / (0.5 * 0.05 ** 2 * np.pi * 64.0))
    # the same with the manager every 16 macro-steps from an empty pool
    # (sphere-amr-dynamic-L3-Re20-D16): the reference's gate is a force
    # history bit-identical to the static run's
    with pylbm.Amr(dict(adesc, manageEvery=16)) as a:
        a.upload_macro(amac)
        a.step(1)
        a.read_forces(1)
        t0 = time.perf_counter()
        a.step(30 * 320 - 1)
        f_dyn = a.read_forces(30 * 320)
        l3d_ms = (time.perf_counter() - t0) * 1e3 / (30 * 320 - 1)
    out["sphere_amr_l3_dynamic"] = float(f_dyn[-1, 0] / (0.5 * 0.05 ** 2 * np.pi * 64.0))
    l3d_same = bool(np.array_equal(f_static.view(np.uint64), f_dyn.view(np.uint64)))
    return {"sphere_re20_d16_cd": out["sphere_re20"], "sphere_re20_d16_cd_reference_measured": 2.9436,
            "sphere_amr_re20_d16_cd": out["sphere_amr"], "sphere_amr_re20_d16_cd_reference_measured": 2.9416,
            "sphere_amr_re20_d16_tiles": int((sb >= 0).sum()), "sphere_amr_re20_d16_ms_per_macro_step": amr_ms,
            "sphere_amr_l3_re20_d16_cd": out["sphere_amr_l3"], "sphere_amr_l3_re20_d16_cd_reference_measured": 2.9015,
            "sphere_amr_l3_re20_d16_tiles": [int((t[1] >= 0).sum()) for t in topo[1:]],
            "sphere_amr_l3_re20_d16_ms_per_macro_step": l3_ms,
            "sphere_amr_dynamic_l3_re20_d16_cd": out["sphere_amr_l3_dynamic"],
            "sphere_amr_dynamic_l3_bit_identical_to_static": l3d_same,
            "sphere_amr_dynamic_l3_ms_per_macro_step": l3d_ms,
            "plate_cd_pinned": out["pinned"], "plate_cd_pinned_reference_measured": 1.2081, "plate_cd_literature": 1.18,
            "plate_cd_half_tow": out["half_tow"], "plate_half_tow_vs_pinned": out["half_tow"] / out["pinned"] - 1.0,
            "plate_half_tow_vs_pinned_reference_measured": -0.0321,
            "sphere_tau06_cd_pinned": out["sphere_pinned"], "sphere_tau06_cd_pinned_reference_measured": 5.8957,
            "sphere_tau06_cd_towed": out["sphere_towed"], "sphere_tau06_cd_towed_reference_measured": 5.8481}


def measure_drag_vs_fp64(threads, steps=3000, window=1000):
    """Drag of every storage format against the FP64 oracle over a settled
    window, independent of --steps: the headline's sphere scaled down to
    96x48x48 (radius 4.5 at (24, 24, 24), sponge 6) at the headline's tau
    0.518 (u0 0.05, Re 75), D3Q19 TRT, `steps` steps, Cd averaged over the
    last `window` steps."""
    import oracle_runner

    desc, mac, meta = sc.sphere(nx=96, ny=48, nz=48, radius=4.5, centre=(24.0, 24.0, 24.0), re=75.0,
                                sponge_width=6)
    q = 0.5 * meta["u0"] ** 2 * meta["area"]
    r = oracle_runner.run(desc, mac, steps=steps, threads=threads, want=("force",))
    cd64 = float(r["force"][-window:, 0].mean() / q)
    out = {"grid": [96, 48, 48], "tau": round(meta["tau"], 4), "steps": steps, "window": window,
           "cd_fp64_oracle": cd64}
    for st in ("f32", "f16", "bf16"):
        with pylbm.Solve(dict(desc, storage=st, autotune=0)) as s:
            s.upload_macro(mac)
            s.step(steps)
            f = s.read_forces(steps)
        cd = float(f[-window:, 0].mean() / q)
        out[st] = {"cd": cd, "drag_dev_vs_fp64_pct": 100.0 * (cd / cd64 - 1.0)}
    return out


def cpu_baseline(steps, threads):
    """The FP64 JS oracle (node, worker threads over x-rows) on `steps` steps
    of the benchmark grid.  `threads` is the CPU share one GPU's job gets on
    the box (16 cores), not the host's whole CPU count, which is reported
    beside it."""
    import oracle_runner

    desc, mac, _ = sc.sphere()
    r = oracle_runner.run(desc, mac, steps=steps, threads=threads, want=())
    s = r["summary"]
    return {
        "value": s["mlups"], "unit": "MLUPS", "cores": s["threads"], "kind": "port",
        "host_logical_cpus": s.get("cpu_count"), "cpu_model": s["cpu_model"],
        "sample": "FP64 JS oracle (node, %d worker threads on %d of the host's %s logical CPUs, %s), %d steps "
                  "of the 256x128x128 D3Q19 TRT sphere grid (%.1f s)" %
                  (s["threads"], s["threads"], s.get("cpu_count"), s["cpu_model"], steps, s["seconds"]),
    }


def measure_amr(steps, manage_every=50, interface="explode", reflux=0):
    """The two-level AMR workload: D3Q27 BGK, 128^3 root, 8^3-cell blocks
    (16^3-cell tiles), seeded vortex rings (radius 20, core 4, circulation
    0.1), |curl u| criterion with hysteresis, manager every 50 macro-steps.
    Cell updates per macro-step count the L0 grid plus two fine substeps of
    every tile interior (the reference's cellUpdatesPerMacroStep).
    `interface`/`reflux` select the coarse/fine coupling (explode/coalesce,
    or interp with or without the seam flux correction)."""
    desc, mac, meta = sc.vortex_rings(n=128, radius=20.0, core=4.0, circulation=0.1, seed=7, jitter=8.0,
                                      tau=0.52, q=27)
    w0 = meta["vorticity_max"]
    adesc = dict(desc, rb=8, maxSlots=2048, manageEvery=manage_every, lead=1,
                 refineAt=float(np.float32(0.25 * w0)), keepAt=float(np.float32(0.15 * w0)),
                 interface=interface, reflux=reflux)
    n0 = 128 ** 3
    interior = 16 ** 3
    with pylbm.Amr(adesc) as a:
        a.upload_macro(mac)
        a.manage_now()
        t_start = a.read_totals()
        updates, seconds, tiles = 0, 0.0, []
        done = 0
        while done < steps:
            k = min(manage_every, steps - done)
            t0 = time.perf_counter()
            a.step(k)
            seconds += time.perf_counter() - t0
            nt = a.info()["active_tiles"]
            tiles.append(nt)
            updates += k * (n0 + 2 * nt * interior)
            done += k
        t_end = a.read_totals()
        info = a.info()
    return {"mlups": updates / seconds / 1e6, "seconds": seconds, "macro_steps": steps,
            "cell_updates": updates, "tiles_min": min(tiles), "tiles_max": max(tiles), "tiles_last": tiles[-1],
            "exhausted": info["exhausted"], "mass_drift_rel": float((t_end[0] - t_start[0]) / t_start[0]),
            "momentum_drift_rel": float(np.abs(t_end[1:] - t_start[1:]).max() / t_start[0]),
            "refine_at": adesc["refineAt"], "keep_at": adesc["keepAt"], "tau": 0.52,
            "interface": interface, "reflux": reflux}


def measure_amr2d(steps, n=1024, manage_every=50):
    """The D2Q9 quadtree workload: a periodic double shear layer (two tanh
    layers of thickness n/80 at u0 0.05, a sin(2 pi x / n) transverse kick of
    5 %), tau 0.52, 8^2-cell blocks (16^2-cell tiles), |curl u| criterion with
    hysteresis at 0.25 / 0.15 of the initial peak, manager every 50
    macro-steps.  Cell updates per macro-step: the root grid plus two fine
    substeps of every tile interior (cellUpdatesPerMacroStep, amr2d.mjs)."""
    u0, delta, eps, tau, rb = 0.05, n / 80.0, 0.05, 0.52, 8
    y, x = np.meshgrid(np.arange(n, dtype=np.float64) + 0.5, np.arange(n, dtype=np.float64), indexing="ij")
    ux = np.where(y <= n / 2, u0 * np.tanh((y - n / 4) / delta), u0 * np.tanh((3 * n / 4 - y) / delta))
    uy = eps * u0 * np.sin(2 * np.pi * x / n)
    mac = np.zeros((n * n, 4), dtype=np.float32)
    mac[:, 0] = 1.0
    mac[:, 1] = ux.reshape(-1)
    mac[:, 2] = uy.reshape(-1)
    w0 = u0 / delta  # peak |du/dy| of a tanh layer
    desc = {"q": 9, "nx": n, "ny": n, "nz": 1, "rb": rb, "maxSlots": 8192, "omega": 1.0 / tau,
            "manageEvery": manage_every, "lead": 1, "refineAt": float(np.float32(0.25 * w0)),
            "keepAt": float(np.float32(0.15 * w0))}
    n0, interior = n * n, (2 * rb) ** 2
    with pylbm.Amr(desc) as a:
        a.upload_macro(mac)
        a.manage_now()
        t_start = a.read_totals()
        updates, seconds, tiles = 0, 0.0, []
        done = 0
        while done < steps:
            k = min(manage_every, steps - done)
            t0 = time.perf_counter()
            a.step(k)
            seconds += time.perf_counter() - t0
            nt = a.info()["active_tiles"]
            tiles.append(nt)
            updates += k * (n0 + 2 * nt * interior)
            done += k
        t_end = a.read_totals()
        info = a.info()
    return {"mlups": updates / seconds / 1e6, "seconds": seconds, "macro_steps": steps, "cell_updates": updates,
            "tiles_min": min(tiles), "tiles_max": max(tiles), "tiles_last": tiles[-1], "exhausted": info["exhausted"],
            "mass_drift_rel": float((t_end[0] - t_start[0]) / t_start[0]),
            "momentum_drift_rel": float(np.abs(t_end[1:] - t_start[1:]).max() / t_start[0]),
            "refine_at": desc["refineAt"], "keep_at": desc["keepAt"], "tau": tau}


def measure_amr_slab(steps, reflux):
    """The seam flux correction where it is valid: a static slab of refined
    blocks (z in the middle half, every x and y), so every seam is a flat face
    -- at a convex corner the fine lattice's channels do not tile the coarse
    one and the correction perturbs rather than conserves
    (common_d3_amr_flux.wgsl:13-32).  Same D3Q27 vortex-ring field and 128^3
    root as the AMR workload, interp interface."""
    desc, mac, _ = sc.vortex_rings(n=128, radius=20.0, core=4.0, circulation=0.1, seed=7, jitter=8.0, tau=0.52, q=27)
    nb = 16
    ids = [(bz * nb + by) * nb + bx for bz in range(4, 12) for by in range(nb) for bx in range(nb)]
    bs = np.full(nb ** 3, -1, dtype=np.int32)
    sb = np.full(len(ids), -1, dtype=np.int32)
    for k, b in enumerate(ids):
        bs[b] = k
        sb[k] = b
    adesc = dict(desc, rb=8, maxSlots=len(ids), manageEvery=0, interface="interp", reflux=reflux)
    with pylbm.Amr(adesc) as a:
        a.upload_topology(bs, sb)
        a.upload_macro(mac)
        t_start = a.read_totals()
        t0 = time.perf_counter()
        a.step(steps)  # returns after the stream drains
        seconds = time.perf_counter() - t0
        t_end = a.read_totals()
    updates = steps * (128 ** 3 + 2 * len(ids) * 16 ** 3)
    return {"mlups": updates / seconds / 1e6, "macro_steps": steps, "tiles": len(ids), "reflux": reflux,
            "mass_drift_rel": float((t_end[0] - t_start[0]) / t_start[0]),
            "momentum_drift_rel": float(np.abs(t_end[1:] - t_start[1:]).max() / t_start[0])}


def enstrophy(mac, n):
    return sc.enstrophy(mac, n)


def measure_turbulence(n=512, steps=2000, n_small=128, threads=16):
    """Config 4: decaying turbulence, E(k) ~ k^4 exp(-2 (k/8)^2), u_rms 0.04,
    tau 0.55, in fp32 and shifted fp16 / bf16 storage.  At n^3 the 16-bit
    runs are compared with the fp32 run; at n_small^3 (the same spectrum, peak
    wavenumber scaled by n_small/n) every format is compared with the FP64
    oracle, which is what shows fp32 standing in for FP64 at n^3."""
    import oracle_runner

    def run(desc, mac, storage, nn):
        with pylbm.Solve(dict(desc, storage=storage, autotune=1)) as s:
            s.upload_macro(mac)
            s.step(20)
            s.sync()
            s.upload_macro(mac)
            t0 = time.perf_counter()
            s.step(steps)
            s.sync()
            secs = time.perf_counter() - t0
            out = s.read_macro()
        return out, nn ** 3 * steps / secs / 1e6

    def errs(m, ref):
        u, r = m.reshape(-1, 4)[:, 1:].astype(np.float64), ref.reshape(-1, 4)[:, 1:].astype(np.float64)
        return float(np.sqrt(((u - r) ** 2).sum() / (r ** 2).sum()))

    def say(*a):
        print("turbulence:", *a, file=sys.stderr, flush=True)

    res = {"n": n, "steps": steps, "n_small": n_small}
    desc, mac = sc.turbulence(n=n)
    say("initial field %d^3 ready" % n)
    ref, ml = run(desc, mac, "f32", n)
    ens_ref = enstrophy(ref, n)
    res["f32"] = {"mlups": ml, "enstrophy": ens_ref, "enstrophy0": enstrophy(mac, n)}
    say("f32", res["f32"])
    for st in ("f16", "bf16"):
        m, ml = run(desc, mac, st, n)
        res[st] = {"mlups": ml, "speedup_vs_f32": ml / res["f32"]["mlups"], "u_rel_l2_vs_f32": errs(m, ref),
                   "enstrophy_rel_vs_f32": enstrophy(m, n) / ens_ref - 1.0}
        say(st, res[st])
        del m
    del ref, mac
    desc, mac = sc.turbulence(n=n_small, kpeak=8.0 * n_small / n)
    say("FP64 oracle, %d^3, %d steps, %d threads" % (n_small, steps, threads))
    o = oracle_runner.run(desc, mac, steps=steps, threads=threads, want=("mac",))
    oref = o["mac"].reshape(-1)
    ens_o = enstrophy(oref, n_small)
    res["small_fp64_oracle"] = {"enstrophy": ens_o, "oracle_mlups": o["summary"]["mlups"]}
    for st in ("f32", "f16", "bf16"):
        m, _ = run(desc, mac, st, n_small)
        res["small_fp64_oracle"][st] = {"u_rel_l2_vs_fp64": errs(m, oref),
                                        "enstrophy_rel_vs_fp64": enstrophy(m, n_small) / ens_o - 1.0}
    return res


def measure_serve(count, rate=50.0, seed=5, steps=2000, budget=200e9, scale=1.0, baseline=True):
    """Config-5 serving: `count` requests with seeded Poisson arrivals at
    `rate`/s (2D cylinder channels nx 256-1024 and 3D boxed spheres 64^3-192^3,
    Re uniform in [50, 500], shifted fp16 storage), submitted at their arrival
    times to the batched serve loop; then the same requests one at a time
    through the plain single-solve path (create, upload, step, read back,
    destroy).  Both paths compute bit-identical results (tests/test_serve_gpu.py)."""
    import threading
    mix = sc.serve_mix(count=count, rate=rate, seed=seed, steps=steps, storage="f16", scale=scale)
    cells = [d["nx"] * d["ny"] * (1 if d["q"] == 9 else d["nz"]) for _, d, _, _, _ in mix]
    updates = float(sum(c * st for c, (_, _, _, _, st) in zip(cells, mix)))
    def serve_run(persistent):
        # 48M resident cells fill the device; 64-step chunks hide admissions
        with pylbm.Server(budget_bytes=budget, chunk_steps=64, autotune=1, target_cells=48e6,
                          persistent=persistent) as sv:
            tickets = []

            def submitter():
                t0 = sv.clock()
                for arrival, d, m, _, st in mix:
                    dt = arrival - (sv.clock() - t0)
                    if dt > 0:
                        time.sleep(dt)
                    tickets.append(sv.submit(d, m, st))

            th = threading.Thread(target=submitter)
            t_start = time.perf_counter()
            th.start()
            th.join()
            sv.drain(-1)
            wall = time.perf_counter() - t_start
            st = sv.stats()
            finite = all(np.isfinite(sv.read_macro(t)).all() for t in tickets[:: max(1, len(tickets) // 16)])
        return wall, st, finite

    wall, st, finite = serve_run(0)
    print("serve: batched stream done in %.2f s" % wall, file=sys.stderr, flush=True)

    def one_at_a_time(storage):
        t = time.perf_counter()
        for _, d, m, _, nsteps in mix:
            with pylbm.Solve(dict(d, autotune=0, storage=storage)) as s:
                s.upload_macro(m)
                s.step(nsteps)
                s.read_macro()
                s.read_forces(nsteps)
        return time.perf_counter() - t

    # the chunk graphs' own device time (HIP events in the serve loop) and
    # the algorithmic bytes they moved: the batched kernel's bandwidth with
    # the launch gaps inside a chunk included and the host work between
    # chunks left out
    device = {"device_seconds": st["device_seconds"], "device_bytes": st["device_bytes"],
              "device_gbs": st["device_bytes"] / st["device_seconds"] / 1e9 if st["device_seconds"] else None,
              "device_busy_frac": st["device_seconds"] / wall,
              "graph_instantiations": st["graph_instantiations"], "graph_updates": st["graph_updates"]}
    served = {"requests": count, "arrival_rate_per_s": rate, "last_arrival_s": mix[-1][0],
              "served_wall_s": wall, "served_solves_per_s": count / wall, "served_mlups": updates / wall / 1e6,
              "p50_latency_s": st["p50_latency"], "p99_latency_s": st["p99_latency"],
              "peak_resident": st["peak_resident"], "peak_pool_bytes": st["peak_resident_bytes"],
              "chunks": st["chunks"], "cell_updates": updates, "finite": bool(finite), "budget_bytes": budget}
    served.update(device)
    if not baseline:
        return served
    seq = one_at_a_time("f16")
    print("serve: one at a time (f16) done in %.2f s" % seq, file=sys.stderr, flush=True)
    seq32 = one_at_a_time("f32")
    return {"requests": count, "arrival_rate_per_s": rate, "last_arrival_s": mix[-1][0],
            "served_wall_s": wall, "served_solves_per_s": count / wall, "served_mlups": updates / wall / 1e6,
            "p50_latency_s": st["p50_latency"], "p99_latency_s": st["p99_latency"],
            "peak_resident": st["peak_resident"], "peak_pool_bytes": st["peak_resident_bytes"],
            "chunks": st["chunks"], "one_at_a_time_wall_s": seq, "one_at_a_time_solves_per_s": count / seq,
            "one_at_a_time_mlups": updates / seq / 1e6, "speedup_solves_per_s": seq / wall,
            "one_at_a_time_f32_wall_s": seq32, "speedup_vs_one_at_a_time_f32": seq32 / wall,
            "cell_updates": updates, "finite": bool(finite), "budget_bytes": budget, **device}


def measure_persistent_chunks(n=192, steps=128, chunk=64):
    """The serve loop's two chunk forms on one resident request (a shifted-fp16
    boxed sphere of n^3 cells, nothing else queued): a captured graph of
    `chunk` batched launches, and one cooperative persistent launch that
    steps every tile `chunk` times with a grid barrier between steps.  Device
    time per step from the serve loop's HIP events around each chunk."""
    desc, mac, _ = sc.boxed_sphere(n=n, re=200.0, storage="f16")
    cells = n ** 3
    out = {"workload": "D3Q19 f16 boxed sphere %d^3, %d steps, %d-step chunks" % (n, steps, chunk)}
    for persistent in (0, 1):
        print("persistent chunks: mode %d start" % persistent, file=sys.stderr, flush=True)
        with pylbm.Server(budget_bytes=8 << 30, chunk_steps=chunk, persistent=persistent) as sv:
            t = sv.submit(desc, mac, steps)
            r = sv.wait(t, 120)
            st = sv.stats()
            mac_out = sv.read_macro(t)
        print("persistent chunks: mode %d done, %.3f s device" % (persistent, st["device_seconds"]),
              file=sys.stderr, flush=True)
        key = "persistent" if persistent else "graph"
        out[key] = {"status": r["status"], "device_ms_per_step": st["device_seconds"] / steps * 1e3,
                    "device_gbs": st["device_bytes"] / st["device_seconds"] / 1e9 if st["device_seconds"] else None,
                    "mlups": cells * steps / st["device_seconds"] / 1e6 if st["device_seconds"] else None,
                    "chunks": st["chunks"]}
        out[key + "_mac"] = mac_out
    a, b = out.pop("graph_mac"), out.pop("persistent_mac")
    out["bit_identical"] = bool(np.array_equal(a.view(np.uint32), b.view(np.uint32)))
    g, p = out["graph"]["device_ms_per_step"], out["persistent"]["device_ms_per_step"]
    # the persistent launch's extra time per step is its grid barrier
    out["barrier_us_per_step"] = (p - g) * 1e3 if p and g else None
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-16bit", action="store_true")
    ap.add_argument("--amr-steps", type=int, default=2000)
    ap.add_argument("--no-amr", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=256)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-serve-baseline", action="store_true",
                    help="serve leg without the one-at-a-time comparison runs (for profiling the served path)")
    ap.add_argument("--only-serve", action="store_true",
                    help="run the serve leg alone and print its object (profiling; not the benchmark line)")
    ap.add_argument("--no-amr-interp", action="store_true",
                    help="AMR leg with the explode/coalesce interface only")
    ap.add_argument("--no-towed", action="store_true", help="skip the moving-body (towed sphere) leg")
    ap.add_argument("--turbulence", action="store_true",
                    help="also run the 512^3 decaying-turbulence precision study (several minutes)")
    ap.add_argument("--tile", default=None,
                    help="fixed tile bx,by,flags (flags bit0 XCD swizzle, bit1 non-temporal, bit2 linear, bits 4-7 "
                         "waves-per-SIMD cap) "
                         "instead of autotuning")
    args = ap.parse_args()
    tile = tuple(int(v) for v in args.tile.split(",")) if args.tile else None
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: the dense solve is single-device here "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.only_serve:
        r = measure_serve(args.serve_requests, baseline=not args.no_serve_baseline)
        print(json.dumps({"serve_config5": r}), flush=True)
        print(json.dumps({"serve_chunk_forms": measure_persistent_chunks()}), flush=True)
        return

    dev = pylbm.device_info()
    print("device: %s, %d CUs, %.0f GB" % (dev["name"], dev["compute_units"], dev["hbm_bytes"] / 1e9),
          file=sys.stderr, flush=True)
    copy_gbs, copy_ms = pylbm.copy_bandwidth(1 << 30, 20)
    print("copy probe: %.1f GB/s" % copy_gbs, file=sys.stderr, flush=True)
    main_run = measure("f32", args.steps, args.warmup, 50, tile=tile)
    print("fp32: %s" % json.dumps(main_run), file=sys.stderr, flush=True)
    extra = {}
    if not args.no_16bit:
        per = measure("f32", min(args.steps, 1000), args.warmup, 50, periodic=True)
        print("fp32 periodic: %s" % json.dumps(per), file=sys.stderr, flush=True)
        extra["periodic_f32"] = per
        per["speedup_vs_f32"] = per["mlups"] / main_run["mlups"]
    if not args.no_16bit:
        for st in ("f16", "bf16"):
            r = measure(st, args.steps, args.warmup, 50)
            r["speedup_vs_f32"] = r["mlups"] / main_run["mlups"]
            extra[st + "_store"] = r
            print("%s: %s" % (st, json.dumps(r)), file=sys.stderr, flush=True)

    drag64 = None
    if not args.no_16bit and not args.no_cpu:
        drag64 = measure_drag_vs_fp64(min(16, os.cpu_count() or 1))
        print("drag vs FP64: %s" % json.dumps(drag64), file=sys.stderr, flush=True)

    towed = plate = None
    if not args.no_towed:
        towed = measure_towed(args.steps, args.warmup, 50)
        print("towed sphere: %s" % json.dumps(towed), file=sys.stderr, flush=True)
        plate = measure_plate()
        print("plate: %s" % json.dumps(plate), file=sys.stderr, flush=True)

    extra_amr_interp = None
    if not args.no_amr:
        r = measure_amr(args.amr_steps)
        extra_amr = r
        print("amr: %s" % json.dumps(r), file=sys.stderr, flush=True)
        extra_amr2d = measure_amr2d(args.amr_steps)
        print("amr 2d: %s" % json.dumps(extra_amr2d), file=sys.stderr, flush=True)
        if not args.no_amr_interp:
            extra_amr_interp = {"dynamic": measure_amr(args.amr_steps, interface="interp", reflux=0)}
            print("amr interp: %s" % json.dumps(extra_amr_interp["dynamic"]), file=sys.stderr, flush=True)
            for rf in (0, 1):
                r = measure_amr_slab(min(args.amr_steps, 500), rf)
                extra_amr_interp["slab_reflux" if rf else "slab_no_reflux"] = r
                print("amr interp slab reflux=%d: %s" % (rf, json.dumps(r)), file=sys.stderr, flush=True)
    else:
        extra_amr = extra_amr2d = None

    extra_serve = None
    if not args.no_serve:
        extra_serve = measure_serve(args.serve_requests, baseline=not args.no_serve_baseline)
        print("serve: %s" % json.dumps(extra_serve), file=sys.stderr, flush=True)

    extra_turb = None
    extra_persistent = None
    if not args.no_serve:
        extra_persistent = measure_persistent_chunks()
    if args.turbulence:
        extra_turb = measure_turbulence()
        print("turbulence: %s" % json.dumps(extra_turb), file=sys.stderr, flush=True)

    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        # the box gives one GPU's job 16 cores, whatever os.cpu_count() shows
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(args.cpu_steps, threads)
        print("cpu: %s" % json.dumps(cpu), file=sys.stderr, flush=True)

    achieved = main_run["achieved_gbs"]
    out = {
        "metric": METRIC,
        "value": round(main_run["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_run["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded freestream initial field)",
        "config": {"workload": "D3Q19 TRT flow past a pinned sphere, 256x128x128, R=12, Re=200, "
                               "sponge 12, per-step FP64 drag/lift, fp32 storage",
                   "cells": main_run["ncells"], "tile": main_run["tile"]},
        "pct_hbm_peak": round(100.0 * achieved / HBM_PEAK_GBS, 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_step<19, StoreF32, general>", "kernel_ms": round(main_run["kernel_ms"], 5),
                     "bytes_per_launch": main_run["bytes_per_launch"]},
        "cpu_baseline": cpu,
        "drag_cd_last1000": main_run["cd_last1000"],
        "copy_probe_gbs": round(copy_gbs, 1),
        "frac_of_copy_probe": round(achieved / copy_gbs, 4),
    }
    for k, v in extra.items():
        out[k] = {"mlups": round(v["mlups"], 1), "speedup_vs_f32": round(v["speedup_vs_f32"], 3),
                  "kernel_speedup_vs_f32": round(main_run["kernel_ms"] / v["kernel_ms"], 3),
                  "kernel_ms": round(v["kernel_ms"], 5), "achieved_gbs": round(v["achieved_gbs"], 1),
                  "pct_hbm_peak": round(100.0 * v["achieved_gbs"] / HBM_PEAK_GBS, 2),
                  "cd_last1000": v["cd_last1000"],
                  "drag_dev_vs_f32_pct": (100.0 * (v["cd_last1000"] / main_run["cd_last1000"] - 1.0)
                                          if v["cd_last1000"] is not None and main_run["cd_last1000"] else None)}
    if drag64:
        out["drag_vs_fp64_sphere_d3q19_96x48x48"] = drag64
    if towed:
        out["towed_sphere_d3q19_256x128x128"] = {
            "mlups": round(towed["mlups"], 1), "ms_per_step": round(towed["ms_per_step"], 5),
            "overhead_vs_pinned": round(towed["ms_per_step"] / main_run["ms_per_step"] - 1.0, 4),
            "step_kernel_ms": round(towed["step_kernel_ms"], 5), "re": towed["re"], "tau": round(towed["tau"], 4),
            "cd_towed_last1000": towed["cd_last1000"], "cd_pinned_last1000": towed["cd_pinned_last1000"],
            "cd_towed_vs_pinned": round(towed["cd_last1000"] / towed["cd_pinned_last1000"] - 1.0, 4),
            "travel_cells": round(towed["travel_cells"], 3), "finite": towed["finite"]}
    if plate:
        out["reference_body_cases"] = {k: (round(v, 5) if isinstance(v, float) else v) for k, v in plate.items()}
    if extra_turb:
        out["turbulence_d3q19_512"] = extra_turb
    if extra_serve:
        out["serve_config5"] = {k: (float("%.5g" % v) if isinstance(v, float) else v) for k, v in extra_serve.items()}
    if extra_persistent:
        out["serve_chunk_forms"] = {k: ({kk: (float("%.5g" % vv) if isinstance(vv, float) else vv)
                                         for kk, vv in v.items()} if isinstance(v, dict) else
                                        (float("%.5g" % v) if isinstance(v, float) else v))
                                    for k, v in extra_persistent.items()}
    def short(d):
        return {k: (float("%.5g" % v) if isinstance(v, float) else v) for k, v in d.items()}

    if extra_amr:
        out["amr_d3q27_vortex_rings_128"] = short(extra_amr)
    if extra_amr2d:
        out["amr_d2q9_shear_layer_1024"] = short(extra_amr2d)
    if extra_amr_interp:
        out["amr_d3q27_vortex_rings_128_interp_interface"] = {k: short(v) for k, v in extra_amr_interp.items()}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
