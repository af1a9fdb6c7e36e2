"""GPU tests of the batched serve loop (lbm_serve_* through the C ABI).

- a request's result is bit-for-bit the same alone and in a mix of uneven
  2D/3D, fp32/fp16 requests that arrive while others run;
- a served solve equals the plain single-solve path bit for bit (same tile
  shape, same rounding, same force summation order), so the serve loop
  inherits the single-solve parity with the FP64 oracle;
- admission respects the byte budget and a request larger than the budget
  is refused.
"""
import os
import sys
import threading
import time

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "getting-air_amd"))

import pylbm  # noqa: E402
from pylbm import scenarios as sc  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def mix():
    return [
        sc.cylinder2d(nx=256, re=120.0, storage="f16"),
        sc.boxed_sphere(n=48, re=80.0, storage="f16"),
        sc.cylinder2d(nx=320, re=60.0, storage="f32"),
        sc.boxed_sphere(n=32, re=150.0, storage="f32"),
        sc.cylinder2d(nx=192, re=300.0, storage="bf16"),
        sc.boxed_sphere(n=64, re=200.0, storage="f16"),
    ]


def test_alone_equals_in_the_mix_bit_for_bit():
    reqs = mix()
    target = 1  # the 48^3 f16 sphere
    desc, mac, _ = reqs[target]
    say("alone")
    with pylbm.Server(budget_bytes=2 << 30, chunk_steps=24) as sv:
        t = sv.submit(desc, mac, 150)
        sv.wait(t, 120)
        alone_mac, alone_f = sv.read_macro(t), sv.read_forces(t)
    say("mix")
    with pylbm.Server(budget_bytes=2 << 30, chunk_steps=24) as sv:
        tickets = {}
        # three up front, the rest while those run (they join mid-flight)
        for k in (0, 2, 3):
            tickets[k] = sv.submit(reqs[k][0], reqs[k][1], 100 + 10 * k)
        time.sleep(0.05)
        for k in (1, 4, 5):
            tickets[k] = sv.submit(reqs[k][0], reqs[k][1], 150 if k == target else 90)
        sv.drain(300)
        mix_mac, mix_f = sv.read_macro(tickets[target]), sv.read_forces(tickets[target])
        st = sv.stats()
        for k, t in tickets.items():
            r = sv.wait(t, 1)
            assert r["status"] == 0 and np.isfinite(sv.read_macro(t)).all(), k
    say("stats", {k: st[k] for k in ("completed", "peak_resident", "chunks", "p50_latency", "p99_latency")})
    assert st["completed"] == 6 and st["peak_resident"] >= 3
    assert np.array_equal(alone_mac.view(np.uint32), mix_mac.view(np.uint32))
    assert alone_f.shape == (150, 6) and np.array_equal(alone_f, mix_f)


@pytest.mark.parametrize("k", [0, 1, 2, 5])
def test_served_equals_the_plain_solve(k):
    desc, mac, _ = mix()[k]
    steps = 70
    with pylbm.Server(budget_bytes=2 << 30, chunk_steps=32) as sv:
        t = sv.submit(desc, mac, steps)
        sv.wait(t, 120)
        smac, sf = sv.read_macro(t), sv.read_forces(t)
    d = dict(desc, autotune=0)
    with pylbm.Solve(d) as s:
        s.upload_macro(mac)
        s.step(steps)
        pmac, pf = s.read_macro(), s.read_forces(steps)
    say(k, desc["q"], desc.get("storage"), "max |d mac|", float(np.abs(smac - pmac).max()))
    assert np.array_equal(smac.view(np.uint32), pmac.view(np.uint32))
    assert np.array_equal(sf, pf)


def test_persistent_chunks_equal_graph_chunks():
    """The persistent form of a chunk (one cooperative launch per lattice
    class, a grid barrier between steps) gives every request of the uneven
    mix the same field and force history, bit for bit, as the captured graph
    of batched launches; requests join while chunks run in both."""
    reqs = mix()
    steps = [100, 150, 120, 130, 90, 150]
    out = {}
    for persistent in (0, 1):
        say("persistent", persistent)
        with pylbm.Server(budget_bytes=2 << 30, chunk_steps=24, persistent=persistent) as sv:
            tickets = {}
            for k in (0, 2, 3):
                tickets[k] = sv.submit(reqs[k][0], reqs[k][1], steps[k])
            time.sleep(0.05)
            for k in (1, 4, 5):
                tickets[k] = sv.submit(reqs[k][0], reqs[k][1], steps[k])
            sv.drain(300)
            st = sv.stats()
            res = {}
            for k, t in tickets.items():
                assert sv.wait(t, 1)["status"] == 0, (persistent, k)
                res[k] = (sv.read_macro(t), sv.read_forces(t))
        say("stats", {k: st[k] for k in ("completed", "chunks", "device_seconds", "graph_instantiations")})
        assert st["completed"] == 6
        if persistent:
            assert st["graph_instantiations"] == 0 and st["graph_updates"] == 0
        out[persistent] = res
    for k in range(6):
        (a, fa), (b, fb) = out[0][k], out[1][k]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k
        assert np.array_equal(fa, fb), k


def test_budget_admission_and_refusal():
    reqs = [sc.boxed_sphere(n=48, re=100.0, storage="f32") for _ in range(5)]
    # the footprint as the library computes it, from the refusal of a budget
    # that is too small; then a budget that holds two of these solves at once
    with pylbm.Server(budget_bytes=1 << 20) as sv:
        with pytest.raises(pylbm.LbmError) as e:
            sv.submit(reqs[0][0], reqs[0][1], 10)
        assert e.value.code == -3 and "budget" in str(e.value)
        foot = int(str(e.value).split("needs ")[1].split(" ")[0])
    budget = int(2.5 * foot)
    with pylbm.Server(budget_bytes=budget, chunk_steps=16) as sv:
        ts = [sv.submit(d, m, 40) for d, m, _ in reqs]
        sv.drain(300)
        st = sv.stats()
        for t in ts:
            assert sv.wait(t, 1)["status"] == 0
    say("footprint", foot, "budget", budget, "peak resident", st["peak_resident"], st["peak_resident_bytes"])
    assert st["peak_resident"] == 2
    assert st["peak_resident_bytes"] <= budget
    assert st["completed"] == 5 and st["p99_latency"] >= st["p50_latency"] > 0


def test_alone_equals_mix_all_submitted_at_once():
    # the JS host test's mix: a small f32 TRT sphere, a D2Q9 TGV and a second
    # f32 sphere, all submitted before the first chunk
    sphere = sc.sphere(nx=48, ny=24, nz=24, radius=4.0, centre=(14.0, 12.0, 12.0), re=80.0)
    tgv = sc.tgv2d(n=64)
    sphere2 = sc.sphere(nx=32, ny=32, nz=32, radius=3.0, centre=(10.0, 16.0, 16.0), re=60.0)
    with pylbm.Server(budget_bytes=1 << 30, chunk_steps=16) as sv:
        t = sv.submit(sphere[0], sphere[1], 60)
        sv.wait(t, 120)
        a = sv.read_macro(t)
    with pylbm.Server(budget_bytes=1 << 30, chunk_steps=16) as sv:
        sv.submit(tgv[0], tgv[1], 50)
        t = sv.submit(sphere[0], sphere[1], 60)
        sv.submit(sphere2[0], sphere2[1], 40)
        sv.drain(120)
        b = sv.read_macro(t)
    diff = np.nonzero(a.view(np.uint32) != b.view(np.uint32))[0]
    say("differing", diff.size, "first", diff[:8], "max", float(np.abs(a - b).max()) if diff.size else 0.0)
    assert diff.size == 0


def test_joining_while_a_chunk_runs():
    # a request admitted while another request's chunk is in flight takes no
    # steps in that chunk: it must still run exactly its own step count
    sphere = sc.sphere(nx=48, ny=24, nz=24, radius=4.0, centre=(14.0, 12.0, 12.0), re=80.0)
    long_run = sc.tgv2d(n=256)
    steps = 60
    with pylbm.Server(budget_bytes=1 << 30, chunk_steps=8) as sv:
        tl = sv.submit(long_run[0], long_run[1], 4000)
        t_end = time.time() + 60
        while sv.stats()["chunks"] < 2 and time.time() < t_end:
            time.sleep(0.001)
        t = sv.submit(sphere[0], sphere[1], steps)
        r = sv.wait(t, 120)
        served = sv.read_macro(t)
        rl = sv.wait(tl, 120)
    say("target admitted", r["t_admit"], "long request finished", rl["t_finish"])
    assert r["t_admit"] < rl["t_finish"], "the target did not overlap a running chunk"
    d = dict(sphere[0], autotune=0)
    with pylbm.Solve(d) as s:
        s.upload_macro(sphere[1])
        s.step(steps)
        plain = s.read_macro()
    diff = np.nonzero(served.view(np.uint32) != plain.view(np.uint32))[0]
    say("differing", diff.size, "max", float(np.abs(served - plain).max()) if diff.size else 0.0)
    assert diff.size == 0


def test_concurrent_submitters():
    reqs = mix()
    results = {}
    with pylbm.Server(budget_bytes=2 << 30, chunk_steps=16) as sv:
        def client(k):
            d, m, _ = reqs[k]
            t = sv.submit(d, m, 50)
            results[k] = sv.wait(t, 300)
        th = [threading.Thread(target=client, args=(k,)) for k in range(len(reqs))]
        for x in th:
            x.start()
        for x in th:
            x.join()
        st = sv.stats()
    assert len(results) == len(reqs) and all(r["status"] == 0 for r in results.values())
    assert st["completed"] == len(reqs)


def test_config5_mix_at_its_sizes_served_equals_plain_and_matches_the_oracle():
    """Config 5's request shapes (fp16 storage, fp32 collision): 2D cylinder
    channels from nx 256 to 1024 and 3D boxed spheres from 64^3 to 192^3, Re
    from the seeded mix, submitted in a staggered stream with uneven step
    counts.  Every served result equals the plain single-solve path bit for
    bit (field and force history); the two smallest members are also scored
    against the FP64 oracle with the shifted-fp16 bounds of the dense tests
    (u within 3e-2, mean drag of the last 20 steps within 5e-3)."""
    import oracle_runner

    mixed = sc.serve_mix(count=4, seed=21, steps=0, storage="f16")
    reqs = [sc.cylinder2d(nx=256, re=90.0, storage="f16"), sc.cylinder2d(nx=1024, re=400.0, storage="f16"),
            sc.boxed_sphere(n=64, re=120.0, storage="f16"), sc.boxed_sphere(n=192, re=350.0, storage="f16")]
    reqs += [(d, m, meta) for _, d, m, meta, _ in mixed]
    steps = [160, 96, 200, 64, 120, 80, 144, 112]
    say("mix:", [(d["q"], d["nx"], d["ny"], d.get("nz", 1), round(meta["re"])) for d, _, meta in reqs])
    served = {}
    with pylbm.Server(budget_bytes=16 << 30, chunk_steps=32) as sv:
        tickets = []
        for k, (d, m, _) in enumerate(reqs):
            tickets.append(sv.submit(d, m, steps[k]))
            time.sleep(0.01)
        sv.drain(600)
        st = sv.stats()
        for k, t in enumerate(tickets):
            assert sv.wait(t, 1)["status"] == 0, k
            served[k] = (sv.read_macro(t), sv.read_forces(t))
    say("stats", {k: st[k] for k in ("completed", "peak_resident", "chunks", "p50_latency", "p99_latency", "mlups",
                                     "device_seconds", "device_bytes", "graph_instantiations", "graph_updates")})
    assert st["completed"] == len(reqs) and st["peak_resident"] >= 2
    # the chunk graphs' algorithmic bytes are every member's steps x 2 Q s
    cells = [d["nx"] * d["ny"] * (1 if d["q"] == 9 else d["nz"]) for d, _, _ in reqs]
    want_bytes = sum(c * n * 2 * d["q"] * 2 for c, n, (d, _, _) in zip(cells, steps, reqs))
    assert st["device_bytes"] == want_bytes and st["device_seconds"] > 0
    assert st["graph_instantiations"] >= 1 and st["graph_instantiations"] + st["graph_updates"] >= 2
    for k, (d, m, meta) in enumerate(reqs):
        with pylbm.Solve(dict(d, autotune=0)) as s:
            s.upload_macro(m)
            s.step(steps[k])
            pmac, pf = s.read_macro(), s.read_forces(steps[k])
        smac, sf = served[k]
        say(k, "served == plain:", np.array_equal(smac.view(np.uint32), pmac.view(np.uint32)), np.array_equal(sf, pf))
        assert np.array_equal(smac.view(np.uint32), pmac.view(np.uint32)), k
        assert np.array_equal(sf, pf), k
    for k in (0, 2):
        d, m, meta = reqs[k]
        r = oracle_runner.run(dict(d, storage="f32"), m, steps=steps[k], threads=8, want=("mac", "force"))
        g = served[k][0].reshape(-1, 4).astype(np.float64)
        eu = float(np.linalg.norm(g[:, 1:] - r["mac"][:, 1:]) / np.linalg.norm(r["mac"][:, 1:]))
        drag = abs(served[k][1][-20:, 0].mean() / r["force"][-20:, 0].mean() - 1.0)
        say(k, "vs FP64 oracle: u %.3e, drag %.3e" % (eu, drag))
        assert eu < 3e-2 and drag < 5e-3, (k, eu, drag)
