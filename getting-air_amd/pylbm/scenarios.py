"""Workload definitions: solve descriptors plus initial macroscopic fields.

Each function returns (desc, mac) where desc is the descriptor dict accepted
by pylbm.Solve (and by the oracle) and mac is float32 [rho, ux, uy, uz] per
cell, cell = (z * ny + y) * nx + x.
"""
import math

import numpy as np

CS2 = 1.0 / 3.0


def f32(x):
    return float(np.float32(x))


def trt_omega_minus(omega, lam=3.0 / 16.0):
    """Odd TRT rate for a magic parameter Lambda (plans/TRT.md:129-133):
    tau_minus = 1/2 + Lambda / (tau_plus - 1/2)."""
    tau = 1.0 / f32(omega)
    return 1.0 / (0.5 + lam / (tau - 0.5))


def tgv2d(n=256, tau=0.6, u0=0.05):
    """D2Q9 Taylor-Green vortex, periodic n x n:
    u = u0 (sin kx cos ky, -cos kx sin ky), p = (u0^2/4)(cos 2kx + cos 2ky),
    rho = 1 + p / cs^2 (the field of main-tgv.js:69-90 with this sign
    convention)."""
    k = 2.0 * math.pi / n
    x = np.arange(n, dtype=np.float64)[None, :]
    y = np.arange(n, dtype=np.float64)[:, None]
    ux = u0 * np.sin(k * x) * np.cos(k * y)
    uy = -u0 * np.cos(k * x) * np.sin(k * y)
    p = (u0 * u0 / 4.0) * (np.cos(2 * k * x) + np.cos(2 * k * y))
    mac = np.zeros((n, n, 4), dtype=np.float32)
    mac[..., 0] = 1.0 + p / CS2
    mac[..., 1] = ux
    mac[..., 2] = uy
    desc = {"q": 9, "nx": n, "ny": n, "nz": 1, "omega": f32(1.0 / tau)}
    return desc, mac.reshape(-1)


def tgv2d_analytic(n, tau, u0, t):
    """Analytic TGV velocity at step t (decay exp(-2 nu k^2 t))."""
    nu = (tau - 0.5) / 3.0
    k = 2.0 * math.pi / n
    d = math.exp(-2.0 * nu * k * k * t)
    x = np.arange(n, dtype=np.float64)[None, :]
    y = np.arange(n, dtype=np.float64)[:, None]
    return (u0 * np.sin(k * x) * np.cos(k * y) * d).reshape(-1), (-u0 * np.cos(k * x) * np.sin(k * y) * d).reshape(-1)


def sphere(nx=256, ny=128, nz=128, radius=12.0, centre=(64.0, 64.0, 64.0), u0=0.05, re=200.0,
           trt=True, sponge_width=None, storage="f32"):
    """D3Q19 flow past a pinned sphere with halfway bounce-back and an ALBC
    sponge on every face relaxing toward the freestream (SCENARIOS.sphere,
    d3-scenarios.mjs:350-391): nu = u0 D / Re, tau = 3 nu + 1/2, sponge width
    max(6, round(D/2)), freestream seeded everywhere with
    rho = 1 - |u|^2 / (2 cs^2)."""
    d_ = 2.0 * radius
    nu = u0 * d_ / re
    tau = 3.0 * nu + 0.5
    omega = f32(1.0 / tau)
    if sponge_width is None:
        sponge_width = max(6, int(round(d_ / 2.0)))
    desc = {
        "q": 19, "nx": nx, "ny": ny, "nz": nz, "omega": omega,
        "sponge": {"width": float(sponge_width), "axes": [1, 1, 1], "u": [u0, 0.0, 0.0]},
        "body": {"a": float(radius), "centre": [float(c) for c in centre]},
        "solidEq": 1, "storage": storage,
    }
    if trt:
        desc["omegaMinus"] = trt_omega_minus(omega)
    n = nx * ny * nz
    mac = np.empty((n, 4), dtype=np.float32)
    mac[:, 0] = 1.0 - (u0 * u0) / (2.0 * CS2)
    mac[:, 1] = u0
    mac[:, 2] = 0.0
    mac[:, 3] = 0.0
    meta = {"area": math.pi * radius * radius, "u0": u0, "nu": nu, "tau": tau, "re": re}
    return desc, mac.reshape(-1), meta


SQRT3_UP = 1.7320509  # d3-amr.mjs:120, sqrt(3) rounded up
BLOCK_BB_DEPTH = 3    # d3-amr.mjs:125


def near_body_want(sdf, margin):
    """Whether any point of a block's box [lo, hi) comes within `margin` of
    the body (nearBodyWant, d3-amr.mjs:139-159): the SDF is 1-Lipschitz, so
    phi(centre) - h sqrt(3) bounds the box minimum; an ambiguous box splits
    into octants down to depth 3 and is accepted at the bottom."""
    def near(c, h, depth):
        phi = sdf(c)
        if phi - h * SQRT3_UP > margin:
            return False
        if phi <= margin or depth == 0:
            return True
        q = h / 2
        return any(near([c[0] + (q if k & 1 else -q), c[1] + (q if k & 2 else -q), c[2] + (q if k & 4 else -q)],
                        q, depth - 1) for k in range(8))
    return lambda lo, hi: near([(lo[k] + hi[k]) / 2 for k in range(3)], (hi[0] - lo[0]) / 2, BLOCK_BB_DEPTH)


def refine_near_sphere(nx, ny, nz, rb, centre, radius, margin=2.0, max_slots=None):
    """refineNearBody with a sphere's SDF (d3-amr.mjs:80-106): blocks in
    ascending id order take slots 0, 1, ...; returns (block_slot,
    slot_to_block) sized for max_slots (default: the count)."""
    want = near_body_want(lambda c: math.sqrt(sum((c[k] - centre[k]) ** 2 for k in range(3))) - radius, margin)
    nbx, nby, nbz = nx // rb, ny // rb, nz // rb
    blocks = []
    for b in range(nbx * nby * nbz):
        bx, by, bz = b % nbx, (b // nbx) % nby, b // (nbx * nby)
        lo = [bx * rb, by * rb, bz * rb]
        if want(lo, [v + rb for v in lo]):
            blocks.append(b)
    slots = len(blocks) if max_slots is None else max_slots
    if len(blocks) > slots:
        raise ValueError("refinement needs %d tiles, max_slots %d" % (len(blocks), slots))
    bs = np.full(nbx * nby * nbz, -1, np.int32)
    sb = np.full(slots, -1, np.int32)
    for k, b in enumerate(blocks):
        bs[b], sb[k] = k, b
    return bs, sb


def sphere_hierarchy(nx, ny, nz, rb, levels, centre, radius, margin=2.0):
    """refineHierarchy for a sphere (d3-amr.mjs:603-651): the finest level
    wants every block within `margin` L0 cells (nearBodyWant) and cascade21
    closes the coarser ones.  Runs the package's JS module under Node
    (js/hierarchy-cli.mjs).  Returns [None, (block_slot, slot_to_block), ...]
    per level, every slot array padded with -1 to the largest level's count."""
    import json
    import os
    import subprocess
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "js", "hierarchy-cli.mjs")
    arg = json.dumps({"dims": [nx, ny, nz], "rb": rb, "levels": levels, "centre": list(map(float, centre)),
                      "radius": float(radius), "margin": float(margin)})
    out = json.loads(subprocess.run(["node", cli, arg], check=True, capture_output=True, text=True).stdout)
    slots = max(lv["activeSlots"] for lv in out["byLevel"][1:])
    maps = [None]
    for lv in out["byLevel"][1:]:
        sb = np.full(slots, -1, np.int32)
        sb[:lv["activeSlots"]] = lv["slotToBlock"]
        maps.append((np.array(lv["blockSlot"], np.int32), sb))
    return maps


def sphere_amr(nx=192, ny=128, nz=128, radius=8.0, centre=(48.0, 64.0, 64.0), u0=0.05, re=20.0, rb=4,
               sponge_width=8, margin=2.0, levels=2):
    """The pinned sphere of `sphere` (BGK on L0) on `levels` levels: the body
    on the finest level alone, the blocks within `margin` L0 cells of it
    refined at the finest level and the coarser levels closed around them
    (main-3d.js:482-530, ?refine=body, margin 2; refineHierarchy at depth),
    explode interface, static topology.  Returns (AMR descriptor dict, L0
    seed, topology, meta) with topology (block_slot, slot_to_block) for two
    levels and [None, (block_slot, slot_to_block) per level ...] for more."""
    desc, mac, meta = sphere(nx=nx, ny=ny, nz=nz, radius=radius, centre=centre, u0=u0, re=re, trt=False,
                             sponge_width=sponge_width)
    if levels == 2:
        topo = refine_near_sphere(nx, ny, nz, rb, centre, radius, margin)
        slots = len(topo[1])
    else:
        topo = sphere_hierarchy(nx, ny, nz, rb, levels, centre, radius, margin)
        slots = len(topo[1][1])
    amr = {"q": 19, "nx": nx, "ny": ny, "nz": nz, "rb": rb, "maxSlots": slots, "omega": desc["omega"],
           "sponge": desc["sponge"], "body": desc["body"], "levels": levels}
    return amr, mac, topo, meta


def plate(n=16, re=200.0, u_t=0.05, aspect=0.125, tow=0.0, stream=-0.05, i_star=0.17):
    """The blunt square plate normal to the flow with prescribed relative
    motion (SCENARIOS.card with tow or stream, span 1, tilt 0;
    d3-scenarios.mjs:703-903): chord n, thickness aspect * n, Re on the chord,
    a rounded box with r = 0 turned by 120 degrees about (1, 1, 1) so that its
    thickness lies along x, placed with its faces on cell faces (the centre on
    a cell corner or centre as the half-extent is integer or not), in an 8n
    cube with a sponge of width max(6, n/2) toward `stream`.  `tow` moves the
    body (force measured, never applied; moving window on x anchored at the
    start), `stream` moves the fluid; only stream - tow is physical.  BGK.
    Returns (desc, mac, rigid, meta); meta["sign"] * F_x / (1/2 U^2 A) is Cd."""
    a = n / 2.0
    b = n / 2.0
    c = aspect * n / 2.0
    nu = 2.0 * u_t * a / re
    tau = nu / CS2 + 0.5
    rho_b = max(1.05, i_star * 2.0 * a ** 3 / (c * (a * a + c * c)))
    d = [8 * n, 8 * n, int(round(max(8 * n, 2 * n + 6 * n)))]
    towed, streamed = tow != 0.0, stream != 0.0
    rel_sign = math.copysign(1.0, tow - stream)
    x0 = d[0] - 2 * n if rel_sign >= 0 else 2 * n

    def snap(centre, h):
        return round(centre) + 0.5 if float(h).is_integer() else float(round(centre))

    place = [snap(x0, c), snap(d[1] / 2.0, a), snap(d[2] / 2.0, b)]
    rigid = {"shape": {"kind": 2, "a": a, "b": b, "c": c, "r": 0.0}, "x": place, "q": [0.5, 0.5, 0.5, 0.5],
             "v": [tow, 0.0, 0.0], "density": rho_b, "pinned": streamed and not towed, "noFluidForce": towed,
             "window": [1, 0, 0] if towed else [0, 0, 0], "anchor": [math.floor(v) for v in place]}
    desc = {"q": 19, "nx": d[0], "ny": d[1], "nz": d[2], "omega": f32(1.0 / tau),
            "sponge": {"width": float(max(6, round(n / 2.0))), "axes": [1, 1, 1], "u": [stream, 0.0, 0.0]},
            "solidEq": 1}
    ncell = d[0] * d[1] * d[2]
    mac = np.empty((ncell, 4), dtype=np.float32)
    mac[:, 0] = 1.0 - stream * stream / (2.0 * CS2)
    mac[:, 1] = stream
    mac[:, 2] = 0.0
    mac[:, 3] = 0.0
    u_rel = abs(stream - tow)
    meta = {"tau": tau, "area": 4.0 * a * b, "u_rel": u_rel, "sign": math.copysign(1.0, stream - tow),
            "convective": max(1, round(n / u_rel))}
    return desc, mac.reshape(-1), rigid, meta


def fall(n=12, re=14.4, u_t=0.04, tow=0.0, stream=-0.025, rho_b=4.0):
    """A sphere of diameter n with prescribed relative motion (SCENARIOS.fall
    with tow or stream, d3-scenarios.mjs:422-600): nu = u_t n / Re, a 16n x 6n
    x 6n box, the sphere 2n from the downstream face, sponge of width
    max(6, n/2) toward `stream`, moving window on x for a towed body anchored
    at its start.  BGK.  Returns (desc, mac, rigid, meta) like plate()."""
    nu = u_t * n / re
    tau = nu / CS2 + 0.5
    d = [16 * n, 6 * n, 6 * n]
    towed, streamed = tow != 0.0, stream != 0.0
    rel_sign = math.copysign(1.0, tow - stream)
    x0 = d[0] - 2 * n if rel_sign >= 0 else 2 * n
    place = [float(x0), d[1] / 2.0, d[2] / 2.0]
    rigid = {"shape": {"kind": 0, "a": n / 2.0}, "x": place, "v": [tow, 0.0, 0.0], "density": rho_b,
             "pinned": streamed and not towed, "noFluidForce": towed,
             "window": [0, 0, 0] if (streamed and not towed) else [1, 0, 0],
             "anchor": [math.floor(v) for v in place]}
    desc = {"q": 19, "nx": d[0], "ny": d[1], "nz": d[2], "omega": f32(1.0 / tau),
            "sponge": {"width": float(max(6, round(n / 2.0))), "axes": [1, 1, 1], "u": [stream, 0.0, 0.0]},
            "solidEq": 1}
    ncell = d[0] * d[1] * d[2]
    mac = np.empty((ncell, 4), dtype=np.float32)
    mac[:, 0] = 1.0 - stream * stream / (2.0 * CS2)
    mac[:, 1] = stream
    mac[:, 2] = 0.0
    mac[:, 3] = 0.0
    u_rel = abs(stream - tow)
    meta = {"tau": tau, "area": math.pi * (n / 2.0) ** 2, "u_rel": u_rel, "sign": math.copysign(1.0, stream - tow),
            "convective": max(1, round(n / u_rel))}
    return desc, mac.reshape(-1), rigid, meta


def beltrami(n=32, q=19, tau=0.7, amp=(0.02, 0.015, 0.01)):
    """Periodic ABC (Beltrami) flow, curl u = k u, exact decay
    u(t) = u0 exp(-nu k^2 t) (d3-scenarios.mjs:19-27)."""
    k = 2.0 * math.pi / n
    a, b, c = amp
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ux = a * np.sin(k * z) + c * np.cos(k * y)
    uy = b * np.sin(k * x) + a * np.cos(k * z)
    uz = c * np.sin(k * y) + b * np.cos(k * x)
    mac = np.stack([1.0 - 1.5 * (ux * ux + uy * uy + uz * uz), ux, uy, uz], axis=-1).astype(np.float32)
    desc = {"q": q, "nx": n, "ny": n, "nz": n, "omega": f32(1.0 / tau)}
    return desc, mac.reshape(-1)


def random_field(nx, ny, nz, q=19, seed=1, amp=0.03, tau=0.62, modes=4):
    """Smooth seeded periodic field: each of rho, ux, uy, uz is a sum of a few
    random low-wavenumber Fourier modes (so it does not decay to round-off
    within a short parity run)."""
    rng = np.random.default_rng(seed)
    if q == 9:
        nz = 1
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    comps = []
    for c in range(4):
        v = np.zeros(x.shape)
        for _ in range(modes):
            kx, ky = rng.integers(0, 3, 2)
            kz = 0 if q == 9 else rng.integers(0, 3)
            ph = rng.random() * 2.0 * math.pi
            v += np.cos(2 * math.pi * (kx * x / nx + ky * y / ny + kz * z / nz) + ph)
        comps.append(v / modes)
    mac = np.empty(x.shape + (4,), dtype=np.float32)
    mac[..., 0] = 1.0 + 0.01 * comps[0]
    for j in range(3):
        mac[..., 1 + j] = amp * comps[1 + j]
    if q == 9:
        mac[..., 3] = 0.0
    desc = {"q": q, "nx": nx, "ny": ny, "nz": nz, "omega": f32(1.0 / tau)}
    return desc, mac.reshape(-1)


def turbulence(n=512, seed=2024, urms=0.04, kpeak=8.0, tau=0.55, q=19, storage="f32"):
    """Decaying turbulence in a periodic box: random-phase solenoidal field
    with E(k) ~ k^4 exp(-2 (k/kpeak)^2), scaled to u_rms, rho from the
    incompressible pressure estimate rho = 1 - |u|^2/(2 cs^2).  The reference
    has no such initialiser; this one is seeded and deterministic."""
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(n, d=1.0 / n)
    kr = np.fft.rfftfreq(n, d=1.0 / n)
    KZ, KY, KX = np.meshgrid(kx, kx, kr, indexing="ij")
    k2 = KX * KX + KY * KY + KZ * KZ
    kk = np.sqrt(k2)
    k2[0, 0, 0] = 1.0
    # amplitude per mode so that the shell energy follows E(k): |u_k|^2 ~ E(k) / k^2
    amp = np.sqrt(np.where(kk > 0, kk ** 4 * np.exp(-2.0 * (kk / kpeak) ** 2) / (4.0 * math.pi * k2), 0.0))
    comps = []
    for _ in range(3):
        ph = rng.random(KX.shape) * 2.0 * math.pi
        comps.append(amp * np.exp(1j * ph))
    ux_k, uy_k, uz_k = comps
    # project onto divergence-free modes
    dot = (KX * ux_k + KY * uy_k + KZ * uz_k) / k2
    ux_k = ux_k - KX * dot
    uy_k = uy_k - KY * dot
    uz_k = uz_k - KZ * dot
    u = [np.fft.irfftn(c, s=(n, n, n)) for c in (ux_k, uy_k, uz_k)]
    rms = math.sqrt(sum(float(np.mean(c * c)) for c in u) / 3.0)
    scale = urms / rms
    mac = np.empty((n, n, n, 4), dtype=np.float32)
    usq = np.zeros((n, n, n))
    for j, c in enumerate(u):
        c *= scale
        usq += c * c
        mac[..., 1 + j] = c
    mac[..., 0] = 1.0 - usq / (2.0 * CS2)
    desc = {"q": q, "nx": n, "ny": n, "nz": n, "omega": f32(1.0 / tau), "storage": storage}
    return desc, mac.reshape(-1)


def enstrophy(mac, n, dtype=np.float64):
    """0.5 <|curl u|^2> with central differences on the periodic n^3 box
    (main-3d.js:3627-3630), u from the [rho, ux, uy, uz] field.  dtype
    float32 halves the memory of a 512^3 evaluation (the mean is still
    accumulated in float64)."""
    m = mac.reshape(n, n, n, 4)
    u = [m[..., k].astype(dtype) for k in (1, 2, 3)]  # ux, uy, uz

    def d(c, axis):  # central difference of u[c]; axis 0 = z, 1 = y, 2 = x
        return (np.roll(u[c], -1, axis) - np.roll(u[c], 1, axis)) * dtype(0.5)

    acc = 0.0
    # wx = duz/dy - duy/dz, wy = dux/dz - duz/dx, wz = duy/dx - dux/dy
    for ca, a, cb, b in ((2, 1, 1, 0), (0, 0, 2, 2), (1, 2, 0, 1)):
        w = d(ca, a) - d(cb, b)
        acc += float(np.mean(w * w, dtype=np.float64))
    return 0.5 * acc


def mulberry32(seed):
    """Seeded 32-bit PRNG, the generator the reference seeds its
    perturbations with (d3-scenarios.mjs:207-216); same sequence as the JS
    host's mulberry32."""
    state = [seed & 0xFFFFFFFF]

    def imul(a, b):
        return (a * b) & 0xFFFFFFFF

    def nxt():
        state[0] = (state[0] + 0x6D2B79F5) & 0xFFFFFFFF
        t = state[0]
        t = imul(t ^ (t >> 15), t | 1)
        t ^= (t + imul(t ^ (t >> 7), t | 61)) & 0xFFFFFFFF
        return ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296.0

    return nxt


def vortex_rings(n=128, radius=20.0, core=4.0, circulation=0.1, seed=7, jitter=8.0, tau=0.6, q=27, rings=8):
    """Seeded random vortex-ring field in a periodic n^3 box.

    Up to eight rings (the first `rings` in x-fastest order) sit on the 2x2x2
    lattice of centres n/4 + (n/2) {0, 1}^3, each centre moved by up to +-`jitter` cells per axis from mulberry32(seed), the
    ring axis cycling x, y, z.  Each ring is a Gaussian (Lamb-Oseen) core of
    radius `core` carrying `circulation` around a circle of radius `radius`:
    the vorticity field w = Gamma / (pi a^2) exp(-d^2 / a^2) t_hat, with d
    the distance to the ring's centre line and t_hat its tangent.  The
    velocity is the periodic Biot-Savart solution u = curl(-lap^-1 w),
    computed spectrally, so it is divergence-free; rho = 1 - |u|^2 / (2 cs^2).
    The reference has no such initialiser; this one is new and deterministic.
    """
    rnd = mulberry32(seed)
    ax = np.arange(n, dtype=np.float64)
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    w = [np.zeros((n, n, n)) for _ in range(3)]
    k = 0
    for cz in (0, 1):
        for cy in (0, 1):
            for cx in (0, 1):
                c = [n / 4 + n / 2 * v + jitter * (2.0 * rnd() - 1.0) for v in (cx, cy, cz)]
                if k >= rings:
                    continue
                axis = k % 3
                k += 1
                # periodic offsets from the centre, nearest image
                d = [(G - c[j] + n / 2) % n - n / 2 for j, G in enumerate((X, Y, Z))]
                a1, a2 = (axis + 1) % 3, (axis + 2) % 3
                rho_p = np.sqrt(d[a1] ** 2 + d[a2] ** 2)
                dist2 = (rho_p - radius) ** 2 + d[axis] ** 2
                mag = circulation / (math.pi * core * core) * np.exp(-dist2 / (core * core))
                inv = 1.0 / np.maximum(rho_p, 1e-12)
                # tangent of the circle around `axis`: (-d2, d1) / rho in the (a1, a2) plane
                w[a1] += -mag * d[a2] * inv
                w[a2] += mag * d[a1] * inv
    kf = np.fft.fftfreq(n, d=1.0 / n) * (2.0 * math.pi / n)
    kr = np.fft.rfftfreq(n, d=1.0 / n) * (2.0 * math.pi / n)
    KZ, KY, KX = np.meshgrid(kf, kf, kr, indexing="ij")
    k2 = KX * KX + KY * KY + KZ * KZ
    k2[0, 0, 0] = 1.0
    wk = [np.fft.rfftn(c) for c in w]
    # psi = w / k^2 (lap psi = -w), u = curl psi = i k x psi
    px, py, pz = (c / k2 for c in wk)
    uk = [1j * (KY * pz - KZ * py), 1j * (KZ * px - KX * pz), 1j * (KX * py - KY * px)]
    u = [np.fft.irfftn(c, s=(n, n, n)) for c in uk]
    mac = np.empty((n, n, n, 4), dtype=np.float32)
    usq = u[0] ** 2 + u[1] ** 2 + u[2] ** 2
    mac[..., 0] = 1.0 - usq / (2.0 * CS2)
    for j in range(3):
        mac[..., 1 + j] = u[j]
    wmax = float(np.sqrt(w[0] ** 2 + w[1] ** 2 + w[2] ** 2).max())
    desc = {"q": q, "nx": n, "ny": n, "nz": n, "omega": f32(1.0 / tau)}
    return desc, mac.reshape(-1), {"vorticity_max": wmax}


def cylinder2d(nx=256, u0=0.1, re=100.0, trt=True, storage="f32"):
    """D2Q9 channel of nx x nx/4 cells with a pinned cylinder of diameter
    D = ny/8 at (ny/2 + D, ny/2 + 1/2) (slightly off the mid-line so the wake
    can shed), halfway bounce-back, freestream u0 seeded everywhere and a sponge
    of width max(6, D) on the x faces relaxing toward it (the channel pages'
    setup, main-channel.js): nu = u0 D / Re, tau = 3 nu + 1/2."""
    ny = nx // 4
    d_ = ny / 8.0
    nu = u0 * d_ / re
    tau = 3.0 * nu + 0.5
    omega = f32(1.0 / tau)
    desc = {
        "q": 9, "nx": nx, "ny": ny, "nz": 1, "omega": omega,
        "sponge": {"width": float(max(6, int(round(d_)))), "axes": [1, 0, 0], "u": [u0, 0.0, 0.0]},
        "body": {"a": d_ / 2.0, "centre": [ny / 2.0 + d_, ny / 2.0 + 0.5]},
        "solidEq": 1, "storage": storage,
    }
    if trt:
        desc["omegaMinus"] = trt_omega_minus(omega)
    n = nx * ny
    mac = np.empty((n, 4), dtype=np.float32)
    mac[:, 0] = 1.0 - (u0 * u0) / (2.0 * CS2)
    mac[:, 1] = u0
    mac[:, 2] = 0.0
    mac[:, 3] = 0.0
    meta = {"area": d_, "u0": u0, "nu": nu, "tau": tau, "re": re}
    return desc, mac.reshape(-1), meta


def cylinder_open(re, d=16, u0=0.05, trt=True, storage="f32"):
    """D2Q9 circular cylinder of diameter d in an unconfined uniform crossflow
    (benchmarks/cylinder.json: open domain, low blockage): a 60 d x 30 d box
    (blockage 1/30), the cylinder at (15 d, 15 d + 1/2) (half a cell off the
    mid-line so the wake can shed), halfway bounce-back, freestream u0 seeded
    everywhere and a sponge of width 2 d on all four faces relaxing toward it:
    nu = u0 d / Re, tau = 3 nu + 1/2; Cd = 2 Fx / (u0^2 d), St = f d / u0."""
    nx, ny = 60 * d, 30 * d
    nu = u0 * d / re
    tau = 3.0 * nu + 0.5
    omega = f32(1.0 / tau)
    desc = {
        "q": 9, "nx": nx, "ny": ny, "nz": 1, "omega": omega,
        "sponge": {"width": float(2 * d), "axes": [1, 1, 0], "u": [u0, 0.0, 0.0]},
        "body": {"a": d / 2.0, "centre": [15.0 * d, 15.0 * d + 0.5]},
        "solidEq": 1, "storage": storage,
    }
    if trt:
        desc["omegaMinus"] = trt_omega_minus(omega)
    mac = np.empty((nx * ny, 4), dtype=np.float32)
    mac[:, 0] = 1.0 - (u0 * u0) / (2.0 * CS2)
    mac[:, 1] = u0
    mac[:, 2] = 0.0
    mac[:, 3] = 0.0
    meta = {"area": float(d), "u0": u0, "nu": nu, "tau": tau, "re": re, "d": d}
    return desc, mac.reshape(-1), meta


def boxed_sphere(n=64, u0=0.08, re=100.0, trt=True, storage="f32"):
    """D3Q19 sphere of diameter n/4 in an n^3 box at (n/3, n/2, n/2), sponge
    on every face toward the freestream (the boxed-sphere serving domain)."""
    radius = n / 8.0
    return sphere(nx=n, ny=n, nz=n, radius=radius, centre=(n / 3.0, n / 2.0, n / 2.0), u0=u0, re=re, trt=trt,
                  storage=storage)


def serve_mix(count=256, rate=50.0, seed=5, steps=2000, storage="f16", scale=1.0):
    """Seeded serving workload: Poisson arrivals at `rate` per second; each
    request is, with equal odds, a 2D cylinder channel with nx uniform in
    [256, 1024] (multiples of 32) or a 3D boxed sphere with n uniform in
    [64, 192] (multiples of 16); Re uniform in [50, 500].  `scale` shrinks the
    grids for tests.  Returns [(arrival_s, desc, mac, meta, steps)]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(count):
        t += rng.exponential(1.0 / rate)
        re = float(rng.uniform(50.0, 500.0))
        if rng.random() < 0.5:
            nx = int(round(rng.uniform(256, 1024) * scale / 32.0)) * 32
            nx = max(nx, 64)
            desc, mac, meta = cylinder2d(nx=nx, re=re, storage=storage)
        else:
            n = int(round(rng.uniform(64, 192) * scale / 16.0)) * 16
            n = max(n, 32)
            desc, mac, meta = boxed_sphere(n=n, re=re, storage=storage)
        out.append((t, desc, mac, meta, steps))
    return out
