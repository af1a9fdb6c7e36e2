"""ctypes mirror of include/lbm.h, used by the Python tests and bench.py.

Loads getting-air_amd/lib/liblbm.so (built in-tree by ``make``) and raises
immediately if it is missing: there is no CPU fallback behind this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "liblbm.so")

ABI_VERSION = 7
BGK, TRT = 0, 1
STORE = {"f32": 0, "f16": 1, "bf16": 2}
BODY_NONE, BODY_SPHERE, BODY_CIRCLE = 0, 1, 2

# Every symbol include/lbm.h declares (checked by tests/test_abi.py).
EXPORTS = (
    "lbm_last_error", "lbm_abi_version", "lbm_device_info", "lbm_solve_create",
    "lbm_solve_destroy", "lbm_upload_macro", "lbm_upload_f", "lbm_read_f", "lbm_step",
    "lbm_step_timed", "lbm_sync", "lbm_read_macro", "lbm_read_forces", "lbm_read_totals",
    "lbm_read_classes", "lbm_solve_info", "lbm_set_tile", "lbm_time_step_kernel",
    "lbm_probe_copy_bandwidth", "lbm_state_size", "lbm_save_state", "lbm_load_state",
    "lbm_amr_create", "lbm_amr_destroy", "lbm_amr_upload_topology", "lbm_amr_read_topology",
    "lbm_amr_upload_macro", "lbm_amr_step", "lbm_amr_manage_now", "lbm_amr_read_macro",
    "lbm_amr_read_pool_macro", "lbm_amr_read_f", "lbm_amr_read_criterion", "lbm_amr_read_active",
    "lbm_amr_read_totals", "lbm_amr_info", "lbm_amr_read_forces", "lbm_amr_upload_level_topology",
    "lbm_amr_read_level_f", "lbm_amr_read_level_macro", "lbm_amr_read_level_topology",
    "lbm_serve_create", "lbm_serve_destroy", "lbm_serve_submit", "lbm_serve_wait", "lbm_serve_read_macro",
    "lbm_serve_read_forces", "lbm_serve_release", "lbm_serve_drain", "lbm_serve_read_stats", "lbm_serve_clock",
    "lbm_serve_submit_state", "lbm_serve_read_state", "lbm_rigid_attach", "lbm_rigid_read",
    "lbm_totals_noalloc",
)
SHAPE = {"sphere": 0, "spheroid": 1, "roundbox": 2}
GHOST = 2


class Body(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int32), ("radius", ctypes.c_float), ("centre", ctypes.c_float * 3)]


class Desc(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32), ("q", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("collision", ctypes.c_int32), ("omega", ctypes.c_float), ("omega_minus", ctypes.c_float),
        ("force", ctypes.c_float * 3), ("walls", ctypes.c_int32 * 3), ("wall_u", ctypes.c_float * 2),
        ("sponge_width", ctypes.c_float), ("sponge_axes", ctypes.c_int32 * 3), ("sponge_u", ctypes.c_float * 3),
        ("body", Body), ("solid_eq", ctypes.c_int32), ("storage", ctypes.c_int32), ("autotune", ctypes.c_int32),
        ("diffuse", ctypes.c_int32), ("chi_eps", ctypes.c_float),
    ]


class AmrDesc(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32), ("q", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("rb", ctypes.c_int32), ("max_slots", ctypes.c_int32), ("collision", ctypes.c_int32),
        ("omega", ctypes.c_float), ("omega_minus", ctypes.c_float), ("force", ctypes.c_float * 3),
        ("manage_every", ctypes.c_int32), ("refine_at", ctypes.c_float), ("keep_at", ctypes.c_float),
        ("lead", ctypes.c_int32), ("explode_linear", ctypes.c_int32), ("orphans", ctypes.c_int32),
        ("interface_interp", ctypes.c_int32), ("reflux", ctypes.c_int32),
        ("sponge_width", ctypes.c_float), ("sponge_axes", ctypes.c_int32 * 3), ("sponge_u", ctypes.c_float * 3),
        ("body", Body), ("explode_dirs", ctypes.c_int32), ("criterion", ctypes.c_int32),
        ("levels", ctypes.c_int32), ("ladder_octaves", ctypes.c_float),
    ]


class RigidBody(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int32), ("pinned", ctypes.c_int32),
                ("a", ctypes.c_double), ("b", ctypes.c_double), ("c", ctypes.c_double), ("r", ctypes.c_double),
                ("density", ctypes.c_double), ("x", ctypes.c_double * 3), ("q", ctypes.c_double * 4),
                ("v", ctypes.c_double * 3), ("omega", ctypes.c_double * 3), ("gravity", ctypes.c_double * 3),
                ("v_max", ctypes.c_double), ("o_max", ctypes.c_double),
                ("planar", ctypes.c_int32), ("no_fluid_force", ctypes.c_int32), ("swept_force", ctypes.c_int32),
                ("rho_w_local", ctypes.c_int32), ("window", ctypes.c_int32 * 3), ("reserved", ctypes.c_int32),
                ("window_anchor", ctypes.c_double * 3)]


class RigidState(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double * 3), ("q", ctypes.c_double * 4), ("v", ctypes.c_double * 3),
                ("L", ctypes.c_double * 3), ("omega", ctypes.c_double * 3), ("d", ctypes.c_double * 3),
                ("force", ctypes.c_double * 3), ("torque", ctypes.c_double * 3),
                ("mass", ctypes.c_double), ("inertia", ctypes.c_double * 3)]


def make_rigid(p):
    """Rigid-body parameters (the oracle's naming: shape {kind, a, b, c, r}, x,
    q, v, omega, density, pinned, gravity, vMax, oMax, planar, noFluidForce,
    swept, rhoWLocal, window, anchor) -> RigidBody."""
    sh = p["shape"]
    kind = sh["kind"] if isinstance(sh["kind"], int) else SHAPE[sh["kind"]]
    b = RigidBody()
    b.shape = kind
    b.pinned = 1 if p.get("pinned") else 0
    b.a = float(sh["a"])
    b.b = float(sh.get("b", sh["a"]))
    b.c = float(sh.get("c", sh["a"]))
    b.r = float(sh.get("r", 0.0))
    b.density = float(p.get("density", 1.0))
    b.x[:] = [float(v) for v in p["x"]]
    b.q[:] = [float(v) for v in p.get("q", (1.0, 0.0, 0.0, 0.0))]
    b.v[:] = [float(v) for v in p.get("v", (0.0, 0.0, 0.0))]
    b.omega[:] = [float(v) for v in p.get("omega", (0.0, 0.0, 0.0))]
    b.gravity[:] = [float(v) for v in p.get("gravity", (0.0, 0.0, 0.0))]
    b.v_max = float(p.get("vMax", 0.2))
    b.o_max = float(p.get("oMax", 0.5))
    b.planar = 1 if p.get("planar") else 0
    b.no_fluid_force = 1 if p.get("noFluidForce") else 0
    b.swept_force = 0 if p.get("swept") is False else 1
    b.rho_w_local = 0 if p.get("rhoWLocal") is False else 1
    b.window[:] = [1 if v else 0 for v in p.get("window", (0, 0, 0))]
    b.window_anchor[:] = [float(v) for v in p.get("anchor", (0.0, 0.0, 0.0))]
    return b


class ServeDesc(ctypes.Structure):
    _fields_ = [("abi_version", ctypes.c_int32), ("max_resident", ctypes.c_int32), ("budget_bytes", ctypes.c_double),
                ("chunk_steps", ctypes.c_int32), ("autotune", ctypes.c_int32), ("target_cells", ctypes.c_double),
                ("persistent", ctypes.c_int32)]


class ServeResult(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_int64), ("status", ctypes.c_int32), ("steps", ctypes.c_int32),
                ("cells", ctypes.c_uint32), ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_finish", ctypes.c_double), ("force", ctypes.c_double * 6)]


class ServeStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("failed", ctypes.c_int64),
                ("resident", ctypes.c_int64), ("queued", ctypes.c_int64), ("cell_updates", ctypes.c_double),
                ("first_submit", ctypes.c_double), ("last_finish", ctypes.c_double),
                ("p50_latency", ctypes.c_double), ("p99_latency", ctypes.c_double), ("mlups", ctypes.c_double),
                ("pool_bytes", ctypes.c_double), ("peak_resident_bytes", ctypes.c_double),
                ("peak_resident", ctypes.c_int32), ("chunks", ctypes.c_int32),
                ("device_seconds", ctypes.c_double), ("device_bytes", ctypes.c_double),
                ("graph_instantiations", ctypes.c_int64), ("graph_updates", ctypes.c_int64)]


class LbmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("lbm error %d: %s" % (code, msg))
        self.code = code


_lib = None


def load():
    """Returns the loaded liblbm.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("liblbm.so not built at %s: run __graft_entry__.build()" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        lib.lbm_last_error.restype = ctypes.c_char_p
        lib.lbm_solve_create.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(P)]
        lib.lbm_solve_destroy.argtypes = [P]
        lib.lbm_upload_macro.argtypes = [P, P]
        lib.lbm_upload_f.argtypes = [P, P]
        lib.lbm_read_f.argtypes = [P, P]
        lib.lbm_step.argtypes = [P, ctypes.c_int32]
        lib.lbm_step_timed.argtypes = [P, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
        lib.lbm_sync.argtypes = [P]
        lib.lbm_read_macro.argtypes = [P, P]
        lib.lbm_read_forces.argtypes = [P, P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        lib.lbm_read_totals.argtypes = [P, P]
        lib.lbm_totals_noalloc.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, P, P, P]
        lib.lbm_read_classes.argtypes = [P, P]
        lib.lbm_solve_info.argtypes = [P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(ctypes.c_int64)]
        lib.lbm_set_tile.argtypes = [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        lib.lbm_time_step_kernel.argtypes = [P, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
        lib.lbm_state_size.argtypes = [P, ctypes.POINTER(ctypes.c_size_t)]
        lib.lbm_save_state.argtypes = [P, P, ctypes.c_size_t]
        lib.lbm_rigid_attach.argtypes = [P, ctypes.POINTER(RigidBody)]
        lib.lbm_rigid_read.argtypes = [P, ctypes.POINTER(RigidState)]
        lib.lbm_load_state.argtypes = [P, P, ctypes.c_size_t]
        lib.lbm_probe_copy_bandwidth.argtypes = [ctypes.c_size_t, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                                 ctypes.POINTER(ctypes.c_double)]
        lib.lbm_device_info.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_int32)]
        A = ctypes.POINTER(AmrDesc)
        i32p, i64p = ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)
        lib.lbm_amr_create.argtypes = [A, ctypes.POINTER(P)]
        lib.lbm_amr_destroy.argtypes = [P]
        lib.lbm_amr_upload_topology.argtypes = [P, P, P]
        lib.lbm_amr_read_topology.argtypes = [P, P, P]
        lib.lbm_amr_read_level_topology.argtypes = [P, ctypes.c_int32, P, P]
        lib.lbm_amr_upload_macro.argtypes = [P, P]
        lib.lbm_amr_step.argtypes = [P, ctypes.c_int32]
        lib.lbm_amr_manage_now.argtypes = [P]
        lib.lbm_amr_read_macro.argtypes = [P, P]
        lib.lbm_amr_read_pool_macro.argtypes = [P, P]
        lib.lbm_amr_read_f.argtypes = [P, P, P]
        lib.lbm_amr_read_criterion.argtypes = [P, P, P]
        lib.lbm_amr_read_active.argtypes = [P, P]
        lib.lbm_amr_read_totals.argtypes = [P, P]
        lib.lbm_amr_info.argtypes = [P, i64p, i32p, i32p]
        lib.lbm_amr_read_forces.argtypes = [P, P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        lib.lbm_amr_upload_level_topology.argtypes = [P, ctypes.c_int32, P, P]
        lib.lbm_amr_read_level_f.argtypes = [P, ctypes.c_int32, P]
        lib.lbm_amr_read_level_macro.argtypes = [P, ctypes.c_int32, P]
        lib.lbm_serve_create.argtypes = [ctypes.POINTER(ServeDesc), ctypes.POINTER(P)]
        lib.lbm_serve_destroy.argtypes = [P]
        lib.lbm_serve_submit.argtypes = [P, ctypes.POINTER(Desc), P, ctypes.c_int32, i64p]
        lib.lbm_serve_wait.argtypes = [P, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(ServeResult)]
        lib.lbm_serve_read_macro.argtypes = [P, ctypes.c_int64, P]
        lib.lbm_serve_read_forces.argtypes = [P, ctypes.c_int64, P, ctypes.c_int32, i32p]
        lib.lbm_serve_release.argtypes = [P, ctypes.c_int64]
        lib.lbm_serve_drain.argtypes = [P, ctypes.c_double]
        lib.lbm_serve_read_stats.argtypes = [P, ctypes.POINTER(ServeStats)]
        lib.lbm_serve_submit_state.argtypes = [P, ctypes.POINTER(Desc), P, P, ctypes.c_size_t, ctypes.c_int32,
                                                ctypes.c_int32, i64p]
        lib.lbm_serve_read_state.argtypes = [P, ctypes.c_int64, P, ctypes.POINTER(ctypes.c_size_t)]
        lib.lbm_serve_clock.argtypes = [P]
        lib.lbm_serve_clock.restype = ctypes.c_double
        _lib = lib
    return _lib


def _check(rc):
    if rc != 0:
        raise LbmError(rc, load().lbm_last_error().decode())


def make_desc(d):
    """Builds a Desc from the oracle-style dict used throughout the tests:
    {q, nx, ny, nz, omega, omegaMinus, force, walls, wallU,
     sponge: {width, axes, u}, body: {a, centre}, solidEq, storage, autotune,
     diffuse, chiEps}."""
    q = int(d["q"])
    desc = Desc()
    desc.abi_version = ABI_VERSION
    desc.q = q
    desc.nx, desc.ny = int(d["nx"]), int(d["ny"])
    desc.nz = 1 if q == 9 else int(d["nz"])
    desc.omega = float(d["omega"])
    if d.get("omegaMinus") is not None:
        desc.collision = TRT
        desc.omega_minus = float(d["omegaMinus"])
    for k, v in enumerate((list(d.get("force", [0, 0, 0])) + [0, 0, 0])[:3]):
        desc.force[k] = float(v)
    for k, v in enumerate((list(d.get("walls", [0, 0, 0])) + [0, 0, 0])[:3]):
        desc.walls[k] = int(bool(v))
    for k, v in enumerate((list(d.get("wallU", [0, 0])) + [0, 0])[:2]):
        desc.wall_u[k] = float(v)
    sp = d.get("sponge") or {}
    desc.sponge_width = float(sp.get("width", 0.0))
    axes = sp.get("axes", [1, 1, 0] if q == 9 else [1, 1, 1])
    for k in range(3):
        desc.sponge_axes[k] = int(bool(axes[k]))
        desc.sponge_u[k] = float((list(sp.get("u", [0, 0, 0])) + [0, 0, 0])[k])
    body = d.get("body")
    if body:
        desc.body.shape = BODY_CIRCLE if q == 9 else BODY_SPHERE
        desc.body.radius = float(body["a"])
        c = list(body["centre"]) + [0.0]
        for k in range(3):
            desc.body.centre[k] = float(c[k])
    desc.solid_eq = int(d.get("solidEq", 1))
    desc.storage = STORE[d.get("storage", "f32")]
    desc.autotune = int(d.get("autotune", 0))
    desc.diffuse = int(bool(d.get("diffuse", 0)))
    desc.chi_eps = float(d.get("chiEps", 0.0))
    return desc


def copy_bandwidth(nbytes=1 << 30, reps=20):
    """(GB/s, ms per copy) of a plain device-to-device streaming copy."""
    gbs, ms = ctypes.c_double(), ctypes.c_double()
    _check(load().lbm_probe_copy_bandwidth(int(nbytes), int(reps), ctypes.byref(gbs), ctypes.byref(ms)))
    return gbs.value, ms.value


def device_info():
    lib = load()
    name = ctypes.create_string_buffer(256)
    hbm = ctypes.c_double()
    cus = ctypes.c_int32()
    _check(lib.lbm_device_info(name, 256, ctypes.byref(hbm), ctypes.byref(cus)))
    return {"name": name.value.decode(), "hbm_bytes": hbm.value, "compute_units": cus.value}


class Solve:
    """One dense lattice on the GPU (owns an lbm_solve handle)."""

    def __init__(self, desc_dict):
        self.lib = load()
        self.desc_dict = dict(desc_dict)
        self.desc = make_desc(desc_dict)
        self.q = self.desc.q
        self.shape = (self.desc.nz, self.desc.ny, self.desc.nx)
        self.ncells = self.desc.nx * self.desc.ny * self.desc.nz
        h = ctypes.c_void_p()
        _check(self.lib.lbm_solve_create(ctypes.byref(self.desc), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.lbm_solve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload_macro(self, mac4):
        m = np.ascontiguousarray(mac4, dtype=np.float32).reshape(-1)
        if m.size != 4 * self.ncells:
            raise ValueError("mac needs %d floats, got %d" % (4 * self.ncells, m.size))
        _check(self.lib.lbm_upload_macro(self.h, m.ctypes.data))

    def upload_f(self, f):
        a = np.ascontiguousarray(f, dtype=np.float32).reshape(-1)
        if a.size != self.q * self.ncells:
            raise ValueError("f needs %d floats" % (self.q * self.ncells))
        _check(self.lib.lbm_upload_f(self.h, a.ctypes.data))

    def read_f(self):
        out = np.empty(self.q * self.ncells, dtype=np.float32)
        _check(self.lib.lbm_read_f(self.h, out.ctypes.data))
        return out

    def step(self, n):
        _check(self.lib.lbm_step(self.h, int(n)))

    def step_timed(self, n):
        ms = ctypes.c_double()
        _check(self.lib.lbm_step_timed(self.h, int(n), ctypes.byref(ms)))
        return ms.value

    def sync(self):
        _check(self.lib.lbm_sync(self.h))

    def read_macro(self):
        out = np.empty(4 * self.ncells, dtype=np.float32)
        _check(self.lib.lbm_read_macro(self.h, out.ctypes.data))
        return out

    def read_forces(self, max_records=1 << 16):
        out = np.empty(6 * max_records, dtype=np.float64)
        cnt = ctypes.c_int32()
        _check(self.lib.lbm_read_forces(self.h, out.ctypes.data, int(max_records), ctypes.byref(cnt)))
        return out[: 6 * cnt.value].reshape(-1, 6)

    def read_totals(self):
        out = np.empty(4, dtype=np.float64)
        _check(self.lib.lbm_read_totals(self.h, out.ctypes.data))
        return out

    def read_classes(self):
        out = np.empty(self.ncells, dtype=np.uint8)
        _check(self.lib.lbm_read_classes(self.h, out.ctypes.data))
        return out

    def info(self):
        step = ctypes.c_int64()
        bx, by, flags = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        nbytes = ctypes.c_int64()
        _check(self.lib.lbm_solve_info(self.h, ctypes.byref(step), ctypes.byref(bx), ctypes.byref(by),
                                       ctypes.byref(flags), ctypes.byref(nbytes)))
        return {"step": step.value, "block_x": bx.value, "block_y": by.value, "flags": flags.value,
                "device_bytes": nbytes.value}

    def save_state(self):
        """Exact checkpoint (lbm_state_header + populations in the storage format), bytes."""
        size = ctypes.c_size_t()
        _check(self.lib.lbm_state_size(self.h, ctypes.byref(size)))
        out = np.empty(size.value, dtype=np.uint8)
        _check(self.lib.lbm_save_state(self.h, out.ctypes.data, out.size))
        return out.tobytes()

    def load_state(self, state):
        b = np.frombuffer(bytes(state), dtype=np.uint8)
        _check(self.lib.lbm_load_state(self.h, b.ctypes.data, b.size))

    def set_tile(self, bx, by, flags):
        """flags: bit 0 XCD swizzle, bit 1 non-temporal, bit 2 linear tiles, bits 4-7 waves-per-SIMD cap."""
        _check(self.lib.lbm_set_tile(self.h, int(bx), int(by), int(flags)))

    def time_step_kernel(self, n):
        """Mean device milliseconds of one fused stream-collide launch."""
        ms = ctypes.c_double()
        _check(self.lib.lbm_time_step_kernel(self.h, int(n), ctypes.byref(ms)))
        return ms.value

    def attach_rigid(self, params):
        """Attaches a moving rigid body (see make_rigid for the parameters)."""
        b = make_rigid(params)
        _check(self.lib.lbm_rigid_attach(self.h, ctypes.byref(b)))

    def read_rigid(self):
        st = RigidState()
        _check(self.lib.lbm_rigid_read(self.h, ctypes.byref(st)))
        return {k: (list(getattr(st, k)) if k != "mass" else st.mass) for k, _ in RigidState._fields_}


def make_amr_desc(d):
    """AmrDesc from a dict {q, nx, ny, nz, rb, maxSlots, omega, omegaMinus,
    force, manageEvery, refineAt, keepAt, lead, explodeLinear, orphans,
    interface ('explode' or 'interp'), reflux, sponge: {width, axes, u} on
    L0, body: {a, centre} a pinned sphere solved on the fine level,
    explodeDirs: directions per explode thread (0: automatic),
    criterion: "vorticity" (|curl u|, default) or "q", levels: 2 (default),
    3 or 4 including L0, ladderOctaves: octaves of rotation rate each
    further level's field rung costs (default 1)}."""
    a = AmrDesc()
    a.abi_version = ABI_VERSION
    a.q = int(d["q"])
    a.nx, a.ny, a.nz = int(d["nx"]), int(d["ny"]), int(d["nz"])
    a.rb = int(d["rb"])
    a.max_slots = int(d["maxSlots"])
    a.omega = float(d["omega"])
    if d.get("omegaMinus") is not None:
        a.collision = TRT
        a.omega_minus = float(d["omegaMinus"])
    for k, v in enumerate((list(d.get("force", [0, 0, 0])) + [0, 0, 0])[:3]):
        a.force[k] = float(v)
    a.manage_every = int(d.get("manageEvery", 0))
    a.refine_at = float(d.get("refineAt", 1e30))
    a.keep_at = float(d.get("keepAt", 1e30))
    a.lead = int(d.get("lead", 0))
    a.explode_linear = int(bool(d.get("explodeLinear", 1)))
    a.orphans = int(bool(d.get("orphans", 1)))
    iface = d.get("interface", "explode")
    if iface not in ("explode", "interp"):
        raise ValueError("interface must be 'explode' or 'interp'")
    a.interface_interp = int(iface == "interp")
    a.reflux = int(bool(d.get("reflux", 0)))
    sp = d.get("sponge") or {}
    a.sponge_width = float(sp.get("width", 0.0))
    axes = sp.get("axes", [1, 1, 1])
    for k in range(3):
        a.sponge_axes[k] = int(bool(axes[k]))
        a.sponge_u[k] = float((list(sp.get("u", [0, 0, 0])) + [0, 0, 0])[k])
    a.explode_dirs = int(d.get("explodeDirs", 0))
    a.criterion = {"vorticity": 0, "q": 1}[d.get("criterion", "vorticity")]
    a.levels = int(d.get("levels", 2))
    a.ladder_octaves = float(d.get("ladderOctaves", 0.0))
    body = d.get("body")
    if body:
        a.body.shape = BODY_SPHERE
        a.body.radius = float(body["a"])
        for k in range(3):
            a.body.centre[k] = float(body["centre"][k])
    return a


class Amr:
    """Block-pool AMR solve on the GPU (owns an lbm_amr handle): L0 plus
    `levels - 1` pool levels."""

    def __init__(self, desc_dict):
        self.lib = load()
        self.desc_dict = dict(desc_dict)
        self.desc = make_amr_desc(desc_dict)
        d = self.desc
        self.q, self.rb, self.max_slots = d.q, d.rb, d.max_slots
        self.ncells = d.nx * d.ny * d.nz
        # q = 9: the D2Q9 quadtree, one cell deep, tiles of fb x fb cells
        self.dim = 2 if d.q == 9 else 3
        self.nblocks = (d.nx // d.rb) * (d.ny // d.rb) * (d.nz // d.rb if self.dim == 3 else 1)
        self.fb = 2 * d.rb + 2 * GHOST
        self.pool_plane = d.max_slots * self.fb ** self.dim
        self.levels = d.levels or 2
        h = ctypes.c_void_p()
        _check(self.lib.lbm_amr_create(ctypes.byref(self.desc), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.lbm_amr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload_topology(self, block_slot, slot_to_block):
        bs = np.ascontiguousarray(block_slot, dtype=np.int32)
        sb = np.ascontiguousarray(slot_to_block, dtype=np.int32)
        if bs.size != self.nblocks or sb.size != self.max_slots:
            raise ValueError("topology sizes %d/%d, want %d/%d" % (bs.size, sb.size, self.nblocks, self.max_slots))
        _check(self.lib.lbm_amr_upload_topology(self.h, bs.ctypes.data, sb.ctypes.data))

    def level_blocks(self, level):
        """Blocks of pool level `level` (over the parent level's global grid)."""
        s = 1 << (level - 1)
        d = self.desc
        return (d.nx * s // d.rb) * (d.ny * s // d.rb) * (d.nz * s // d.rb)

    def upload_level_topology(self, level, block_slot, slot_to_block):
        bs = np.ascontiguousarray(block_slot, dtype=np.int32)
        sb = np.ascontiguousarray(slot_to_block, dtype=np.int32)
        nb = self.level_blocks(level)
        # a level outside 1 .. levels - 1 goes to the library, which refuses it
        # before reading the maps
        if 1 <= level < self.levels and (bs.size != nb or sb.size != self.max_slots):
            raise ValueError("level %d topology sizes %d/%d, want %d/%d" % (level, bs.size, sb.size, nb,
                                                                             self.max_slots))
        _check(self.lib.lbm_amr_upload_level_topology(self.h, int(level), bs.ctypes.data, sb.ctypes.data))

    def read_level_f(self, level):
        out = np.empty(self.q * self.pool_plane, dtype=np.float32)
        _check(self.lib.lbm_amr_read_level_f(self.h, int(level), out.ctypes.data))
        return out

    def read_level_macro(self, level):
        out = np.empty(4 * self.pool_plane, dtype=np.float32)
        _check(self.lib.lbm_amr_read_level_macro(self.h, int(level), out.ctypes.data))
        return out

    def read_topology(self):
        bs = np.empty(self.nblocks, dtype=np.int32)
        sb = np.empty(self.max_slots, dtype=np.int32)
        _check(self.lib.lbm_amr_read_topology(self.h, bs.ctypes.data, sb.ctypes.data))
        return bs, sb

    def read_level_topology(self, level):
        """Level `level`'s (block_slot, slot_to_block) as the manager left them."""
        bs = np.empty(self.level_blocks(level), dtype=np.int32)
        sb = np.empty(self.max_slots, dtype=np.int32)
        _check(self.lib.lbm_amr_read_level_topology(self.h, int(level), bs.ctypes.data, sb.ctypes.data))
        return bs, sb

    def upload_macro(self, mac4):
        m = np.ascontiguousarray(mac4, dtype=np.float32).reshape(-1)
        if m.size != 4 * self.ncells:
            raise ValueError("mac needs %d floats, got %d" % (4 * self.ncells, m.size))
        _check(self.lib.lbm_amr_upload_macro(self.h, m.ctypes.data))

    def step(self, n):
        _check(self.lib.lbm_amr_step(self.h, int(n)))

    def manage_now(self):
        _check(self.lib.lbm_amr_manage_now(self.h))

    def read_macro(self):
        out = np.empty(4 * self.ncells, dtype=np.float32)
        _check(self.lib.lbm_amr_read_macro(self.h, out.ctypes.data))
        return out

    def read_pool_macro(self):
        out = np.empty(4 * self.pool_plane, dtype=np.float32)
        _check(self.lib.lbm_amr_read_pool_macro(self.h, out.ctypes.data))
        return out

    def read_f(self):
        f0 = np.empty(self.q * self.ncells, dtype=np.float32)
        fp = np.empty(self.q * self.pool_plane, dtype=np.float32)
        _check(self.lib.lbm_amr_read_f(self.h, f0.ctypes.data, fp.ctypes.data))
        return f0, fp

    def read_criterion(self):
        crit = np.empty(self.nblocks, dtype=np.float32)
        want = np.empty(self.nblocks, dtype=np.uint8)
        _check(self.lib.lbm_amr_read_criterion(self.h, crit.ctypes.data, want.ctypes.data))
        return crit, want

    def read_active(self):
        out = np.empty(self.max_slots + 1, dtype=np.int32)
        _check(self.lib.lbm_amr_read_active(self.h, out.ctypes.data))
        return out[1: 1 + out[0]].copy()

    def read_totals(self):
        out = np.empty(4, dtype=np.float64)
        _check(self.lib.lbm_amr_read_totals(self.h, out.ctypes.data))
        return out

    def read_forces(self, max_records=1 << 16):
        """The fine-level body's force records, one per macro-step."""
        out = np.empty(6 * max_records, dtype=np.float64)
        cnt = ctypes.c_int32()
        _check(self.lib.lbm_amr_read_forces(self.h, out.ctypes.data, int(max_records), ctypes.byref(cnt)))
        return out[: 6 * cnt.value].reshape(-1, 6)

    def info(self):
        step, act, exh = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        _check(self.lib.lbm_amr_info(self.h, ctypes.byref(step), ctypes.byref(act), ctypes.byref(exh)))
        return {"step": step.value, "active_tiles": act.value, "exhausted": exh.value}


class Server:
    """Batched serve loop on the GPU (owns an lbm_serve handle)."""

    def __init__(self, budget_bytes, max_resident=0, chunk_steps=32, autotune=0, target_cells=0.0, persistent=0):
        self.lib = load()
        d = ServeDesc()
        d.abi_version = ABI_VERSION
        d.max_resident = int(max_resident)
        d.budget_bytes = float(budget_bytes)
        d.chunk_steps = int(chunk_steps)
        d.autotune = int(autotune)
        d.target_cells = float(target_cells)
        d.persistent = int(persistent)
        h = ctypes.c_void_p()
        _check(self.lib.lbm_serve_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self.cells = {}

    def close(self):
        if self.h:
            self.lib.lbm_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, desc_dict, mac4, steps, state=None, keep_state=False):
        """Queue a solve from [rho, u] per cell (mac4) or from a saved state
        (bytes from Solve.save_state / read_state; mac4 None)."""
        desc = make_desc(desc_dict)
        n = desc.nx * desc.ny * desc.nz
        t = ctypes.c_int64()
        if state is None:
            m = np.ascontiguousarray(mac4, dtype=np.float32).reshape(-1)
            if m.size != 4 * n:
                raise ValueError("mac needs %d floats, got %d" % (4 * n, m.size))
            _check(self.lib.lbm_serve_submit_state(self.h, ctypes.byref(desc), m.ctypes.data, None, 0, int(steps),
                                                   int(bool(keep_state)), ctypes.byref(t)))
        else:
            b = np.frombuffer(bytes(state), dtype=np.uint8)
            _check(self.lib.lbm_serve_submit_state(self.h, ctypes.byref(desc), None, b.ctypes.data, b.size,
                                                   int(steps), int(bool(keep_state)), ctypes.byref(t)))
        self.cells[t.value] = n
        return t.value

    def read_state(self, ticket):
        """Final state of a request submitted with keep_state (bytes)."""
        size = ctypes.c_size_t()
        _check(self.lib.lbm_serve_read_state(self.h, int(ticket), None, ctypes.byref(size)))
        out = np.empty(size.value, dtype=np.uint8)
        _check(self.lib.lbm_serve_read_state(self.h, int(ticket), out.ctypes.data, ctypes.byref(size)))
        return out.tobytes()

    def wait(self, ticket, timeout=-1.0):
        r = ServeResult()
        _check(self.lib.lbm_serve_wait(self.h, int(ticket), float(timeout), ctypes.byref(r)))
        return {"ticket": r.ticket, "status": r.status, "steps": r.steps, "cells": r.cells,
                "t_submit": r.t_submit, "t_admit": r.t_admit, "t_finish": r.t_finish, "force": list(r.force)}

    def read_macro(self, ticket):
        out = np.empty(4 * self.cells[ticket], dtype=np.float32)
        _check(self.lib.lbm_serve_read_macro(self.h, int(ticket), out.ctypes.data))
        return out

    def read_forces(self, ticket, max_records=1 << 16):
        out = np.empty(6 * max_records, dtype=np.float64)
        cnt = ctypes.c_int32()
        _check(self.lib.lbm_serve_read_forces(self.h, int(ticket), out.ctypes.data, int(max_records),
                                              ctypes.byref(cnt)))
        return out[: 6 * cnt.value].reshape(-1, 6)

    def release(self, ticket):
        _check(self.lib.lbm_serve_release(self.h, int(ticket)))
        self.cells.pop(ticket, None)

    def drain(self, timeout=-1.0):
        _check(self.lib.lbm_serve_drain(self.h, float(timeout)))

    def stats(self):
        st = ServeStats()
        _check(self.lib.lbm_serve_read_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServeStats._fields_}

    def clock(self):
        return self.lib.lbm_serve_clock(self.h)
