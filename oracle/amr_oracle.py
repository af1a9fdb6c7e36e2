"""FP64 CPU oracle for the two-level block pool (dense L0 + one level of
fine tiles).  TEST INFRASTRUCTURE ONLY: tests/ run it as the checker; the
engine never imports it and has no CPU fallback.

It restates, in float64 and on unshifted populations f, the reference's
explode/coalesce AMR macro-step:

  L0 step (covered)   shaders/common_d3_step.wgsl:224-400, periodic and
                      without a body; cells of refined blocks are not solved
                      (M4.1a, common_d3_step.wgsl skip on blockSlot >= 0)
  pool fine step      shaders/common_d3_amr_step1.wgsl:160-405
  explode             shaders/common_d3_amr_explode.wgsl:98-214
  coalesce + orphans  shaders/common_d3_amr_coalesce.wgsl:126-285
  interp fill         shaders/common_d3_amr_interp.wgsl:109-185 with
                      interpCoarseToFine3, common_d3_pool.wgsl:172-220
  average / drain     shaders/common_d3_amr_average.wgsl:75-154
  criterion           shaders/common_d3_criterion.wgsl:102-158 with the
                      vorticity magnitude (d3-criterion.mjs omegaOfGrad)
  manager             shaders/common_d3_manage.wgsl:326-546, slots handed
                      out in ascending block id (d3-amr.mjs:75-99)
  active list         shaders/amr_active_list.wgsl:41-70
  macro-step order    main-3d.js:3122-3335; per-level tau main-3d.js:1188-1193
  interp interface    ?interface=interp, main-3d.js:1266-1277 and :3344-3357:
                      coarse step first, ring re-interpolated from the coarse
                      state at t (GHOST_ONLY, TIME_BLEND 0), ring cells collide
                      (COLLIDE_RING = 1), every tile restricted after the two
                      substeps
  flux / reflux       shaders/common_d3_amr_flux.wgsl:80-194 and
                      shaders/common_d3_amr_reflux.wgsl:60-165 (?reflux=1)

One deliberate difference from the reference's manager pass, shared with
the engine: before the manager runs, EVERY tile in use is restricted onto
the L0 cells under it (the reference restricts only the dying ones,
DYING_ONLY).  Between manager passes L0 does not solve under a tile, so
without this a fresh tile next to a surviving one would interpolate cells
that nothing maintains.

Arrays: L0 populations f0[q, z, y, x]; pool[q, slot, fz, fy, fx] with the
tile edge FB = 2 RB + 2 GHOST; macroscopic fields mac0[4, z, y, x] and
mac_pool[4, slot, fz, fy, fx] as [rho, ux, uy, uz].

Parity pins (tests/test_amr_oracle.py): with no tile the macro-step is the
dense FP64 JS oracle's step (oracle/lbm_oracle.js) to rounding; with every
block refined the pool is a dense fine lattice at twice the resolution and
matches the JS oracle run at 2N with tau_f; the composite mass and momentum
are conserved to float64 rounding across explode, coalesce and orphans, with
uniform and linear explosion, on flat faces, edges and corners; a uniform
flow stays uniform; restriction after interpolation is the identity on
fields linear in space; the criterion meets the closed forms of
test-d3-criterion.js (pure shear |w| = gamma, solid rotation |w| = 2 Omega).
The interp interface is not conservative; with the flux correction the
composite mass and momentum are conserved to float64 rounding (the claim of
common_d3_amr_flux.wgsl:8-11), and the same check fails without it.
"""
import math

import numpy as np

GHOST = 2

_E27 = [(0, 0, 0),
        (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
        (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1),
        (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1),
        (1, 1, 1), (-1, -1, -1), (1, 1, -1), (-1, -1, 1), (1, -1, 1), (-1, 1, -1),
        (1, -1, -1), (-1, 1, 1)]
# f32 weights by shell with the exact-sum ulp adjustment (lattice-3d.mjs:85-111)
_W_F32 = {19: (0.33333337306976318, 0.055555541068315506, 0.027777781710028648),
          27: (0.29629626870155334, 0.074074082076549530, 0.018518516793847084, 0.0046296296641230583)}
_W_EXACT = {19: (1 / 3, 1 / 18, 1 / 36), 27: (8 / 27, 2 / 27, 1 / 54, 1 / 216)}


class Tables:
    """Shell-ordered D3Q19/D3Q27 set, opp(i) = i +- 1 (lattice-3d.mjs:146-210)."""

    def __init__(self, q, exact_weights=False):
        if q not in (19, 27):
            raise ValueError("q must be 19 or 27")
        e = np.array(_E27[:q], dtype=np.int64)
        self.q = q
        self.ex, self.ey, self.ez = e[:, 0], e[:, 1], e[:, 2]
        shell = (e * e).sum(1)
        W = (_W_EXACT if exact_weights else _W_F32)[q]
        self.w = np.array([W[s] for s in shell], dtype=np.float64)
        self.opp = np.array([0] + [i + 1 if i % 2 else i - 1 for i in range(1, q)], dtype=np.int64)


def feq(T, rho, ux, uy, uz):
    """w rho (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u.u) for every direction."""
    usq = ux * ux + uy * uy + uz * uz
    out = np.empty((T.q,) + np.shape(rho))
    for i in range(T.q):
        eu = T.ex[i] * ux + T.ey[i] * uy + T.ez[i] * uz
        out[i] = T.w[i] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq)
    return out


def moments(T, f, F=(0.0, 0.0, 0.0)):
    """rho and u = (sum e f + F/2) / rho (Guo, common_d3_step.wgsl)."""
    rho = f.sum(0)
    mx = np.tensordot(T.ex.astype(np.float64), f, 1)
    my = np.tensordot(T.ey.astype(np.float64), f, 1)
    mz = np.tensordot(T.ez.astype(np.float64), f, 1)
    inv = 1.0 / np.maximum(rho, 1e-6)
    return rho, (mx + 0.5 * F[0]) * inv, (my + 0.5 * F[1]) * inv, (mz + 0.5 * F[2]) * inv


def collide(T, f, rho, ux, uy, uz, omega, omega_m, F):
    """BGK / TRT with Guo forcing.  TRT in the delta form of
    plans/TRT.md:333-349: the even part relaxes at omega, the odd at omega_m;
    the Guo source S_i = w (3 (e - u) + 9 (e.u) e) . F splits the same way
    (common_d3_step.wgsl:340-360)."""
    eq = feq(T, rho, ux, uy, uz)
    S = np.empty_like(f)
    for i in range(T.q):
        eu = T.ex[i] * ux + T.ey[i] * uy + T.ez[i] * uz
        eF = T.ex[i] * F[0] + T.ey[i] * F[1] + T.ez[i] * F[2]
        uF = ux * F[0] + uy * F[1] + uz * F[2]
        S[i] = T.w[i] * (3.0 * (eF - uF) + 9.0 * eu * eF)
    o = T.opp
    fp, fm = 0.5 * (f + f[o]), 0.5 * (f - f[o])
    ep, em = 0.5 * (eq + eq[o]), 0.5 * (eq - eq[o])
    sp, sm = 0.5 * (S + S[o]), 0.5 * (S - S[o])
    return (f - omega * (fp - ep) - omega_m * (fm - em)
            + (1.0 - 0.5 * omega) * sp + (1.0 - 0.5 * omega_m) * sm)


def vorticity_magnitude(mac, dtype=np.float64):
    """|curl u| per L0 cell from [rho, ux, uy, uz] planes, central
    differences with periodic wrap (d3-criterion.mjs gradU / omegaOfGrad).
    dtype=float32 rounds every operation to binary32 in the order
    ((a - b) * 0.5) - ((c - d) * 0.5), sqrt((wx^2 + wy^2) + wz^2), which is
    the engine kernel's order, so the result is bit-identical to it."""
    u = [np.asarray(mac[k], dtype=dtype) for k in (1, 2, 3)]
    h = dtype(0.5)

    def d(comp, axis):
        # axis: 0 = z, 1 = y, 2 = x; value(x + 1) - value(x - 1)
        return (np.roll(u[comp], -1, axis) - np.roll(u[comp], 1, axis)) * h

    wx = d(2, 1) - d(1, 0)
    wy = d(0, 0) - d(2, 2)
    wz = d(1, 2) - d(0, 1)
    return np.sqrt((wx * wx + wy * wy) + wz * wz)


def q_criterion(mac, dtype=np.float64):
    """Q = 1/2 (|Omega|^2 - |S|^2) per L0 cell (qAt, common_d3_criterion.wgsl:
    102-120; qOfGrad, d3-criterion.mjs), J[j][i] = du_i/dx_j by central
    differences with periodic wrap.  dtype=float32 rounds every operation to
    binary32 in the kernel's order: J = (u(x+1) - u(x-1)) * 0.5, then over i
    (outer) and j (inner) S = 0.5 (J[j][i] + J[i][j]), O = 0.5 (J[j][i] -
    J[i][j]), s2 += S S, om2 += O O, and Q = 0.5 (om2 - s2)."""
    u = [np.asarray(mac[k], dtype=dtype) for k in (1, 2, 3)]
    h = dtype(0.5)
    axes = (2, 1, 0)  # d/dx, d/dy, d/dz on [z, y, x] arrays
    J = [[(np.roll(u[i], -1, axes[j]) - np.roll(u[i], 1, axes[j])) * h for i in range(3)] for j in range(3)]
    s2 = np.zeros_like(u[0])
    om2 = np.zeros_like(u[0])
    for i in range(3):
        for j in range(3):
            a, b = J[j][i], J[i][j]
            S = h * (a + b)
            O = h * (a - b)
            s2 = s2 + S * S
            om2 = om2 + O * O
    return h * (om2 - s2)


def block_criterion(vm, rb, lead):
    """Per block max of vm over the block's L0 box dilated by `lead` cells
    (common_d3_criterion.wgsl reduceQ), block id (bz nby + by) nbx + bx."""
    nz, ny, nx = vm.shape
    nbx, nby, nbz = nx // rb, ny // rb, nz // rb
    out = np.empty(nbx * nby * nbz, dtype=vm.dtype)
    r = np.arange(-lead, rb + lead)
    for bz in range(nbz):
        iz = (bz * rb + r) % nz
        for by in range(nby):
            iy = (by * rb + r) % ny
            for bx in range(nbx):
                ix = (bx * rb + r) % nx
                out[(bz * nby + by) * nbx + bx] = vm[np.ix_(iz, iy, ix)].max()
    return out


def decide(block_slot, crit, refine_at, keep_at, geom=None):
    """Want flag with hysteresis: a held block stays while crit >= keep_at,
    an unrefined one refines when crit >= refine_at (common_d3_manage.wgsl
    decide, common_refine.wgsl:54-79 ladder at two levels).  `geom`
    (geometry_mask): bit 0 forces a block in, whatever the field says
    (geom || field, common_d3_manage.wgsl:443-462); bit 1 keeps the field
    criterion out of a block."""
    held = block_slot >= 0
    field = np.where(held, crit >= keep_at, crit >= refine_at)
    if geom is None:
        return field.astype(np.uint8)
    return (((geom & 1) != 0) | (field & ((geom & 2) == 0))).astype(np.uint8)


SQRT3_UP = 1.7320509  # sqrt(3) rounded up (d3-amr.mjs:120, common_d3_manage.wgsl:178)


def near_body(sdf, lo, hi, margin, depth=3):
    """Whether any point of the box [lo, hi) comes within `margin` of the
    body: the Lipschitz branch and bound of nearBodyWant (d3-amr.mjs:139-159)
    and boxNearBody (common_d3_manage.wgsl:226-255).  phi(centre) - h sqrt(3)
    bounds the box minimum from below; an ambiguous box splits into octants,
    down to `depth`, where it is accepted."""
    def near(c, h, d):
        phi = sdf(c)
        if phi - h * SQRT3_UP > margin:
            return False
        if phi <= margin or d == 0:
            return True
        q = h / 2
        for k in range(8):
            cc = [c[0] + (q if k & 1 else -q), c[1] + (q if k & 2 else -q), c[2] + (q if k & 4 else -q)]
            if near(cc, q, d - 1):
                return True
        return False
    return near([(lo[k] + hi[k]) / 2 for k in range(3)], (hi[0] - lo[0]) / 2, depth)


def geometry_mask(nbx, nby, nbz, rb, centre, radius, margin, sponge_width=0.0, sponge_axes=(0, 0, 0), dims=None):
    """Per block: bit 0 the body's geometry wants it refined (blockWanted,
    common_d3_manage.wgsl:257-302, a pinned body: no look-ahead term), bit 1
    it reaches into the sponge band, so the field criterion may not refine it
    (the fine step does not relax toward the far field)."""
    out = np.zeros(nbx * nby * nbz, dtype=np.uint8)
    sdf = lambda c: math.sqrt(sum((c[k] - centre[k]) ** 2 for k in range(3))) - radius
    for b in range(out.size):
        bc = (b % nbx, (b // nbx) % nby, b // (nbx * nby))
        lo = [bc[k] * rb for k in range(3)]
        if radius is not None and near_body(sdf, lo, [v + rb for v in lo], margin):
            out[b] |= 1
        if sponge_width > 0:
            for k in range(3):
                if sponge_axes[k] and (lo[k] < sponge_width or (dims[k] - 1) - (lo[k] + rb - 1) < sponge_width):
                    out[b] |= 2
    return out


def manage_slots(block_slot, slot_to_block, want):
    """Coarsen, then give the k-th newly wanted block (ascending id) the k-th
    free slot (ascending).  Returns (block_slot, slot_to_block, dying pairs,
    fresh pairs, exhausted count)."""
    bs = np.array(block_slot, dtype=np.int32)
    sb = np.array(slot_to_block, dtype=np.int32)
    dying = [(b, int(bs[b])) for b in range(bs.size) if bs[b] >= 0 and not want[b]]
    for b, s in dying:
        sb[s] = -1
        bs[b] = -1
    free = [s for s in range(sb.size) if sb[s] < 0]
    fresh, k, exhausted = [], 0, 0
    for b in range(bs.size):
        if want[b] and bs[b] < 0:
            if k < len(free):
                s = free[k]
                k += 1
                bs[b], sb[s] = s, b
                fresh.append((b, s))
            else:
                exhausted += 1
    return bs, sb, dying, fresh, exhausted


def active_list(slot_to_block):
    """Ascending slots in use (amr_active_list.wgsl:41-70)."""
    return np.nonzero(np.asarray(slot_to_block) >= 0)[0].astype(np.int32)


class AmrOracle:
    """desc keys: q, nx, ny, nz, rb, maxSlots, omega, omegaMinus (TRT on L0
    if present), force, manageEvery, refineAt, keepAt, lead, explodeLinear,
    orphans, exactWeights, body ({a, centre}: a pinned sphere on the fine
    level, in L0 units)."""

    def __init__(self, desc):
        self.T = T = Tables(int(desc["q"]), bool(desc.get("exactWeights", False)))
        self.q = T.q
        self.nx, self.ny, self.nz = int(desc["nx"]), int(desc["ny"]), int(desc["nz"])
        self.rb = int(desc["rb"])
        self.rb2 = 2 * self.rb
        self.fb = self.rb2 + 2 * GHOST
        self.nbx, self.nby, self.nbz = self.nx // self.rb, self.ny // self.rb, self.nz // self.rb
        self.nblocks = self.nbx * self.nby * self.nbz
        self.max_slots = int(desc["maxSlots"])
        self.omega = float(desc["omega"])
        om = desc.get("omegaMinus")
        self.omega_m = self.omega if om is None else float(om)
        self.F = tuple(float(v) for v in (list(desc.get("force", [0, 0, 0])) + [0, 0, 0])[:3])
        tc = 1.0 / self.omega
        tf = 2.0 * tc - 0.5
        self.omega_f = 1.0 / tf
        self.c2f = 0.5 * (tf - 1.0) / (tc - 1.0)
        self.f2c = 2.0 * (tc - 1.0) / (tf - 1.0)
        self.manage_every = int(desc.get("manageEvery", 0))
        self.refine_at = float(desc.get("refineAt", 1e30))
        self.keep_at = float(desc.get("keepAt", 1e30))
        self.lead = int(desc.get("lead", 0))
        self.criterion_kind = desc.get("criterion", "vorticity")
        self.linear = bool(desc.get("explodeLinear", 1))
        self.orphans = bool(desc.get("orphans", 1))
        self.interp_iface = desc.get("interface", "explode") == "interp"
        self.reflux_on = bool(desc.get("reflux", 0))
        if self.reflux_on and not self.interp_iface:
            raise ValueError("reflux needs the interp interface")
        self.flux_acc = None
        body = desc.get("body")
        self.body = None
        if body:
            self.body = (np.array(body["centre"], dtype=np.float32), np.float32(body["a"]))
        self.geom = None
        if body and self.manage_every > 0:
            sp = desc.get("sponge") or {}
            self.geom = geometry_mask(self.nbx, self.nby, self.nbz, self.rb, [float(v) for v in self.body[0]],
                                      float(self.body[1]), float(desc.get("bodyMargin", 2.0)),
                                      float(sp.get("width", 0.0)), sp.get("axes", (0, 0, 0)),
                                      (self.nx, self.ny, self.nz))
        self.forces = []
        self.block_slot = np.full(self.nblocks, -1, dtype=np.int32)
        self.slot_to_block = np.full(self.max_slots, -1, dtype=np.int32)
        n = self.fb
        self.pool = np.zeros((self.q, self.max_slots, n, n, n))
        self.mac_pool = np.zeros((4, self.max_slots, n, n, n))
        self.f0 = None
        self.mac0 = None
        self.step_count = 0
        self.crit = np.zeros(self.nblocks, dtype=np.float32)
        self.want = np.zeros(self.nblocks, dtype=np.uint8)
        self.exhausted = 0
        self.last_dying, self.last_fresh = [], []

    # --- geometry -----------------------------------------------------------
    def block_xyz(self, b):
        return b % self.nbx, (b // self.nbx) % self.nby, b // (self.nbx * self.nby)

    def block_id(self, bx, by, bz):
        return (bz * self.nby + by) * self.nbx + bx

    def covered(self, cx, cy, cz):
        """Is L0 cell (cx, cy, cz) (arrays, wrapped) under a tile."""
        cx, cy, cz = np.asarray(cx) % self.nx, np.asarray(cy) % self.ny, np.asarray(cz) % self.nz
        return self.block_slot[self.block_id(cx // self.rb, cy // self.rb, cz // self.rb)] >= 0

    def covered_mask(self):
        z, y, x = np.meshgrid(np.arange(self.nz), np.arange(self.ny), np.arange(self.nx), indexing="ij")
        return self.covered(x, y, z)

    def local_in_tile(self, p, b, nf):
        """Tile-local index of global fine coordinate p relative to block
        coordinate b, nearest periodic image (common_d3_pool.wgsl:75-85)."""
        d = np.asarray(p) - b * self.rb2
        d = np.where(d < -(nf // 2), d + nf, d)
        d = np.where(d > nf // 2, d - nf, d)
        return d + GHOST

    def tile_positions(self, slot):
        """L0-unit centres (float32) of every fine cell of tile `slot`, ring
        included: origin - 1/4 + (j - GHOST) / 2 (fineToCoarseUnit,
        d3-amr.mjs:45-47)."""
        bx, by, bz = self.block_xyz(int(self.slot_to_block[slot]))
        j = np.arange(self.fb, dtype=np.float32)
        h = np.float32(0.5) * (j - np.float32(GHOST)) - np.float32(0.25)
        z, y, x = np.meshgrid(np.float32(bz * self.rb) + h, np.float32(by * self.rb) + h,
                              np.float32(bx * self.rb) + h, indexing="ij")
        return x, y, z

    def phi(self, x, y, z):
        """Signed distance to the body in binary32, as the dense
        classification computes it (common_d3_geometry.wgsl:119)."""
        c, a = self.body
        dx, dy, dz = (np.asarray(v, dtype=np.float32) - c[k] for k, v in enumerate((x, y, z)))
        return np.sqrt(dx * dx + dy * dy + dz * dz) - a

    def body_force(self):
        """The macro-step's one force pass (common_d3_force_pool.wgsl,
        bounce-back branch, body at rest): every interior fluid cell adds
        -e_i 2 f_opp(i) for each direction whose geometric source is solid,
        from the pool state at t; the total is weighted dx^2 = 1/4 and the
        torque taken about the centre in L0 units."""
        T, lo, hi = self.T, GHOST, GHOST + self.rb2
        c = self.body[0].astype(np.float64)
        out = np.zeros(6)
        for slot in active_list(self.slot_to_block):
            x, y, z = (v[lo:hi, lo:hi, lo:hi] for v in self.tile_positions(int(slot)))
            fluid = self.phi(x, y, z) >= 0
            fsum = np.zeros((3,) + x.shape)
            for i in range(1, T.q):
                e = (T.ex[i], T.ey[i], T.ez[i])
                src = fluid & (self.phi(x - np.float32(0.5 * e[0]), y - np.float32(0.5 * e[1]),
                                        z - np.float32(0.5 * e[2])) < 0)
                m = 2.0 * self.pool[T.opp[i], slot, lo:hi, lo:hi, lo:hi]
                for k in range(3):
                    fsum[k] -= np.where(src, e[k] * m, 0.0)
            fsum *= 0.25
            r = (x.astype(np.float64) - c[0], y.astype(np.float64) - c[1], z.astype(np.float64) - c[2])
            out[:3] += fsum.reshape(3, -1).sum(1)
            out[3] += (r[1] * fsum[2] - r[2] * fsum[1]).sum()
            out[4] += (r[2] * fsum[0] - r[0] * fsum[2]).sum()
            out[5] += (r[0] * fsum[1] - r[1] * fsum[0]).sum()
        self.forces.append(out)

    # --- set-up -------------------------------------------------------------
    def upload_topology(self, block_slot, slot_to_block):
        if self.f0 is not None:
            self.restrict_active()
        self.block_slot = np.array(block_slot, dtype=np.int32)
        self.slot_to_block = np.array(slot_to_block, dtype=np.int32)
        if self.f0 is not None:
            self.interp([(int(self.slot_to_block[s]), int(s)) for s in active_list(self.slot_to_block)])

    def upload_macro(self, mac4):
        """mac4: float32 (N, 4) [rho, ux, uy, uz]; L0 equilibrium, then every
        tile filled by interpolation (main-3d.js:2973-3035)."""
        m = np.asarray(mac4, dtype=np.float32).astype(np.float64).reshape(self.nz, self.ny, self.nx, 4)
        self.mac0 = np.moveaxis(m, -1, 0).copy()
        self.f0 = feq(self.T, self.mac0[0], self.mac0[1], self.mac0[2], self.mac0[3])
        self.step_count = 0
        self.interp([(int(self.slot_to_block[s]), int(s)) for s in active_list(self.slot_to_block)])

    # --- passes -------------------------------------------------------------
    def l0_step(self):
        """Pull-stream + collide on every unrefined L0 cell.  Cells under a
        tile are not solved; they are set to NaN so that any later read of
        them that the scheme does not allow poisons the result."""
        T = self.T
        fs = np.empty_like(self.f0)
        for i in range(T.q):
            fs[i] = np.roll(self.f0[i], (T.ez[i], T.ey[i], T.ex[i]), (0, 1, 2))
        rho, ux, uy, uz = moments(T, fs, self.F)
        out = collide(T, fs, rho, ux, uy, uz, self.omega, self.omega_m, self.F)
        cov = self.covered_mask()
        out[:, cov] = np.nan
        for k, v in enumerate((rho, ux, uy, uz)):
            self.mac0[k] = np.where(cov, self.mac0[k], v)
        self.f0 = out

    def _tile_sources(self, slot):
        """Per direction, the flat pool index each cell of tile `slot` pulls
        from (DIRECT_GHOST, common_d3_amr_step1.wgsl:210-312): a source
        outside the interior reads the interior of the tile of the block it
        lies in when that block is refined, else this tile's ring (clamped)."""
        n, lo, hi = self.fb, GHOST, GHOST + self.rb2
        b = int(self.slot_to_block[slot])
        bx, by, bz = self.block_xyz(b)
        fz, fy, fx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        srcs = []
        for i in range(self.q):
            s = [fx - self.T.ex[i], fy - self.T.ey[i], fz - self.T.ez[i]]
            o = [np.where(v < lo, -1, np.where(v >= hi, 1, 0)) for v in s]
            nb = self.block_id((bx + o[0]) % self.nbx, (by + o[1]) % self.nby, (bz + o[2]) % self.nbz)
            ns = self.block_slot[nb]
            other = ((o[0] != 0) | (o[1] != 0) | (o[2] != 0)) & (ns >= 0)
            src_slot = np.where(other, ns, slot)
            c = [np.where(other, s[k] - o[k] * self.rb2, np.clip(s[k], 0, n - 1)) for k in range(3)]
            assert all((v >= 0).all() and (v < n).all() for v in c)
            srcs.append(((src_slot * n + c[2]) * n + c[1]) * n + c[0])
        return srcs

    def pool_step(self):
        """One fine substep of every tile: gather, moments into mac_pool,
        interior BGK at omega_f with the unchanged body force; ring cells
        advect only with explode/coalesce (COLLIDE_RING = 0) and collide too
        with the interp interface (COLLIDE_RING = 1,
        common_d3_amr_step1.wgsl:369-372, main-3d.js:1395)."""
        T, n = self.T, self.fb
        flat = self.pool.reshape(self.q, -1)
        new = self.pool.copy()
        lo, hi = GHOST, GHOST + self.rb2
        r = np.arange(n)
        inside = (r >= lo) & (r < hi)
        interior = inside[:, None, None] & inside[None, :, None] & inside[None, None, :]
        for slot in active_list(self.slot_to_block):
            srcs = self._tile_sources(int(slot))
            g = np.stack([flat[i][srcs[i]] for i in range(self.q)])
            if self.body is not None:
                # halfway bounce-back off the body at rest: a direction whose
                # geometric source is solid takes this cell's own opposite
                # population (common_d3_amr_step1.wgsl:258-276)
                x, y, z = self.tile_positions(int(slot))
                solid = self.phi(x, y, z) < 0
                for i in range(1, T.q):
                    hit = self.phi(x - np.float32(0.5 * T.ex[i]), y - np.float32(0.5 * T.ey[i]),
                                   z - np.float32(0.5 * T.ez[i])) < 0
                    g[i] = np.where(hit, self.pool[T.opp[i], slot], g[i])
            rho, ux, uy, uz = moments(T, g, self.F)
            for k, v in enumerate((rho, ux, uy, uz)):
                self.mac_pool[k, slot] = v
            col = collide(T, g, rho, ux, uy, uz, self.omega_f, self.omega_f, self.F)
            new[:, slot] = col if self.interp_iface else np.where(interior, col, g)
            if self.body is not None:
                # the solid interior held at feq(1, 0) (:387-403)
                new[:, slot] = np.where(solid, T.w.reshape(-1, 1, 1, 1), new[:, slot])
                for k, v in enumerate((1.0, 0.0, 0.0, 0.0)):
                    self.mac_pool[k, slot] = np.where(solid, v, self.mac_pool[k, slot])
        self.pool = new

    def explode(self):
        """Coarse state -> ring of every tile (common_d3_amr_explode.wgsl:98-214)."""
        T, n = self.T, self.fb
        lo, hi = GHOST, GHOST + self.rb2
        NF = (2 * self.nx, 2 * self.ny, 2 * self.nz)
        f0 = self.f0
        for slot in active_list(self.slot_to_block):
            b = int(self.slot_to_block[slot])
            bx, by, bz = self.block_xyz(b)
            fz, fy, fx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
            ring = ~((fx >= lo) & (fx < hi) & (fy >= lo) & (fy < hi) & (fz >= lo) & (fz < hi))
            fx, fy, fz = fx[ring], fy[ring], fz[ring]
            gx = (bx * self.rb2 + fx - GHOST) % NF[0]
            gy = (by * self.rb2 + fy - GHOST) % NF[1]
            gz = (bz * self.rb2 + fz - GHOST) % NF[2]
            vx, vy, vz = gx // 2, gy // 2, gz // 2
            keep = ~self.covered(vx, vy, vz)
            fx, fy, fz, gx, gy, gz, vx, vy, vz = (a[keep] for a in (fx, fy, fz, gx, gy, gz, vx, vy, vz))
            okx = ~self.covered(vx - 1, vy, vz) & ~self.covered(vx + 1, vy, vz)
            oky = ~self.covered(vx, vy - 1, vz) & ~self.covered(vx, vy + 1, vz)
            okz = ~self.covered(vx, vy, vz - 1) & ~self.covered(vx, vy, vz + 1)
            d = [((g - 2 * v) - 0.5) * 0.5 for g, v in ((gx, vx), (gy, vy), (gz, vz))]

            def fc(i, x, y, z):
                return f0[i, z % self.nz, y % self.ny, x % self.nx]

            for i in range(T.q):
                ex, ey, ez = int(T.ex[i]), int(T.ey[i]), int(T.ez[i])
                tcov = self.covered(vx + ex, vy + ey, vz + ez)
                base = fc(i, vx, vy, vz)
                # in-orphan: the child's own one-fine-step destination is in
                # this tile's refined interior although v + e_i is not
                dx, dy, dz = (gx + ex) % NF[0], (gy + ey) % NF[1], (gz + ez) % NF[2]
                into = (self.covered(dx // 2, dy // 2, dz // 2) & (dx // self.rb2 == bx) & (dy // self.rb2 == by)
                        & (dz // self.rb2 == bz))
                val = np.where(into, base, 0.0)
                if self.linear and i > 0:
                    G = [np.where(okx, 0.5 * (fc(i, vx + 1, vy, vz) - fc(i, vx - 1, vy, vz)), 0.0),
                         np.where(oky, 0.5 * (fc(i, vx, vy + 1, vz) - fc(i, vx, vy - 1, vz)), 0.0),
                         np.where(okz, 0.5 * (fc(i, vx, vy, vz + 1) - fc(i, vx, vy, vz - 1)), 0.0)]
                    ee = float(ex * ex + ey * ey + ez * ez)
                    proj = (G[0] * ex + G[1] * ey + G[2] * ez) / ee
                    G = [G[0] - ex * proj, G[1] - ey * proj, G[2] - ez * proj]
                    lin = base + (d[0] * G[0] + d[1] * G[1] + d[2] * G[2])
                else:
                    lin = base
                self.pool[i, slot, fz, fy, fx] = np.where(tcov, lin, val)

    def coalesce(self):
        """Ring outflux -> time-t L0 slots of covered cells, mac projection,
        orphan correction (common_d3_amr_coalesce.wgsl:126-285)."""
        T = self.T
        NF = (2 * self.nx, 2 * self.ny, 2 * self.nz)
        f0 = self.f0
        for slot in active_list(self.slot_to_block):
            b = int(self.slot_to_block[slot])
            bx, by, bz = self.block_xyz(b)
            r = np.arange(self.rb)
            cz, cy, cx = (a.ravel() for a in np.meshgrid(bz * self.rb + r, by * self.rb + r, bx * self.rb + r,
                                                         indexing="ij"))
            f0x, f0y, f0z = GHOST + 2 * (cx - bx * self.rb), GHOST + 2 * (cy - by * self.rb), GHOST + 2 * (cz - bz * self.rb)
            rs = np.zeros(cx.size)
            m = [np.zeros(cx.size) for _ in range(3)]
            for k in range(8):
                ox, oy, oz = k & 1, (k >> 1) & 1, (k >> 2) & 1
                rho = self.mac_pool[0, slot, f0z + oz, f0y + oy, f0x + ox]
                rs += rho
                for a in range(3):
                    m[a] += rho * self.mac_pool[1 + a, slot, f0z + oz, f0y + oy, f0x + ox]
            inv = 1.0 / np.maximum(rs, 1e-6)
            self.mac0[0, cz, cy, cx] = rs * 0.125
            for a in range(3):
                self.mac0[1 + a, cz, cy, cx] = m[a] * inv
            for i in range(T.q):
                ex, ey, ez = int(T.ex[i]), int(T.ey[i]), int(T.ez[i])
                open_ = ~self.covered(cx + ex, cy + ey, cz + ez)
                tx, ty, tz = cx[open_] + ex, cy[open_] + ey, cz[open_] + ez
                s = np.zeros(tx.size)
                for k in range(8):
                    px, py, pz = 2 * tx + (k & 1), 2 * ty + ((k >> 1) & 1), 2 * tz + ((k >> 2) & 1)
                    lx = self.local_in_tile(px, bx, NF[0])
                    ly = self.local_in_tile(py, by, NF[1])
                    lz = self.local_in_tile(pz, bz, NF[2])
                    # the target is one coarse cell outside the block: its
                    # children are this tile's ring
                    assert all(((v >= 0) & (v < self.fb)).all() for v in (lx, ly, lz))
                    s += self.pool[i, slot, lz, ly, lx]
                f0[i, cz[open_], cy[open_], cx[open_]] = s * 0.125
        if not self.orphans:
            return
        cov = self.covered_mask()
        z, y, x = np.meshgrid(np.arange(self.nz), np.arange(self.ny), np.arange(self.nx), indexing="ij")
        near = np.zeros_like(cov)
        for i in range(1, T.q):
            near |= self.covered(x + T.ex[i], y + T.ey[i], z + T.ez[i])
        cz, cy, cx = np.nonzero(near & ~cov)
        for i in range(1, T.q):
            ex, ey, ez = int(T.ex[i]), int(T.ey[i]), int(T.ez[i])
            tx, ty, tz = cx + ex, cy + ey, cz + ez
            sel = ~self.covered(tx, ty, tz)
            s = np.zeros(cx.size)
            nin = np.zeros(cx.size)
            for k in range(8):
                ox, oy, oz = k & 1, (k >> 1) & 1, (k >> 2) & 1
                px, py, pz = (2 * tx + ox) % NF[0], (2 * ty + oy) % NF[1], (2 * tz + oz) % NF[2]
                qx, qy, qz = (px - ex) % NF[0], (py - ey) % NF[1], (pz - ez) % NF[2]
                qbx, qby, qbz = qx // self.rb2, qy // self.rb2, qz // self.rb2
                qs = self.block_slot[self.block_id(qbx, qby, qbz)]
                has = qs >= 0
                lx = self.local_in_tile(px, qbx, NF[0])
                ly = self.local_in_tile(py, qby, NF[1])
                lz = self.local_in_tile(pz, qbz, NF[2])
                idx = np.nonzero(has)[0]
                if idx.size:
                    assert ((lx[idx] >= 0) & (lx[idx] < self.fb) & (ly[idx] >= 0) & (ly[idx] < self.fb)
                            & (lz[idx] >= 0) & (lz[idx] < self.fb)).all()
                    s[idx] += self.pool[i, qs[idx], lz[idx], ly[idx], lx[idx]]
                hx, hy, hz = (2 * cx + ox + ex) % NF[0], (2 * cy + oy + ey) % NF[1], (2 * cz + oz + ez) % NF[2]
                nin += self.covered(hx // 2, hy // 2, hz // 2)
            fc = f0[i, cz, cy, cx]
            net = s - nin * fc
            f0[i, cz[sel], cy[sel], cx[sel]] = (fc + net * 0.125)[sel]

    def _coarse_sample(self, x, y, z, f0=None):
        f = (self.f0 if f0 is None else f0)[:, z % self.nz, y % self.ny, x % self.nx]
        rho, ux, uy, uz = moments(self.T, f)
        return rho, ux, uy, uz, f - feq(self.T, rho, ux, uy, uz)

    def interp(self, pairs, f0=None, ring_only=False):
        """Fill whole tiles (ring included) from L0: trilinear rho, u and
        fneq x 1/2 (tau_f - 1)/(tau_c - 1) (common_d3_pool.wgsl:172-220).
        ring_only: the ring cells alone (GHOST_ONLY = 1,
        common_d3_amr_interp.wgsl:4-9), from the coarse populations f0."""
        n = self.fb
        src = self.f0 if f0 is None else f0
        lo, hi = GHOST, GHOST + self.rb2
        for b, slot in pairs:
            bx, by, bz = self.block_xyz(b)
            fz, fy, fx = (a.ravel() for a in np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"))
            if ring_only:
                ring = ~((fx >= lo) & (fx < hi) & (fy >= lo) & (fy < hi) & (fz >= lo) & (fz < hi))
                fz, fy, fx = fz[ring], fy[ring], fx[ring]
            P = [o * self.rb - 0.25 + 0.5 * (j - GHOST) for o, j in ((bx, fx), (by, fy), (bz, fz))]
            c0 = [np.floor(p).astype(np.int64) for p in P]
            t = [p - c for p, c in zip(P, c0)]
            rho = np.zeros(fx.size)
            u = [np.zeros(fx.size) for _ in range(3)]
            fneq = np.zeros((self.q, fx.size))
            for k in range(8):
                i, j, kk = k & 1, (k >> 1) & 1, (k >> 2) & 1
                w = (t[0] if i else 1 - t[0]) * (t[1] if j else 1 - t[1]) * (t[2] if kk else 1 - t[2])
                sr, sx, sy, sz, sf = self._coarse_sample(c0[0] + i, c0[1] + j, c0[2] + kk, src)
                rho += w * sr
                u[0] += w * sx
                u[1] += w * sy
                u[2] += w * sz
                fneq += w * sf
            out = feq(self.T, rho, u[0], u[1], u[2]) + self.c2f * fneq
            self.pool[:, slot, fz, fy, fx] = out

    def average(self, pairs):
        """8-child restriction of each listed tile onto its L0 cells: rho
        mean, mass-weighted u, fneq mean x 2 (tau_c - 1)/(tau_f - 1)
        (common_d3_amr_average.wgsl:75-154)."""
        T = self.T
        for b, slot in pairs:
            bx, by, bz = self.block_xyz(b)
            r = np.arange(self.rb)
            lz, ly, lx = (a.ravel() for a in np.meshgrid(r, r, r, indexing="ij"))
            rs = np.zeros(lx.size)
            m = [np.zeros(lx.size) for _ in range(3)]
            sneq = np.zeros((self.q, lx.size))
            for c in range(8):
                f = self.pool[:, slot, GHOST + 2 * lz + ((c >> 2) & 1), GHOST + 2 * ly + ((c >> 1) & 1),
                              GHOST + 2 * lx + (c & 1)]
                rho, ux, uy, uz = moments(T, f)
                rs += rho
                m[0] += rho * ux
                m[1] += rho * uy
                m[2] += rho * uz
                sneq += f - feq(T, rho, ux, uy, uz)
            rho = rs * 0.125
            inv = 1.0 / np.maximum(rs, 1e-6)
            ux, uy, uz = m[0] * inv, m[1] * inv, m[2] * inv
            cx, cy, cz = bx * self.rb + lx, by * self.rb + ly, bz * self.rb + lz
            self.f0[:, cz, cy, cx] = feq(T, rho, ux, uy, uz) + self.f2c * (sneq * 0.125)
            for k, v in enumerate((rho, ux, uy, uz)):
                self.mac0[k, cz, cy, cx] = v

    def restrict_active(self):
        self.average([(int(self.slot_to_block[s]), int(s)) for s in active_list(self.slot_to_block)])

    def criterion(self):
        """Per block |curl u| (or Q) maximum, binary32 as the engine computes it."""
        field = q_criterion if self.criterion_kind == "q" else vorticity_magnitude
        return block_criterion(field(self.mac0.astype(np.float32), np.float32), self.rb, self.lead)

    def manage(self, crit=None):
        """Criterion -> restrict every tile -> decide -> coarsen -> refine ->
        fill.  `crit` overrides the criterion (the engine's own values, to
        follow its integer decisions exactly)."""
        self.crit = self.criterion() if crit is None else np.asarray(crit, dtype=np.float32)
        self.restrict_active()
        self.want = decide(self.block_slot, self.crit, np.float32(self.refine_at), np.float32(self.keep_at),
                           self.geom)
        bs, sb, dying, fresh, exh = manage_slots(self.block_slot, self.slot_to_block, self.want)
        self.block_slot, self.slot_to_block = bs, sb
        self.last_dying, self.last_fresh = dying, fresh
        self.exhausted += exh
        self.interp(fresh)

    def macro_step(self, crit=None):
        if self.manage_every > 0 and self.step_count % self.manage_every == 0:
            self.manage(crit)
        if self.interp_iface:
            self._macro_step_interp()
        else:
            if self.body is not None:
                self.body_force()
            self.explode()
            self.pool_step()
            self.pool_step()
            self.coalesce()
            self.l0_step()
        self.step_count += 1

    def _macro_step_interp(self):
        """main-3d.js:3344-3357: coarse step, ring from the coarse state at t,
        [flux], substep A, [flux], substep B, average, [reflux]."""
        ft = self.f0.copy()
        self.l0_step()
        pairs = [(int(self.slot_to_block[s]), int(s)) for s in active_list(self.slot_to_block)]
        if not pairs:
            return
        self.interp(pairs, f0=ft, ring_only=True)
        if self.reflux_on:
            acc = self.flux()
        self.pool_step()
        if self.reflux_on:
            acc = acc + self.flux()
        self.pool_step()
        self.average(pairs)
        if self.reflux_on:
            self.reflux(ft, acc)

    def flux_cells(self):
        """Unrefined L0 cells with a refined block among the blocks their 26
        neighbours span (common_d3_amr_flux.wgsl:101-122): (x, y, z) arrays."""
        cov = self.covered_mask()
        near = np.zeros_like(cov)
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    near |= np.roll(cov, (-dz, -dy, -dx), (0, 1, 2))
        z, y, x = np.nonzero(near & ~cov)
        return x, y, z

    def _slot_of_fine(self, p):
        b = [p[k] // self.rb2 for k in range(3)]
        return self.block_slot[self.block_id(b[0], b[1], b[2])], b

    def flux(self):
        """(Fout - Fin)/8 over the 8 fine cells of every flux cell, from the
        pool state the next substep streams (common_d3_amr_flux.wgsl:131-193):
        array [4, ncells_listed] of [dm, dpx, dpy, dpz]."""
        T = self.T
        x, y, z = self.flux_cells()
        NF = (2 * self.nx, 2 * self.ny, 2 * self.nz)
        flat = self.pool.reshape(self.q, self.max_slots, self.fb, self.fb, self.fb)
        dm = np.zeros(x.size)
        dp = np.zeros((3, x.size))
        for oc in range(8):
            p = [2 * x + (oc & 1), 2 * y + ((oc >> 1) & 1), 2 * z + ((oc >> 2) & 1)]
            for i in range(T.q):
                e = (T.ex[i], T.ey[i], T.ez[i])
                q = [(p[k] - e[k]) % NF[k] for k in range(3)]
                so, bq = self._slot_of_fine(q)
                sel = so >= 0
                if sel.any():
                    l = [self.local_in_tile(q[k][sel], bq[k][sel], NF[k]) for k in range(3)]
                    f = flat[i, so[sel], l[2], l[1], l[0]]
                    dm[sel] += f
                    for k in range(3):
                        dp[k][sel] += f * e[k]
                t = [(p[k] + e[k]) % NF[k] for k in range(3)]
                si, bt = self._slot_of_fine(t)
                sel = si >= 0
                if sel.any():
                    l = [self.local_in_tile(p[k][sel], bt[k][sel], NF[k]) for k in range(3)]
                    f = flat[i, si[sel], l[2], l[1], l[0]]
                    dm[sel] -= f
                    for k in range(3):
                        dp[k][sel] -= f * e[k]
        return np.vstack([dm, dp]) * 0.125

    def reflux(self, ft, acc):
        """delta = acc + (coarse_in - coarse_out) applied to the unrefined
        flux cells as f_i += w_i (d_rho + 3 e_i . d_mom), then their mac
        republished (common_d3_amr_reflux.wgsl:105-164)."""
        T = self.T
        x, y, z = self.flux_cells()
        dm, dp = acc[0].copy(), acc[1:].copy()
        for i in range(T.q):
            e = (T.ex[i], T.ey[i], T.ez[i])
            sx, sy, sz = (x - e[0]) % self.nx, (y - e[1]) % self.ny, (z - e[2]) % self.nz
            sel = self.covered(sx, sy, sz)
            v = ft[i, sz, sy, sx]
            dm -= np.where(sel, v, 0.0)
            for k in range(3):
                dp[k] -= np.where(sel, v * e[k], 0.0)
            sel = self.covered(x + e[0], y + e[1], z + e[2])
            v = ft[i, z, y, x]
            dm += np.where(sel, v, 0.0)
            for k in range(3):
                dp[k] += np.where(sel, v * e[k], 0.0)
        self.flux_acc = np.vstack([dm, dp])
        rho = np.zeros(x.size)
        m = np.zeros((3, x.size))
        for i in range(T.q):
            e = (T.ex[i], T.ey[i], T.ez[i])
            fo = self.f0[i, z, y, x] + T.w[i] * (dm + 3.0 * (e[0] * dp[0] + e[1] * dp[1] + e[2] * dp[2]))
            self.f0[i, z, y, x] = fo
            rho += fo
            for k in range(3):
                m[k] += fo * e[k]
        den = np.maximum(rho, 1e-6)
        self.mac0[0, z, y, x] = rho
        for k in range(3):
            self.mac0[1 + k, z, y, x] = m[k] / den + self.F[k] / (2.0 * den)

    # --- read-outs ----------------------------------------------------------
    def totals(self):
        """[mass, px, py, pz] of unrefined L0 cells plus fine interiors x 1/8."""
        T = self.T
        cov = self.covered_mask()
        e = np.stack([np.ones(T.q), T.ex, T.ey, T.ez]).astype(np.float64)
        f0 = np.where(cov[None], 0.0, self.f0)
        out = e @ f0.reshape(T.q, -1).sum(1)
        lo, hi = GHOST, GHOST + self.rb2
        for s in active_list(self.slot_to_block):
            inner = self.pool[:, s, lo:hi, lo:hi, lo:hi].reshape(T.q, -1).sum(1)
            out = out + 0.125 * (e @ inner)
        return out

    def l0_f(self):
        return self.f0.reshape(self.q, -1)

    def mac(self):
        """(N, 4) [rho, ux, uy, uz] of L0."""
        return np.moveaxis(self.mac0, 0, -1).reshape(-1, 4)
