"""FP64 CPU oracle for the two-level D2Q9 quadtree pool (a dense L0 grid and
one level of fine tiles), explode/coalesce interface.  TEST INFRASTRUCTURE
ONLY: tests/ run it as the checker; the engine never imports it and has no
CPU fallback.

It is the two-level oracle of oracle/amr_oracle.py restated in two
dimensions, as the reference's 2D pool is the same scheme on a quadtree
(amr2d.mjs; shaders amr_step1.wgsl, amr_explode.wgsl, amr_coalesce.wgsl,
common_interp*.wgsl, common_average.wgsl):

  L0 step (covered)   D2Q9 BGK (TRT with omegaMinus), periodic; cells of
                      refined blocks are not solved
  pool fine step      pull with DIRECT_GHOST (a source in a refined
                      neighbour's block reads that tile's interior), BGK at
                      tau_f = 2 tau_c - 1/2 on interior cells, the ring only
                      advects
  explode             coarse state -> ring: the covered-target directions
                      injected (uniform, or linear with the gradient
                      projected perpendicular to e_i), in-orphans kept,
                      everything else zero
  coalesce + orphans  the ring's outflux averaged into the covered seam
                      cells (4 children), the moments of every covered cell
                      projected from its 4 children, the orphan correction
                      on unrefined cells next to the refined region
  interp fill         bilinear rho, u and fneq x 1/2 (tau_f - 1)/(tau_c - 1)
  average / drain     4-child restriction, fneq x 2 (tau_c - 1)/(tau_f - 1)
  criterion           |curl u| = |d_x u_y - d_y u_x| per block (box dilated
                      by `lead`), binary32 in the engine's order
  manager             decide with hysteresis, ascending-id slots, restrict
                      every tile first, fill the new ones

A fine cell is a quarter of a coarse cell, so the composite totals weight
the fine interiors by 1/4.

Arrays: L0 f0[q, y, x]; pool[q, slot, fy, fx] with FB = 2 RB + 2 GHOST;
mac0[3, y, x] and mac_pool[3, slot, fy, fx] as [rho, ux, uy].

Parity pins (tests/test_amr2d_oracle.py): with no tile the macro-step is a
dense D2Q9 step; with every block refined the pool is a dense fine lattice at
twice the resolution; the composite mass and momentum are conserved to
float64 rounding with uniform and linear explosion on flat faces and
corners, and the same check fails without the orphan correction; a uniform
flow stays uniform; restriction after interpolation is the identity on
linear fields; the criterion meets its closed forms.
"""
import numpy as np

from amr_oracle import active_list, decide, manage_slots, vorticity_magnitude

GHOST = 2
EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=np.int64)
EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=np.int64)
OPP = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6], dtype=np.int64)
# the f32 weights (lattice-2d.mjs:119-125) or the exact ones
W_F32 = np.array([0.44444444775581360] + [0.11111111193895340] * 4 + [0.027777777984738350] * 4)
W_EXACT = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def feq(w, rho, ux, uy):
    usq = ux * ux + uy * uy
    out = np.empty((9,) + np.shape(rho))
    for i in range(9):
        eu = EX[i] * ux + EY[i] * uy
        out[i] = w[i] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq)
    return out


def moments(f):
    rho = f.sum(0)
    inv = 1.0 / np.maximum(rho, 1e-6)
    return rho, np.tensordot(EX.astype(np.float64), f, 1) * inv, np.tensordot(EY.astype(np.float64), f, 1) * inv


def collide(w, f, rho, ux, uy, omega, omega_m):
    """BGK / TRT (delta form: even part at omega, odd part at omega_m)."""
    eq = feq(w, rho, ux, uy)
    fp, fm = 0.5 * (f + f[OPP]), 0.5 * (f - f[OPP])
    ep, em = 0.5 * (eq + eq[OPP]), 0.5 * (eq - eq[OPP])
    return f - omega * (fp - ep) - omega_m * (fm - em)


def block_criterion(vm, rb, lead):
    """Per block max of vm over the block's box dilated by `lead` (periodic)."""
    ny, nx = vm.shape
    nbx, nby = nx // rb, ny // rb
    out = np.empty(nbx * nby, dtype=vm.dtype)
    r = np.arange(-lead, rb + lead)
    for by in range(nby):
        iy = (by * rb + r) % ny
        for bx in range(nbx):
            ix = (bx * rb + r) % nx
            out[by * nbx + bx] = vm[np.ix_(iy, ix)].max()
    return out


def vorticity_2d(mac):
    """|curl u| of a [rho, ux, uy] field in binary32, as the engine computes
    it: the 3D formula on a grid one cell deep, where every z difference is
    zero (its rounding included: sqrt((0 + 0) + wz^2))."""
    m = np.zeros((4, 1) + mac.shape[1:], dtype=np.float32)
    m[:3, 0] = mac.astype(np.float32)
    return vorticity_magnitude(m, np.float32)[0]


class Amr2DOracle:
    """desc keys: q (9), nx, ny, rb, maxSlots, omega, omegaMinus, manageEvery,
    refineAt, keepAt, lead, explodeLinear, orphans, exactWeights, sponge
    ({width, axes}: only the manager's exclusion band)."""

    def __init__(self, desc):
        assert int(desc["q"]) == 9
        self.q = 9
        self.w = W_EXACT if desc.get("exactWeights") else W_F32
        self.nx, self.ny = int(desc["nx"]), int(desc["ny"])
        self.rb = int(desc["rb"])
        self.rb2 = 2 * self.rb
        self.fb = self.rb2 + 2 * GHOST
        self.nbx, self.nby = self.nx // self.rb, self.ny // self.rb
        self.nblocks = self.nbx * self.nby
        self.max_slots = int(desc["maxSlots"])
        self.omega = float(desc["omega"])
        om = desc.get("omegaMinus")
        self.omega_m = self.omega if om is None else float(om)
        tc = 1.0 / self.omega
        tf = 2.0 * tc - 0.5
        self.omega_f = 1.0 / tf
        self.c2f = 0.5 * (tf - 1.0) / (tc - 1.0)
        self.f2c = 2.0 * (tc - 1.0) / (tf - 1.0)
        self.manage_every = int(desc.get("manageEvery", 0))
        self.refine_at = float(np.float32(desc.get("refineAt", 1e30)))
        self.keep_at = float(np.float32(desc.get("keepAt", 1e30)))
        self.lead = int(desc.get("lead", 0))
        self.linear = bool(desc.get("explodeLinear", 1))
        self.orphans = bool(desc.get("orphans", 1))
        sp = desc.get("sponge") or {}
        self.geom = self._band_mask(float(np.float32(sp.get("width", 0.0))), sp.get("axes", (1, 1, 1)))
        self.block_slot = np.full(self.nblocks, -1, dtype=np.int32)
        self.slot_to_block = np.full(self.max_slots, -1, dtype=np.int32)
        n = self.fb
        self.pool = np.zeros((9, self.max_slots, n, n))
        self.mac_pool = np.zeros((3, self.max_slots, n, n))
        self.f0 = None
        self.mac0 = None
        self.step_count = 0
        self.crit = None
        self.want = None
        self.exhausted = 0
        self.last_dying, self.last_fresh = [], []

    # --- geometry -----------------------------------------------------------
    def _band_mask(self, width, axes):
        """bit 1: the block reaches into the sponge band (the field may not
        refine it)."""
        out = np.zeros(self.nblocks, dtype=np.uint8)
        if width <= 0:
            return out
        n = (self.nx, self.ny)
        for b in range(self.nblocks):
            bc = (b % self.nbx, b // self.nbx)
            for k in range(2):
                lo = bc[k] * self.rb
                if axes[k] and (lo < width or lo + self.rb > n[k] - width):
                    out[b] |= 2
        return out

    def block_xy(self, b):
        return b % self.nbx, b // self.nbx

    def block_id(self, bx, by):
        return by * self.nbx + bx

    def covered(self, cx, cy):
        cx, cy = np.asarray(cx) % self.nx, np.asarray(cy) % self.ny
        return self.block_slot[self.block_id(cx // self.rb, cy // self.rb)] >= 0

    def covered_mask(self):
        y, x = np.meshgrid(np.arange(self.ny), np.arange(self.nx), indexing="ij")
        return self.covered(x, y)

    def local_in_tile(self, p, b, nf):
        d = np.asarray(p) - b * self.rb2
        d = np.where(d < -(nf // 2), d + nf, d)
        d = np.where(d > nf // 2, d - nf, d)
        return d + GHOST

    def tiles(self):
        return [(int(self.slot_to_block[s]), int(s)) for s in active_list(self.slot_to_block)]

    # --- set-up -------------------------------------------------------------
    def upload_topology(self, block_slot, slot_to_block):
        if self.f0 is not None:
            self.average(self.tiles())
        self.block_slot = np.array(block_slot, dtype=np.int32)
        self.slot_to_block = np.array(slot_to_block, dtype=np.int32)
        if self.f0 is not None:
            self.interp(self.tiles())

    def upload_macro(self, mac4):
        """mac4: float32 (N, 4) [rho, ux, uy, uz]; L0 equilibrium, then every
        tile filled by interpolation."""
        m = np.asarray(mac4, dtype=np.float32).astype(np.float64).reshape(self.ny, self.nx, 4)
        self.mac0 = np.moveaxis(m[..., :3], -1, 0).copy()
        self.f0 = feq(self.w, self.mac0[0], self.mac0[1], self.mac0[2])
        self.step_count = 0
        self.interp(self.tiles())

    # --- passes -------------------------------------------------------------
    def l0_step(self):
        fs = np.empty_like(self.f0)
        for i in range(9):
            fs[i] = np.roll(self.f0[i], (EY[i], EX[i]), (0, 1))
        rho, ux, uy = moments(fs)
        out = collide(self.w, fs, rho, ux, uy, self.omega, self.omega_m)
        cov = self.covered_mask()
        out[:, cov] = np.nan
        for k, v in enumerate((rho, ux, uy)):
            self.mac0[k] = np.where(cov, self.mac0[k], v)
        self.f0 = out

    def _tile_sources(self, slot):
        n, lo, hi = self.fb, GHOST, GHOST + self.rb2
        bx, by = self.block_xy(int(self.slot_to_block[slot]))
        fy, fx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        srcs = []
        for i in range(9):
            s = [fx - EX[i], fy - EY[i]]
            o = [np.where(v < lo, -1, np.where(v >= hi, 1, 0)) for v in s]
            ns = self.block_slot[self.block_id((bx + o[0]) % self.nbx, (by + o[1]) % self.nby)]
            other = ((o[0] != 0) | (o[1] != 0)) & (ns >= 0)
            src_slot = np.where(other, ns, slot)
            c = [np.where(other, s[k] - o[k] * self.rb2, np.clip(s[k], 0, n - 1)) for k in range(2)]
            srcs.append((src_slot * n + c[1]) * n + c[0])
        return srcs

    def pool_step(self):
        n = self.fb
        flat = self.pool.reshape(9, -1)
        new = self.pool.copy()
        lo, hi = GHOST, GHOST + self.rb2
        r = np.arange(n)
        inside = (r >= lo) & (r < hi)
        interior = inside[:, None] & inside[None, :]
        for slot in active_list(self.slot_to_block):
            srcs = self._tile_sources(int(slot))
            g = np.stack([flat[i][srcs[i]] for i in range(9)])
            rho, ux, uy = moments(g)
            for k, v in enumerate((rho, ux, uy)):
                self.mac_pool[k, slot] = v
            col = collide(self.w, g, rho, ux, uy, self.omega_f, self.omega_f)
            new[:, slot] = np.where(interior, col, g)
        self.pool = new

    def explode(self):
        n, lo, hi = self.fb, GHOST, GHOST + self.rb2
        NF = (2 * self.nx, 2 * self.ny)
        f0 = self.f0
        for slot in active_list(self.slot_to_block):
            bx, by = self.block_xy(int(self.slot_to_block[slot]))
            fy, fx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
            ring = ~((fx >= lo) & (fx < hi) & (fy >= lo) & (fy < hi))
            fx, fy = fx[ring], fy[ring]
            gx = (bx * self.rb2 + fx - GHOST) % NF[0]
            gy = (by * self.rb2 + fy - GHOST) % NF[1]
            vx, vy = gx // 2, gy // 2
            keep = ~self.covered(vx, vy)
            fx, fy, gx, gy, vx, vy = (a[keep] for a in (fx, fy, gx, gy, vx, vy))
            okx = ~self.covered(vx - 1, vy) & ~self.covered(vx + 1, vy)
            oky = ~self.covered(vx, vy - 1) & ~self.covered(vx, vy + 1)
            d = [((g - 2 * v) - 0.5) * 0.5 for g, v in ((gx, vx), (gy, vy))]

            def fc(i, x, y):
                return f0[i, y % self.ny, x % self.nx]

            for i in range(9):
                ex, ey = int(EX[i]), int(EY[i])
                tcov = self.covered(vx + ex, vy + ey)
                base = fc(i, vx, vy)
                dx, dy = (gx + ex) % NF[0], (gy + ey) % NF[1]
                into = self.covered(dx // 2, dy // 2) & (dx // self.rb2 == bx) & (dy // self.rb2 == by)
                val = np.where(into, base, 0.0)
                if self.linear and i > 0:
                    G = [np.where(okx, 0.5 * (fc(i, vx + 1, vy) - fc(i, vx - 1, vy)), 0.0),
                         np.where(oky, 0.5 * (fc(i, vx, vy + 1) - fc(i, vx, vy - 1)), 0.0)]
                    ee = float(ex * ex + ey * ey)
                    proj = (G[0] * ex + G[1] * ey) / ee
                    G = [G[0] - ex * proj, G[1] - ey * proj]
                    lin = base + (d[0] * G[0] + d[1] * G[1])
                else:
                    lin = base
                self.pool[i, slot, fy, fx] = np.where(tcov, lin, val)

    def coalesce(self):
        NF = (2 * self.nx, 2 * self.ny)
        f0 = self.f0
        for slot in active_list(self.slot_to_block):
            bx, by = self.block_xy(int(self.slot_to_block[slot]))
            r = np.arange(self.rb)
            cy, cx = (a.ravel() for a in np.meshgrid(by * self.rb + r, bx * self.rb + r, indexing="ij"))
            f0x, f0y = GHOST + 2 * (cx - bx * self.rb), GHOST + 2 * (cy - by * self.rb)
            rs = np.zeros(cx.size)
            m = [np.zeros(cx.size) for _ in range(2)]
            for k in range(4):
                ox, oy = k & 1, (k >> 1) & 1
                rho = self.mac_pool[0, slot, f0y + oy, f0x + ox]
                rs += rho
                for a in range(2):
                    m[a] += rho * self.mac_pool[1 + a, slot, f0y + oy, f0x + ox]
            inv = 1.0 / np.maximum(rs, 1e-6)
            self.mac0[0, cy, cx] = rs * 0.25
            for a in range(2):
                self.mac0[1 + a, cy, cx] = m[a] * inv
            for i in range(9):
                ex, ey = int(EX[i]), int(EY[i])
                open_ = ~self.covered(cx + ex, cy + ey)
                tx, ty = cx[open_] + ex, cy[open_] + ey
                s = np.zeros(tx.size)
                for k in range(4):
                    px, py = 2 * tx + (k & 1), 2 * ty + ((k >> 1) & 1)
                    lx = self.local_in_tile(px, bx, NF[0])
                    ly = self.local_in_tile(py, by, NF[1])
                    assert all(((v >= 0) & (v < self.fb)).all() for v in (lx, ly))
                    s += self.pool[i, slot, ly, lx]
                f0[i, cy[open_], cx[open_]] = s * 0.25
        if not self.orphans:
            return
        cov = self.covered_mask()
        y, x = np.meshgrid(np.arange(self.ny), np.arange(self.nx), indexing="ij")
        near = np.zeros_like(cov)
        for i in range(1, 9):
            near |= self.covered(x + EX[i], y + EY[i])
        cy, cx = np.nonzero(near & ~cov)
        for i in range(1, 9):
            ex, ey = int(EX[i]), int(EY[i])
            tx, ty = cx + ex, cy + ey
            sel = ~self.covered(tx, ty)
            s = np.zeros(cx.size)
            nin = np.zeros(cx.size)
            for k in range(4):
                ox, oy = k & 1, (k >> 1) & 1
                px, py = (2 * tx + ox) % NF[0], (2 * ty + oy) % NF[1]
                qx, qy = (px - ex) % NF[0], (py - ey) % NF[1]
                qbx, qby = qx // self.rb2, qy // self.rb2
                qs = self.block_slot[self.block_id(qbx, qby)]
                has = qs >= 0
                lx = self.local_in_tile(px, qbx, NF[0])
                ly = self.local_in_tile(py, qby, NF[1])
                idx = np.nonzero(has)[0]
                if idx.size:
                    s[idx] += self.pool[i, qs[idx], ly[idx], lx[idx]]
                hx, hy = (2 * cx + ox + ex) % NF[0], (2 * cy + oy + ey) % NF[1]
                nin += self.covered(hx // 2, hy // 2)
            fc = f0[i, cy, cx]
            net = s - nin * fc
            f0[i, cy[sel], cx[sel]] = (fc + net * 0.25)[sel]

    def interp(self, pairs):
        """Fill whole tiles (ring included) from L0: bilinear rho, u and fneq
        x c2f."""
        n = self.fb
        for b, slot in pairs:
            bx, by = self.block_xy(b)
            fy, fx = (a.ravel() for a in np.meshgrid(np.arange(n), np.arange(n), indexing="ij"))
            P = [o * self.rb - 0.25 + 0.5 * (j - GHOST) for o, j in ((bx, fx), (by, fy))]
            c0 = [np.floor(p).astype(np.int64) for p in P]
            t = [p - c for p, c in zip(P, c0)]
            rho = np.zeros(fx.size)
            u = [np.zeros(fx.size) for _ in range(2)]
            fneq = np.zeros((9, fx.size))
            for k in range(4):
                i, j = k & 1, (k >> 1) & 1
                wgt = (t[0] if i else 1 - t[0]) * (t[1] if j else 1 - t[1])
                f = self.f0[:, (c0[1] + j) % self.ny, (c0[0] + i) % self.nx]
                sr, sx, sy = moments(f)
                rho += wgt * sr
                u[0] += wgt * sx
                u[1] += wgt * sy
                fneq += wgt * (f - feq(self.w, sr, sx, sy))
            self.pool[:, slot, fy, fx] = feq(self.w, rho, u[0], u[1]) + self.c2f * fneq

    def average(self, pairs):
        """4-child restriction of each listed tile onto its L0 cells."""
        for b, slot in pairs:
            bx, by = self.block_xy(b)
            r = np.arange(self.rb)
            ly, lx = (a.ravel() for a in np.meshgrid(r, r, indexing="ij"))
            rs = np.zeros(lx.size)
            m = [np.zeros(lx.size) for _ in range(2)]
            sneq = np.zeros((9, lx.size))
            for c in range(4):
                f = self.pool[:, slot, GHOST + 2 * ly + ((c >> 1) & 1), GHOST + 2 * lx + (c & 1)]
                rho, ux, uy = moments(f)
                rs += rho
                m[0] += rho * ux
                m[1] += rho * uy
                sneq += f - feq(self.w, rho, ux, uy)
            rho = rs * 0.25
            inv = 1.0 / np.maximum(rs, 1e-6)
            ux, uy = m[0] * inv, m[1] * inv
            cx, cy = bx * self.rb + lx, by * self.rb + ly
            self.f0[:, cy, cx] = feq(self.w, rho, ux, uy) + self.f2c * (sneq * 0.25)
            for k, v in enumerate((rho, ux, uy)):
                self.mac0[k, cy, cx] = v

    def criterion(self):
        return block_criterion(vorticity_2d(self.mac0), self.rb, self.lead)

    def manage(self, crit=None):
        """Criterion -> restrict every tile -> decide -> coarsen -> refine ->
        fill (`crit` overrides the criterion to follow the engine's integer
        decisions exactly)."""
        self.crit = self.criterion() if crit is None else np.asarray(crit, dtype=np.float32)
        self.average(self.tiles())
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
        self.explode()
        self.pool_step()
        self.pool_step()
        self.coalesce()
        self.l0_step()
        self.step_count += 1

    # --- read-outs ----------------------------------------------------------
    def totals(self):
        """[mass, px, py] of unrefined L0 cells plus fine interiors x 1/4."""
        cov = self.covered_mask()
        e = np.stack([np.ones(9), EX, EY]).astype(np.float64)
        out = e @ np.where(cov[None], 0.0, self.f0).reshape(9, -1).sum(1)
        lo, hi = GHOST, GHOST + self.rb2
        for s in active_list(self.slot_to_block):
            out = out + 0.25 * (e @ self.pool[:, s, lo:hi, lo:hi].reshape(9, -1).sum(1))
        return out

    def mac(self):
        """(N, 4) [rho, ux, uy, 0] of L0."""
        m = np.zeros((self.ny * self.nx, 4))
        m[:, :3] = np.moveaxis(self.mac0, 0, -1).reshape(-1, 3)
        return m
