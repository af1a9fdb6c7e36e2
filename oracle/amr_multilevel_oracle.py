"""FP64 CPU oracle for a block-structured hierarchy of two to four levels
(dense L0 and pool levels 1 .. L - 1) with the explode/coalesce interface and
a static topology.  TEST INFRASTRUCTURE ONLY: tests/ run it as the checker;
the engine never imports it and has no CPU fallback.

It extends oracle/amr_oracle.py's restatement one rung at a time.  Level m
refines level m - 1 exactly as level 1 refines L0: its blocks are RB^3 cells
of the parent level's global grid (L0's times 2^(m-1)), its tiles have the
same FB^3 shape, and tau_m = 2 tau_(m-1) - 1/2.  What differs at m >= 2 is
only where a parent cell's data lives (the parent level's tile that owns it,
common_d3_parent_pool.wgsl:37-59), so explode and coalesce run here against
a dense view of the parent level's tile interiors (absent cells are NaN) and
write back into those tiles; the interpolation fill reads each parent cell
in the parent tile that owns it, falling back to the ring of the parent tile
that holds the child block where no parent tile owns it.

  schedule            advanceLevel, main-3d.js:3169-3180: explode level m +
                      1 from level m, two substeps of level m + 1, coalesce
                      it back, then level m's own substep
  covered cells       coveredByChild, common_d3_amr_step1.wgsl:107-143: an
                      interior cell under a finer tile is not solved (its
                      populations are the child's outbox; here NaN, so that a
                      read the scheme does not allow poisons the result) and
                      its moments are the child's coalesced projection
  body                on the finest level alone, positions mapped to L0 by
                      L0_SCALE = 2^-(m-1), L0_OFFSET = -(1 - 2^-(m-1))/2
                      (common_d3_amr_step1.wgsl:64-84), force weight dx^2 =
                      4^-m (common_d3_force_pool.wgsl DX_WEIGHT)
  gradient presence   explode's linear gradient needs both parent neighbours
                      present, else it is uniform along that axis
                      (parentPresent, common_d3_parent_pool.wgsl:31-35)

With two levels every pass is AmrOracle's, and tests/test_amr_oracle.py
checks that the two classes agree bit for bit.  Deeper hierarchies are pinned
by the conservation of the composite mass and momentum (explode/coalesce are
exactly conservative), a uniform flow staying uniform, and by a three-level
hierarchy whose third level has no tile reproducing the two-level run.
"""
import numpy as np

import math

from amr_oracle import (GHOST, Tables, active_list, block_criterion, collide, feq, manage_slots, moments, near_body,
                        q_criterion, vorticity_magnitude)


class Level:
    """Pool level m >= 1 over its parent grid (nx, ny, nz cells)."""

    def __init__(self, m, q, nx, ny, nz, rb, max_slots, omega_f, c2f, f2c):
        self.m = m
        self.q = q
        self.nx, self.ny, self.nz = nx, ny, nz
        self.rb, self.rb2 = rb, 2 * rb
        self.fb = 2 * rb + 2 * GHOST
        self.nbx, self.nby, self.nbz = nx // rb, ny // rb, nz // rb
        self.nblocks = self.nbx * self.nby * self.nbz
        self.max_slots = max_slots
        self.omega_f = omega_f
        self.c2f = c2f
        self.f2c = f2c
        self.scale = 2.0 ** -(m - 1)
        self.offset = -0.5 * (1.0 - self.scale)
        self.dx2 = 4.0 ** -m
        self.block_slot = np.full(self.nblocks, -1, dtype=np.int32)
        self.slot_to_block = np.full(max_slots, -1, dtype=np.int32)
        n = self.fb
        self.pool = np.zeros((q, max_slots, n, n, n))
        self.mac_pool = np.zeros((4, max_slots, n, n, n))

    def block_xyz(self, b):
        return b % self.nbx, (b // self.nbx) % self.nby, b // (self.nbx * self.nby)

    def block_id(self, bx, by, bz):
        return (bz * self.nby + by) * self.nbx + bx

    def covered(self, cx, cy, cz):
        """Is parent cell (cx, cy, cz) (arrays, wrapped) under a tile of this level."""
        cx, cy, cz = np.asarray(cx) % self.nx, np.asarray(cy) % self.ny, np.asarray(cz) % self.nz
        return self.block_slot[self.block_id(cx // self.rb, cy // self.rb, cz // self.rb)] >= 0

    def covered_mask(self):
        z, y, x = np.meshgrid(np.arange(self.nz), np.arange(self.ny), np.arange(self.nx), indexing="ij")
        return self.covered(x, y, z)

    def local_in_tile(self, p, b, nf):
        d = np.asarray(p) - b * self.rb2
        d = np.where(d < -(nf // 2), d + nf, d)
        d = np.where(d > nf // 2, d - nf, d)
        return d + GHOST

    def tiles(self):
        return [(int(self.slot_to_block[s]), int(s)) for s in active_list(self.slot_to_block)]

    def tile_positions(self, slot):
        """L0-unit centres (float32, exact dyadic values) of every cell of
        tile `slot`: (origin - 1/4 + (j - GHOST)/2) * scale + offset."""
        bx, by, bz = self.block_xyz(int(self.slot_to_block[slot]))
        j = np.arange(self.fb, dtype=np.float32)
        h = np.float32(0.5) * (j - np.float32(GHOST)) - np.float32(0.25)
        sc, of = np.float32(self.scale), np.float32(self.offset)
        z, y, x = np.meshgrid((np.float32(bz * self.rb) + h) * sc + of, (np.float32(by * self.rb) + h) * sc + of,
                              (np.float32(bx * self.rb) + h) * sc + of, indexing="ij")
        return x, y, z


class MultiLevelOracle:
    """desc keys: q, nx, ny, nz, rb, maxSlots, omega, omegaMinus (TRT on L0 if
    present), force, levels (2..4), explodeLinear, orphans, exactWeights,
    body ({a, centre}: a pinned sphere on the finest level, L0 units)."""

    def __init__(self, desc):
        self.T = T = Tables(int(desc["q"]), bool(desc.get("exactWeights", False)))
        self.q = T.q
        self.nx, self.ny, self.nz = int(desc["nx"]), int(desc["ny"]), int(desc["nz"])
        self.rb = int(desc["rb"])
        self.levels = int(desc.get("levels", 2))
        self.max_slots = int(desc["maxSlots"])
        self.omega = float(desc["omega"])
        om = desc.get("omegaMinus")
        self.omega_m = self.omega if om is None else float(om)
        self.F = tuple(float(v) for v in (list(desc.get("force", [0, 0, 0])) + [0, 0, 0])[:3])
        self.linear = bool(desc.get("explodeLinear", 1))
        self.orphans = bool(desc.get("orphans", 1))
        body = desc.get("body")
        self.body = (np.array(body["centre"], dtype=np.float32), np.float32(body["a"])) if body else None
        self.lv = [None]
        tau = 1.0 / self.omega
        for m in range(1, self.levels):
            tf = 2.0 * tau - 0.5
            s = 1 << (m - 1)
            self.lv.append(Level(m, self.q, self.nx * s, self.ny * s, self.nz * s, self.rb, self.max_slots,
                                 1.0 / tf, 0.5 * (tf - 1.0) / (tau - 1.0), 2.0 * (tau - 1.0) / (tf - 1.0)))
            tau = tf
        self.forces = []
        self.f0 = None
        self.mac0 = None
        # the manager
        self.manage_every = int(desc.get("manageEvery", 0))
        self.refine_at = float(np.float32(desc.get("refineAt", 1e30)))
        self.keep_at = float(np.float32(desc.get("keepAt", 1e30)))
        self.lead = int(desc.get("lead", 0))
        self.criterion_kind = desc.get("criterion", "vorticity")
        oct_ = float(np.float32(desc.get("ladderOctaves", 0.0)))
        self.ladder_octaves = oct_ if oct_ > 0 else 1.0
        sp = desc.get("sponge") or {}
        self.sponge_width = float(np.float32(sp.get("width", 0.0)))
        self.sponge_axes = [int(bool(v)) for v in sp.get("axes", [1, 1, 1])]
        self.geom = [None] + [self.level_mask(m) for m in range(1, self.levels)] if self.manage_every > 0 else None
        self.step_count = 0
        self.exhausted = 0
        self.band = 0
        self.crit = None
        self.wants = None
        self.last_fresh = None

    # --- set-up -------------------------------------------------------------
    def upload_topology(self, level, block_slot, slot_to_block):
        L = self.lv[level]
        L.block_slot = np.array(block_slot, dtype=np.int32)
        L.slot_to_block = np.array(slot_to_block, dtype=np.int32)

    def upload_macro(self, mac4):
        m = np.asarray(mac4, dtype=np.float32).astype(np.float64).reshape(self.nz, self.ny, self.nx, 4)
        self.mac0 = np.moveaxis(m, -1, 0).copy()
        self.f0 = feq(self.T, self.mac0[0], self.mac0[1], self.mac0[2], self.mac0[3])
        for lvl in range(1, self.levels):
            self.interp(lvl)

    # --- parent access --------------------------------------------------------
    def _present(self, m):
        """Parent-grid mask of the cells that exist at level m - 1."""
        if m == 1:
            return np.ones((self.nz, self.ny, self.nx), dtype=bool)
        P = self.lv[m - 1]
        C = self.lv[m]
        z, y, x = np.meshgrid(np.arange(C.nz), np.arange(C.ny), np.arange(C.nx), indexing="ij")
        r2 = P.rb2
        return P.block_slot[P.block_id(x // r2, y // r2, z // r2)] >= 0

    def _parent_view(self, m):
        """Dense (f, mac) of level m's parent over its grid; a pool parent's
        cells outside its tiles are NaN."""
        if m == 1:
            return self.f0, self.mac0
        P, C = self.lv[m - 1], self.lv[m]
        f = np.full((self.q, C.nz, C.ny, C.nx), np.nan)
        mac = np.full((4, C.nz, C.ny, C.nx), np.nan)
        lo, hi = GHOST, GHOST + P.rb2
        for b, s in P.tiles():
            bx, by, bz = P.block_xyz(b)
            sl = (slice(bz * P.rb2, (bz + 1) * P.rb2), slice(by * P.rb2, (by + 1) * P.rb2),
                  slice(bx * P.rb2, (bx + 1) * P.rb2))
            f[(slice(None),) + sl] = P.pool[:, s, lo:hi, lo:hi, lo:hi]
            mac[(slice(None),) + sl] = P.mac_pool[:, s, lo:hi, lo:hi, lo:hi]
        return f, mac

    def _write_parent(self, m, f, mac):
        """Writes a pool parent's view back into its tile interiors."""
        if m == 1:
            return
        P = self.lv[m - 1]
        lo, hi = GHOST, GHOST + P.rb2
        for b, s in P.tiles():
            bx, by, bz = P.block_xyz(b)
            sl = (slice(bz * P.rb2, (bz + 1) * P.rb2), slice(by * P.rb2, (by + 1) * P.rb2),
                  slice(bx * P.rb2, (bx + 1) * P.rb2))
            P.pool[:, s, lo:hi, lo:hi, lo:hi] = f[(slice(None),) + sl]
            P.mac_pool[:, s, lo:hi, lo:hi, lo:hi] = mac[(slice(None),) + sl]

    # --- passes -------------------------------------------------------------
    def interp(self, m, owner=True, pairs=None):
        """Fills every tile of level m from its parent: trilinear rho, u and
        fneq x c2f (common_d3_pool.wgsl:172-220); a pool parent cell is read
        in the parent tile that owns it (parentIndex, common_d3_parent_pool.
        wgsl:37-59), or, where none does, in the ring of the parent tile that
        holds the block.  owner=False reads the holding tile's ring alone
        (the variant that the tests show to be the same at a seed).  pairs:
        the (block, slot) tiles to fill (default every tile)."""
        C = self.lv[m]
        n = C.fb
        for b, slot in (C.tiles() if pairs is None else pairs):
            bx, by, bz = C.block_xyz(b)
            fz, fy, fx = (a.ravel() for a in np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"))
            Pp = [o * C.rb - 0.25 + 0.5 * (j - GHOST) for o, j in ((bx, fx), (by, fy), (bz, fz))]
            c0 = [np.floor(p).astype(np.int64) for p in Pp]
            t = [p - c for p, c in zip(Pp, c0)]
            rho = np.zeros(fx.size)
            u = [np.zeros(fx.size) for _ in range(3)]
            fneq = np.zeros((self.q, fx.size))
            if m >= 2:
                P = self.lv[m - 1]
                pb = [(o * C.rb) // P.rb2 for o in (bx, by, bz)]
                ps = int(P.block_slot[P.block_id(*pb)])
                assert ps >= 0, "a tile's parent block holds no tile"
            for k in range(8):
                i, j, kk = k & 1, (k >> 1) & 1, (k >> 2) & 1
                w = (t[0] if i else 1 - t[0]) * (t[1] if j else 1 - t[1]) * (t[2] if kk else 1 - t[2])
                cx, cy, cz = c0[0] + i, c0[1] + j, c0[2] + kk
                if m == 1:
                    f = self.f0[:, cz % self.nz, cy % self.ny, cx % self.nx]
                else:
                    # the owning parent tile's interior; the holding tile's
                    # ring where no parent tile owns the cell
                    vx, vy, vz = cx % C.nx, cy % C.ny, cz % C.nz
                    os_ = P.block_slot[P.block_id(vx // P.rb2, vy // P.rb2, vz // P.rb2)]
                    own = os_ >= 0
                    lx, ly, lz = (c - p * P.rb2 + GHOST for c, p in zip((cx, cy, cz), pb))
                    assert all(((v >= 0) & (v < P.fb)).all() for v in (lx, ly, lz))
                    f = P.pool[:, ps, lz, ly, lx]
                    if owner and own.any():
                        ox, oy, oz = (v % P.rb2 + GHOST for v in (vx, vy, vz))
                        f = np.where(own[None], P.pool[:, np.where(own, os_, 0), oz, oy, ox], f)
                sr, sx, sy, sz = moments(self.T, f)
                rho += w * sr
                u[0] += w * sx
                u[1] += w * sy
                u[2] += w * sz
                fneq += w * (f - feq(self.T, sr, sx, sy, sz))
            C.pool[:, slot, fz, fy, fx] = feq(self.T, rho, u[0], u[1], u[2]) + C.c2f * fneq

    def l0_step(self):
        T = self.T
        fs = np.empty_like(self.f0)
        for i in range(T.q):
            fs[i] = np.roll(self.f0[i], (T.ez[i], T.ey[i], T.ex[i]), (0, 1, 2))
        rho, ux, uy, uz = moments(T, fs, self.F)
        out = collide(T, fs, rho, ux, uy, uz, self.omega, self.omega_m, self.F)
        cov = self.lv[1].covered_mask() if self.levels > 1 else np.zeros(rho.shape, dtype=bool)
        out[:, cov] = np.nan
        for k, v in enumerate((rho, ux, uy, uz)):
            self.mac0[k] = np.where(cov, self.mac0[k], v)
        self.f0 = out

    def _phi(self, x, y, z):
        c, a = self.body
        dx, dy, dz = (np.asarray(v, dtype=np.float32) - c[k] for k, v in enumerate((x, y, z)))
        return np.sqrt(dx * dx + dy * dy + dz * dz) - a

    def _tile_sources(self, L, slot):
        n, lo, hi = L.fb, GHOST, GHOST + L.rb2
        bx, by, bz = L.block_xyz(int(L.slot_to_block[slot]))
        fz, fy, fx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        srcs = []
        for i in range(self.q):
            s = [fx - self.T.ex[i], fy - self.T.ey[i], fz - self.T.ez[i]]
            o = [np.where(v < lo, -1, np.where(v >= hi, 1, 0)) for v in s]
            nb = L.block_id((bx + o[0]) % L.nbx, (by + o[1]) % L.nby, (bz + o[2]) % L.nbz)
            ns = L.block_slot[nb]
            other = ((o[0] != 0) | (o[1] != 0) | (o[2] != 0)) & (ns >= 0)
            src_slot = np.where(other, ns, slot)
            c = [np.where(other, s[k] - o[k] * L.rb2, np.clip(s[k], 0, n - 1)) for k in range(3)]
            srcs.append(((src_slot * n + c[2]) * n + c[1]) * n + c[0])
        return srcs

    def pool_step(self, m):
        """One substep of level m (AmrOracle.pool_step), leaving the interior
        cells a finer tile covers unsolved (NaN) with their coalesced moments."""
        T, L = self.T, self.lv[m]
        n = L.fb
        flat = L.pool.reshape(self.q, -1)
        new = L.pool.copy()
        lo, hi = GHOST, GHOST + L.rb2
        r = np.arange(n)
        inside = (r >= lo) & (r < hi)
        interior = inside[:, None, None] & inside[None, :, None] & inside[None, None, :]
        child = self.lv[m + 1] if m + 1 < self.levels else None
        has_body = self.body is not None and m == self.levels - 1
        for slot in active_list(L.slot_to_block):
            srcs = self._tile_sources(L, int(slot))
            g = np.stack([flat[i][srcs[i]] for i in range(self.q)])
            if has_body:
                x, y, z = L.tile_positions(int(slot))
                solid = self._phi(x, y, z) < 0
                sc = np.float32(0.5 * L.scale)
                for i in range(1, T.q):
                    hit = self._phi(x - sc * np.float32(T.ex[i]), y - sc * np.float32(T.ey[i]),
                                    z - sc * np.float32(T.ez[i])) < 0
                    g[i] = np.where(hit, L.pool[T.opp[i], slot], g[i])
            rho, ux, uy, uz = moments(T, g, self.F)
            cov = np.zeros((n, n, n), dtype=bool)
            if child is not None:
                bx, by, bz = L.block_xyz(int(L.slot_to_block[slot]))
                fz, fy, fx = np.meshgrid(r, r, r, indexing="ij")
                gx, gy, gz = bx * L.rb2 + fx - GHOST, by * L.rb2 + fy - GHOST, bz * L.rb2 + fz - GHOST
                cov = interior & child.covered(gx, gy, gz)
            for k, v in enumerate((rho, ux, uy, uz)):
                L.mac_pool[k, slot] = np.where(cov, L.mac_pool[k, slot], v)
            col = collide(T, g, rho, ux, uy, uz, L.omega_f, L.omega_f, self.F)
            out = np.where(interior, col, g)
            if has_body:
                out = np.where(solid, T.w.reshape(-1, 1, 1, 1), out)
                for k, v in enumerate((1.0, 0.0, 0.0, 0.0)):
                    L.mac_pool[k, slot] = np.where(solid, v, L.mac_pool[k, slot])
            new[:, slot] = np.where(cov, np.nan, out)
        L.pool = new

    def explode(self, m):
        """Level m's ring from its parent's state (AmrOracle.explode against
        the parent view; a gradient axis also needs both parent neighbours
        present)."""
        T, C = self.T, self.lv[m]
        n, lo, hi = C.fb, GHOST, GHOST + C.rb2
        NF = (2 * C.nx, 2 * C.ny, 2 * C.nz)
        f0, _ = self._parent_view(m)
        pres = self._present(m)

        def present(x, y, z):
            return pres[z % C.nz, y % C.ny, x % C.nx]

        for slot in active_list(C.slot_to_block):
            b = int(C.slot_to_block[slot])
            bx, by, bz = C.block_xyz(b)
            fz, fy, fx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
            ring = ~((fx >= lo) & (fx < hi) & (fy >= lo) & (fy < hi) & (fz >= lo) & (fz < hi))
            fx, fy, fz = fx[ring], fy[ring], fz[ring]
            gx = (bx * C.rb2 + fx - GHOST) % NF[0]
            gy = (by * C.rb2 + fy - GHOST) % NF[1]
            gz = (bz * C.rb2 + fz - GHOST) % NF[2]
            vx, vy, vz = gx // 2, gy // 2, gz // 2
            keep = ~C.covered(vx, vy, vz)
            fx, fy, fz, gx, gy, gz, vx, vy, vz = (a[keep] for a in (fx, fy, fz, gx, gy, gz, vx, vy, vz))
            assert present(vx, vy, vz).all(), "a ring cell's parent cell lies in no parent tile"

            def ok(ax):
                d = [0, 0, 0]
                d[ax] = 1
                a = (vx - d[0], vy - d[1], vz - d[2])
                p = (vx + d[0], vy + d[1], vz + d[2])
                return ~C.covered(*a) & ~C.covered(*p) & present(*a) & present(*p)

            okx, oky, okz = ok(0), ok(1), ok(2)
            d = [((g - 2 * v) - 0.5) * 0.5 for g, v in ((gx, vx), (gy, vy), (gz, vz))]

            def fc(i, x, y, z):
                return f0[i, z % C.nz, y % C.ny, x % C.nx]

            for i in range(T.q):
                ex, ey, ez = int(T.ex[i]), int(T.ey[i]), int(T.ez[i])
                tcov = C.covered(vx + ex, vy + ey, vz + ez)
                base = fc(i, vx, vy, vz)
                dx, dy, dz = (gx + ex) % NF[0], (gy + ey) % NF[1], (gz + ez) % NF[2]
                into = (C.covered(dx // 2, dy // 2, dz // 2) & (dx // C.rb2 == bx) & (dy // C.rb2 == by)
                        & (dz // C.rb2 == bz))
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
                C.pool[i, slot, fz, fy, fx] = np.where(tcov, lin, val)

    def coalesce(self, m):
        """Level m's ring outflux into its parent's state at the same time
        level, mac projection, orphans (AmrOracle.coalesce against the parent
        view, written back into a pool parent's tiles)."""
        T, C = self.T, self.lv[m]
        NF = (2 * C.nx, 2 * C.ny, 2 * C.nz)
        f0, mac0 = self._parent_view(m)
        pres = self._present(m)
        for slot in active_list(C.slot_to_block):
            b = int(C.slot_to_block[slot])
            bx, by, bz = C.block_xyz(b)
            r = np.arange(C.rb)
            cz, cy, cx = (a.ravel() for a in np.meshgrid(bz * C.rb + r, by * C.rb + r, bx * C.rb + r,
                                                         indexing="ij"))
            f0x, f0y, f0z = GHOST + 2 * (cx - bx * C.rb), GHOST + 2 * (cy - by * C.rb), GHOST + 2 * (cz - bz * C.rb)
            rs = np.zeros(cx.size)
            mm = [np.zeros(cx.size) for _ in range(3)]
            for k in range(8):
                ox, oy, oz = k & 1, (k >> 1) & 1, (k >> 2) & 1
                rho = C.mac_pool[0, slot, f0z + oz, f0y + oy, f0x + ox]
                rs += rho
                for a in range(3):
                    mm[a] += rho * C.mac_pool[1 + a, slot, f0z + oz, f0y + oy, f0x + ox]
            inv = 1.0 / np.maximum(rs, 1e-6)
            mac0[0, cz, cy, cx] = rs * 0.125
            for a in range(3):
                mac0[1 + a, cz, cy, cx] = mm[a] * inv
            for i in range(T.q):
                ex, ey, ez = int(T.ex[i]), int(T.ey[i]), int(T.ez[i])
                open_ = ~C.covered(cx + ex, cy + ey, cz + ez)
                tx, ty, tz = cx[open_] + ex, cy[open_] + ey, cz[open_] + ez
                s = np.zeros(tx.size)
                for k in range(8):
                    px, py, pz = 2 * tx + (k & 1), 2 * ty + ((k >> 1) & 1), 2 * tz + ((k >> 2) & 1)
                    lx = C.local_in_tile(px, bx, NF[0])
                    ly = C.local_in_tile(py, by, NF[1])
                    lz = C.local_in_tile(pz, bz, NF[2])
                    s += C.pool[i, slot, lz, ly, lx]
                f0[i, cz[open_], cy[open_], cx[open_]] = s * 0.125
        if self.orphans:
            cov = C.covered_mask()
            z, y, x = np.meshgrid(np.arange(C.nz), np.arange(C.ny), np.arange(C.nx), indexing="ij")
            near = np.zeros_like(cov)
            for i in range(1, T.q):
                near |= C.covered(x + T.ex[i], y + T.ey[i], z + T.ez[i])
            cz, cy, cx = np.nonzero(near & ~cov & pres)
            for i in range(1, T.q):
                ex, ey, ez = int(T.ex[i]), int(T.ey[i]), int(T.ez[i])
                tx, ty, tz = cx + ex, cy + ey, cz + ez
                sel = ~C.covered(tx, ty, tz)
                s = np.zeros(cx.size)
                nin = np.zeros(cx.size)
                for k in range(8):
                    ox, oy, oz = k & 1, (k >> 1) & 1, (k >> 2) & 1
                    px, py, pz = (2 * tx + ox) % NF[0], (2 * ty + oy) % NF[1], (2 * tz + oz) % NF[2]
                    qx, qy, qz = (px - ex) % NF[0], (py - ey) % NF[1], (pz - ez) % NF[2]
                    qbx, qby, qbz = qx // C.rb2, qy // C.rb2, qz // C.rb2
                    qs = C.block_slot[C.block_id(qbx, qby, qbz)]
                    has = qs >= 0
                    lx = C.local_in_tile(px, qbx, NF[0])
                    ly = C.local_in_tile(py, qby, NF[1])
                    lz = C.local_in_tile(pz, qbz, NF[2])
                    idx = np.nonzero(has)[0]
                    if idx.size:
                        s[idx] += C.pool[i, qs[idx], lz[idx], ly[idx], lx[idx]]
                    hx, hy, hz = (2 * cx + ox + ex) % NF[0], (2 * cy + oy + ey) % NF[1], (2 * cz + oz + ez) % NF[2]
                    nin += C.covered(hx // 2, hy // 2, hz // 2)
                fc = f0[i, cz, cy, cx]
                net = s - nin * fc
                f0[i, cz[sel], cy[sel], cx[sel]] = (fc + net * 0.125)[sel]
        self._write_parent(m, f0, mac0)

    def body_force(self):
        """The macro-step's force pass on the finest level's state at t: every
        interior fluid cell adds -e_i 2 f_opp(i) per solid geometric source,
        weighted dx^2 = 4^-m, torque about the centre in L0 units."""
        T, L = self.T, self.lv[self.levels - 1]
        lo, hi = GHOST, GHOST + L.rb2
        c = self.body[0].astype(np.float64)
        out = np.zeros(6)
        sc = np.float32(0.5 * L.scale)
        for slot in active_list(L.slot_to_block):
            x, y, z = (v[lo:hi, lo:hi, lo:hi] for v in L.tile_positions(int(slot)))
            fluid = self._phi(x, y, z) >= 0
            fsum = np.zeros((3,) + x.shape)
            for i in range(1, T.q):
                e = (T.ex[i], T.ey[i], T.ez[i])
                src = fluid & (self._phi(x - sc * np.float32(e[0]), y - sc * np.float32(e[1]),
                                         z - sc * np.float32(e[2])) < 0)
                mval = 2.0 * L.pool[T.opp[i], slot, lo:hi, lo:hi, lo:hi]
                for k in range(3):
                    fsum[k] -= np.where(src, e[k] * mval, 0.0)
            fsum *= L.dx2
            rr = (x.astype(np.float64) - c[0], y.astype(np.float64) - c[1], z.astype(np.float64) - c[2])
            out[:3] += fsum.reshape(3, -1).sum(1)
            out[3] += (rr[1] * fsum[2] - rr[2] * fsum[1]).sum()
            out[4] += (rr[2] * fsum[0] - rr[0] * fsum[2]).sum()
            out[5] += (rr[0] * fsum[1] - rr[1] * fsum[0]).sum()
        self.forces.append(out)

    def advance(self, m):
        if m + 1 < self.levels and active_list(self.lv[m + 1].slot_to_block).size:
            self.explode(m + 1)
            self.advance(m + 1)
            self.advance(m + 1)
            self.coalesce(m + 1)
        self.pool_step(m)

    def macro_step(self, crit=None):
        if self.manage_every > 0 and self.step_count % self.manage_every == 0:
            self.manage(crit)
        self.step_count += 1
        if self.body is not None:
            self.body_force()
        self.explode(1)
        self.advance(1)
        self.advance(1)
        self.coalesce(1)
        self.l0_step()

    # --- the manager ----------------------------------------------------------
    def level_mask(self, m):
        """Per block of level m: bit 0 (the finest level, with a body) the
        body's geometry wants it (nearBodyWant at margin 2 on the block's L0
        box, scaled by 2^-(m-1) as refineHierarchy scales it, d3-amr.mjs:
        603-651); bit 1 the field may not refine it: its L0 box comes within
        the sponge band dilated by 3 rb (1 - 2^-(m-1)) L0 cells, the most the
        2:1 closure can carry a level-m want toward level 1."""
        L = self.lv[m]
        sc = 2.0 ** -(m - 1)
        edge = self.rb * sc
        dil = 3.0 * self.rb * (1.0 - sc)
        dims = (self.nx, self.ny, self.nz)
        out = np.zeros(L.nblocks, dtype=np.uint8)
        body = self.body is not None and m == self.levels - 1
        if body:
            c = [float(v) for v in self.body[0]]
            r = float(self.body[1])
            sdf = lambda p: math.sqrt(sum((p[k] - c[k]) ** 2 for k in range(3))) - r
        for b in range(L.nblocks):
            bc = L.block_xyz(b)
            lo = [bc[k] * edge for k in range(3)]
            if body and near_body(sdf, lo, [v + edge for v in lo], 2.0):
                out[b] |= 1
            if self.sponge_width > 0:
                for k in range(3):
                    if self.sponge_axes[k] and (lo[k] < self.sponge_width + dil
                                                or lo[k] + edge > dims[k] - self.sponge_width - dil):
                        out[b] |= 2
        return out

    def criterion(self):
        """Per level-1 block maximum of |curl u| (or Q) over the L0 field,
        binary32 as the engine computes it."""
        field = q_criterion if self.criterion_kind == "q" else vorticity_magnitude
        return block_criterion(field(self.mac0.astype(np.float32), np.float32), self.rb, self.lead)

    def rung(self, m):
        """Level m's (refine, keep) thresholds: the ladder desiredLevelAt
        (common_refine.wgsl:54-79) asks whether the criterion's level reaches
        m, i.e. rate >= base x 2^(octaves (m - 1)); Q is a squared rate."""
        f = 2.0 ** (self.ladder_octaves * (2.0 if self.criterion_kind == "q" else 1.0) * (m - 1))
        return np.float32(self.refine_at * f), np.float32(self.keep_at * f)

    def decide(self, crit):
        """Every level's want (level_wants), closed (close)."""
        return self.close(self.level_wants(crit))

    def level_wants(self, crit):
        """Each level's own want: the body's geometry first, then the field
        on the level's rung, a held block judged against the keep threshold
        (common_d3_manage.wgsl decide)."""
        wants = [None]
        crit = np.asarray(crit, dtype=np.float32)
        L1 = self.lv[1]
        for m in range(1, self.levels):
            L = self.lv[m]
            refine, keep = self.rung(m)
            g = self.geom[m]
            b = np.arange(L.nblocks)
            bx, by, bz = L.block_xyz(b)
            sh = m - 1
            c = crit[((bz >> sh) * L1.nby + (by >> sh)) * L1.nbx + (bx >> sh)]
            field = np.where(L.block_slot >= 0, c >= keep, c >= refine) & ((g & 2) == 0)
            wants.append((((g & 1) != 0) | field).astype(np.uint8))
        return wants

    def close(self, wants):
        """The 2:1 closure finest-first: octet completion at m >= 2
        (completeOctets), then the parent gather over the child blocks
        [2p - 1, 2p + 2]^3 (balance, common_d3_manage.wgsl:318-372), which
        ORs into the parent level's own want."""
        wants = [None] + [np.array(w, dtype=np.uint8) for w in wants[1:]]
        for m in range(self.levels - 1, 1, -1):
            L, P = self.lv[m], self.lv[m - 1]
            w = wants[m].reshape(L.nbz, L.nby, L.nbx)
            o = w.reshape(L.nbz // 2, 2, L.nby // 2, 2, L.nbx // 2, 2).max(axis=(1, 3, 5))
            w = np.repeat(np.repeat(np.repeat(o, 2, 0), 2, 1), 2, 2)
            wants[m] = w.reshape(-1).copy()
            gathered = np.zeros((P.nbz, P.nby, P.nbx), dtype=np.uint8)
            for dz in (-1, 0, 1, 2):
                for dy in (-1, 0, 1, 2):
                    for dx in (-1, 0, 1, 2):
                        # parent p takes child 2p + d (periodic on the child grid)
                        sub = np.roll(w, (-dz, -dy, -dx), (0, 1, 2))[::2, ::2, ::2]
                        gathered |= sub
            wants[m - 1] = wants[m - 1] | gathered.reshape(-1)
        return wants

    def _average(self, m, pairs):
        """Restriction of level m's listed tiles onto their parent cells
        (AmrOracle.average against the parent level: rho mean, mass-weighted
        u, fneq mean x 2 (tau_(m-1) - 1)/(tau_m - 1)), a pool parent's cells
        written in the tile that owns them."""
        T, C = self.T, self.lv[m]
        P = self.lv[m - 1] if m >= 2 else None
        for b, slot in pairs:
            bx, by, bz = C.block_xyz(b)
            r = np.arange(C.rb)
            lz, ly, lx = (a.ravel() for a in np.meshgrid(r, r, r, indexing="ij"))
            rs = np.zeros(lx.size)
            mm = [np.zeros(lx.size) for _ in range(3)]
            sneq = np.zeros((self.q, lx.size))
            for k in range(8):
                f = C.pool[:, slot, GHOST + 2 * lz + ((k >> 2) & 1), GHOST + 2 * ly + ((k >> 1) & 1),
                           GHOST + 2 * lx + (k & 1)]
                rho, ux, uy, uz = moments(T, f)
                rs += rho
                mm[0] += rho * ux
                mm[1] += rho * uy
                mm[2] += rho * uz
                sneq += f - feq(T, rho, ux, uy, uz)
            rho = rs * 0.125
            inv = 1.0 / np.maximum(rs, 1e-6)
            u = [v * inv for v in mm]
            out = feq(T, rho, u[0], u[1], u[2]) + C.f2c * (sneq * 0.125)
            cx, cy, cz = bx * C.rb + lx, by * C.rb + ly, bz * C.rb + lz
            if P is None:
                self.f0[:, cz, cy, cx] = out
                for k, v in enumerate((rho, u[0], u[1], u[2])):
                    self.mac0[k, cz, cy, cx] = v
            else:
                ps = P.block_slot[P.block_id(cx // P.rb2, cy // P.rb2, cz // P.rb2)]
                assert (ps >= 0).all(), "a tile's parent cells lie in parent tiles"
                ox, oy, oz = cx % P.rb2 + GHOST, cy % P.rb2 + GHOST, cz % P.rb2 + GHOST
                P.pool[:, ps, oz, oy, ox] = out
                for k, v in enumerate((rho, u[0], u[1], u[2])):
                    P.mac_pool[k, ps, oz, oy, ox] = v

    def manage(self, crit=None):
        """The manager over the hierarchy (main-3d.js:3251-3321): criterion,
        every tile restricted into its parent finest-first, every level's
        want and the closure, each level's slots (coarsen, then the k-th new
        block in ascending id takes the k-th free slot), the new tiles filled
        from their parents coarsest-first.  `crit` overrides the criterion
        (the engine's own values, to follow its decisions exactly)."""
        self.crit = self.criterion() if crit is None else np.asarray(crit, dtype=np.float32)
        for m in range(self.levels - 1, 0, -1):
            self._average(m, self.lv[m].tiles())
        self.wants = self.decide(self.crit)
        self.band += int((self.wants[1] & ((self.geom[1] & 2) != 0)).sum())
        fresh = [None]
        for m in range(1, self.levels):
            L = self.lv[m]
            bs, sb, _, fr, exh = manage_slots(L.block_slot, L.slot_to_block, self.wants[m])
            L.block_slot, L.slot_to_block = bs, sb
            fresh.append(fr)
            self.exhausted += exh
        for m in range(1, self.levels):
            if fresh[m]:
                self.interp(m, pairs=fresh[m])
        self.last_fresh = fresh

    # --- read-back ------------------------------------------------------------
    def totals(self):
        """[mass, px, py, pz]: unrefined L0 cells plus, per level m, the
        interior cells no finer tile covers, weighted 8^-m."""
        T = self.T
        e = np.stack([np.ones(T.q), T.ex, T.ey, T.ez]).astype(np.float64)
        cov = self.lv[1].covered_mask()
        out = e @ np.where(cov[None], 0.0, self.f0).reshape(T.q, -1).sum(1)
        for m in range(1, self.levels):
            L = self.lv[m]
            child = self.lv[m + 1] if m + 1 < self.levels else None
            lo, hi = GHOST, GHOST + L.rb2
            r = np.arange(L.rb2)
            for b, s in L.tiles():
                inner = L.pool[:, s, lo:hi, lo:hi, lo:hi]
                if child is not None:
                    bx, by, bz = L.block_xyz(b)
                    lz, ly, lx = np.meshgrid(r, r, r, indexing="ij")
                    c = child.covered(bx * L.rb2 + lx, by * L.rb2 + ly, bz * L.rb2 + lz)
                    inner = np.where(c[None], 0.0, inner)
                out = out + 8.0 ** -m * (e @ inner.reshape(T.q, -1).sum(1))
        return out

    def level_f(self, m):
        """Level m's populations in the engine's readback layout [q][slots * FB^3]."""
        return self.lv[m].pool.reshape(self.q, -1)

    def level_mac(self, m):
        return self.lv[m].mac_pool.reshape(4, -1)

    def mac(self):
        return np.moveaxis(self.mac0, 0, -1).reshape(-1, 4)
