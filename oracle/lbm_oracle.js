'use strict';
// FP64 CPU ORACLE for the dense lattice-Boltzmann step.  TEST INFRASTRUCTURE
// ONLY: tests/, the smoke check and bench.py's cpu_baseline leg run it as the
// checker.  The engine never calls it and never falls back to it.
//
// It restates, in double precision, the fused pull-stream + collide kernels of
// the reference:
//   3D  shaders/common_d3_step.wgsl:206-222 (initEq) and :224-400 (step)
//   2D  shaders/lbm_step.wgsl:137-270 (main)
//   TRT plans/TRT.md:333-349 (delta form, sec 8.0 item 1) and :406-415
//   MEM force shaders/common_d3_force.wgsl:171-283, shaders/lbm_force.wgsl:91-208
// The solid classification (which cells are inside the body, which cells can
// have a solid neighbour) is integer output and is computed with emulated f32
// arithmetic, operation by operation, exactly as the HIP classification
// kernel does, so the two agree bit for bit.
//
// The pinned body's velocity is zero, so Ladd's moving-wall term vanishes on
// its links.  A moving rigid body (rigid_oracle.js) passes its own link masks,
// body velocity and window through stepSlab's `R`.  Moving walls (the 2D
// Couette form, shaders/common_walls.wgsl:27-31) keep their correction term.
//
// Parity pins for this file live in tests/js/test-oracle.js (closed forms,
// invariants and the sponge profile's table) and tests/test_amr_oracle.py
// (the AMR oracle's no-refinement and all-refined limits against this one).

// --- lattice tables (restated; lattice-2d.mjs:97-145, lattice-3d.mjs:146-200)
function d3Tables(Q) {
  // rest, faces, edges, corners; each velocity followed by its negation.
  const list = [[0, 0, 0],
    [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
    [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1],
    [1, 0, -1], [-1, 0, 1], [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1],
    [1, 1, 1], [-1, -1, -1], [1, 1, -1], [-1, -1, 1], [1, -1, 1], [-1, 1, -1],
    [1, -1, -1], [-1, 1, 1]].slice(0, Q);
  // f32 weights with the exact-sum ulp adjustment (lattice-3d.mjs:85-111).
  const W = Q === 19
    ? [0.33333337306976318, 0.055555541068315506, 0.027777781710028648]
    : [0.29629626870155334, 0.074074082076549530, 0.018518516793847084, 0.0046296296641230583];
  return list.map((e) => ({ e: e, w: W[e[0] * e[0] + e[1] * e[1] + e[2] * e[2]] }));
}
function d2Tables() {
  const list = [[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]];
  const W = [0.44444444775581360, 0.11111111193895340, 0.027777777984738350];
  return list.map((e) => ({ e: [e[0], e[1], 0], w: W[e[0] * e[0] + e[1] * e[1]] }));
}

// `exactWeights` swaps the f32 weights for the double-precision fractions.
// The invariant tests use it: with exact weights the moment conditions hold to
// double rounding, so mass and momentum budgets close to ~1e-14 and any
// missing or mis-signed term in the collision shows up at the size of the
// term itself.  The engine and the parity runs always use the f32 weights.
const EXACT_W = { 9: [4 / 9, 1 / 9, 1 / 36], 19: [1 / 3, 1 / 18, 1 / 36], 27: [8 / 27, 2 / 27, 1 / 54, 1 / 216] };

function tables(Q, exactWeights) {
  let rows = Q === 9 ? d2Tables() : d3Tables(Q);
  if (exactWeights) {
    rows = rows.map((r) => ({ e: r.e, w: EXACT_W[Q][r.e[0] * r.e[0] + r.e[1] * r.e[1] + r.e[2] * r.e[2]] }));
  }
  const t = {
    Q: Q,
    ex: new Int32Array(Q), ey: new Int32Array(Q), ez: new Int32Array(Q),
    w: new Float64Array(Q), opp: new Int32Array(Q),
  };
  rows.forEach((r, i) => { t.ex[i] = r.e[0]; t.ey[i] = r.e[1]; t.ez[i] = r.e[2]; t.w[i] = r.w; });
  for (let i = 0; i < Q; i++) {
    for (let j = 0; j < Q; j++) {
      if (t.ex[j] === -t.ex[i] && t.ey[j] === -t.ey[i] && t.ez[j] === -t.ez[i]) t.opp[i] = j;
    }
  }
  return t;
}

const F32 = Math.fround;
const CS2_F32 = F32(1 / 3);
const CLS_SOLID = 1;
const CLS_NEAR = 2;

// Signed distance of the pinned body at integer cell (x, y, z), in emulated
// f32, one rounding per operation with no contraction.
//   3D sphere: length(p - c) - a   (common_d3_geometry.wgsl:119, :174-187)
//   2D circle: the nearest periodic image of p - c, then length - a
//              (common_geometry.wgsl:94-108, circle branch)
function phiF32(desc, x, y, z) {
  const b = desc.body;
  let dx = F32(x - b.centre[0]);
  let dy = F32(y - b.centre[1]);
  if (desc.q === 9) {
    const W = desc.nx, H = desc.ny;
    dx = F32(dx - F32(W * roundHalfEven(F32(dx / W))));
    dy = F32(dy - F32(H * roundHalfEven(F32(dy / H))));
    return F32(F32(Math.sqrt(F32(F32(dx * dx) + F32(dy * dy)))) - b.a);
  }
  const dz = F32(z - b.centre[2]);
  const s = F32(F32(F32(dx * dx) + F32(dy * dy)) + F32(dz * dz));
  return F32(F32(Math.sqrt(s)) - b.a);
}

function roundHalfEven(v) {
  const r = Math.round(v);
  if (Math.abs(v - Math.trunc(v)) === 0.5) return 2 * Math.round(v / 2);
  return r;
}

// Per-cell class byte: bit 0 solid (phi < 0), bit 1 near (phi < 2).  Cells
// with phi >= 2 cannot have a solid streaming source (every link is at most
// sqrt(3) long and the distance is 1-Lipschitz), so only near cells test
// their sources (common_d3_step.wgsl:243-276 tests every cell; the result is
// the same).
function classify(desc) {
  const n = desc.nx * desc.ny * desc.nz;
  const cls = new Uint8Array(n);
  if (!desc.body) return cls;
  for (let z = 0; z < desc.nz; z++) {
    for (let y = 0; y < desc.ny; y++) {
      for (let x = 0; x < desc.nx; x++) {
        const phi = phiF32(desc, x, y, z);
        let c = 0;
        if (phi < 0) c |= CLS_SOLID;
        if (phi < 2) c |= CLS_NEAR;
        cls[(z * desc.ny + y) * desc.nx + x] = c;
      }
    }
  }
  return cls;
}

// Normalises a descriptor: defaults, derived TRT delta, per-dimension flags.
function normalise(d) {
  const q = d.q;
  const dim = q === 9 ? 2 : 3;
  const out = {
    q: q, dim: dim, nx: d.nx, ny: d.ny, nz: dim === 2 ? 1 : d.nz,
    omega: F32(d.omega),
    omegaMinus: d.omegaMinus === undefined || d.omegaMinus === null ? F32(d.omega) : F32(d.omegaMinus),
    force: (d.force || [0, 0, 0]).map(F32),
    walls: (d.walls || [0, 0, 0]).slice(),
    wallU: (d.wallU || [0, 0]).map(F32),
    sponge: {
      width: F32((d.sponge && d.sponge.width) || 0),
      axes: (d.sponge && d.sponge.axes) ? d.sponge.axes.slice() : (dim === 2 ? [1, 1, 0] : [1, 1, 1]),
      u: ((d.sponge && d.sponge.u) || [0, 0, 0]).map(F32),
    },
    body: d.body ? { a: F32(d.body.a), centre: d.body.centre.map(F32) } : null,
    solidEq: d.solidEq === undefined ? 1 : d.solidEq,
    // the diffuse penalty coupling (the reference's default: no ?bounceback,
    // common_d3_step.wgsl:288-300; CHI_EPS main-3d.js:1026)
    diffuse: !!d.diffuse,
    chiEps: F32(d.chiEps === undefined || !(d.chiEps > 0) ? 1.5 : d.chiEps),
    storage: d.storage || 'f32',
    exactWeights: !!d.exactWeights,
  };
  if (dim === 2) out.sponge.axes[2] = 0;
  while (out.force.length < 3) out.force.push(0);
  out.dw = F32(out.omegaMinus - out.omega);
  return out;
}

// Sponge weight (common_d3_sponge.wgsl:10-14; the 2D form common_sponge.wgsl:7-11
// takes the minimum of two distances): cubic smoothstep of
// clamp(1 - min(dx, dy, dz) / width, 0, 1), the minimum over the axes (not a
// product), a masked axis passing 1e30.  Called through `hooks` so that the
// pin in tests/js/test-oracle.js can plant a wrong profile and watch it fail.
function spongeWeight3(dx, dy, dz, sw) {
  if (!(sw > 0)) return 0;
  let t = 1 - Math.min(dx, Math.min(dy, dz)) / sw;
  t = t < 0 ? 0 : (t > 1 ? 1 : t);
  return t * t * (3 - 2 * t);
}
const hooks = { spongeWeight3: spongeWeight3 };

// One fused pull-stream + collide over the x-rows r = z * NY + y in [r0, r1).
// `fin`/`fout` are plane-major: f[i * N + (z * NY + y) * NX + x].  `mac`
// receives [rho, ux, uy, uz] per cell when non-null.  `R` (optional) is a
// moving rigid body: { links[cell] (bit i: the pull source of direction i is
// solid), us(x, y, z) (body velocity at the cell), rhoWLocal, winCoord(x, y,
// z) (sponge coordinates) } (common_d3_step.wgsl:224-400 with HAS_BODY and
// USE_BOUNCEBACK).
// The diffuse coupling's solid fraction (chiFromPhiEps3,
// common_d3_geometry.wgsl:251; chiFromPhiEps, common_geometry.wgsl):
// 1/2 (1 - tanh(clamp(phi / eps, -20, 20))).
function chiOf(phi, eps) {
  const t = Math.max(-20, Math.min(20, phi / eps));
  return 0.5 * (1 - Math.tanh(t));
}

function stepSlab(S, T, cls, fin, fout, mac, r0, r1, R) {
  const Q = T.Q, NX = S.nx, NY = S.ny, NZ = S.nz, N = NX * NY * NZ;
  const ex = T.ex, ey = T.ey, ez = T.ez, w = T.w, opp = T.opp;
  const om = S.omega, dw = S.dw;
  const Fx0 = S.force[0], Fy0 = S.force[1], Fz0 = S.force[2];
  const wallX = S.walls[0], wallY = S.walls[1], wallZ = S.walls[2];
  const anyWall = wallX || wallY || wallZ;
  const sw = S.sponge.width, sax = S.sponge.axes, su = S.sponge.u;
  const uFarSq = su[0] * su[0] + su[1] * su[1] + su[2] * su[2];
  const hasBody = !!S.body || !!R;
  const diffuse = S.diffuse && !!S.body && !R;
  // under the diffuse coupling there is no bounce-back and no SOLID_EQ
  // (common_d3_step.wgsl:244-255, :388)
  const solidEq = hasBody && S.solidEq && !diffuse;
  const f = new Float64Array(Q);
  const fo = new Float64Array(Q);
  const fTarget = new Float64Array(Q);
  for (let i = 0; i < Q; i++) {
    const eu = ex[i] * su[0] + ey[i] * su[1] + ez[i] * su[2];
    fTarget[i] = w[i] * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * uFarSq);
  }
  const srcGuo = 1 - 0.5 * om;
  const rowBase = new Int32Array(Q);   // wrapped source row start, per direction
  const rowOut = new Uint8Array(Q);    // source row lies beyond a y or z wall
  for (let row = r0; row < r1; row++) {
    const z = Math.floor(row / NY), y = row - z * NY;
    {
      for (let i = 0; i < Q; i++) {
        const sy = y - ey[i], sz = z - ez[i];
        const wy = sy < 0 ? sy + NY : (sy >= NY ? sy - NY : sy);
        const wz = sz < 0 ? sz + NZ : (sz >= NZ ? sz - NZ : sz);
        rowBase[i] = (wz * NY + wy) * NX;
        rowOut[i] = ((wallY && (sy < 0 || sy >= NY)) ? 2 : 0) | ((wallZ && (sz < 0 || sz >= NZ)) ? 4 : 0);
      }
      for (let x = 0; x < NX; x++) {
        const cell = (z * NY + y) * NX + x;
        const c = cls[cell];
        // body velocity and wall density of a moving body, where needed
        let us = null, rhoW = 1;
        const lm = R ? R.links[cell] : 0;
        if (R && (lm || (c & CLS_SOLID))) us = R.us(x, y, z);
        if (lm && R.rhoWLocal) { rhoW = 0; for (let i = 0; i < Q; i++) rhoW += fin[i * N + cell]; }
        // 1. pull-stream with body bounce-back, then wall bounce-back.
        for (let i = 0; i < Q; i++) {
          const sx = x - ex[i];
          const src = rowBase[i] + (sx < 0 ? sx + NX : (sx >= NX ? sx - NX : sx));
          if (R ? ((lm >>> i) & 1) : (!diffuse && c !== 0 && (c & CLS_NEAR) && (cls[src] & CLS_SOLID))) {
            // Ladd's moving-wall term 2 w rho_w (e . u_body) / cs2
            const corr = R ? 2 * w[i] * rhoW * (ex[i] * us[0] + ey[i] * us[1] + ez[i] * us[2]) * 3 : 0;
            f[i] = fin[opp[i] * N + cell] + corr;
          } else if (anyWall && (rowOut[i] || (wallX && (sx < 0 || sx >= NX)))) {
            // Halfway wall; y-walls may move along x (common_walls.wgsl:27-31).
            let corr = 0;
            if ((rowOut[i] & 2) && ex[i] !== 0) {
              const uw = (y - ey[i]) < 0 ? S.wallU[0] : S.wallU[1];
              corr = 2 * w[i] * ex[i] * uw / CS2_F32;
            }
            f[i] = fin[opp[i] * N + cell] + corr;
          } else {
            f[i] = fin[i * N + src];
          }
        }
        // 2. moments
        let rho = 0, mx = 0, my = 0, mz = 0;
        for (let i = 0; i < Q; i++) {
          rho += f[i]; mx += f[i] * ex[i]; my += f[i] * ey[i]; mz += f[i] * ez[i];
        }
        const rhoDen = Math.max(rho, 1e-6);
        // 3. Guo forcing: the uniform body force, plus under the diffuse
        // coupling the penalty rho chi (U_s - u*) toward the body's velocity
        // (zero for a pinned body), u* = m / rho (common_d3_step.wgsl:288-300);
        // under bounce-back chi = 0.
        let Fx = Fx0, Fy = Fy0, Fz = Fz0;
        if (diffuse) {
          const chi = chiOf(phiF32(S, x, y, z), S.chiEps);
          Fx += rho * chi * (0 - mx / rhoDen);
          Fy += rho * chi * (0 - my / rhoDen);
          Fz += rho * chi * (0 - mz / rhoDen);
        }
        const ux = mx / rhoDen + Fx / (2 * rhoDen);
        const uy = my / rhoDen + Fy / (2 * rhoDen);
        const uz = mz / rhoDen + Fz / (2 * rhoDen);
        const usq = ux * ux + uy * uy + uz * uz;
        const solid = solidEq && (c & CLS_SOLID);
        if (mac) {
          if (solid && R) {
            mac[4 * cell] = 1; mac[4 * cell + 1] = us[0]; mac[4 * cell + 2] = us[1]; mac[4 * cell + 3] = us[2];
          } else if (solid && S.dim === 3) {
            mac[4 * cell] = 1; mac[4 * cell + 1] = 0; mac[4 * cell + 2] = 0; mac[4 * cell + 3] = 0;
          } else {
            mac[4 * cell] = rho; mac[4 * cell + 1] = ux; mac[4 * cell + 2] = uy; mac[4 * cell + 3] = uz;
          }
        }
        if (solid) {
          // SOLID_EQ: feq(1, u_body); u_body = 0 for a pinned body.
          if (R) {
            const uu = us[0] * us[0] + us[1] * us[1] + us[2] * us[2];
            for (let i = 0; i < Q; i++) {
              const eu = ex[i] * us[0] + ey[i] * us[1] + ez[i] * us[2];
              fout[i * N + cell] = w[i] * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * uu);
            }
          } else {
            for (let i = 0; i < Q; i++) fout[i * N + cell] = w[i];
          }
          continue;
        }
        // sponge weight (common_d3_sponge.wgsl:10-14, common_sponge.wgsl:7-11)
        let sp = 0;
        if (sw > 0) {
          // window coordinates under a moving window (common_d3_step.wgsl:332-337)
          const wp = R ? R.winCoord(x, y, z) : [x, y, z];
          const dX = sax[0] ? Math.min(wp[0], NX - 1 - wp[0]) : 1e30;
          const dY = sax[1] ? Math.min(wp[1], NY - 1 - wp[1]) : 1e30;
          const dZ = sax[2] ? Math.min(wp[2], NZ - 1 - wp[2]) : 1e30;
          sp = hooks.spongeWeight3(dX, dY, dZ, sw);
        }
        // 4. collide: BGK line plus the TRT delta terms.
        for (let i = 0; i < Q; i++) {
          const eu = ex[i] * ux + ey[i] * uy + ez[i] * uz;
          const feq = w[i] * rho * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * usq);
          const eF = ex[i] * Fx + ey[i] * Fy + ez[i] * Fz;
          const uF = ux * Fx + uy * Fy + uz * Fz;
          const Si = srcGuo * w[i] * (3 * (eF - uF) + 9 * eu * eF);
          let fc = f[i] - om * (f[i] - feq) + Si;
          if (dw !== 0) {
            const fMinus = 0.5 * (f[i] - f[opp[i]]);
            const feqMinus = 3 * w[i] * rho * eu;
            const sMinus = 3 * w[i] * eF;
            fc = fc - dw * (fMinus - feqMinus) - 0.5 * dw * sMinus;
          }
          fo[i] = sp > 0 ? fc * (1 - sp) + fTarget[i] * sp : fc;
        }
        for (let i = 0; i < Q; i++) fout[i * N + cell] = fo[i];
      }
    }
  }
}

// Momentum-exchange force and torque on the pinned body from pre-stream
// populations `fin`, over the x-rows r = z * NY + y in [r0, r1):
// [fx, fy, fz, tx, ty, tz].  Fluid cells only; a link counts when its
// streaming source is solid (common_d3_force.wgsl:192-216,
// lbm_force.wgsl:128-147).
function forceSlab(S, T, cls, fin, r0, r1) {
  const out = [0, 0, 0, 0, 0, 0];
  if (!S.body) return out;
  if (S.diffuse) return diffuseForceSlab(S, T, fin, r0, r1);
  const Q = T.Q, NX = S.nx, NY = S.ny, NZ = S.nz, N = NX * NY * NZ;
  const cx = S.body.centre[0], cy = S.body.centre[1], cz = S.dim === 3 ? S.body.centre[2] : 0;
  for (let row = r0; row < r1; row++) {
    const z = Math.floor(row / NY), y = row - z * NY;
    {
      for (let x = 0; x < NX; x++) {
        const cell = (z * NY + y) * NX + x;
        const c = cls[cell];
        if (!(c & CLS_NEAR) || (c & CLS_SOLID)) continue;
        let fx = 0, fy = 0, fz = 0;
        for (let i = 1; i < Q; i++) {
          let sx = x - T.ex[i], sy = y - T.ey[i], sz = z - T.ez[i];
          sx = sx < 0 ? sx + NX : (sx >= NX ? sx - NX : sx);
          sy = sy < 0 ? sy + NY : (sy >= NY ? sy - NY : sy);
          sz = sz < 0 ? sz + NZ : (sz >= NZ ? sz - NZ : sz);
          if (!(cls[(sz * NY + sy) * NX + sx] & CLS_SOLID)) continue;
          const m = 2 * fin[T.opp[i] * N + cell];
          fx -= T.ex[i] * m; fy -= T.ey[i] * m; fz -= T.ez[i] * m;
        }
        let rx = x - cx, ry = y - cy;
        const rz = z - cz;
        if (S.dim === 2) {
          rx -= NX * roundHalfEven(rx / NX);
          ry -= NY * roundHalfEven(ry / NY);
        }
        out[0] += fx; out[1] += fy; out[2] += fz;
        out[3] += ry * fz - rz * fy;
        out[4] += rz * fx - rx * fz;
        out[5] += rx * fy - ry * fx;
      }
    }
  }
  return out;
}

// Force and torque of the diffuse coupling on the pinned body, from the
// pre-stream populations: every cell with chi >= 1e-6 gathers rho and
// u* = m / rho from its periodic pull sources (as the step does) and the body
// feels the negative of the penalty, rho chi (u* - U_s), U_s = 0
// (common_d3_force.wgsl:242-268).
function diffuseForceSlab(S, T, fin, r0, r1) {
  const out = [0, 0, 0, 0, 0, 0];
  const Q = T.Q, NX = S.nx, NY = S.ny, NZ = S.nz, N = NX * NY * NZ;
  const cx = S.body.centre[0], cy = S.body.centre[1], cz = S.dim === 3 ? S.body.centre[2] : 0;
  for (let row = r0; row < r1; row++) {
    const z = Math.floor(row / NY), y = row - z * NY;
    for (let x = 0; x < NX; x++) {
      const chi = chiOf(phiF32(S, x, y, z), S.chiEps);
      if (!(chi >= 1e-6)) continue;
      let rho = 0, mx = 0, my = 0, mz = 0;
      for (let i = 0; i < Q; i++) {
        let sx = x - T.ex[i], sy = y - T.ey[i], sz = z - T.ez[i];
        sx = sx < 0 ? sx + NX : (sx >= NX ? sx - NX : sx);
        sy = sy < 0 ? sy + NY : (sy >= NY ? sy - NY : sy);
        sz = sz < 0 ? sz + NZ : (sz >= NZ ? sz - NZ : sz);
        const fi = fin[i * N + (sz * NY + sy) * NX + sx];
        rho += fi; mx += fi * T.ex[i]; my += fi * T.ey[i]; mz += fi * T.ez[i];
      }
      const den = Math.max(rho, 1e-6);
      const fx = rho * chi * (mx / den), fy = rho * chi * (my / den), fz = rho * chi * (mz / den);
      let rx = x - cx, ry = y - cy;
      const rz = z - cz;
      if (S.dim === 2) {
        rx -= NX * roundHalfEven(rx / NX);
        ry -= NY * roundHalfEven(ry / NY);
      }
      out[0] += fx; out[1] += fy; out[2] += fz;
      out[3] += ry * fz - rz * fy;
      out[4] += rz * fx - rx * fz;
      out[5] += rx * fy - ry * fx;
    }
  }
  return out;
}

class DenseOracle {
  constructor(desc) {
    this.S = normalise(desc);
    this.T = tables(this.S.q, this.S.exactWeights);
    const S = this.S;
    this.N = S.nx * S.ny * S.nz;
    this.cls = classify(S);
    this.f = new Float64Array(this.T.Q * this.N);
    this.g = new Float64Array(this.T.Q * this.N);
    this.mac = new Float64Array(4 * this.N);
    this.forceHistory = [];
    this.stepCount = 0;
  }

  // f = feq(mac) per cell (common_d3_step.wgsl:206-222).
  initEq(mac4) {
    const T = this.T, N = this.N;
    for (let c = 0; c < N; c++) {
      const rho = mac4[4 * c], ux = mac4[4 * c + 1], uy = mac4[4 * c + 2], uz = mac4[4 * c + 3];
      const usq = ux * ux + uy * uy + uz * uz;
      for (let i = 0; i < T.Q; i++) {
        const eu = T.ex[i] * ux + T.ey[i] * uy + T.ez[i] * uz;
        this.f[i * N + c] = T.w[i] * rho * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * usq);
      }
      this.mac[4 * c] = rho; this.mac[4 * c + 1] = ux; this.mac[4 * c + 2] = uy; this.mac[4 * c + 3] = uz;
    }
    this.stepCount = 0;
    this.forceHistory = [];
  }

  uploadF(f) { this.f.set(f); }

  step(n) {
    for (let s = 0; s < n; s++) {
      const rows = this.S.nz * this.S.ny;
      if (this.S.body) this.forceHistory.push(forceSlab(this.S, this.T, this.cls, this.f, 0, rows));
      stepSlab(this.S, this.T, this.cls, this.f, this.g, this.mac, 0, rows);
      const t = this.f; this.f = this.g; this.g = t;
      this.stepCount++;
    }
  }

  // FP64 totals over the current populations: [mass, px, py, pz].
  totals() {
    const T = this.T, N = this.N;
    const out = [0, 0, 0, 0];
    for (let i = 0; i < T.Q; i++) {
      let s = 0;
      for (let c = 0; c < N; c++) s += this.f[i * N + c];
      out[0] += s; out[1] += s * T.ex[i]; out[2] += s * T.ey[i]; out[3] += s * T.ez[i];
    }
    return out;
  }
}

module.exports = {
  tables, classify, phiF32, normalise, stepSlab, forceSlab, DenseOracle, chiOf,
  CLS_SOLID, CLS_NEAR, roundHalfEven, spongeWeight3, hooks,
};
