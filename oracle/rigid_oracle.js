'use strict';
// FP64 CPU ORACLE for a moving rigid body in a dense 3D solve.  TEST
// INFRASTRUCTURE ONLY: tests/ run it as the checker; the engine never calls
// it and never falls back to it.
//
// Restates, in double precision:
//   shapes     d3-body.mjs:56-101 (sphere, spheroid by 3 Newton steps on the
//              meridional ellipse, rounded box), :134-156 (circumradius,
//              principal inertia), :103-118 (volume)
//   algebra    d3-body.mjs:161-218 (quaternions [w,x,y,z], expMap,
//              omegaFromL), :299-307 (makeBodyState)
//   integrator d3-body.mjs:234-297 stepFreeBody, with the kernel's flags
//              from d3_physics.wgsl:56-158 (pinned records the force and
//              holds; NO_FLUID_FORCE; PLANAR; magnitude clamps)
//   window     d3-window.mjs:96-162 (wrapDelta, windowOffset, windowCoord,
//              wrapPosition), common_d3_window.wgsl
//   pose       common_d3_geometry.wgsl: bodyDelta3, get_phi3,
//              get_phi3Ahead (one step ahead by v and the cached omega),
//              bodyVelocity3
//   force      common_d3_force.wgsl:171-222: momentum exchange over the
//              bounce-back links with Ladd's moving-wall term at the local
//              wall density, plus the swept-cell term
//   step       common_d3_step.wgsl:224-400 with a body: moving-wall
//              bounce-back at the local wall density, the solid interior
//              held at feq(1, u_body), the sponge measured in window
//              coordinates (lbm_oracle.js stepSlab does the arithmetic)
// Per step, in the reference's order (main-3d.js:3109-3148): the force on
// the pre-stream populations at the current pose, the integrator, the
// classification at the new pose, then the fluid step at the new pose.
//
// Parity pins: tests/js/test-rigid.js (the integrator's exact invariants
// and closed forms, the tennis-racket flip, the SDFs against brute force,
// the co-moving fixed point, the momentum budget that the swept-cell term
// closes).

const O = require('./lbm_oracle.js');

const SHAPE = { SPHERE: 0, SPHEROID: 1, ROUNDBOX: 2 };

function def(v, d) { return v === undefined || v === null ? d : v; }

// --- signed distances in the body frame -----------------------------------
function sdSphere(p, a) { return Math.hypot(p[0], p[1], p[2]) - a; }

function sdRoundBox(p, a, b, c, r) {
  const qx = Math.abs(p[0]) - (a - r), qy = Math.abs(p[1]) - (b - r), qz = Math.abs(p[2]) - (c - r);
  const outside = Math.hypot(Math.max(qx, 0), Math.max(qy, 0), Math.max(qz, 0));
  const inside = Math.min(Math.max(qx, Math.max(qy, qz)), 0);
  return outside + inside - r;
}

// Closest point of the meridional ellipse (a, c) by Newton on its angle
// parameter, seeded at atan2(y a, x c), three steps, clamped to [0, pi/2].
function sdSpheroid(p, a, c) {
  const r0 = Math.hypot(p[0], p[1]);
  if (Math.abs(a - c) <= 1e-6 * Math.max(a, c)) return Math.hypot(r0, p[2]) - a;
  const x = Math.abs(r0), y = Math.abs(p[2]);
  const inside = Math.sqrt((x * x) / (a * a) + (y * y) / (c * c)) < 1;
  let t = Math.atan2(y * a, x * c);
  for (let k = 0; k < 3; k++) {
    const ct = Math.cos(t), st = Math.sin(t);
    const ex = x - a * ct, ey = y - c * st;
    const F = ex * (-a * st) + ey * (c * ct);
    const Fp = -(a * a * st * st + c * c * ct * ct) + ex * (-a * ct) + ey * (-c * st);
    t -= F / (Math.abs(Fp) < 1e-9 ? -1e-9 : Fp);
    t = Math.min(Math.max(t, 0), Math.PI / 2);
  }
  const d = Math.hypot(x - a * Math.cos(t), y - c * Math.sin(t));
  return inside ? -d : d;
}

function sdfBody(p, s) {
  if (s.kind === SHAPE.SPHERE) return sdSphere(p, s.a);
  if (s.kind === SHAPE.SPHEROID) return sdSpheroid(p, s.a, s.c);
  return sdRoundBox(p, s.a, s.b, s.c, s.r);
}

function bodyVolume(s) {
  if (s.kind === SHAPE.SPHERE) return (4 / 3) * Math.PI * s.a ** 3;
  if (s.kind === SHAPE.SPHEROID) return (4 / 3) * Math.PI * s.a * s.a * s.c;
  return 8 * s.a * s.b * s.c;
}

function circumradius(s) {
  if (s.kind === SHAPE.SPHERE) return s.a;
  if (s.kind === SHAPE.SPHEROID) return Math.max(s.a, s.c);
  return Math.hypot(s.a, s.b, s.c);
}

function principalInertia(s, m) {
  if (s.kind === SHAPE.SPHERE) { const i = 0.4 * m * s.a * s.a; return [i, i, i]; }
  if (s.kind === SHAPE.SPHEROID) {
    const ix = m * (s.a * s.a + s.c * s.c) / 5;
    return [ix, ix, 2 * m * s.a * s.a / 5];
  }
  return [m * (s.b * s.b + s.c * s.c) / 3, m * (s.a * s.a + s.c * s.c) / 3, m * (s.a * s.a + s.b * s.b) / 3];
}

// --- quaternions [w, x, y, z] ---------------------------------------------
function qMul(p, q) {
  return [
    p[0] * q[0] - p[1] * q[1] - p[2] * q[2] - p[3] * q[3],
    p[0] * q[1] + p[1] * q[0] + p[2] * q[3] - p[3] * q[2],
    p[0] * q[2] - p[1] * q[3] + p[2] * q[0] + p[3] * q[1],
    p[0] * q[3] + p[1] * q[2] - p[2] * q[1] + p[3] * q[0],
  ];
}
function qConj(q) { return [q[0], -q[1], -q[2], -q[3]]; }
function qNormalize(q) {
  const n = Math.hypot(q[0], q[1], q[2], q[3]);
  if (!(n > 1e-12)) return [1, 0, 0, 0];
  return [q[0] / n, q[1] / n, q[2] / n, q[3] / n];
}
function qFromAxisAngle(axis, angle) {
  const n = Math.hypot(axis[0], axis[1], axis[2]) || 1;
  const s = Math.sin(angle / 2) / n;
  return [Math.cos(angle / 2), axis[0] * s, axis[1] * s, axis[2] * s];
}
function qRotate(q, v) { const r = qMul(qMul(q, [0, v[0], v[1], v[2]]), qConj(q)); return [r[1], r[2], r[3]]; }
function qRotateInv(q, v) { const r = qMul(qMul(qConj(q), [0, v[0], v[1], v[2]]), q); return [r[1], r[2], r[3]]; }
function expMap(w, dt) {
  const m = Math.hypot(w[0], w[1], w[2]);
  return m > 1e-12 ? qFromAxisAngle(w, m * dt) : [1, 0, 0, 0];
}
function omegaFromL(q, I, L) {
  const Lb = qRotateInv(q, L);
  return qRotate(q, [Lb[0] / I[0], Lb[1] / I[1], Lb[2] / I[2]]);
}
function rotationalEnergy(q, I, L) {
  const w = omegaFromL(q, I, L);
  return 0.5 * (L[0] * w[0] + L[1] * w[1] + L[2] * w[2]);
}

// --- moving window (per axis: n > 0 on a windowed axis, else 0) -----------
function wrapDelta(d, n) { return n > 0 ? d - n * Math.round(d / n) : d; }
function wrapPosition(c, n) { return n > 0 ? c - n * Math.floor(c / n) : c; }
function windowOffset(c, anchor, n) { return n > 0 ? Math.floor(c) - anchor : 0; }
function windowCoord(p, off, n) {
  if (!(n > 0)) return p;
  const w = p - off;
  return w - n * Math.floor(w / n);
}

// --- body state -------------------------------------------------------------
// `p`: { shape: {kind, a, b, c, r}, x, q, v, omega, density, pinned, gravity,
//        vMax, oMax, planar, noFluidForce, swept, rhoWLocal, window: [0|1]*3,
//        anchor: [3] }; `dims` sizes the windowed axes.
function makeBody(p, dims) {
  const shape = { kind: p.shape.kind, a: p.shape.a, b: def(p.shape.b, p.shape.a), c: def(p.shape.c, p.shape.a),
                  r: def(p.shape.r, 0) };
  const density = def(p.density, 1);
  const mass = density * bodyVolume(shape);
  const I = principalInertia(shape, mass);
  const q = (p.q || [1, 0, 0, 0]).slice();
  const omega = (p.omega || [0, 0, 0]).slice();
  const Lb = qRotateInv(q, omega);
  const L = qRotate(q, [Lb[0] * I[0], Lb[1] * I[1], Lb[2] * I[2]]);
  const win = p.window || [0, 0, 0];
  return {
    shape, mass, I, x: p.x.slice(), q, v: (p.v || [0, 0, 0]).slice(), L, w: omega, d: [0, 0, 0],
    force: [0, 0, 0], torque: [0, 0, 0],
    pinned: !!p.pinned, gravity: (p.gravity || [0, 0, 0]).slice(),
    vMax: def(p.vMax, 0.2), oMax: def(p.oMax, 0.5), planar: !!p.planar, noFluidForce: !!p.noFluidForce,
    swept: p.swept === undefined ? true : !!p.swept, rhoWLocal: p.rhoWLocal === undefined ? true : !!p.rhoWLocal,
    winN: [0, 1, 2].map((k) => (win[k] ? dims[k] : 0)), anchor: (p.anchor || [0, 0, 0]).slice(),
    rc: circumradius(shape),
  };
}

function bodyDelta(p, B) {
  return [wrapDelta(p[0] - B.x[0], B.winN[0]), wrapDelta(p[1] - B.x[1], B.winN[1]),
          wrapDelta(p[2] - B.x[2], B.winN[2])];
}
function phi(p, B) { return sdfBody(qRotateInv(B.q, bodyDelta(p, B)), B.shape); }
function phiAhead(p, B) {
  const c = [B.x[0] + B.v[0], B.x[1] + B.v[1], B.x[2] + B.v[2]];
  const q = qMul(expMap(B.w, 1), B.q);
  const d = [wrapDelta(p[0] - c[0], B.winN[0]), wrapDelta(p[1] - c[1], B.winN[1]), wrapDelta(p[2] - c[2], B.winN[2])];
  return sdfBody(qRotateInv(q, d), B.shape);
}
function bodyVelocity(p, B) {
  const r = bodyDelta(p, B), w = B.w;
  return [B.v[0] + (w[1] * r[2] - w[2] * r[1]), B.v[1] + (w[2] * r[0] - w[0] * r[2]),
          B.v[2] + (w[0] * r[1] - w[1] * r[0])];
}

// The integrator (d3_physics.wgsl main, d3-body.mjs stepFreeBody with dt = 1).
function physics(B, force, torque) {
  B.force = force.slice(); B.torque = torque.slice();
  if (B.pinned) return;
  const use = B.noFluidForce ? 0 : 1;
  const m = B.mass;
  let v = [B.v[0] + (use * force[0] / m + B.gravity[0]), B.v[1] + (use * force[1] / m + B.gravity[1]),
           B.v[2] + (use * force[2] / m + B.gravity[2])];
  if (B.planar) v[2] = 0;
  const vmag = Math.hypot(v[0], v[1], v[2]);
  if (vmag > B.vMax) v = v.map((c) => c * (B.vMax / vmag));
  let L = [B.L[0] + use * torque[0], B.L[1] + use * torque[1], B.L[2] + use * torque[2]];
  if (B.planar) { L[0] = 0; L[1] = 0; }
  let w = omegaFromL(B.q, B.I, L);
  if (B.planar) w = [0, 0, w[2]];
  const wmag = Math.hypot(w[0], w[1], w[2]);
  if (wmag > B.oMax) { const k = B.oMax / wmag; w = w.map((c) => c * k); L = L.map((c) => c * k); }
  const qHalf = qNormalize(qMul(expMap(w, 0.5), B.q));
  let wHalf = omegaFromL(qHalf, B.I, L);
  if (B.planar) wHalf = [0, 0, wHalf[2]];
  const qn = qNormalize(qMul(expMap(wHalf, 1), B.q));
  B.d = [B.d[0] + v[0], B.d[1] + v[1], B.d[2] + v[2]];
  B.x = [0, 1, 2].map((k) => wrapPosition(B.x[k] + v[k], B.winN[k]));
  B.v = v; B.L = L; B.w = wHalf; B.q = qn;
}

// Cells a body can touch: within its circumradius plus 2 of the centre
// (window-wrapped delta), so that every solid cell and every cell with a
// solid pull source is inside.
function nearBody(p, B) {
  const r = bodyDelta(p, B);
  return Math.hypot(r[0], r[1], r[2]) <= B.rc + 2;
}

// Solid bit per cell and, for fluid cells, the mask of directions whose pull
// source (at the unwrapped neighbour coordinate) is solid.
function classify(S, T, B) {
  const N = S.nx * S.ny * S.nz;
  const cls = new Uint8Array(N);
  const links = new Uint32Array(N);
  for (let z = 0; z < S.nz; z++) {
    for (let y = 0; y < S.ny; y++) {
      for (let x = 0; x < S.nx; x++) {
        const p = [x, y, z];
        if (!nearBody(p, B)) continue;
        const cell = (z * S.ny + y) * S.nx + x;
        if (phi(p, B) < 0) { cls[cell] = O.CLS_SOLID; continue; }
        let m = 0;
        for (let i = 1; i < T.Q; i++) {
          if (phi([x - T.ex[i], y - T.ey[i], z - T.ez[i]], B) < 0) m |= 1 << i;
        }
        links[cell] = m >>> 0;
        if (m) cls[cell] = O.CLS_NEAR;
      }
    }
  }
  return { cls, links };
}

// Force and torque on the body from the pre-stream populations `fin` at the
// current pose (cls/links of that pose): [fx, fy, fz, tx, ty, tz].
function force(S, T, B, cls, links, fin) {
  const N = S.nx * S.ny * S.nz, Q = T.Q;
  const out = [0, 0, 0, 0, 0, 0];
  for (let z = 0; z < S.nz; z++) {
    for (let y = 0; y < S.ny; y++) {
      for (let x = 0; x < S.nx; x++) {
        const p = [x, y, z];
        if (!nearBody(p, B)) continue;
        const cell = (z * S.ny + y) * S.nx + x;
        const solidNow = (cls[cell] & O.CLS_SOLID) !== 0;
        let fx = 0, fy = 0, fz = 0;
        if (!solidNow && links[cell]) {
          const us = bodyVelocity(p, B);
          let rhoW = 1;
          if (B.rhoWLocal) { rhoW = 0; for (let i = 0; i < Q; i++) rhoW += fin[i * N + cell]; }
          for (let i = 1; i < Q; i++) {
            if (!((links[cell] >>> i) & 1)) continue;
            const eu = T.ex[i] * us[0] + T.ey[i] * us[1] + T.ez[i] * us[2];
            const m = 2 * fin[T.opp[i] * N + cell] + 2 * T.w[i] * rhoW * eu * 3;
            fx -= T.ex[i] * m; fy -= T.ey[i] * m; fz -= T.ez[i] * m;
          }
        }
        if (B.swept && solidNow !== (phiAhead(p, B) < 0)) {
          let mx = 0, my = 0, mz = 0;
          for (let i = 0; i < Q; i++) {
            const fi = fin[i * N + cell];
            mx += fi * T.ex[i]; my += fi * T.ey[i]; mz += fi * T.ez[i];
          }
          const s = solidNow ? -1 : 1;
          fx += s * mx; fy += s * my; fz += s * mz;
        }
        if (fx === 0 && fy === 0 && fz === 0) continue;
        const r = bodyDelta(p, B);
        out[0] += fx; out[1] += fy; out[2] += fz;
        out[3] += r[1] * fz - r[2] * fy;
        out[4] += r[2] * fx - r[0] * fz;
        out[5] += r[0] * fy - r[1] * fx;
      }
    }
  }
  return out;
}

// The step's view of the body (lbm_oracle.js stepSlab's `R`).
function stepContext(B, links) {
  const off = [0, 1, 2].map((k) => windowOffset(B.x[k], B.anchor[k], B.winN[k]));
  return {
    links, rhoWLocal: B.rhoWLocal,
    us: (x, y, z) => bodyVelocity([x, y, z], B),
    winCoord: (x, y, z) => [windowCoord(x, off[0], B.winN[0]), windowCoord(y, off[1], B.winN[1]),
                            windowCoord(z, off[2], B.winN[2])],
  };
}

// A dense 3D oracle with one rigid body (the descriptor's own `body` null).
class RigidOracle {
  constructor(desc, bodyParams) {
    this.dense = new O.DenseOracle(Object.assign({}, desc, { body: null }));
    this.S = this.dense.S;
    this.T = this.dense.T;
    this.B = makeBody(bodyParams, [this.S.nx, this.S.ny, this.S.nz]);
    const c = classify(this.S, this.T, this.B);
    this.cls = c.cls; this.links = c.links;
    this.forceHistory = [];
    this.bodyHistory = [];
  }

  initEq(mac4) { this.dense.initEq(mac4); this.forceHistory = []; this.bodyHistory = []; }

  get f() { return this.dense.f; }
  get mac() { return this.dense.mac; }

  step(n) {
    const S = this.S, T = this.T, B = this.B, D = this.dense;
    const rows = S.nz * S.ny;
    for (let s = 0; s < n; s++) {
      const F = force(S, T, B, this.cls, this.links, D.f);
      this.forceHistory.push(F);
      physics(B, F.slice(0, 3), F.slice(3, 6));
      const c = classify(S, T, B);
      this.cls = c.cls; this.links = c.links;
      O.stepSlab(S, T, this.cls, D.f, D.g, D.mac, 0, rows, stepContext(B, this.links));
      const t = D.f; D.f = D.g; D.g = t;
      D.stepCount++;
      this.bodyHistory.push(snapshot(B));
    }
  }
}

function snapshot(B) {
  return { x: B.x.slice(), q: B.q.slice(), v: B.v.slice(), L: B.L.slice(), w: B.w.slice(), d: B.d.slice(),
           force: B.force.slice(), torque: B.torque.slice() };
}

module.exports = {
  SHAPE, sdfBody, bodyVolume, circumradius, principalInertia,
  qMul, qConj, qNormalize, qFromAxisAngle, qRotate, qRotateInv, expMap, omegaFromL, rotationalEnergy,
  wrapDelta, wrapPosition, windowOffset, windowCoord,
  makeBody, phi, phiAhead, bodyVelocity, physics, classify, force, stepContext, RigidOracle, snapshot,
};
