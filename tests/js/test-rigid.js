'use strict';
// Pins the rigid-body oracle (oracle/rigid_oracle.js) against exact
// invariants, closed forms and independent instruments, chosen so that a
// dropped term, a wrong sign or a transposed quaternion fails one of them
// (the checks of the reference's tools/test-d3-body.js, restated, plus fluid
// invariants of the coupling):
//   - |L| exactly conserved by a free body, rotational energy to truncation
//     order, |q| normalised
//   - the intermediate principal axis flips (tennis-racket theorem), the
//     minor and major axes do not: a sign error in the gyroscopic coupling
//     conserves |L| and never flips
//   - constant acceleration gives v_n = n g and x_n = g n (n + 1) / 2
//   - magnitude clamps keep the direction; the planar projection leaves the
//     in-plane dynamics equal to an in-plane-loaded body
//   - every SDF is a true distance (|grad phi| = 1) and agrees with a
//     brute-force nearest-surface-point search
//   - a uniform flow feq(1, V) with the body moving at V is a fixed point of
//     the whole scheme (the moving-wall term and the solid interior's
//     feq(1, u_body) are exactly what keep it uniform)
//   - a pinned rigid sphere reproduces the pinned-sphere oracle (the tow = 0
//     leg), and for a towed body the fluid's momentum change balances the
//     reported force (the swept-cell term is what closes the budget)
// Run: node tests/js/test-rigid.js   (exit code 1 on any failure)

const R = require('../../oracle/rigid_oracle.js');
const O = require('../../oracle/lbm_oracle.js');

let failures = 0;
function ok(name, fn) {
  try {
    const msg = fn();
    console.log('ok   ' + name + (msg ? '  (' + msg + ')' : ''));
  } catch (e) {
    failures++;
    console.log('FAIL ' + name + ': ' + (e && e.message ? e.message : e));
  }
}
function assert(c, m) { if (!c) throw new Error(m); }
function close(a, b, tol, what) {
  assert(Math.abs(a - b) <= tol, `${what}: got ${a}, want ${b} (|diff| ${Math.abs(a - b).toExponential(2)})`);
}
const dot = (a, b) => a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
const norm = (a) => Math.hypot(a[0], a[1], a[2]);

// A free body without clamps (stepFreeBody's defaults are no clamps).
function freeBody(p) {
  return R.makeBody(Object.assign({ x: [0, 0, 0], vMax: Infinity, oMax: Infinity }, p), [0, 0, 0]);
}

// --- integrator -------------------------------------------------------------
ok('omegaFromL inverts the L built from omega', () => {
  const B = freeBody({ shape: { kind: R.SHAPE.ROUNDBOX, a: 4, b: 3, c: 1, r: 0 },
                       q: R.qFromAxisAngle([1, 1, 0], 0.4), omega: [0.01, -0.02, 0.03] });
  const w = R.omegaFromL(B.q, B.I, B.L);
  for (let i = 0; i < 3; i++) close(w[i], B.w[i], 1e-12, `omega ${i}`);
});

ok('free body: |L| exact, energy within 2% over 20000 steps, |q| = 1', () => {
  const B = freeBody({ shape: { kind: R.SHAPE.ROUNDBOX, a: 6, b: 4, c: 1, r: 0 }, omega: [0.02, 0.01, 0.005] });
  const L0 = norm(B.L), E0 = R.rotationalEnergy(B.q, B.I, B.L);
  for (let i = 0; i < 20000; i++) R.physics(B, [0, 0, 0], [0, 0, 0]);
  close(norm(B.L), L0, 1e-12 * L0, '|L|');
  const E = R.rotationalEnergy(B.q, B.I, B.L);
  assert(Math.abs(E - E0) / E0 < 0.02, `energy drift ${((E - E0) / E0 * 100).toFixed(3)}%`);
  close(Math.hypot(B.q[0], B.q[1], B.q[2], B.q[3]), 1, 1e-9, '|q|');
  return `energy drift ${((E - E0) / E0 * 100).toExponential(2)}%`;
});

function alignment(shape, omega, axis, steps) {
  const B = freeBody({ shape, omega });
  const a0 = R.qRotate(B.q, axis);
  let minDot = 1;
  for (let i = 0; i < steps; i++) {
    R.physics(B, [0, 0, 0], [0, 0, 0]);
    if (i % 100 === 0) minDot = Math.min(minDot, dot(R.qRotate(B.q, axis), a0));
  }
  return minDot;
}

ok('the intermediate axis flips, the minor and major axes do not', () => {
  const shape = { kind: R.SHAPE.ROUNDBOX, a: 6, b: 3, c: 1, r: 0 };
  const I = R.principalInertia(shape, 1);
  assert(I[0] < I[1] && I[1] < I[2], 'moments ordered');
  const mid = alignment(shape, [1e-4, 0.02, 1e-4], [0, 1, 0], 400000);
  assert(mid < -0.9, `intermediate axis never flipped (min alignment ${mid.toFixed(3)})`);
  const minor = alignment(shape, [0.02, 1e-4, 0], [1, 0, 0], 400000);
  const major = alignment(shape, [0, 1e-4, 0.02], [0, 0, 1], 400000);
  assert(minor > 0.9 && major > 0.9, `stable axes moved: minor ${minor.toFixed(3)}, major ${major.toFixed(3)}`);
  return `alignment min: intermediate ${mid.toFixed(3)}, minor ${minor.toFixed(3)}, major ${major.toFixed(3)}`;
});

ok('constant acceleration: v_n = n g, x_n = g n (n + 1) / 2', () => {
  const g = [0, -1e-4, 0];
  const B = freeBody({ shape: { kind: R.SHAPE.SPHERE, a: 2 }, gravity: g });
  const n = 1000;
  for (let i = 0; i < n; i++) R.physics(B, [0, 0, 0], [0, 0, 0]);
  close(B.v[1], n * g[1], 1e-12 * Math.abs(n * g[1]), 'v');
  close(B.x[1], g[1] * n * (n + 1) / 2, Math.abs(g[1]) * 1e-6, 'x');
  close(B.d[1], B.x[1], 1e-15, 'displacement without a window');
});

ok('clamps keep the direction', () => {
  const B = freeBody({ shape: { kind: R.SHAPE.SPHERE, a: 2 }, v: [0.3, 0.4, 0] });
  B.vMax = 0.1;
  R.physics(B, [0, 0, 0], [0, 0, 0]);
  close(norm(B.v), 0.1, 1e-12, '|v|');
  close(B.v[0] / B.v[1], 0.75, 1e-12, 'direction');
  const C = freeBody({ shape: { kind: R.SHAPE.ROUNDBOX, a: 3, b: 2, c: 1, r: 0 }, omega: [0.3, 0.1, -0.2] });
  C.oMax = 0.05;
  const L0 = C.L.slice();
  R.physics(C, [0, 0, 0], [0, 0, 0]);
  const k = C.L[0] / L0[0];
  close(C.L[1] / L0[1], k, 1e-12, 'L scaled along its own axis');
  close(C.L[2] / L0[2], k, 1e-12, 'L scaled along its own axis');
});

ok('planar: v_z, L_x, L_y stay zero and the in-plane motion is unchanged', () => {
  const shape = { kind: R.SHAPE.ROUNDBOX, a: 16, b: 40, c: 2, r: 0 };
  const q = R.qFromAxisAngle([1, 1, 1], 2 * Math.PI / 3);
  const s = freeBody({ shape, q, density: 2.7, gravity: [1e-3, 0, 0] });
  s.planar = true;
  const u = freeBody({ shape, q, density: 2.7, gravity: [1e-3, 0, 0] });
  for (let i = 0; i < 50; i++) {
    const f = [Math.sin(i), Math.cos(2 * i), 0.7 * Math.sin(3 * i)];
    const t = [0.2 * Math.cos(i), 0.1 * Math.sin(i), 0.05 * Math.cos(4 * i)];
    R.physics(s, f, t);
    R.physics(u, [f[0], f[1], 0], [0, 0, t[2]]);
    assert(s.v[2] === 0 && s.L[0] === 0 && s.L[1] === 0 && s.w[0] === 0 && s.w[1] === 0, 'planar components');
  }
  close(s.v[0], u.v[0], 1e-15, 'v_x'); close(s.v[1], u.v[1], 1e-15, 'v_y'); close(s.L[2], u.L[2], 1e-15, 'L_z');
  for (let k = 0; k < 4; k++) close(s.q[k], u.q[k], 1e-12, `q${k}`);
});

ok('the moving window wraps the position and keeps the displacement', () => {
  const B = R.makeBody({ shape: { kind: R.SHAPE.SPHERE, a: 2 }, x: [30, 8, 8], v: [0.7, 0, 0], vMax: 1,
                         window: [1, 0, 0] }, [32, 16, 16]);
  for (let i = 0; i < 10; i++) R.physics(B, [0, 0, 0], [0, 0, 0]);
  close(B.d[0], 7, 1e-12, 'displacement');
  close(B.x[0], 37 - 32, 1e-12, 'wrapped position');
  // a cell across the wrap is the body's near image
  close(R.phi([1, 8, 8], B), Math.abs(1 - 5) - 2, 1e-12, 'phi across the wrap');
});

// --- signed distances -------------------------------------------------------
const SHAPES = [
  ['sphere', { kind: R.SHAPE.SPHERE, a: 7, b: 7, c: 7, r: 0 }],
  ['spheroid oblate', { kind: R.SHAPE.SPHEROID, a: 9, b: 9, c: 3, r: 0 }],
  ['spheroid prolate', { kind: R.SHAPE.SPHEROID, a: 3, b: 3, c: 11, r: 0 }],
  ['roundbox plate', { kind: R.SHAPE.ROUNDBOX, a: 10, b: 6, c: 1.5, r: 0.4 }],
  ['roundbox cube', { kind: R.SHAPE.ROUNDBOX, a: 5, b: 5, c: 5, r: 0 }],
];

// Points on (spheroids) or inside-and-on (rounded boxes: the inner box's
// surface grown by r in sampled directions) the body.  None lies outside,
// so the nearest one is never closer than the true surface.
function samples(s, n) {
  const out = [];
  if (s.kind !== R.SHAPE.ROUNDBOX) {
    const c = s.kind === R.SHAPE.SPHERE ? s.a : s.c;
    for (let i = 0; i <= n; i++) {
      const th = Math.PI * i / n;
      for (let j = 0; j < 2 * n; j++) {
        const ph = Math.PI * j / n;
        out.push([s.a * Math.sin(th) * Math.cos(ph), s.a * Math.sin(th) * Math.sin(ph), c * Math.cos(th)]);
      }
    }
    return out;
  }
  const h = [s.a - s.r, s.b - s.r, s.c - s.r];
  const dirs = [];
  const m = s.r > 0 ? 6 : 0;
  for (let i = 0; i <= m; i++) {
    for (let j = 0; j < Math.max(1, 2 * m); j++) {
      const th = m ? Math.PI * i / m : 0, ph = m ? Math.PI * j / m : 0;
      dirs.push([Math.sin(th) * Math.cos(ph), Math.sin(th) * Math.sin(ph), Math.cos(th)]);
    }
  }
  for (let ax = 0; ax < 3; ax++) {
    for (const sg of [-1, 1]) {
      for (let i = 0; i <= n; i++) {
        for (let j = 0; j <= n; j++) {
          const p = [0, 0, 0];
          const u = (ax + 1) % 3, v = (ax + 2) % 3;
          p[ax] = sg * h[ax];
          p[u] = -h[u] + 2 * h[u] * i / n;
          p[v] = -h[v] + 2 * h[v] * j / n;
          if (s.r > 0) for (const d of dirs) out.push([p[0] + s.r * d[0], p[1] + s.r * d[1], p[2] + s.r * d[2]]);
          else out.push(p);
        }
      }
    }
  }
  return out;
}

for (const [name, shape] of SHAPES) {
  ok(`${name}: |grad phi| = 1`, () => {
    const h = 1e-4;
    let worst = 0;
    for (const p of [[13, 2, 1], [0, 0, 14], [8, 8, 8], [20, -3, 2], [-11, 5, -6], [2, 12, 9]]) {
      const g = [0, 1, 2].map((i) => {
        const a = p.slice(), b = p.slice();
        a[i] += h; b[i] -= h;
        return (R.sdfBody(a, shape) - R.sdfBody(b, shape)) / (2 * h);
      });
      worst = Math.max(worst, Math.abs(norm(g) - 1));
    }
    assert(worst < 2e-3, `|grad phi| off by ${worst.toExponential(2)}`);
  });
  ok(`${name}: agrees with a brute-force nearest-surface search`, () => {
    const pts = samples(shape, 120);
    let worstRel = 0;
    for (const p of [[15, 3, 2], [0, 0, 18], [9, 9, 9], [-13, 4, -5]]) {
      let best = Infinity;
      for (const q of pts) best = Math.min(best, Math.hypot(p[0] - q[0], p[1] - q[1], p[2] - q[2]));
      const got = R.sdfBody(p, shape);
      assert(got > 0, 'outside');
      assert(got <= best + 1e-9, `analytic ${got} exceeds brute force ${best}`);
      assert(best - got < 0.05 * Math.max(1, got), `analytic ${got} vs brute force ${best}`);
      worstRel = Math.max(worstRel, (best - got) / Math.max(1, got));
    }
    assert(R.sdfBody([0, 0, 0], shape) < 0, 'centre inside');
    return `worst gap ${(worstRel * 100).toFixed(3)}%`;
  });
}

ok('rotation: a plate turned 90 degrees about z has its long axis along y', () => {
  const shape = { kind: R.SHAPE.ROUNDBOX, a: 10, b: 2, c: 1, r: 0 };
  const q = R.qFromAxisAngle([0, 0, 1], Math.PI / 2);
  const B = R.makeBody({ shape, x: [0, 0, 0], q }, [0, 0, 0]);
  assert(R.phi([0, 9, 0], B) < 0, 'long axis along +y');
  assert(R.phi([9, 0, 0], B) > 0, 'and no longer along +x');
});

// --- fluid coupling ----------------------------------------------------------
function uniformMac(N, u) {
  const m = new Float64Array(4 * N);
  for (let c = 0; c < N; c++) { m[4 * c] = 1; m[4 * c + 1] = u[0]; m[4 * c + 2] = u[1]; m[4 * c + 3] = u[2]; }
  return m;
}

ok('a uniform flow with the body moving at the same velocity stays uniform', () => {
  const V = [0.03, -0.01, 0.02];
  // exact weights: with the f32 weights sum w differs from 1 by ~1e-8 in
  // double, and the wall density then carries that into every corrected link
  const desc = { q: 19, nx: 24, ny: 16, nz: 16, omega: 1.2, exactWeights: true };
  const o = new R.RigidOracle(desc, { shape: { kind: R.SHAPE.SPHEROID, a: 4, c: 2.5 }, x: [9.3, 7.6, 8.2],
                                      q: R.qFromAxisAngle([0.2, 1, 0.4], 0.7), v: V, noFluidForce: true });
  const N = 24 * 16 * 16;
  o.initEq(uniformMac(N, V));
  o.step(12);
  const T = o.T;
  let worst = 0;
  const uu = dot(V, V);
  for (let i = 0; i < T.Q; i++) {
    const eu = T.ex[i] * V[0] + T.ey[i] * V[1] + T.ez[i] * V[2];
    const feq = T.w[i] * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * uu);
    for (let c = 0; c < N; c++) worst = Math.max(worst, Math.abs(o.f[i * N + c] - feq));
  }
  assert(worst < 1e-13, `field moved by ${worst.toExponential(2)}`);
  close(o.B.x[0], 9.3 + 12 * V[0], 1e-12, 'prescribed trajectory');
  return `max |f - feq| ${worst.toExponential(2)} after 12 steps`;
});

ok('a pinned rigid sphere reproduces the pinned-sphere oracle', () => {
  const desc = { q: 19, nx: 20, ny: 16, nz: 16, omega: 1.1, sponge: { width: 3, u: [0.03, 0, 0] } };
  const N = 20 * 16 * 16;
  const mac = uniformMac(N, [0.03, 0, 0]);
  const a = new O.DenseOracle(Object.assign({}, desc, { body: { a: 3.5, centre: [8, 8, 8] } }));
  a.initEq(mac); a.step(20);
  const b = new R.RigidOracle(desc, { shape: { kind: R.SHAPE.SPHERE, a: 3.5 }, x: [8, 8, 8], pinned: true });
  b.initEq(mac); b.step(20);
  let df = 0, dF = 0;
  for (let k = 0; k < a.f.length; k++) df = Math.max(df, Math.abs(a.f[k] - b.f[k]));
  for (let s = 0; s < 20; s++) for (let j = 0; j < 6; j++) dF = Math.max(dF, Math.abs(a.forceHistory[s][j] - b.forceHistory[s][j]));
  assert(df < 1e-14 && dF < 1e-12, `f differs by ${df.toExponential(2)}, force by ${dF.toExponential(2)}`);
  return `|df| ${df.toExponential(2)}, |dF| ${dF.toExponential(2)}`;
});

// Fluid momentum over the cells that are fluid at the current pose.
function fluidMomentum(o) {
  const T = o.T, N = o.f.length / T.Q;
  const m = [0, 0, 0];
  for (let c = 0; c < N; c++) {
    if (o.cls[c] & O.CLS_SOLID) continue;
    for (let i = 0; i < T.Q; i++) {
      const fi = o.f[i * N + c];
      m[0] += fi * T.ex[i]; m[1] += fi * T.ey[i]; m[2] += fi * T.ez[i];
    }
  }
  return m;
}

function budget(swept) {
  const desc = { q: 19, nx: 28, ny: 16, nz: 16, omega: 1.3, exactWeights: true };
  const N = 28 * 16 * 16;
  const o = new R.RigidOracle(desc, { shape: { kind: R.SHAPE.SPHERE, a: 3.6 }, x: [9.2, 8.1, 7.9],
                                      v: [0.04, 0.006, 0], noFluidForce: true, swept });
  o.initEq(uniformMac(N, [0, 0, 0]));
  o.step(30);  // let the wake form
  let res = 0, mag = 0;
  for (let s = 0; s < 40; s++) {
    const p0 = fluidMomentum(o);
    o.step(1);
    const p1 = fluidMomentum(o);
    const F = o.forceHistory[o.forceHistory.length - 1];
    for (let k = 0; k < 3; k++) { res += Math.abs(p1[k] - p0[k] + F[k]); mag += Math.abs(F[k]); }
  }
  return res / mag;
}

ok('towed body: fluid momentum change balances the force; without the swept term it does not', () => {
  const withSwept = budget(true), without = budget(false);
  // What stays is the force's link set (pose t) against the one the step
  // bounces on (pose t + 1); dropping the swept term triples the residual.
  assert(withSwept < 0.05, `budget residual ${(withSwept * 100).toFixed(2)}% of |F| with the swept term`);
  assert(without > 0.05 && without > 2.5 * withSwept,
         `swept term changes the residual only from ${(without * 100).toFixed(2)}% to ${(withSwept * 100).toFixed(2)}%`);
  return `residual ${(withSwept * 100).toFixed(3)}% of |F|, ${(without * 100).toFixed(1)}% without the swept term`;
});

if (failures) { console.log(failures + ' failure(s)'); process.exit(1); }
console.log('all rigid-body oracle checks passed');
