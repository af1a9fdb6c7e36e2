'use strict';
// Pins the FP64 oracle (oracle/lbm_oracle.js) against closed forms and exact
// invariants, chosen so that a dropped term, a wrong sign or a transposed
// table fails at least one of them:
//   - 2D Taylor-Green decay (analytic field and energy decay, main-tgv.js:65-90,
//     tolerances of benchmarks/tgv.json: field L2rel <= 1e-2, rate <= 1%)
//   - 3D Beltrami (ABC) exact decay u(t) = u0 exp(-nu k^2 t) (d3-scenarios.mjs:19-27)
//   - exact mass conservation in a periodic box (sum w = 1 exactly)
//   - exact momentum input N*F per step under Guo forcing, BGK and TRT
//     (plans/TRT.md sec 8.0 item 6)
//   - TRT at Lambda = 3/16 reproduces plane Poiseuille exactly with halfway
//     bounce-back walls; BGK reproduces plane Couette exactly
//   - steady momentum balance: the MEM force on a pinned sphere in a periodic,
//     uniformly forced box equals the total force applied to the fluid
//   - the f32-emulated classification of a sphere at an integer centre equals
//     the exact integer test |p - c|^2 < a^2
//   - the sponge profile read back cell by cell after one step from rest
//     equals the smoothstep table of common_d3_sponge.wgsl (and a linear ramp
//     or a per-axis product planted in the oracle fails that check)
//   - D3Q19 / D3Q27 plane Couette between two moving y walls is exact
// Run: node tests/js/test-oracle.js   (exit code 1 on any failure)

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

function uniformMac(N, rho, u) {
  const m = new Float64Array(4 * N);
  for (let c = 0; c < N; c++) { m[4 * c] = rho; m[4 * c + 1] = u[0]; m[4 * c + 2] = u[1]; m[4 * c + 3] = u[2]; }
  return m;
}

ok('2D TGV: field within 1e-2 and energy decay rate within 1% of analytic', () => {
  const N = 48, tau = 0.8, u0 = 0.02, T = 600;
  const nu = (tau - 0.5) / 3, k = 2 * Math.PI / N;
  const o = new O.DenseOracle({ q: 9, nx: N, ny: N, nz: 1, omega: 1 / tau });
  const mac = new Float64Array(4 * N * N);
  for (let y = 0; y < N; y++) for (let x = 0; x < N; x++) {
    const c = y * N + x;
    // u = u0 (sin kx cos ky, -cos kx sin ky); p = (u0^2/4)(cos 2kx + cos 2ky)
    mac[4 * c + 1] = u0 * Math.sin(k * x) * Math.cos(k * y);
    mac[4 * c + 2] = -u0 * Math.cos(k * x) * Math.sin(k * y);
    mac[4 * c] = 1 + 3 * (u0 * u0 / 4) * (Math.cos(2 * k * x) + Math.cos(2 * k * y));
  }
  o.initEq(mac);
  const ke = () => { let s = 0; for (let c = 0; c < N * N; c++) s += o.mac[4 * c + 1] ** 2 + o.mac[4 * c + 2] ** 2; return s; };
  o.step(100);
  const ke1 = ke();
  o.step(T - 100);
  const ke2 = ke();
  const decay = Math.exp(-nu * 2 * k * k * T);
  let num = 0, den = 0;
  for (let y = 0; y < N; y++) for (let x = 0; x < N; x++) {
    const c = y * N + x;
    const ax = u0 * Math.sin(k * x) * Math.cos(k * y) * decay, ay = -u0 * Math.cos(k * x) * Math.sin(k * y) * decay;
    num += (o.mac[4 * c + 1] - ax) ** 2 + (o.mac[4 * c + 2] - ay) ** 2;
    den += ax * ax + ay * ay;
  }
  const l2 = Math.sqrt(num / den);
  const rate = Math.log(ke1 / ke2) / (T - 100);
  const rateExact = 4 * nu * k * k;
  assert(l2 < 1e-2, 'L2rel ' + l2);
  assert(Math.abs(rate / rateExact - 1) < 0.01, 'rate error ' + (rate / rateExact - 1));
  return 'L2rel ' + l2.toExponential(2) + ', rate err ' + (rate / rateExact - 1).toExponential(2);
});

for (const Q of [19, 27]) {
  ok('D3Q' + Q + ' Beltrami: energy decays at 2 nu k^2 within 2%', () => {
    const N = 24, tau = 0.7, T = 120;
    const nu = (tau - 0.5) / 3, k = 2 * Math.PI / N, A = 0.02, B = 0.015, C = 0.01;
    const o = new O.DenseOracle({ q: Q, nx: N, ny: N, nz: N, omega: 1 / tau });
    const n3 = N * N * N;
    const mac = new Float64Array(4 * n3);
    for (let z = 0; z < N; z++) for (let y = 0; y < N; y++) for (let x = 0; x < N; x++) {
      const c = (z * N + y) * N + x;
      const ux = A * Math.sin(k * z) + C * Math.cos(k * y);
      const uy = B * Math.sin(k * x) + A * Math.cos(k * z);
      const uz = C * Math.sin(k * y) + B * Math.cos(k * x);
      mac[4 * c] = 1 - (ux * ux + uy * uy + uz * uz) * 1.5;
      mac[4 * c + 1] = ux; mac[4 * c + 2] = uy; mac[4 * c + 3] = uz;
    }
    o.initEq(mac);
    const ke = () => { let s = 0; for (let c = 0; c < n3; c++) s += o.mac[4 * c + 1] ** 2 + o.mac[4 * c + 2] ** 2 + o.mac[4 * c + 3] ** 2; return s; };
    o.step(20);
    const k1 = ke();
    o.step(T - 20);
    const k2 = ke();
    const rate = Math.log(k1 / k2) / (T - 20);
    const exact = 2 * nu * k * k;
    assert(Math.abs(rate / exact - 1) < 0.02, 'rate error ' + (rate / exact - 1));
    return 'rate err ' + (rate / exact - 1).toExponential(2);
  });
}

for (const Q of [9, 19, 27]) {
  ok('Q' + Q + ' periodic box conserves mass to 1e-13 and gains exactly N*F momentum per step (BGK, TRT)', () => {
    for (const trt of [false, true]) {
      const nx = 10, ny = 8, nz = Q === 9 ? 1 : 6, N = nx * ny * nz;
      const F = [2e-5, -1e-5, Q === 9 ? 0 : 3e-5];
      const omega = 1 / 0.62;
      const desc = { q: Q, nx: nx, ny: ny, nz: nz, omega: omega, force: F, exactWeights: true };
      if (trt) desc.omegaMinus = 1 / (0.5 + 0.1875 / (1 / Math.fround(omega) - 0.5));
      const o = new O.DenseOracle(desc);
      const mac = new Float64Array(4 * N);
      let seed = 7;
      const rnd = () => { seed = (seed * 1103515245 + 12345) % 2147483648; return seed / 2147483648 - 0.5; };
      for (let c = 0; c < N; c++) { mac[4 * c] = 1 + 0.01 * rnd(); mac[4 * c + 1] = 0.03 * rnd(); mac[4 * c + 2] = 0.03 * rnd(); mac[4 * c + 3] = Q === 9 ? 0 : 0.03 * rnd(); }
      o.initEq(mac);
      const t0 = o.totals();
      const steps = 40;
      o.step(steps);
      const t1 = o.totals();
      assert(Math.abs(t1[0] - t0[0]) / t0[0] < 1e-13, 'mass drift ' + (t1[0] - t0[0]));
      for (let a = 0; a < 3; a++) {
        const expect = N * Math.fround(F[a]) * steps;
        assert(Math.abs((t1[1 + a] - t0[1 + a]) - expect) < 1e-12, 'momentum axis ' + a + ' got ' + (t1[1 + a] - t0[1 + a]) + ' want ' + expect + (trt ? ' (TRT)' : ''));
      }
    }
  });
}

function poiseuilleCase(Q, trt) {
  const H = 8, F = 1e-6, tau = 0.9;
  const nx = 2, nz = Q === 9 ? 1 : 2;
  const omega = Math.fround(1 / tau);
  const desc = { q: Q, nx: nx, ny: H, nz: nz, omega: omega, force: [F, 0, 0], walls: [0, 1, 0], exactWeights: true };
  if (trt) desc.omegaMinus = 1 / (0.5 + 0.1875 / (1 / omega - 0.5));
  const o = new O.DenseOracle(desc);
  o.initEq(uniformMac(nx * H * nz, 1, [0, 0, 0]));
  o.step(6000);
  const nu = (1 / omega - 0.5) / 3;
  let worst = 0;
  for (let y = 0; y < H; y++) {
    const exact = Math.fround(F) / (2 * nu) * (y + 0.5) * (H - 0.5 - y);
    const got = o.mac[4 * (y * nx) + 1];
    worst = Math.max(worst, Math.abs(got / exact - 1));
  }
  return worst;
}

ok('TRT Lambda=3/16 plane Poiseuille with halfway walls is exact (D2Q9, D3Q19)', () => {
  const e9 = poiseuilleCase(9, true), e19 = poiseuilleCase(19, true);
  assert(e9 < 1e-6 && e19 < 1e-6, 'max rel err Q9 ' + e9 + ' Q19 ' + e19);
  const bgk = poiseuilleCase(9, false);
  // BGK's halfway wall slips by a tau-dependent amount: the check must see it.
  assert(bgk > 1e-3, 'BGK unexpectedly exact: ' + bgk);
  return 'TRT ' + e9.toExponential(1) + ' / ' + e19.toExponential(1) + ', BGK ' + bgk.toExponential(1);
});

ok('BGK plane Couette with a moving y wall is exact (D2Q9)', () => {
  const H = 10, U = 0.01, nx = 2;
  const o = new O.DenseOracle({ q: 9, nx: nx, ny: H, nz: 1, omega: 1 / 0.8, walls: [0, 1, 0], wallU: [0, U], exactWeights: true });
  o.initEq(uniformMac(nx * H, 1, [0, 0, 0]));
  o.step(8000);
  let worst = 0;
  for (let y = 0; y < H; y++) {
    const exact = Math.fround(U) * (y + 0.5) / H;
    worst = Math.max(worst, Math.abs(o.mac[4 * (y * nx) + 1] - exact) / U);
  }
  assert(worst < 1e-7, 'max err / U ' + worst);
  return worst.toExponential(1);
});

ok('sphere classification equals the exact integer test at an integer centre', () => {
  const d = O.normalise({ q: 19, nx: 40, ny: 36, nz: 32, omega: 1, body: { a: 12, centre: [20, 18, 16] } });
  const cls = O.classify(d);
  for (let z = 0; z < 32; z++) for (let y = 0; y < 36; y++) for (let x = 0; x < 40; x++) {
    const r2 = (x - 20) ** 2 + (y - 18) ** 2 + (z - 16) ** 2;
    const c = cls[(z * 36 + y) * 40 + x];
    assert(((c & 1) !== 0) === (r2 < 144), 'solid bit at ' + [x, y, z]);
    assert(((c & 2) !== 0) === (r2 < 196), 'near bit at ' + [x, y, z]);
  }
});

ok('MEM force on a pinned sphere balances the applied body force at steady state (D3Q19, TRT)', () => {
  const n = 12, F = 2e-6, omega = Math.fround(1 / 0.8);
  const o = new O.DenseOracle({
    q: 19, nx: n, ny: n, nz: n, omega: omega, omegaMinus: 1 / (0.5 + 0.1875 / (1 / omega - 0.5)),
    force: [F, 0, 0], body: { a: 2.6, centre: [6, 6, 6] },
  });
  o.initEq(uniformMac(n * n * n, 1, [0, 0, 0]));
  o.step(4000);
  let fluid = 0;
  for (let c = 0; c < o.N; c++) if (!(o.cls[c] & 1)) fluid++;
  const want = Math.fround(F) * fluid;
  const got = o.forceHistory[o.forceHistory.length - 1];
  const err = Math.abs(got[0] / want - 1);
  assert(err < 1e-4, 'Fx ' + got[0] + ' want ' + want);
  assert(Math.abs(got[1]) < 1e-9 * want && Math.abs(got[2]) < 1e-9 * want, 'lateral ' + got[1] + ', ' + got[2]);
  return 'rel err ' + err.toExponential(2);
});

ok('uniform flow at the sponge velocity stays uniform under the sponge (D3Q19)', () => {
  const n = 10, u = [Math.fround(0.05), 0, 0];
  const o = new O.DenseOracle({ q: 19, nx: n, ny: n, nz: n, omega: 1 / 0.6, sponge: { width: 3, u: u }, exactWeights: true });
  o.initEq(uniformMac(n * n * n, 1, u));
  o.step(20);
  let worst = 0;
  for (let c = 0; c < o.N; c++) worst = Math.max(worst, Math.abs(o.mac[4 * c + 1] - Math.fround(0.05)), Math.abs(o.mac[4 * c] - 1));
  assert(worst < 1e-14, 'deviation ' + worst);
});

// The sponge profile.  From rest (f = w_i) streaming and the collision leave
// every population at w_i (the f32 weights sum to 1 exactly, so rho = 1), so
// after one step f_i = (1 - s) w_i + s feq_i(1, U) and s can be read back
// per cell.  It must follow the reference's spongeWeight3 (smoothstep of
// clamp(1 - d / W), common_d3_sponge.wgsl:10-14) with d the distance to the
// nearest enabled face, x = 0 and x = NX - 1 both being faces
// (common_d3_step.wgsl:317-320).  At width 5 the smoothstep takes the values
// tabled here by hand for d = 0..5 (t = 1, .8, .6, .4, .2, 0).
const SPONGE_TABLE_W5 = [1, 0.896, 0.648, 0.352, 0.104, 0];
function spongeReadBack(axes) {
  const NX = 20, NY = 16, NZ = 12, U = [Math.fround(0.05), 0, 0];
  const o = new O.DenseOracle({ q: 19, nx: NX, ny: NY, nz: NZ, omega: 1 / 0.6, sponge: { width: 5, axes: axes, u: U } });
  const N = NX * NY * NZ;
  o.initEq(uniformMac(N, 1, [0, 0, 0]));
  o.step(1);
  const w1 = o.T.w[1];  // direction 1 is e = (1, 0, 0)
  const target = w1 * (1 + 3 * U[0] + 4.5 * U[0] * U[0] - 1.5 * U[0] * U[0]);
  const s = new Float64Array(N);
  for (let c = 0; c < N; c++) s[c] = (o.f[1 * N + c] - w1) / (target - w1);
  return { s: s, at: (x, y, z) => s[(z * NY + y) * NX + x], NX: NX, NY: NY, NZ: NZ };
}
function spongeProfileError(axes) {
  const r = spongeReadBack(axes);
  let worst = 0;
  for (let z = 0; z < r.NZ; z++) {
    for (let y = 0; y < r.NY; y++) {
      for (let x = 0; x < r.NX; x++) {
        const d = Math.min(axes[0] ? Math.min(x, r.NX - 1 - x) : 99, axes[1] ? Math.min(y, r.NY - 1 - y) : 99,
          axes[2] ? Math.min(z, r.NZ - 1 - z) : 99);
        worst = Math.max(worst, Math.abs(r.at(x, y, z) - SPONGE_TABLE_W5[Math.min(d, 5)]));
      }
    }
  }
  return { worst: worst, r: r };
}

ok('sponge profile: one step from rest gives the smoothstep table at every cell, faces at 0 and N-1, min at corners', () => {
  const all = spongeProfileError([1, 1, 1]);
  assert(all.worst < 1e-14, 'all axes: worst deviation ' + all.worst);
  const r = all.r;
  // single cells named explicitly: the far face x = NX-1, a corner (min, not a
  // product: 0.896 * 0.896 = 0.803 would be a product), the interior
  const cases = [[0, 8, 6, 1], [19, 8, 6, 1], [18, 8, 6, 0.896], [1, 1, 6, 0.896], [2, 8, 6, 0.648],
    [17, 13, 6, 0.648], [3, 8, 6, 0.352], [4, 8, 6, 0.104], [5, 8, 6, 0], [10, 8, 5, 0], [1, 2, 11, 1]];
  for (const [x, y, z, want] of cases) {
    assert(Math.abs(r.at(x, y, z) - want) < 1e-14, 's(' + x + ',' + y + ',' + z + ') = ' + r.at(x, y, z) + ' want ' + want);
  }
  const masked = spongeProfileError([1, 0, 1]);
  assert(masked.worst < 1e-14, 'y masked: worst deviation ' + masked.worst);
  assert(Math.abs(masked.r.at(10, 0, 6)) < 1e-14 && Math.abs(r.at(10, 0, 6) - 1) < 1e-14, 'masked y face');
  return 'worst ' + Math.max(all.worst, masked.worst).toExponential(1);
});

ok('sponge profile pin has teeth: a linear ramp and a per-axis product both fail it', () => {
  const real = O.hooks.spongeWeight3;
  const clampT = (d, sw) => { const t = 1 - d / sw; return t < 0 ? 0 : (t > 1 ? 1 : t); };
  const smooth = (t) => t * t * (3 - 2 * t);
  const mutants = {
    linear: (dx, dy, dz, sw) => clampT(Math.min(dx, dy, dz), sw),
    product: (dx, dy, dz, sw) => 1 - (1 - smooth(clampT(dx, sw))) * (1 - smooth(clampT(dy, sw))) * (1 - smooth(clampT(dz, sw))),
    productOfWeights: (dx, dy, dz, sw) => smooth(clampT(dx, sw)) * smooth(clampT(dy, sw)) * smooth(clampT(dz, sw)),
  };
  const out = [];
  try {
    for (const name of Object.keys(mutants)) {
      O.hooks.spongeWeight3 = mutants[name];
      const e = spongeProfileError([1, 1, 1]).worst;
      assert(e > 1e-2, name + ' mutant passes the pin (deviation ' + e + ')');
      out.push(name + ' ' + e.toFixed(3));
    }
  } finally {
    O.hooks.spongeWeight3 = real;
  }
  return out.join(', ');
});

// Moving y-walls in 3D (this build's extension of the reference's stationary
// halfway walls, common_d3_step.wgsl:264-269, with the 2D moving-wall term of
// common_walls.wgsl:27-31): BGK plane Couette between a bottom wall moving at
// U0 and a top wall at U1 is linear, u_x(y) = U0 + (U1 - U0)(y + 1/2)/H, and
// halfway bounce-back reproduces it exactly.
for (const q of [19, 27]) {
  ok('BGK plane Couette between two moving y walls is exact (D3Q' + q + ')', () => {
    const H = 10, U0 = Math.fround(-0.004), U1 = Math.fround(0.01), nx = 2, nz = 2;
    const o = new O.DenseOracle({ q: q, nx: nx, ny: H, nz: nz, omega: 1 / 0.8, walls: [0, 1, 0], wallU: [U0, U1], exactWeights: true });
    o.initEq(uniformMac(nx * H * nz, 1, [0, 0, 0]));
    o.step(8000);
    let worst = 0, cross = 0;
    for (let z = 0; z < nz; z++) {
      for (let y = 0; y < H; y++) {
        const c = (z * H + y) * nx;
        const exact = U0 + (U1 - U0) * (y + 0.5) / H;
        worst = Math.max(worst, Math.abs(o.mac[4 * c + 1] - exact) / U1);
        cross = Math.max(cross, Math.abs(o.mac[4 * c + 2]), Math.abs(o.mac[4 * c + 3]));
      }
    }
    assert(worst < 1e-7, 'max err / U ' + worst);
    assert(cross < 1e-12, 'cross-stream velocity ' + cross);
    return worst.toExponential(1);
  });
}

// The diffuse coupling's momentum budget.  In a periodic box without walls,
// streaming conserves momentum and Guo's forcing adds exactly the cell's
// force F = F0 + rho chi (U_s - u*) to it (BGK and TRT alike), so the fluid's
// momentum change per step is N F0 minus the force the body feels.  On a
// 12^3 box around a radius-2 sphere every cell has chi >= 1e-6, so the force
// record covers every penalised cell and the budget closes to round-off.
// A force taken from the post-forcing velocity u instead of u*, a dropped
// rho, or the wrong sign each break it.
for (const trt of [false, true]) {
  ok('diffuse coupling: the force record closes the momentum budget (D3Q19, ' + (trt ? 'TRT' : 'BGK') + ')', () => {
    const n = 12, N = n * n * n, F0 = [Math.fround(2e-5), Math.fround(-1e-5), 0];
    const om = Math.fround(1 / 0.6);
    // exact weights: with the f32 weights the equilibrium's momentum is
    // rho u (1 + ~1e-8), a drift of its own that would hide in the budget
    const desc = { q: 19, nx: n, ny: n, nz: n, omega: om, force: F0, diffuse: 1, chiEps: 1.5, exactWeights: true,
                   body: { a: 2, centre: [6.3, 5.8, 6.1] } };
    if (trt) desc.omegaMinus = Math.fround(1 / (0.5 + (3 / 16) / (1 / om - 0.5)));
    const o = new O.DenseOracle(desc);
    let minChi = 1;
    for (let z = 0; z < n; z++) for (let y = 0; y < n; y++) for (let x = 0; x < n; x++) {
      minChi = Math.min(minChi, O.chiOf(O.phiF32(o.S, x, y, z), o.S.chiEps));
    }
    assert(minChi >= 1e-6, 'a cell has chi < 1e-6: ' + minChi);
    o.initEq(uniformMac(N, 1, [0.03, 0.01, -0.005]));
    let worst = 0, scale = 0;
    for (let s = 0; s < 20; s++) {
      const p0 = o.totals();
      o.step(1);
      const p1 = o.totals();
      const rec = o.forceHistory[o.forceHistory.length - 1];
      for (let k = 0; k < 3; k++) {
        const budget = N * F0[k] - rec[k];
        worst = Math.max(worst, Math.abs((p1[k + 1] - p0[k + 1]) - budget));
        scale = Math.max(scale, Math.abs(rec[k]));
      }
    }
    assert(scale > 0.05, 'force too small to test: ' + scale);
    assert(worst < 1e-10 * scale, 'budget off by ' + worst + ' (force ' + scale + ')');
    return 'off by ' + worst.toExponential(1) + ' of a force ' + scale.toFixed(3);
  });
}

process.exitCode = failures ? 1 : 0;
