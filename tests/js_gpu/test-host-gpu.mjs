// GPU test of the JavaScript host path: lbm-host.mjs -> N-API addon ->
// liblbm.so, checked against the FP64 oracle run in the same process.
// Run: node tests/js_gpu/test-host-gpu.mjs   (needs an MI355X)
import { createRequire } from 'module';
import { DenseSolver, AmrSolver, ServeClient, sphereScenario, tgvScenario, deviceInfo } from '../../getting-air_amd/js/lbm-host.mjs';
import { makePool, refineWhere, refineNearBody, nearBodyWant, refineHierarchy, manageSlots, activeSlotList } from '../../getting-air_amd/js/d3-amr.mjs';
import * as Q2 from '../../getting-air_amd/js/amr2d.mjs';

const require = createRequire(import.meta.url);
const O = require('../../oracle/lbm_oracle.js');
const RO = require('../../oracle/rigid_oracle.js');

let failures = 0;
async function ok(name, fn) {
  try { const m = await fn(); console.log('ok   ' + name + (m ? '  (' + m + ')' : '')); } catch (e) {
    failures++; console.log('FAIL ' + name + ': ' + (e && e.stack ? e.stack : e));
  }
}
function assert(c, m) { if (!c) throw new Error(m); }
function pool_blocks(dims) { return dims[0] * dims[1] * dims[2] / 64; }
function relErr(gpuMac, oracleMac, comp) {
  let num = 0, den = 0;
  const n = oracleMac.length / 4;
  for (let c = 0; c < n; c++) {
    for (const k of comp) {
      const d = gpuMac[4 * c + k] - oracleMac[4 * c + k];
      num += d * d; den += oracleMac[4 * c + k] * oracleMac[4 * c + k];
    }
  }
  return Math.sqrt(num / den);
}

(async () => {
  await ok('device info through the addon', () => {
    const d = deviceInfo();
    assert(d.computeUnits > 0 && d.hbmBytes > 1e9, JSON.stringify(d));
    return d.name + ', ' + d.computeUnits + ' CUs';
  });

  await ok('sphere (D3Q19 TRT, bounce-back, sponge): macro field and drag history match the oracle', async () => {
    const sc = sphereScenario({ R: 4, dims: [48, 24, 24], centre: [14, 12, 12], re: 80, trt: true });
    const s = new DenseSolver(sc);
    await s.reset();
    const run = await s.debugRunAndCollect(40, 10);
    assert(run.history.length === 4 && run.step === 40, 'history ' + run.history.length);
    const body = await s.readBody();
    assert(Number.isFinite(body.cd), 'cd ' + body.cd);
    const mac = await s.readMacro();
    const hist = await s.readForceHistory(40);
    const stats = await s.readStats();
    s.destroy();
    const o = new O.DenseOracle(sc.desc);
    o.initEq(sc.macro());
    o.step(40);
    const eu = relErr(mac, o.mac, [1, 2, 3]), er = relErr(mac, o.mac, [0]);
    assert(eu < 1e-5 && er < 1e-5, 'u ' + eu + ' rho ' + er);
    let dmax = 0, scale = 0;
    for (let s2 = 0; s2 < 40; s2++) scale = Math.max(scale, Math.abs(o.forceHistory[s2][0]));
    for (let s2 = 0; s2 < 40; s2++) {
      for (let k = 0; k < 3; k++) dmax = Math.max(dmax, Math.abs(hist[6 * s2 + k] - o.forceHistory[s2][k]) / scale);
    }
    assert(dmax < 1e-6, 'force deviation ' + dmax);
    const tot = o.totals();
    assert(Math.abs(stats.mass / tot[0] - 1) < 1e-6, 'mass ' + stats.mass + ' vs ' + tot[0]);
    return 'u ' + eu.toExponential(2) + ', drag dev ' + dmax.toExponential(2) + ', Cd ' + body.cd.toFixed(4);
  });

  await ok('TGV (D2Q9 BGK) readField matches the oracle', async () => {
    const sc = tgvScenario({ n: 64, tau: 0.6, u0: 0.05 });
    const s = new DenseSolver(sc);
    await s.reset();
    await s.debugStepSync(100);
    const f = await s.readField();
    s.destroy();
    const o = new O.DenseOracle(sc.desc);
    o.initEq(sc.macro());
    o.step(100);
    let num = 0, den = 0;
    for (let c = 0; c < 64 * 64; c++) {
      num += (f.ux[c] - o.mac[4 * c + 1]) ** 2 + (f.uy[c] - o.mac[4 * c + 2]) ** 2;
      den += o.mac[4 * c + 1] ** 2 + o.mac[4 * c + 2] ** 2;
    }
    const e = Math.sqrt(num / den);
    assert(e < 1e-5, 'u err ' + e);
    return 'u ' + e.toExponential(2);
  });

  await ok('AMR (D3Q27, 2 levels): block maps and active list follow d3-amr.mjs manageSlots bit for bit', async () => {
    const n = 32, rb = 4, k = 2 * Math.PI / n, A = 0.02;
    const macro = new Float32Array(4 * n * n * n);
    for (let z = 0; z < n; z++) for (let y = 0; y < n; y++) for (let x = 0; x < n; x++) {
      const c = 4 * ((z * n + y) * n + x);
      const ux = A * Math.sin(k * x) * Math.cos(k * y) * Math.cos(k * z);
      const uy = -A * Math.cos(k * x) * Math.sin(k * y) * Math.cos(k * z);
      macro[c] = 1 - 1.5 * (ux * ux + uy * uy); macro[c + 1] = ux; macro[c + 2] = uy; macro[c + 3] = 0;
    }
    // thresholds on the scale A k sqrt(3); 64 of the 512 blocks reach 0.9 of it
    const wmax = A * k * Math.sqrt(3);
    const pool = makePool({ dims: [n, n, n], rb: rb, maxSlots: 128 });
    const topo = refineWhere(pool, (b) => b.bx === 3 && b.by === 3);
    const s = new AmrSolver({
      desc: { q: 27, nx: n, ny: n, nz: n, rb: rb, maxSlots: 128, omega: 1 / 0.6, manageEvery: 2, lead: 0,
        refineAt: 0.9 * wmax, keepAt: 0.85 * wmax },
      macro: () => macro, topology: topo,
    });
    await s.reset();
    let prev = await s.debugPoolState();
    assert(prev.activeTiles === topo.activeSlots, 'seeded tiles ' + prev.activeTiles);
    const st0 = await s.readStats();
    let changes = 0;
    for (let step = 0; step < 6; step++) {
      await s.debugStepSync(1);
      if (step % 2 !== 0) continue;
      const cur = await s.debugPoolState();
      for (let b = 0; b < pool.nBlocks; b++) {
        const held = prev.blockSlot[b] >= 0;
        const w = held ? cur.crit[b] >= Math.fround(0.85 * wmax) : cur.crit[b] >= Math.fround(0.9 * wmax);
        assert((w ? 1 : 0) === cur.want[b], 'want of block ' + b + ' at step ' + step);
      }
      const twin = manageSlots(prev.blockSlot, prev.slotToBlock, cur.want);
      for (let b = 0; b < pool.nBlocks; b++) assert(twin.blockSlot[b] === cur.blockSlot[b], 'blockSlot ' + b);
      for (let q = 0; q < 128; q++) assert(twin.slotToBlock[q] === cur.slotToBlock[q], 'slotToBlock ' + q);
      const act = activeSlotList(twin.slotToBlock);
      assert(act.length === cur.active.length && act.every((v, i) => v === cur.active[i]), 'active list');
      changes += twin.coarsened.length + twin.refined.length;
      prev = cur;
    }
    const st = await s.readStats();
    s.destroy();
    assert(changes > 0, 'no topology change exercised');
    const dm = Math.abs(st.mass - st0.mass) / st0.mass;
    assert(dm < 1e-6, 'mass drift ' + dm);
    return changes + ' topology changes, ' + prev.activeTiles + ' tiles, mass drift ' + dm.toExponential(2);
  });

  await ok('serve loop: a sphere gives the same bits alone and in a mix, and matches DenseSolver', async () => {
    const sphere = sphereScenario({ R: 4, dims: [48, 24, 24], centre: [14, 12, 12], re: 80, trt: true });
    const others = [tgvScenario({ n: 64, tau: 0.6, u0: 0.05 }),
      sphereScenario({ R: 3, dims: [32, 32, 32], centre: [10, 16, 16], re: 60, trt: true })];
    const aloneRun = async () => {
      const c = new ServeClient({ budgetBytes: 1 << 30, chunkSteps: 16 });
      const r = await c.result(c.submit(sphere, 60), 120);
      c.destroy();
      return r;
    };
    const ra = await aloneRun();
    const rb = await aloneRun();
    const mixed = new ServeClient({ budgetBytes: 1 << 30, chunkSteps: 16 });
    const t0 = mixed.submit(others[0], 50);
    const t1 = mixed.submit(sphere, 60);
    const t2 = mixed.submit(others[1], 40);
    const [r0, r1, r2] = await Promise.all([mixed.result(t0, 120), mixed.result(t1, 120), mixed.result(t2, 120)]);
    const st = mixed.stats();
    mixed.destroy();
    const s = new DenseSolver(sphere);
    await s.reset();
    await s.debugStepSync(60);
    const dm = await s.readMacro();
    s.destroy();
    const cmp = (x, y) => {
      let n = 0, mx = 0;
      for (let c = 0; c < x.length; c++) if (!Object.is(x[c], y[c])) { n++; mx = Math.max(mx, Math.abs(x[c] - y[c])); }
      return n + ' (max ' + mx.toExponential(2) + ')';
    };
    const report = 'alone/alone ' + cmp(ra.macro, rb.macro) + ', alone/mix ' + cmp(ra.macro, r1.macro) +
      ', alone/dense ' + cmp(ra.macro, dm) + ', mix/dense ' + cmp(r1.macro, dm) +
      ', steps ' + [ra.steps, rb.steps, r1.steps, r0.steps, r2.steps].join('/');
    console.log('     ' + report);
    assert(r0.steps === 50 && r2.steps === 40 && st.completed === 3, JSON.stringify(st));
    assert(cmp(ra.macro, r1.macro).startsWith('0 '), report);
    assert(cmp(ra.forces, r1.forces).startsWith('0 '), 'forces differ: ' + report);
    assert(cmp(r1.macro, dm).startsWith('0 '), report);
    return 'p50 ' + st.p50Latency.toFixed(3) + ' s; ' + report;
  });

  await ok('checkpoint: saveState/loadState and a ServeClient request resumed from it give the straight run bits', async () => {
    const sc = sphereScenario({ R: 4, dims: [48, 24, 24], centre: [14, 12, 12], re: 80, trt: true, storage: 'f16' });
    const a = new DenseSolver(sc);
    await a.reset();
    await a.debugStepSync(50);
    const ref = await a.readMacro();
    a.destroy();
    const b = new DenseSolver(sc);
    await b.reset();
    await b.debugStepSync(20);
    const st = await b.saveState();
    b.destroy();
    const c = new DenseSolver(sc);
    await c.loadState(st);
    assert(c.getStep() === 20, 'step after loadState ' + c.getStep());
    await c.debugStepSync(30);
    const m1 = await c.readMacro();
    c.destroy();
    const sv = new ServeClient({ budgetBytes: 1 << 30, chunkSteps: 8 });
    const r = await sv.result(sv.submit(sc, 30, { state: st, keepState: true }), 120);
    sv.destroy();
    const diff = (x, y) => { let n = 0; for (let k = 0; k < x.length; k++) if (!Object.is(x[k], y[k])) n++; return n; };
    const d1 = diff(m1, ref), d2 = diff(r.macro, ref);
    assert(d1 === 0 && d2 === 0, 'differing values: resumed solver ' + d1 + ', resumed request ' + d2);
    assert(r.state && r.state.byteLength === st.byteLength, 'served request kept its state');
    return 'state ' + st.byteLength + ' bytes, 0 differing values';
  });

  await ok('moving body: a falling spheroid through attachBody/readBody matches the rigid-body oracle', async () => {
    const desc = { q: 19, nx: 40, ny: 24, nz: 24, omega: 1 / 0.7 };
    const N = 40 * 24 * 24;
    const macro = () => { const m = new Float32Array(4 * N); for (let c = 0; c < N; c++) m[4 * c] = 1; return m; };
    const body = { shape: { kind: 1, a: 4, c: 2.5 }, x: [12.4, 12.1, 11.8], q: RO.qFromAxisAngle([0, 1, 0], 0.6),
                   density: 1.5, gravity: [2e-4, 0, 0] };
    const s = new DenseSolver({ desc: desc, params: {}, macro: macro });
    await s.reset();
    s.attachBody(body);
    await s.debugStepSync(80);
    const b = await s.readBody();
    const mac = await s.readMacro();
    s.destroy();
    const o = new RO.RigidOracle(desc, body);
    o.initEq(macro());
    o.step(80);
    const eu = relErr(mac, o.mac, [1, 2, 3]);
    const ex = Math.max(Math.abs(b.cx - o.B.x[0]), Math.abs(b.cy - o.B.x[1]), Math.abs(b.cz - o.B.x[2]));
    const eq = Math.max(...[b.qw, b.qx, b.qy, b.qz].map((v, k) => Math.abs(v - o.B.q[k])));
    assert(eu < 1e-5 && ex < 1e-6 && eq < 1e-6, 'u ' + eu + ' x ' + ex + ' q ' + eq);
    return 'u ' + eu.toExponential(2) + ', x ' + ex.toExponential(2) + ', vx ' + b.vx.toExponential(3);
  });

  await ok('AMR sphere on the fine level (sphere-amr-Re20-D16) through the host', async () => {
    const sc = sphereScenario({ re: 20, R: 8, u0: 0.05, spongeWidth: 8 });
    const pool = makePool({ dims: sc.params.dims, rb: 4, maxSlots: pool_blocks(sc.params.dims) });
    const c = sc.params.centre;
    const topo = refineNearBody(pool, (p) => Math.hypot(p[0] - c[0], p[1] - c[1], p[2] - c[2]) - 8, 2);
    const desc = { q: 19, nx: sc.desc.nx, ny: sc.desc.ny, nz: sc.desc.nz, rb: 4, maxSlots: topo.activeSlots,
      omega: sc.desc.omega, sponge: { width: 8, axes: [1, 1, 1], u: [0.05, 0, 0] }, body: sc.desc.body };
    const s = new AmrSolver({ desc: desc, macro: sc.macro, topology: topo });
    await s.reset();
    await s.debugStepSync(30 * 320);
    const b = await s.readBody();
    const cd = b.fx / (0.5 * 0.05 * 0.05 * sc.params.area);
    assert(topo.activeSlots === 160 && Math.abs(cd / 2.9416 - 1) < 0.005, 'tiles ' + topo.activeSlots + ' Cd ' + cd);
    return topo.activeSlots + ' tiles, Cd ' + cd.toFixed(4) + ' (reference 2.9416)';
  });

  await ok('AMR sphere on the finest of three levels (sphere-amr-L3-Re20-D16) through the host', async () => {
    const sc = sphereScenario({ re: 20, R: 8, u0: 0.05, spongeWidth: 8 });
    const pool = makePool({ dims: sc.params.dims, rb: 4, maxSlots: pool_blocks(sc.params.dims) });
    const c = sc.params.centre;
    const sdf = (p) => Math.hypot(p[0] - c[0], p[1] - c[1], p[2] - c[2]) - 8;
    const h = refineHierarchy(pool, { levels: 3, want: nearBodyWant(sdf, 2) });
    const tiles = [h.byLevel[1].activeSlots, h.byLevel[2].activeSlots];
    const desc = { q: 19, nx: sc.desc.nx, ny: sc.desc.ny, nz: sc.desc.nz, rb: 4, maxSlots: Math.max(...tiles),
      omega: sc.desc.omega, sponge: { width: 8, axes: [1, 1, 1], u: [0.05, 0, 0] }, body: sc.desc.body, levels: 3 };
    const s = new AmrSolver({ desc: desc, macro: sc.macro, hierarchy: h });
    await s.reset();
    await s.debugStepSync(30 * 320);
    const b = await s.readBody();
    s.destroy();
    const cd = b.fx / (0.5 * 0.05 * 0.05 * sc.params.area);
    assert(tiles[0] === 432 && tiles[1] === 1280 && Math.abs(cd / 2.9015 - 1) < 0.005 && Math.abs(cd / 2.9436 - 1) < 0.03,
      'tiles ' + tiles + ' Cd ' + cd);
    return tiles.join(' + ') + ' tiles, Cd ' + cd.toFixed(4) + ' (reference 2.9015)';
  });

  await ok('AMR 2D quadtree (D2Q9, tgv amr-N64-L2 with a refined centre) through the host', async () => {
    const sc = tgvScenario({ n: 64, tau: 0.8, u0: 0.04 });
    const pool = Q2.makePool({ dims: [64, 64], rb: 8, maxSlots: 16 });
    const topo = Q2.refineWhere(pool, (b) => b.bx >= 2 && b.bx < 6 && b.by >= 2 && b.by < 6);
    const desc = { q: 9, nx: 64, ny: 64, nz: 1, rb: 8, maxSlots: 16, omega: sc.desc.omega };
    const s = new AmrSolver({ desc: desc, macro: sc.macro, topology: topo });
    assert(s.nBlocks === pool.nBlocks, 'nBlocks ' + s.nBlocks);
    await s.reset();
    const st = await s.debugPoolState();
    assert(st.activeTiles === 16, 'tiles ' + st.activeTiles);
    for (let b = 0; b < pool.nBlocks; b++) assert(st.blockSlot[b] === topo.blockSlot[b], 'blockSlot ' + b);
    const t = Math.round(0.5 * sc.params.td);
    await s.debugStepSync(t);
    const m = await s.readMacro();
    s.destroy();
    const k = sc.params.kx, dec = Math.exp(-t / sc.params.td);
    let num = 0, den = 0;
    for (let y = 0; y < 64; y++) for (let x = 0; x < 64; x++) {
      const c = y * 64 + x;
      const ax = 0.04 * Math.sin(k * x) * Math.cos(k * y) * dec, ay = -0.04 * Math.cos(k * x) * Math.sin(k * y) * dec;
      num += (m[4 * c + 1] - ax) ** 2 + (m[4 * c + 2] - ay) ** 2;
      den += ax * ax + ay * ay;
    }
    const e = Math.sqrt(num / den);
    assert(e < 0.01, 'L2rel ' + e);
    const mlups = Q2.cellUpdatesPerMacroStep({ rootCells: 64 * 64, rb: 8, activeByLevel: { 1: 16 } });
    return '16 tiles, L2rel ' + e.toExponential(2) + ' at ' + t + ' steps, ' + mlups + ' cell updates per macro-step';
  });

  await ok('AMR 2D quadtree manager through the host: block maps follow manageSlots bit for bit', async () => {
    const sc = tgvScenario({ n: 96, tau: 0.6, u0: 0.05 });
    // |curl u| of the TGV peaks at 2 k u0; thresholds near it select a changing subset as it decays
    const wmax = 2 * (2 * Math.PI / 96) * 0.05;
    const desc = { q: 9, nx: 96, ny: 96, nz: 1, rb: 4, maxSlots: 256, omega: sc.desc.omega, manageEvery: 2, lead: 0,
      refineAt: 0.8 * wmax, keepAt: 0.75 * wmax };
    const s = new AmrSolver({ desc: desc, macro: sc.macro });
    await s.reset();
    let prev = await s.debugPoolState();
    let changes = 0;
    for (let step = 0; step < 40; step++) {
      await s.debugStepSync(1);
      if (step % 2 !== 0) continue;
      const cur = await s.debugPoolState();
      const twin = manageSlots(prev.blockSlot, prev.slotToBlock, cur.want);
      for (let b = 0; b < s.nBlocks; b++) assert(twin.blockSlot[b] === cur.blockSlot[b], 'blockSlot ' + b + ' step ' + step);
      for (let q = 0; q < 256; q++) assert(twin.slotToBlock[q] === cur.slotToBlock[q], 'slotToBlock ' + q);
      const act = activeSlotList(twin.slotToBlock);
      assert(act.length === cur.active.length && act.every((v, i) => v === cur.active[i]), 'active list');
      changes += twin.coarsened.length + twin.refined.length;
      prev = cur;
    }
    s.destroy();
    assert(changes > 0, 'no topology change exercised');
    return changes + ' topology changes over ' + s.nBlocks + ' blocks';
  });

  await ok('invalid descriptor raises with the library message', () => {
    let msg = '';
    try { new DenseSolver({ desc: { q: 19, nx: 4, ny: 4, nz: 4, omega: 3 }, params: {}, macro: null }); } catch (e) { msg = e.message; }
    assert(/omega/.test(msg), 'message: ' + msg);
  });

  await ok('addon: short arrays raise RangeError, calls during a step in flight and after destroy raise', async () => {
    const a = (await import('../../getting-air_amd/js/lbm-host.mjs')).loadAddon();
    const h = a.createSolve({ q: 19, nx: 16, ny: 8, nz: 8, omega: 1.2 });
    const n = 16 * 8 * 8;
    const throwsAs = (fn, cls, re) => {
      let e = null;
      try { fn(); } catch (x) { e = x; }
      assert(e instanceof cls && re.test(e.message), 'expected ' + cls.name + ' ' + re + ', got ' + e);
    };
    throwsAs(() => a.uploadMacro(h, new Float32Array(4 * n - 1)), RangeError, /needed/);
    const mac = new Float32Array(4 * n);
    for (let c = 0; c < n; c++) mac[4 * c] = 1;
    a.uploadMacro(h, mac);
    throwsAs(() => a.readMacro(h, new Float32Array(4 * n - 4)), RangeError, /needed/);
    throwsAs(() => a.readF(h, new Float32Array(19 * n - 1)), RangeError, /needed/);
    throwsAs(() => a.uploadF(h, new Float32Array(18 * n)), RangeError, /needed/);
    throwsAs(() => a.readClasses(h, new Uint8Array(n - 1)), RangeError, /needed/);
    const p = a.step(h, 64);
    throwsAs(() => a.readMacro(h, new Float32Array(4 * n)), Error, /in flight/);
    throwsAs(() => a.destroy(h), Error, /in flight/);
    await p;
    a.readMacro(h, mac);
    assert(Math.abs(mac[0] - 1) < 1e-6, 'rho ' + mac[0]);
    a.destroy(h);
    throwsAs(() => a.readMacro(h, mac), Error, /destroyed/);
    throwsAs(() => a.amrInfo(h), TypeError, /AMR solve handle/);
  });

  process.exitCode = failures ? 1 : 0;
})();
