// GPU test of the JavaScript host path: lbm-host.mjs -> N-API addon ->
// liblbm.so, checked against the FP64 oracle run in the same process.
// Run: node tests/js_gpu/test-host-gpu.mjs   (needs an MI355X)
import { createRequire } from 'module';
import { DenseSolver, sphereScenario, tgvScenario, deviceInfo } from '../../getting-air_amd/js/lbm-host.mjs';

const require = createRequire(import.meta.url);
const O = require('../../oracle/lbm_oracle.js');

let failures = 0;
async function ok(name, fn) {
  try { const m = await fn(); console.log('ok   ' + name + (m ? '  (' + m + ')' : '')); } catch (e) {
    failures++; console.log('FAIL ' + name + ': ' + (e && e.stack ? e.stack : e));
  }
}
function assert(c, m) { if (!c) throw new Error(m); }
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

  await ok('invalid descriptor raises with the library message', () => {
    let msg = '';
    try { new DenseSolver({ desc: { q: 19, nx: 4, ny: 4, nz: 4, omega: 3 }, params: {}, macro: null }); } catch (e) { msg = e.message; }
    assert(/omega/.test(msg), 'message: ' + msg);
  });

  process.exitCode = failures ? 1 : 0;
})();
