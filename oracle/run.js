'use strict';
// Command-line driver for the FP64 oracle (test infrastructure only).
//
//   node oracle/run.js job.json
//
// job.json:
//   { "desc":    { q, nx, ny, nz, omega, omegaMinus, force, walls, wallU,
//                  sponge: {width, axes, u}, body: {a, centre}, solidEq },
//     "macIn":   "mac.f32"     float32 [rho,ux,uy,uz] per cell -> initEq
//     "fIn":     "f.f32"       optional float32 plane-major populations
//     "steps":   n,
//     "threads": T             worker threads over x-rows (default 1)
//     "out":     { "mac": path, "f": path, "force": path, "cls": path, "body": path },
//     "rigid":   optional rigid-body parameters (rigid_oracle.js makeBody); the
//                run is then single-threaded and "body" receives, per step,
//                [x(3), q(4), v(3), L(3), w(3), d(3)] }
//
// Writes float64 little-endian files (mac: 4 per cell, f: Q planes, force:
// 6 per step, cls: uint8 per cell) and prints one JSON line with the wall
// time, the thread count, the CPU model and the FP64 totals.

const fs = require('fs');
const os = require('os');
const path = require('path');
const { Worker, isMainThread, parentPort, workerData } = require('worker_threads');
const O = require('./lbm_oracle.js');
const R = require('./rigid_oracle.js');

function readF32(p) {
  const b = fs.readFileSync(p);
  return new Float32Array(b.buffer, b.byteOffset, b.byteLength / 4);
}

function writeArr(p, arr) {
  fs.writeFileSync(p, Buffer.from(arr.buffer, arr.byteOffset, arr.byteLength));
}

if (!isMainThread) {
  // Worker: owns x-rows [r0, r1); steps on command, replies with its force part.
  const d = workerData;
  const S = O.normalise(d.desc);
  const T = O.tables(S.q, S.exactWeights);
  const cls = new Uint8Array(d.cls);
  const bufs = [new Float64Array(d.fa), new Float64Array(d.fb)];
  const mac = new Float64Array(d.mac);
  parentPort.on('message', (m) => {
    const fin = bufs[m.parity], fout = bufs[1 - m.parity];
    const force = S.body ? O.forceSlab(S, T, cls, fin, d.r0, d.r1) : null;
    O.stepSlab(S, T, cls, fin, fout, m.writeMac ? mac : null, d.r0, d.r1);
    parentPort.postMessage({ force: force });
  });
} else {
  main().catch((e) => { console.error(e && e.stack ? e.stack : e); process.exit(1); });
}

async function main() {
  const job = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
  const base = path.dirname(path.resolve(process.argv[2]));
  const rel = (p) => path.resolve(base, p);
  if (job.rigid) return rigidMain(job, rel);
  const S = O.normalise(job.desc);
  const T = O.tables(S.q, S.exactWeights);
  const N = S.nx * S.ny * S.nz;
  const Q = T.Q;
  const rows = S.nz * S.ny;
  const threads = Math.max(1, Math.min(job.threads || 1, rows));
  const out = job.out || {};

  const fa = new SharedArrayBuffer(8 * Q * N);
  const fb = new SharedArrayBuffer(8 * Q * N);
  const macBuf = new SharedArrayBuffer(8 * 4 * N);
  const clsBuf = new SharedArrayBuffer(N);
  const cls = new Uint8Array(clsBuf);
  cls.set(O.classify(S));
  const f0 = new Float64Array(fa);
  const mac = new Float64Array(macBuf);

  // Initial state: equilibrium of the uploaded macroscopic field, or f itself.
  const o = new O.DenseOracle(job.desc);
  if (job.macIn) o.initEq(readF32(rel(job.macIn)));
  if (job.fIn) o.uploadF(readF32(rel(job.fIn)));
  f0.set(o.f);
  mac.set(o.mac);

  const steps = job.steps | 0;
  const forces = new Float64Array(6 * steps);
  const t0 = process.hrtime.bigint();
  let parity = 0;
  if (threads === 1) {
    const bufs = [f0, new Float64Array(fb)];
    for (let s = 0; s < steps; s++) {
      const fin = bufs[parity], fout = bufs[1 - parity];
      if (S.body) forces.set(O.forceSlab(S, T, cls, fin, 0, rows), 6 * s);
      O.stepSlab(S, T, cls, fin, fout, s === steps - 1 ? mac : null, 0, rows);
      parity = 1 - parity;
    }
  } else {
    const workers = [];
    for (let k = 0; k < threads; k++) {
      const r0 = Math.floor(k * rows / threads), r1 = Math.floor((k + 1) * rows / threads);
      workers.push(new Worker(__filename, { workerData: { desc: job.desc, cls: clsBuf, fa: fa, fb: fb, mac: macBuf, r0: r0, r1: r1 } }));
    }
    for (let s = 0; s < steps; s++) {
      const parts = await Promise.all(workers.map((w) => new Promise((resolve) => {
        w.once('message', (m) => resolve(m.force));
        w.postMessage({ parity: parity, writeMac: s === steps - 1 });
      })));
      if (S.body) {
        // Fixed worker order, so the sum is the same on every run.
        const acc = [0, 0, 0, 0, 0, 0];
        parts.forEach((p) => { for (let j = 0; j < 6; j++) acc[j] += p[j]; });
        forces.set(acc, 6 * s);
      }
      parity = 1 - parity;
    }
    await Promise.all(workers.map((w) => w.terminate()));
  }
  const seconds = Number(process.hrtime.bigint() - t0) / 1e9;

  const fFinal = parity === 0 ? f0 : new Float64Array(fb);
  const totals = [0, 0, 0, 0];
  for (let i = 0; i < Q; i++) {
    let s = 0;
    for (let c = 0; c < N; c++) s += fFinal[i * N + c];
    totals[0] += s; totals[1] += s * T.ex[i]; totals[2] += s * T.ey[i]; totals[3] += s * T.ez[i];
  }
  if (out.mac) writeArr(rel(out.mac), mac);
  if (out.f) writeArr(rel(out.f), fFinal);
  if (out.force) writeArr(rel(out.force), forces);
  if (out.cls) writeArr(rel(out.cls), cls);
  const cpus = os.cpus();
  console.log(JSON.stringify({
    steps: steps, cells: N, q: Q, threads: threads, seconds: seconds,
    mlups: steps > 0 ? N * steps / seconds / 1e6 : 0,
    cpu_model: cpus.length ? cpus[0].model : 'unknown', cpu_count: cpus.length,
    totals: totals,
  }));
}

// A dense 3D solve with one moving rigid body (rigid_oracle.js).
function rigidMain(job, rel) {
  const o = new R.RigidOracle(job.desc, job.rigid);
  const N = o.S.nx * o.S.ny * o.S.nz, Q = o.T.Q;
  if (job.macIn) o.initEq(readF32(rel(job.macIn)));
  const steps = job.steps | 0;
  const t0 = process.hrtime.bigint();
  o.step(steps);
  const seconds = Number(process.hrtime.bigint() - t0) / 1e9;
  const out = job.out || {};
  if (out.mac) writeArr(rel(out.mac), o.mac);
  if (out.f) writeArr(rel(out.f), o.f);
  if (out.cls) writeArr(rel(out.cls), o.cls);
  if (out.force) {
    const F = new Float64Array(6 * steps);
    o.forceHistory.forEach((r, s) => F.set(r, 6 * s));
    writeArr(rel(out.force), F);
  }
  if (out.body) {
    const B = new Float64Array(19 * steps);
    o.bodyHistory.forEach((b, s) => B.set([].concat(b.x, b.q, b.v, b.L, b.w, b.d), 19 * s));
    writeArr(rel(out.body), B);
  }
  const totals = [0, 0, 0, 0];
  for (let i = 0; i < Q; i++) {
    let sum = 0;
    for (let c = 0; c < N; c++) sum += o.f[i * N + c];
    totals[0] += sum; totals[1] += sum * o.T.ex[i]; totals[2] += sum * o.T.ey[i]; totals[3] += sum * o.T.ez[i];
  }
  console.log(JSON.stringify({ steps: steps, cells: N, q: Q, threads: 1, seconds: seconds, totals: totals,
                               mass: o.B.mass, inertia: o.B.I }));
}
