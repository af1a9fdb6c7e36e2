// JavaScript host for the MI355X engine: the test/driver surface of the
// reference's pages (window.__D3, main-3d.js:4944-5016; window.__CYL,
// main-tgv.js:286-296) as a Node class over the N-API addon
// (getting-air_amd/lib/lbm_napi.node -> liblbm.so).  Readbacks return
// Promises, as the reference's do (they await mapAsync there).
//
// Node 12 syntax only.  The addon is mandatory: constructing a solver without
// it throws, there is no CPU fallback.

import { createRequire } from 'module';
import { fileURLToPath } from 'url';
import { dirname, join } from 'path';
import { velocities, weights, SUPPORTED_Q } from './lattice-3d.mjs';

const here = dirname(fileURLToPath(import.meta.url));
let addon = null;

export function loadAddon() {
  if (!addon) {
    const require = createRequire(import.meta.url);
    const path = join(here, '..', 'lib', 'lbm_napi.node');
    try {
      addon = require(path);
    } catch (e) {
      throw new Error('lbm addon not built at ' + path + ' (run make in getting-air_amd): ' + e.message);
    }
  }
  return addon;
}

export const CS2 = 1 / 3;
export function nuFromTau(tau) { return CS2 * (tau - 0.5); }
export function tauFromNu(nu) { return nu / CS2 + 0.5; }
// TRT odd relaxation time for a magic parameter (plans/TRT.md:129-133).
export function tauMinus(tau, lambda) { return 0.5 + lambda / (tau - 0.5); }

// Seeded PRNG (mulberry32), the generator the reference seeds its
// perturbations with (d3-scenarios.mjs:207-216).
export function mulberry32(seed) {
  let a = seed | 0;
  return function next() {
    a = (a + 0x6d2b79f5) | 0;
    let t = a;
    t = Math.imul(t ^ (t >>> 15), t | 1);
    t ^= t + Math.imul(t ^ (t >>> 7), t | 61);
    return ((t ^ (t >>> 14)) >>> 0) / 4294967296;
  };
}

// [rho, ux, uy, uz] per cell with rho = 1 - |u|^2 / (2 cs^2)
// (d3-scenarios.mjs:245-261).
export function seedMacro3(dims, velAt) {
  const nx = dims[0], ny = dims[1], nz = dims[2];
  const out = new Float32Array(4 * nx * ny * nz);
  let c = 0;
  for (let z = 0; z < nz; z++) {
    for (let y = 0; y < ny; y++) {
      for (let x = 0; x < nx; x++, c++) {
        const u = velAt(x, y, z);
        out[4 * c] = 1 - (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) / (2 * CS2);
        out[4 * c + 1] = u[0]; out[4 * c + 2] = u[1]; out[4 * c + 3] = u[2];
      }
    }
  }
  return out;
}

// Flow past a pinned sphere: sponge on every face toward the freestream,
// tau from Re (SCENARIOS.sphere, d3-scenarios.mjs:350-391).  `dims`, `centre`
// and `R` may be given explicitly (the benchmark case is 256x128x128 with
// R = 12 at (64, 64, 64)).
export function sphereScenario(p) {
  const u0 = p.u0 || 0.05, re = p.re || 100;
  const R = p.R || 8, D = 2 * R;
  const dims = p.dims || [Math.round(12 * D), Math.round(8 * D), Math.round(8 * D)];
  const centre = p.centre || [3 * D, dims[1] / 2, dims[2] / 2];
  const nu = u0 * D / re;
  const tau = tauFromNu(nu);
  const desc = {
    q: p.q || 19, nx: dims[0], ny: dims[1], nz: dims[2], omega: 1 / tau,
    sponge: { width: p.spongeWidth || Math.max(6, Math.round(D / 2)), u: [u0, 0, 0] },
    body: { a: R, centre: centre }, solidEq: 1, storage: p.storage || 'f32',
  };
  if (p.trt) desc.omegaMinus = 1 / tauMinus(tau, p.lambda || 3 / 16);
  const params = { scenario: 'sphere', u0: u0, re: re, R: R, D: D, nu: nu, tau: tau, dims: dims, centre: centre,
    area: Math.PI * R * R };
  return { desc: desc, params: params, macro: () => seedMacro3(dims, () => [u0, 0, 0]) };
}

// 2D Taylor-Green vortex u = u0 (sin kx cos ky, -cos kx sin ky),
// rho = 1 + p/cs2 with p = (u0^2/4)(cos 2kx + cos 2ky) (main-tgv.js:69-90).
export function tgvScenario(p) {
  const n = p.n || 256, tau = p.tau || 0.6, u0 = p.u0 || 0.05;
  const k = 2 * Math.PI / n;
  const nu = nuFromTau(tau);
  const desc = { q: 9, nx: n, ny: n, nz: 1, omega: 1 / tau, storage: p.storage || 'f32' };
  const params = { scenario: 'tgv', N: n, W: n, H: n, TAU: tau, U0: u0, nu: nu, kx: k, ky: k, td: 1 / (nu * 2 * k * k) };
  const macro = () => {
    const out = new Float32Array(4 * n * n);
    for (let y = 0; y < n; y++) {
      for (let x = 0; x < n; x++) {
        const c = y * n + x;
        const pr = (u0 * u0 / 4) * (Math.cos(2 * k * x) + Math.cos(2 * k * y));
        out[4 * c] = 1 + pr / CS2;
        out[4 * c + 1] = u0 * Math.sin(k * x) * Math.cos(k * y);
        out[4 * c + 2] = -u0 * Math.cos(k * x) * Math.sin(k * y);
      }
    }
    return out;
  };
  return { desc: desc, params: params, macro: macro };
}

// One dense solve on the GPU with the reference pages' driver surface.
export class DenseSolver {
  constructor(scenario) {
    const a = loadAddon();
    this.addon = a;
    this.desc = scenario.desc;
    this.params = scenario.params;
    this.macroFn = scenario.macro;
    this.NX = this.desc.nx; this.NY = this.desc.ny; this.NZ = this.desc.q === 9 ? 1 : this.desc.nz;
    this.NCELLS = this.NX * this.NY * this.NZ;
    this.Q = this.desc.q;
    this.handle = a.createSolve(this.desc);
    this.step = 0;
    this.lastBody = null;
  }

  destroy() {
    if (this.handle) { this.addon.destroy(this.handle); this.handle = null; }
  }

  // writeBuffer(mac) + initEq (main-3d.js:2973-2994).
  reset() {
    this.addon.uploadMacro(this.handle, this.macroFn());
    this.step = 0;
    return Promise.resolve();
  }

  getStep() { return this.step; }

  getParams() {
    const p = Object.assign({}, this.params);
    p.Q = this.Q; p.NX = this.NX; p.NY = this.NY; p.NZ = this.NZ; p.NCELLS = this.NCELLS;
    p.hasBody = !!this.desc.body; p.bounceback = this.desc.body ? 1 : 0;
    p.storage = this.desc.storage || 'f32';
    p.tile = this.addon.info(this.handle);
    return p;
  }

  async debugStepSync(n) {
    await this.addon.step(this.handle, n | 0);
    this.step += n | 0;
    return { step: this.step, exhausted: false };
  }

  // main-3d.js:3485-3503: one body record per `every` steps.
  async debugRunAndCollect(nSteps, every) {
    const k = Math.max(1, Math.round(every) || 1);
    const history = [];
    let done = 0;
    while (done < nSteps) {
      const n = Math.min(k, nSteps - done);
      await this.debugStepSync(n);
      done += n;
      const b = await this.readBody();
      history.push([this.step, b ? b.fx : NaN, b ? b.fy : NaN, b && b.cd !== undefined ? b.cd : NaN,
        b && b.cl !== undefined ? b.cl : NaN, b && b.cs !== undefined ? b.cs : NaN, b ? b.fz : NaN]);
    }
    return { step: this.step, history: history, exhausted: false };
  }

  async readMacro() {
    const out = new Float32Array(4 * this.NCELLS);
    this.addon.readMacro(this.handle, out);
    return out;
  }

  // 2D velocity field (main-tgv.js:274-285).
  async readField() {
    const m = await this.readMacro();
    const ux = new Array(this.NCELLS), uy = new Array(this.NCELLS);
    for (let c = 0; c < this.NCELLS; c++) { ux[c] = m[4 * c + 1]; uy[c] = m[4 * c + 2]; }
    return { ux: ux, uy: uy };
  }

  async readF() {
    const out = new Float32Array(this.Q * this.NCELLS);
    this.addon.readF(this.handle, out);
    return out;
  }

  // Exact checkpoint (lbm_state_header + populations in the storage format):
  // restore() and stepping on gives the uninterrupted run's bits, which the
  // f32 wire format of readF / uploadF does not for fp32 storage.
  async saveState() { return this.addon.saveState(this.handle); }

  async loadState(state) {
    this.addon.loadState(this.handle, state);
    this.step = Number(new DataView(state.buffer || state, state.byteOffset || 0).getBigInt64(40, true));
  }

  // Kinetic energy, enstrophy (central-difference vorticity), extrema and the
  // extensive budget (main-3d.js:3610-3650).  mass/px/py/pz come from the
  // FP64 device reduction over the populations.
  async readStats() {
    const m = await this.readMacro();
    const NX = this.NX, NY = this.NY, NZ = this.NZ;
    const at = (x, y, z, c) => m[4 * ((((z + NZ) % NZ) * NY + ((y + NY) % NY)) * NX + ((x + NX) % NX)) + 1 + c];
    let ke = 0, ens = 0, maxSp = 0, rhoMin = Infinity, rhoMax = -Infinity, finite = true;
    for (let z = 0; z < NZ; z++) {
      for (let y = 0; y < NY; y++) {
        for (let x = 0; x < NX; x++) {
          const c = (z * NY + y) * NX + x;
          const rho = m[4 * c], ux = m[4 * c + 1], uy = m[4 * c + 2], uz = m[4 * c + 3];
          if (!Number.isFinite(rho + ux + uy + uz)) { finite = false; continue; }
          const sp2 = ux * ux + uy * uy + uz * uz;
          ke += 0.5 * sp2;
          if (sp2 > maxSp) maxSp = sp2;
          if (rho < rhoMin) rhoMin = rho;
          if (rho > rhoMax) rhoMax = rho;
          const wx = (at(x, y + 1, z, 2) - at(x, y - 1, z, 2)) * 0.5 - (at(x, y, z + 1, 1) - at(x, y, z - 1, 1)) * 0.5;
          const wy = (at(x, y, z + 1, 0) - at(x, y, z - 1, 0)) * 0.5 - (at(x + 1, y, z, 2) - at(x - 1, y, z, 2)) * 0.5;
          const wz = (at(x + 1, y, z, 1) - at(x - 1, y, z, 1)) * 0.5 - (at(x, y + 1, z, 0) - at(x, y - 1, z, 0)) * 0.5;
          ens += 0.5 * (wx * wx + wy * wy + wz * wz);
        }
      }
    }
    const t = this.addon.readTotals(this.handle);
    const n = this.NCELLS;
    const nu = this.params.nu !== undefined ? this.params.nu : nuFromTau(1 / this.desc.omega);
    return {
      step: this.step, ke: ke / n, enstrophy: ens / n, dissipation: 2 * nu * (ens / n),
      maxSpeed: Math.sqrt(maxSp), rhoMin: rhoMin, rhoMax: rhoMax, finite: finite,
      mass: t[0], px: t[1], py: t[2], pz: t[3],
    };
  }

  // A moving rigid body (writeBuffer(bodyBuf, packBodyState(...)),
  // main-3d.js:2983-2990): `body` is d3-body.mjs makeBodyState's object
  // ({shape, x, q, v, omega, mass | density}) with the physics overrides
  // {pinned, gravity, vMax, oMax, planar, noFluidForce, swept, rhoWLocal,
  // window, anchor}.  Attach after reset(): the body restarts from `body`.
  attachBody(body) {
    this.addon.rigidAttach(this.handle, body);
    this.rigid = true;
  }

  // Last step's force and torque, with Cd/Cl/Cs = F / (1/2 U^2 A)
  // (main-3d.js:4721-4738); with a rigid body also its BodyState3D fields.
  async readBody() {
    if (this.rigid) {
      const r = this.addon.rigidRead(this.handle);
      const b = { step: this.step, mass: r.mass,
                  fx: r.force[0], fy: r.force[1], fz: r.force[2], tx: r.torque[0], ty: r.torque[1], tz: r.torque[2] };
      ['c', 'v', 'L', 'w', 'd'].forEach((k) => {
        const src = { c: r.x, v: r.v, L: r.L, w: r.omega, d: r.d }[k];
        ['x', 'y', 'z'].forEach((a, i) => { b[k + a] = src[i]; });
      });
      b.qw = r.q[0]; b.qx = r.q[1]; b.qy = r.q[2]; b.qz = r.q[3];
      const u = this.params.u0, A = this.params.area;
      if (u && A) {
        const q = 0.5 * u * u * A;
        b.cd = b.fx / q; b.cl = b.fy / q; b.cs = b.fz / q;
      }
      return b;
    }
    if (!this.desc.body) return null;
    const f = this.addon.readForces(this.handle, 1);
    if (f.length < 6) return null;
    const b = { fx: f[0], fy: f[1], fz: f[2], tx: f[3], ty: f[4], tz: f[5], step: this.step };
    const u = this.params.u0, A = this.params.area;
    if (u && A) {
      const q = 0.5 * u * u * A;
      b.cd = b.fx / q; b.cl = b.fy / q; b.cs = b.fz / q;
    }
    return b;
  }

  // Every force record since reset, [fx, fy, fz, tx, ty, tz] per step.
  async readForceHistory(maxRecords) {
    return this.addon.readForces(this.handle, maxRecords || this.step);
  }

  async readClasses() {
    const out = new Uint8Array(this.NCELLS);
    this.addon.readClasses(this.handle, out);
    return out;
  }
}

// Two-level AMR solve (the reference's ?levels=2 explode/coalesce path,
// main-3d.js:3122-3335) over the addon.  `scenario` is
// { desc: {q, nx, ny, nz, rb, maxSlots, omega, omegaMinus, force,
//          manageEvery, refineAt, keepAt, lead, explodeLinear, orphans},
//   macro: () => Float32Array(4 * cells), topology: optional {blockSlot,
//   slotToBlock} as d3-amr.mjs refineWhere builds them }.
export class AmrSolver {
  constructor(scenario) {
    const a = loadAddon();
    this.addon = a;
    this.desc = scenario.desc;
    this.macroFn = scenario.macro;
    this.topology = scenario.topology || null;
    // refineHierarchy's result (d3-amr.mjs:603-651) for desc.levels >= 3:
    // byLevel[m] holds level m's blockSlot / slotToBlock
    this.hierarchy = scenario.hierarchy || null;
    this.NX = this.desc.nx; this.NY = this.desc.ny; this.NZ = this.desc.nz;
    this.NCELLS = this.NX * this.NY * this.NZ;
    this.Q = this.desc.q;
    this.RB = this.desc.rb;
    this.maxSlots = this.desc.maxSlots;
    this.nBlocks = (this.NX / this.RB) * (this.NY / this.RB) * (this.Q === 9 ? 1 : this.NZ / this.RB);
    this.handle = a.createAmr(this.desc);
    this.step = 0;
  }

  destroy() {
    if (this.handle) { this.addon.amrDestroy(this.handle); this.handle = null; }
  }

  // uploadLevelTopology + writeBuffer(mac) + initEq + interp seed
  // (main-3d.js:822, :2973-3035).
  reset() {
    if (this.topology) this.addon.amrUploadTopology(this.handle, this.topology.blockSlot, this.topology.slotToBlock);
    if (this.hierarchy) {
      for (let m = 1; m < this.hierarchy.levels; m++) {
        const lv = this.hierarchy.byLevel[m];
        // the library wants maxSlots entries of slotToBlock, -1 past the used ones
        const sb = new Int32Array(this.maxSlots).fill(-1);
        sb.set(lv.slotToBlock.subarray(0, Math.min(lv.slotToBlock.length, this.maxSlots)));
        this.addon.amrUploadLevelTopology(this.handle, m, Int32Array.from(lv.blockSlot), sb);
      }
    }
    this.addon.amrUploadMacro(this.handle, this.macroFn());
    this.step = 0;
    return Promise.resolve();
  }

  getStep() { return this.step; }

  async debugStepSync(n) {
    await this.addon.amrStep(this.handle, n | 0);
    this.step += n | 0;
    const info = this.addon.amrInfo(this.handle);
    return { step: this.step, exhausted: info.exhausted > 0 };
  }

  async manageNow() { this.addon.amrManageNow(this.handle); }

  async readMacro() {
    const out = new Float32Array(4 * this.NCELLS);
    this.addon.amrReadMacro(this.handle, out);
    return out;
  }

  // Block maps, criterion and active list (the reference's debugPoolState).
  async debugPoolState() {
    const blockSlot = new Int32Array(this.nBlocks), slotToBlock = new Int32Array(this.maxSlots);
    this.addon.amrReadTopology(this.handle, blockSlot, slotToBlock);
    const crit = new Float32Array(this.nBlocks), want = new Uint8Array(this.nBlocks);
    this.addon.amrReadCriterion(this.handle, crit, want);
    const act = new Int32Array(this.maxSlots + 1);
    this.addon.amrReadActive(this.handle, act);
    const info = this.addon.amrInfo(this.handle);
    return { blockSlot: blockSlot, slotToBlock: slotToBlock, crit: crit, want: want,
      active: act.slice(0, 1 + act[0]), activeTiles: info.activeTiles, exhausted: info.exhausted };
  }

  // FP64 composite mass and momentum (unrefined L0 cells + fine interiors / 8,
  // / 4 on the 2D quadtree).
  async readStats() {
    const t = this.addon.amrReadTotals(this.handle);
    return { step: this.step, mass: t[0], px: t[1], py: t[2], pz: t[3] };
  }

  // The fine-level body's force: the latest record [fx, fy, fz, tx, ty, tz]
  // of the macro-step's one force pass (readBody's force and torque).
  async readBody() {
    const f = this.addon.amrReadForces(this.handle, 1);
    if (f.length < 6) return null;
    return { fx: f[0], fy: f[1], fz: f[2], tx: f[3], ty: f[4], tz: f[5] };
  }
}

// Batched serve loop: many independent solves resident in one device pool,
// admitted against a byte budget and stepped together (lbm_serve_*).
// submit() returns a ticket at once; result() resolves when the solve is done.
export class ServeClient {
  constructor(opts) {
    const a = loadAddon();
    this.addon = a;
    this.handle = a.serveCreate(opts || {});
    this.cells = new Map();
  }

  destroy() {
    if (this.handle) { this.addon.serveDestroy(this.handle); this.handle = null; }
  }

  // scenario: {desc, macro()} as sphereScenario / tgvScenario build them.
  // opts.state: start from a saved state (DenseSolver.saveState or
  // result().state) instead of the scenario's field; opts.keepState: the
  // result carries the final state, so a later request can resume the solve.
  submit(scenario, steps, opts) {
    const d = scenario.desc;
    const o = opts || {};
    const t = (o.state || o.keepState)
      ? this.addon.serveSubmitState(this.handle, d, o.state || scenario.macro(), steps | 0, !!o.keepState)
      : this.addon.serveSubmit(this.handle, d, scenario.macro(), steps | 0);
    this.cells.set(t, d.nx * d.ny * (d.q === 9 ? 1 : d.nz));
    if (o.keepState) this.keep = (this.keep || new Set()).add(t);
    return t;
  }

  // -> {ticket, steps, cells, tSubmit, tAdmit, tFinish, force, macro, forces}
  async result(ticket, timeoutSeconds) {
    const r = await this.addon.serveWait(this.handle, ticket, timeoutSeconds === undefined ? -1 : timeoutSeconds);
    r.macro = this.addon.serveReadMacro(this.handle, ticket, new Float32Array(4 * this.cells.get(ticket)));
    r.forces = this.addon.serveReadForces(this.handle, ticket, r.steps);
    if (this.keep && this.keep.has(ticket)) r.state = this.addon.serveReadState(this.handle, ticket);
    return r;
  }

  release(ticket) { this.addon.serveRelease(this.handle, ticket); this.cells.delete(ticket); }

  stats() { return this.addon.serveStats(this.handle); }
}

export function deviceInfo() { return loadAddon().deviceInfo(); }

export { velocities, weights, SUPPORTED_Q };
