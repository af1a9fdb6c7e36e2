// D3Q19 / D3Q27 velocity sets for the MI355X engine, with the module API of
// the reference's lattice-3d.mjs (lattice-3d.mjs:40-228): the same export
// names, argument lists and return shapes, so a caller written against the
// reference imports this file unchanged.
//
// Ordering contract (lattice-3d.mjs:146-170): direction 0 is rest, then the
// face shell, the edge shell and (Q27 only) the corner shell.  Inside a shell
// the candidate offsets are visited with each axis running +1, -1, 0 (x
// outermost), and every new offset is emitted together with its negation.
// That makes opp[i] = i+1 for odd i and i-1 for even i > 0, and makes D3Q19 a
// prefix of D3Q27.
//
// Weights (lattice-3d.mjs:85-111): each shell weight is the correctly rounded
// f32 of its fraction moved by a small whole number of f32 ulps, so that the
// multiplicity-weighted sum is exactly 1 in f32 arithmetic:
//   D3Q19 ulps [+1, -4, +2], D3Q27 ulps [-1, +1, -1, 0].
// The HIP kernels take their tables from latticeHIP(Q) below, so the device
// and the host read the same f32 values.
//
// Node 12 syntax only (no `??`, no `?.`).

export const SUPPORTED_Q = [19, 27];
export const CS2 = 1 / 3;

const SHELL_FRACTION = { 19: [1 / 3, 1 / 18, 1 / 36], 27: [8 / 27, 2 / 27, 1 / 54, 1 / 216] };
const SHELL_COUNT = { 19: [1, 6, 12], 27: [1, 6, 12, 8] };
const SHELL_ULPS = { 19: [1, -4, 2], 27: [-1, 1, -1, 0] };

// Spacing of f32 values at x (x > 0, normal range).
function ulpOfF32(x) {
  return Math.pow(2, Math.floor(Math.log2(x)) - 23);
}

function shellWeights(Q) {
  const out = [];
  for (let s = 0; s < SHELL_FRACTION[Q].length; s++) {
    const r = Math.fround(SHELL_FRACTION[Q][s]);
    const w = r + SHELL_ULPS[Q][s] * ulpOfF32(r);
    if (Math.fround(w) !== w) throw new Error('D3Q' + Q + ' shell ' + s + ' weight is not an f32');
    out.push(w);
  }
  return out;
}

const SHELL_W = { 19: shellWeights(19), 27: shellWeights(27) };

function checkQ(Q) {
  if (SUPPORTED_Q.indexOf(Q) < 0) throw new Error('unsupported velocity set D3Q' + Q);
}

const VEL_CACHE = {};

export function velocities(Q) {
  checkQ(Q);
  if (!VEL_CACHE[Q]) {
    const steps = [1, -1, 0];
    const neg = (c) => (c === 0 ? 0 : -c);
    const list = [[0, 0, 0]];
    const taken = { '0,0,0': true };
    const lastShell = Q === 19 ? 2 : 3;
    for (let shell = 1; shell <= lastShell; shell++) {
      for (let a = 0; a < 3; a++) {
        for (let b = 0; b < 3; b++) {
          for (let c = 0; c < 3; c++) {
            const e = [steps[a], steps[b], steps[c]];
            if (e[0] * e[0] + e[1] * e[1] + e[2] * e[2] !== shell) continue;
            const k = e.join(',');
            if (taken[k]) continue;
            const m = [neg(e[0]), neg(e[1]), neg(e[2])];
            taken[k] = true;
            taken[m.join(',')] = true;
            list.push(e, m);
          }
        }
      }
    }
    if (list.length !== Q) throw new Error('D3Q' + Q + ': enumerated ' + list.length + ' velocities');
    VEL_CACHE[Q] = list;
  }
  return VEL_CACHE[Q].map((e) => e.slice());
}

export function weights(Q) {
  const w = SHELL_W[Q];
  return velocities(Q).map((e) => w[e[0] * e[0] + e[1] * e[1] + e[2] * e[2]]);
}

export function opposites(Q) {
  const n = velocities(Q).length;
  const o = [0];
  for (let i = 1; i < n; i++) o.push(i & 1 ? i + 1 : i - 1);
  return o;
}

export function shellRanges(Q) {
  const v = velocities(Q);
  const shellOf = v.map((e) => e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
  const span = (s) => {
    let lo = -1, hi = -1;
    for (let i = 0; i < shellOf.length; i++) {
      if (shellOf[i] !== s) continue;
      if (lo < 0) lo = i;
      hi = i + 1;
    }
    return lo < 0 ? [v.length, v.length] : [lo, hi];
  };
  return { rest: span(0), face: span(1), edge: span(2), corner: span(3) };
}

export function weightSum(Q) {
  checkQ(Q);
  let s = 0;
  for (let k = 0; k < SHELL_W[Q].length; k++) s += SHELL_COUNT[Q][k] * SHELL_W[Q][k];
  return s;
}

export function weightProvenance(Q) {
  checkQ(Q);
  return SHELL_W[Q].map((w, s) => ({
    exact: SHELL_FRACTION[Q][s],
    f32: w,
    ulps: SHELL_ULPS[Q][s],
    ulp: ulpOfF32(Math.fround(SHELL_FRACTION[Q][s])),
    multiplicity: SHELL_COUNT[Q][s],
  }));
}

// Second-order BGK equilibrium (lattice-3d.mjs:204-210), double precision.
export function feq(Q, rho, ux, uy, uz, i) {
  const e = velocities(Q)[i];
  const w = weights(Q)[i];
  const eu = e[0] * ux + e[1] * uy + e[2] * uz;
  const uu = ux * ux + uy * uy + uz * uz;
  return w * rho * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * uu);
}

function joinRows(items, perRow, indent) {
  const rows = [];
  for (let i = 0; i < items.length; i += perRow) rows.push(indent + items.slice(i, i + perRow).join(', '));
  return rows.join(',\n');
}

// WGSL form of the tables (declares QN, CS2, ex, ey, ez, wt, opp, feqD3Q),
// kept for callers that still assemble WGSL.  The 17-significant-digit
// literals round-trip the f32 weights exactly.
export function latticeWGSL(Q) {
  const v = velocities(Q), w = weights(Q), o = opposites(Q);
  const col = (k) => joinRows(v.map((e) => String(e[k]).padStart(2, ' ')), 10, '  ');
  return [
    '// D3Q' + Q + ' tables, emitted by lattice-3d.mjs latticeWGSL(' + Q + ').',
    'const QN : u32 = ' + Q + 'u;',
    'const CS2 = 1.0f/3.0f;',
    'const ex = array<i32,' + Q + '>(\n' + col(0) + '\n);',
    'const ey = array<i32,' + Q + '>(\n' + col(1) + '\n);',
    'const ez = array<i32,' + Q + '>(\n' + col(2) + '\n);',
    'const wt = array<f32,' + Q + '>(\n' + joinRows(w.map((x) => x.toPrecision(17) + 'f'), 5, '  ') + '\n);',
    'const opp = array<u32,' + Q + '>(\n' + joinRows(o.map((x) => String(x).padStart(2, ' ') + 'u'), 10, '  ') + '\n);',
    'fn feqD3Q(rho: f32, ux: f32, uy: f32, uz: f32, i: u32) -> f32 {',
    '  let eu = f32(ex[i]) * ux + f32(ey[i]) * uy + f32(ez[i]) * uz;',
    '  return wt[i] * rho * (1f + 3f*eu + 4.5f*eu*eu - 1.5f*(ux*ux + uy*uy + uz*uz));',
    '}',
    '',
  ].join('\n');
}

// HIP/C++ form of the same tables, consumed by the gfx950 kernels
// (getting-air_amd/csrc/lattice_tables.h is this function's output for
// Q = 19 and 27, guarded by a drift test).
export function latticeHIP(Q) {
  const v = velocities(Q), w = weights(Q), o = opposites(Q);
  const ints = (k) => v.map((e) => String(e[k])).join(', ');
  return [
    'template <> struct Lattice<' + Q + '> {',
    '  static constexpr int Q = ' + Q + ';',
    '  static constexpr int D = 3;',
    '  static constexpr int ex[' + Q + '] = {' + ints(0) + '};',
    '  static constexpr int ey[' + Q + '] = {' + ints(1) + '};',
    '  static constexpr int ez[' + Q + '] = {' + ints(2) + '};',
    '  static constexpr int opp[' + Q + '] = {' + o.join(', ') + '};',
    '  static constexpr float w[' + Q + '] = {' + w.map((x) => x.toPrecision(17) + 'f').join(', ') + '};',
    '};',
    '',
  ].join('\n');
}
