// D2Q9 velocity set for the MI355X engine, with the module API of the
// reference's lattice-2d.mjs (lattice-2d.mjs:92-218): CS2, E, EX, EY, WT,
// OPP, weightSum(), velocityDrift(), weightProvenance(), latticeWGSL().
//
// Ordering (lattice-2d.mjs:97-111): 0 rest; 1..4 the axis directions
// +x, +y, -x, -y; 5..8 the diagonals (+,+), (-,+), (-,-), (+,-).  Each shell
// runs counter-clockwise, so the opposite of direction i is the half turn
// inside its shell.  Weights are the correctly rounded f32 of 4/9, 1/9, 1/36
// with no ulp adjustment (lattice-2d.mjs:119-125): for D2Q9 the adjustment
// that would make the sum exactly 1 would instead break the velocity
// identity 2*w1 + 8*w2 - w0 = 0, which the plain rounding satisfies exactly.
//
// Node 12 syntax only.

export const CS2 = 1 / 3;

const ROT = [[1, 0], [0, 1], [-1, 0], [0, -1]];
export const E = [[0, 0]]
  .concat(ROT.map((r) => [r[0], r[1]]))
  .concat(ROT.map((r) => [r[0] - r[1], r[0] + r[1]]));
export const EX = E.map((e) => e[0]);
export const EY = E.map((e) => e[1]);

const FRACTION = [4 / 9, 1 / 9, 1 / 36];
const COUNT = [1, 4, 4];
const SHELL_W = FRACTION.map((x) => Math.fround(x));

export const WT = E.map((e) => SHELL_W[e[0] * e[0] + e[1] * e[1]]);

export const OPP = E.map((e) => {
  for (let j = 0; j < E.length; j++) if (E[j][0] === -e[0] && E[j][1] === -e[1]) return j;
  return -1;
});

export function weightSum() {
  return SHELL_W[0] * COUNT[0] + SHELL_W[1] * COUNT[1] + SHELL_W[2] * COUNT[2];
}

// Relative velocity drift per BGK collision in units of omega:
// 2*w1 + 8*w2 - w0, evaluated exactly in f64 (lattice-2d.mjs:156-162).
export function velocityDrift() {
  return 2 * SHELL_W[1] + 8 * SHELL_W[2] - SHELL_W[0];
}

function ulpOfF32(x) {
  return Math.pow(2, Math.floor(Math.log2(x)) - 23);
}

export function weightProvenance() {
  return SHELL_W.map((w, s) => ({
    exact: FRACTION[s], f32: w, ulps: 0, ulp: ulpOfF32(w), multiplicity: COUNT[s],
  }));
}

export function latticeWGSL() {
  const lit = (x) => x.toPrecision(17) + 'f';
  return [
    '// D2Q9 tables, emitted by lattice-2d.mjs latticeWGSL().',
    'const ex = array<i32,9>(' + EX.map((v) => String(v).padStart(2, ' ')).join(',') + ');',
    'const ey = array<i32,9>(' + EY.map((v) => String(v).padStart(2, ' ')).join(',') + ');',
    'const wt = array<f32,9>(\n' + WT.map((w) => '  ' + lit(w)).join(',\n') + '\n);',
    'const opp = array<u32,9>(' + OPP.map((v) => v + 'u').join(', ') + ');',
    'const CS2 = ' + lit(Math.fround(CS2)) + ';',
    'fn feqD2Q9(rho: f32, ux: f32, uy: f32, i: u32) -> f32 {',
    '  let eu = f32(ex[i]) * ux + f32(ey[i]) * uy;',
    '  return wt[i] * rho * (1f + 3f*eu + 4.5f*eu*eu - 1.5f*(ux*ux+uy*uy));',
    '}',
    '',
  ].join('\n');
}

// HIP/C++ form consumed by the gfx950 kernels (D2Q9 runs as a Q = 9 lattice
// with ez = 0 and a single z plane).
export function latticeHIP() {
  const ints = (a) => a.join(', ');
  return [
    'template <> struct Lattice<9> {',
    '  static constexpr int Q = 9;',
    '  static constexpr int D = 2;',
    '  static constexpr int ex[9] = {' + ints(EX) + '};',
    '  static constexpr int ey[9] = {' + ints(EY) + '};',
    '  static constexpr int ez[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};',
    '  static constexpr int opp[9] = {' + ints(OPP) + '};',
    '  static constexpr float w[9] = {' + WT.map((x) => x.toPrecision(17) + 'f').join(', ') + '};',
    '};',
    '',
  ].join('\n');
}
