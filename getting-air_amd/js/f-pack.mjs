// Packed 16-bit population storage, host side, with the module API of the
// reference's f-pack.mjs (f-pack.mjs:44-124): f32ToF16Bits, f16BitsToF32,
// fWords, packF, unpackF.
//
// The wire format everything outside the GPU speaks stays f32 plane-major
// f[i * ncells + cell] (f-pack.mjs:6-14).  packF/unpackF convert to and from
// the reference's D2Q9 word layout: planes (2p, 2p+1) share u32 word p, low
// half = even plane, word 4 carries plane 8 in both halves; mode 2 stores
// f_i - w_i with the exact-fraction weights the reference uses there.
//
// The MI355X engine stores every Q (9, 19, 27) as tight 16-bit planes of the
// shifted value g_i = f_i - w_i, where w_i are the lattice's f32 weights and
// the subtraction is done in f32 before rounding to nearest even.
// packPlanes16/unpackPlanes16 below are that format bit for bit, for FP16 and
// BF16, so a host can build or inspect device images without a GPU.
//
// All conversions round through f32: the input is first narrowed to f32 and
// then rounded to 16 bits, which is what the GPU conversion of an f32 value
// produces.  Node 12 syntax only.

import { WT as WT9 } from './lattice-2d.mjs';
import { weights } from './lattice-3d.mjs';

const scratch = new DataView(new ArrayBuffer(4));

function f32Bits(v) {
  scratch.setFloat32(0, v);
  return scratch.getUint32(0);
}

function f32FromBits(u) {
  scratch.setUint32(0, u >>> 0);
  return scratch.getFloat32(0);
}

// Rounds the integer m (with `drop` low bits to discard) to nearest, ties to
// even; returns m >> drop rounded.
function roundShift(m, drop) {
  if (drop <= 0) return m;
  if (drop > 31) return 0;
  const kept = Math.floor(m / Math.pow(2, drop));
  const rest = m - kept * Math.pow(2, drop);
  const half = Math.pow(2, drop - 1);
  if (rest > half || (rest === half && (kept % 2) === 1)) return kept + 1;
  return kept;
}

// IEEE binary32 -> binary16 bits, round to nearest even, subnormals kept.
export function f32ToF16Bits(v) {
  const u = f32Bits(v);
  const sign = (u >>> 31) << 15;
  const e = (u >>> 23) & 0xff;
  const frac = u & 0x7fffff;
  if (e === 0xff) return sign | 0x7c00 | (frac ? 0x200 : 0);
  if (e === 0 && frac === 0) return sign;
  // value = sig * 2^(e2), sig an integer (24 bits for normals)
  const sig = e === 0 ? frac : (frac | 0x800000);
  const e2 = (e === 0 ? 1 : e) - 127 - 23;
  // binary16: normal if value >= 2^-14; quantum 2^(E-10) for exponent E
  const lead = Math.floor(Math.log2(sig)) + e2;          // exponent of the leading bit
  if (lead >= -14) {
    let E = lead;
    let m = roundShift(sig, (E - 10) - e2);             // 11-bit significand
    if (m === 0x800) { m = 0x400; E += 1; }             // carry into the exponent
    if (E > 15) return sign | 0x7c00;
    return sign | ((E + 15) << 10) | (m & 0x3ff);
  }
  // subnormal: quantum 2^-24
  const m = roundShift(sig, -24 - e2);
  return sign | m;                                      // m === 0x400 is the smallest normal
}

export function f16BitsToF32(h) {
  const s = (h & 0x8000) ? -1 : 1;
  const E = (h >>> 10) & 0x1f;
  const m = h & 0x3ff;
  if (E === 0x1f) return m ? NaN : s * Infinity;
  if (E === 0) return s * m * Math.pow(2, -24);
  return s * (1024 + m) * Math.pow(2, E - 25);
}

// bfloat16 = the top half of binary32, rounded to nearest even.
export function f32ToBf16Bits(v) {
  const u = f32Bits(v);
  if ((u & 0x7fffffff) > 0x7f800000) return ((u >>> 16) | 0x40) & 0xffff;
  const upper = u >>> 16;
  const lower = u & 0xffff;
  const up = lower > 0x8000 || (lower === 0x8000 && (upper & 1) === 1);
  return (upper + (up ? 1 : 0)) & 0xffff;
}

export function bf16BitsToF32(h) {
  return f32FromBits((h & 0xffff) << 16);
}

// Words per cell of the reference's D2Q9 layout (common_fpack.wgsl fWords).
export function fWords(f16) { return f16 ? 5 : 9; }

const WT_EXACT_9 = [4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36];

// f32 plane-major (9 * ncells) -> the reference's packed D2Q9 words.
// f16 falsy: returns `f` itself.  1: raw halves.  2: halves of f_i - w_i.
export function packF(f, ncells, f16) {
  if (!f16) return f;
  const shift = f16 === 2;
  const out = new Uint32Array(5 * ncells);
  for (let p = 0; p < 5; p++) {
    const lo = 2 * p;
    const hi = lo + 1 < 9 ? lo + 1 : 8;
    const sl = shift ? WT_EXACT_9[lo] : 0;
    const sh = shift ? WT_EXACT_9[hi] : 0;
    for (let c = 0; c < ncells; c++) {
      const a = f32ToF16Bits(f[lo * ncells + c] - sl);
      const b = f32ToF16Bits(f[hi * ncells + c] - sh);
      out[p * ncells + c] = (a | (b << 16)) >>> 0;
    }
  }
  return out;
}

// Inverse of packF; `src` is what the device buffer held.
export function unpackF(src, ncells, f16) {
  if (!f16) return src;
  const shift = f16 === 2;
  const f = new Float32Array(9 * ncells);
  for (let i = 0; i < 9; i++) {
    const p = i >> 1;
    const add = shift ? WT_EXACT_9[i] : 0;
    for (let c = 0; c < ncells; c++) {
      const w = src[p * ncells + c];
      const bits = (i & 1) ? (w >>> 16) : (w & 0xffff);
      f[i * ncells + c] = f16BitsToF32(bits) + add;
    }
  }
  return f;
}

// The engine's f32 weights for a velocity set.
export function latticeWeights(Q) {
  return Q === 9 ? WT9.slice() : weights(Q);
}

// f32 plane-major f (Q * ncells) -> the engine's 16-bit planes of
// g = f - w (f32 subtraction, then round to nearest even).
// format: 'f16' | 'bf16'.
export function packPlanes16(f, ncells, Q, format) {
  const w = latticeWeights(Q);
  const enc = format === 'bf16' ? f32ToBf16Bits : f32ToF16Bits;
  const out = new Uint16Array(Q * ncells);
  for (let i = 0; i < Q; i++) {
    for (let c = 0; c < ncells; c++) out[i * ncells + c] = enc(Math.fround(f[i * ncells + c] - w[i]));
  }
  return out;
}

// Inverse: f = f32(decode(g) + w).
export function unpackPlanes16(planes, ncells, Q, format) {
  const w = latticeWeights(Q);
  const dec = format === 'bf16' ? bf16BitsToF32 : f16BitsToF32;
  const out = new Float32Array(Q * ncells);
  for (let i = 0; i < Q; i++) {
    for (let c = 0; c < ncells; c++) out[i * ncells + c] = Math.fround(dec(planes[i * ncells + c]) + w[i]);
  }
  return out;
}
