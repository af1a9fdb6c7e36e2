// Octree block-pool geometry for the 3D AMR path, with the module API of the
// reference's d3-amr.mjs (d3-amr.mjs:38-685): the same export names,
// arguments and return shapes (plain objects, Int32Array maps, Sets of
// "bx,by,bz" keys).  The HIP manager and pool kernels
// (getting-air_amd/csrc/amr_kernels.hip) implement the same addressing, and
// tests/js/test-d3-amr.mjs checks it by independent routes.
//
// Coordinates: COARSE (L0 cell), BLOCK (RB^3 coarse cells), TILE-LOCAL (fine
// cell in a slot, [0, FB)^3 with FB = 2 RB + 2 GHOST, interior
// [GHOST, GHOST + 2 RB)), GLOBAL FINE ([0, 2 N) per axis), COARSE-UNIT
// (continuous L0 position).  Fine children of coarse cell c sit at c -/+ 1/4.
//
// Slot assignment is deterministic: refineWhere gives slots in ascending
// block id (d3-amr.mjs:75-99).  manageSlots (an extension for the dynamic
// manager) keeps surviving blocks in their slots, frees unwanted ones, and
// hands the lowest free slots to newly wanted blocks in ascending block id --
// the rule the GPU manager reproduces bit for bit.
//
// Node 12 syntax only.

export const GHOST = 2;

export function fineToCoarseUnit(j, origin) {
  return origin - 0.25 + 0.5 * (j - GHOST);
}

export function coarseUnitToFine(p, origin) {
  return Math.round((p - origin + 0.25) * 2) + GHOST;
}

export function makePool(opts) {
  const dims = opts.dims;
  const rb = opts.rb === undefined ? 4 : opts.rb;
  const names = ['NX', 'NY', 'NZ'];
  for (let a = 0; a < 3; a++) {
    if (dims[a] % rb !== 0) throw new Error(names[a] + '=' + dims[a] + ' is not a multiple of RB=' + rb);
  }
  const nb = [dims[0] / rb, dims[1] / rb, dims[2] / rb];
  const FB = 2 * rb + 2 * GHOST;
  const nBlocks = nb[0] * nb[1] * nb[2];
  return {
    dims: dims.slice(), rb: rb, GHOST: GHOST, FB: FB, nb: nb, nBlocks: nBlocks,
    tileCells: FB * FB * FB,
    maxSlots: opts.maxSlots === undefined || opts.maxSlots === null ? nBlocks : opts.maxSlots,
    blockId: (bx, by, bz) => (bz * nb[1] + by) * nb[0] + bx,
    blockOf: (id) => [id % nb[0], Math.floor(id / nb[0]) % nb[1], Math.floor(id / (nb[0] * nb[1]))],
    cellIndex: (slot, fx, fy, fz) => ((slot * FB + fz) * FB + fy) * FB + fx,
    isInterior: (j) => j >= GHOST && j < GHOST + 2 * rb,
  };
}

export function refineWhere(pool, predicate) {
  const blockSlot = new Int32Array(pool.nBlocks).fill(-1);
  const slotToBlock = new Int32Array(pool.maxSlots).fill(-1);
  let used = 0;
  for (let id = 0; id < pool.nBlocks; id++) {
    const b = pool.blockOf(id);
    const lo = [b[0] * pool.rb, b[1] * pool.rb, b[2] * pool.rb];
    const hi = [lo[0] + pool.rb, lo[1] + pool.rb, lo[2] + pool.rb];
    const mid = [lo[0] + pool.rb / 2, lo[1] + pool.rb / 2, lo[2] + pool.rb / 2];
    if (!predicate({ bx: b[0], by: b[1], bz: b[2], id: id, lo: lo, hi: hi, mid: mid })) continue;
    if (used >= pool.maxSlots) throw new Error('refinement needs more than maxSlots=' + pool.maxSlots + ' tiles');
    blockSlot[id] = used;
    slotToBlock[used] = id;
    used++;
  }
  return { blockSlot: blockSlot, slotToBlock: slotToBlock, activeSlots: used };
}

export function refineNearBody(pool, sdf, margin) {
  return refineWhere(pool, nearBodyWant(sdf, margin));
}

// sqrt(3) rounded up, shared with the device criterion (d3-amr.mjs:114).
const SQRT3_UP = 1.7320509;
const BLOCK_BB_DEPTH = 3;

// Does any point of the block come within `margin` of the body?  The SDF is
// 1-Lipschitz, so phi(centre) - h sqrt(3) bounds the box minimum; ambiguous
// boxes split into octants down to depth 3 (d3-amr.mjs:114-159).
export function nearBodyWant(sdf, margin) {
  function near(c, h, depth) {
    const phi = sdf(c);
    if (phi - h * SQRT3_UP > margin) return false;
    if (phi <= margin) return true;
    if (depth === 0) return true;
    const q = h / 2;
    for (let k = 0; k < 8; k++) {
      const cc = [c[0] + ((k & 1) ? q : -q), c[1] + ((k & 2) ? q : -q), c[2] + ((k & 4) ? q : -q)];
      if (near(cc, q, depth - 1)) return true;
    }
    return false;
  }
  return (box) => near(
    [(box.lo[0] + box.hi[0]) / 2, (box.lo[1] + box.hi[1]) / 2, (box.lo[2] + box.hi[2]) / 2],
    (box.hi[0] - box.lo[0]) / 2, BLOCK_BB_DEPTH);
}

// Where does a tile-local source index land?  Inside the own interior, in a
// same-level neighbour tile's interior, or (null) nowhere refined, in which
// case the gather falls back to the own ring (d3-amr.mjs:173-210).
export function resolveSource(pool, blockSlot, blockXYZ, src) {
  const span = 2 * pool.rb;
  const local = [0, 0, 0];
  const step = [0, 0, 0];
  for (let a = 0; a < 3; a++) {
    if (src[a] < GHOST) { step[a] = -1; local[a] = src[a] + span; }
    else if (src[a] >= GHOST + span) { step[a] = 1; local[a] = src[a] - span; }
    else { local[a] = src[a]; }
  }
  if (!step[0] && !step[1] && !step[2]) return { slot: null, own: true, fx: local[0], fy: local[1], fz: local[2] };
  const n = [0, 1, 2].map((a) => (blockXYZ[a] + step[a] + pool.nb[a]) % pool.nb[a]);
  const slot = blockSlot[pool.blockId(n[0], n[1], n[2])];
  if (slot < 0) return null;
  return { slot: slot, own: false, fx: local[0], fy: local[1], fz: local[2] };
}

export function toGlobalFine(pool, blockXYZ, local) {
  return [0, 1, 2].map((a) => blockXYZ[a] * 2 * pool.rb + local[a] - GHOST);
}

export function fromGlobalFine(pool, g) {
  const span = 2 * pool.rb;
  const block = [0, 0, 0], local = [0, 0, 0];
  for (let a = 0; a < 3; a++) {
    const n = 2 * pool.dims[a];
    const w = ((g[a] % n) + n) % n;
    block[a] = Math.floor(w / span);
    local[a] = w - block[a] * span + GHOST;
  }
  return { block: block, local: local };
}

export function storageRatio(pool) {
  return Math.pow(pool.FB / pool.rb, 3);
}

const FACES = [[0, -1], [0, 1], [1, -1], [1, 1], [2, -1], [2, 1]];
const OFFSETS27 = [];
for (let dz = -1; dz <= 1; dz++) for (let dy = -1; dy <= 1; dy++) for (let dx = -1; dx <= 1; dx++) OFFSETS27.push([dx, dy, dz]);

function keyOf(x, y, z) { return x + ',' + y + ',' + z; }
function parseKey(k) { return k.split(',').map(Number); }
function wrapBlock(v, n) { return ((v % n) + n) % n; }

// 2:1 face balance over leaf blocks, periodic block grid
// (d3-amr.mjs:269-346).  levelSets[m] is a Set of "bx,by,bz" at level m.
export function check21Balance(levelSets, nbAt, opts) {
  const levels = opts.levels;
  const has = (m, x, y, z) => m >= 1 && m < levels && !!levelSets[m] && levelSets[m].has(keyOf(x, y, z));
  const hasChild = (m, x, y, z) => has(m + 1, 2 * x, 2 * y, 2 * z);
  const ancestorDepth = (m, x, y, z) => {
    for (let lv = m; lv >= 1; lv--) {
      if (has(lv, x, y, z)) return lv;
      x = Math.floor(x / 2); y = Math.floor(y / 2); z = Math.floor(z / 2);
    }
    return 0;
  };
  const faceMaxDepth = (m, x, y, z, a, s) => {
    if (!hasChild(m, x, y, z)) return m;
    let d = m;
    const want = s > 0 ? 1 : 0;
    for (let k = 0; k < 8; k++) {
      const q = [k & 1, (k >> 1) & 1, (k >> 2) & 1];
      if (q[a] !== want) continue;
      d = Math.max(d, faceMaxDepth(m + 1, 2 * x + q[0], 2 * y + q[1], 2 * z + q[2], a, s));
    }
    return d;
  };
  const violations = [];
  for (let m = 1; m < levels; m++) {
    if (!levelSets[m]) continue;
    const nb = nbAt(m);
    levelSets[m].forEach((key) => {
      const b = parseKey(key);
      if (hasChild(m, b[0], b[1], b[2])) return;
      FACES.forEach((f) => {
        const a = f[0], s = f[1];
        const n = b.slice();
        n[a] = wrapBlock(n[a] + s, nb[a]);
        const nDepth = has(m, n[0], n[1], n[2]) ? faceMaxDepth(m, n[0], n[1], n[2], a, -s) : ancestorDepth(m, n[0], n[1], n[2]);
        if (Math.abs(m - nDepth) > 1) violations.push({ level: m, block: b, neighbour: n, nDepth: nDepth, axis: a, dir: s });
      });
    });
  }
  return { violations: violations, levels: levels, counts: levelSets.map((s) => (s ? s.size : 0)) };
}

// Every block at level m >= 2 needs a parent at m-1 under each of its 27
// neighbours (the ring's explode source) (d3-amr.mjs:348-390).
export function checkRingParentCoverage(levelSets, nbAt, opts) {
  const violations = [];
  let required = 0;
  for (let m = 2; m < opts.levels; m++) {
    const nb = nbAt(m);
    const parent = levelSets[m - 1] || new Set();
    (levelSets[m] || new Set()).forEach((key) => {
      const b = parseKey(key);
      OFFSETS27.forEach((d) => {
        const n = [0, 1, 2].map((a) => wrapBlock(b[a] + d[a], nb[a]));
        const p = keyOf(n[0] >> 1, n[1] >> 1, n[2] >> 1);
        required++;
        if (!parent.has(p) && violations.length < 32) violations.push({ level: m, block: b, offset: d, missingParent: p });
      });
    });
  }
  return { violations: violations, required: required };
}

export function checkGeometryCoverage(pool, blockSlot, sdf, margin) {
  const sample = [];
  let required = 0, nViolations = 0;
  for (let z = 0; z < pool.dims[2]; z++) {
    for (let y = 0; y < pool.dims[1]; y++) {
      for (let x = 0; x < pool.dims[0]; x++) {
        if (sdf([x, y, z]) > margin) continue;
        required++;
        const id = pool.blockId(Math.floor(x / pool.rb), Math.floor(y / pool.rb), Math.floor(z / pool.rb));
        if (blockSlot[id] < 0) {
          nViolations++;
          if (sample.length < 32) sample.push({ cell: [x, y, z], block: id });
        }
      }
    }
  }
  return { violations: sample, nViolations: nViolations, required: required };
}

// Closes a want hierarchy under octet completion and parent coverage of the
// 27-neighbourhood, deepest level first (d3-amr.mjs:464-543).  At levels = 2
// it is the identity.
export function cascade21(wantSets, nbAt, opts) {
  const levels = opts.levels;
  const sets = [null];
  for (let m = 1; m < levels; m++) sets.push(new Set(wantSets[m] || []));
  const forced = [];
  const ensure = (m, b, because) => {
    const p = [b[0] >> 1, b[1] >> 1, b[2] >> 1];
    const members = m >= 2 ? 8 : 1;
    for (let k = 0; k < members; k++) {
      const c = m >= 2 ? [2 * p[0] + (k & 1), 2 * p[1] + ((k >> 1) & 1), 2 * p[2] + ((k >> 2) & 1)] : b;
      const key = keyOf(c[0], c[1], c[2]);
      if (sets[m].has(key)) continue;
      sets[m].add(key);
      forced.push({ level: m, block: c, octet: m >= 2 ? p : null, because: because });
    }
  };
  for (let m = levels - 1; m >= 2; m--) {
    Array.from(sets[m]).forEach((key) => ensure(m, parseKey(key), { siblingOf: key }));
    const nb = nbAt(m);
    sets[m].forEach((key) => {
      const b = parseKey(key);
      OFFSETS27.forEach((d) => {
        const n = [0, 1, 2].map((a) => wrapBlock(b[a] + d[a], nb[a]));
        const self = d[0] === 0 && d[1] === 0 && d[2] === 0;
        ensure(m - 1, [n[0] >> 1, n[1] >> 1, n[2] >> 1],
          self ? { parentOf: b.slice(), level: m } : { neighbourOf: b.slice(), level: m, offset: d.slice() });
      });
    });
  }
  return { sets: sets, forced: forced, counts: sets.map((s, m) => (m === 0 ? null : s.size)) };
}

export function poolAtLevel(pool, m, opts) {
  if (m < 1) throw new Error('level ' + m + ' has no pool (level 0 is the dense grid)');
  const s = Math.pow(2, m - 1);
  return makePool({ dims: pool.dims.map((d) => d * s), rb: pool.rb, maxSlots: opts ? opts.maxSlots : undefined });
}

export function parentOfBlock(b) { return [b[0] >> 1, b[1] >> 1, b[2] >> 1]; }
export function octantOfBlock(b) { return [b[0] & 1, b[1] & 1, b[2] & 1]; }
export function octantOrigin(rb, q) { return GHOST + q * rb; }

// Static refinement hierarchy: the want predicate at the finest level, closed
// by cascade21, then ascending-id slots per level (d3-amr.mjs:603-651).
export function refineHierarchy(pool, opts) {
  const levels = opts.levels;
  if (levels < 2) return { levels: 0, byLevel: [null], sets: [null] };
  const deepest = levels - 1;
  const view = poolAtLevel(pool, deepest);
  const scale = Math.pow(2, deepest - 1);
  const finest = new Set();
  for (let id = 0; id < view.nBlocks; id++) {
    const b = view.blockOf(id);
    const lo = [b[0] * view.rb, b[1] * view.rb, b[2] * view.rb];
    const hi = lo.map((c) => c + view.rb);
    const mid = lo.map((c) => c + view.rb / 2);
    const toL0 = (v) => v.map((c) => c / scale);
    if (opts.want({ level: deepest, bx: b[0], by: b[1], bz: b[2], id: id, lo: toL0(lo), hi: toL0(hi), mid: toL0(mid) })) {
      finest.add(keyOf(b[0], b[1], b[2]));
    }
  }
  const wants = [null];
  for (let m = 1; m < levels; m++) wants.push(m === deepest ? finest : new Set());
  const closed = cascade21(wants, (m) => poolAtLevel(pool, m).nb, { levels: levels });
  const byLevel = [null];
  for (let m = 1; m < levels; m++) {
    const lv = poolAtLevel(pool, m, { maxSlots: opts.maxSlots ? opts.maxSlots[m] : undefined });
    const alloc = refineWhere(lv, (bb) => closed.sets[m].has(keyOf(bb.bx, bb.by, bb.bz)));
    byLevel.push({ level: m, pool: lv, blockSlot: alloc.blockSlot, slotToBlock: alloc.slotToBlock, activeSlots: alloc.activeSlots });
  }
  return { levels: levels, byLevel: byLevel, sets: closed.sets, forced: closed.forced };
}

export function cellAtLevel(p, m, dims, rb) {
  const s = Math.pow(2, m);
  const g = p.map((c, a) => {
    const n = dims[a] * s;
    return ((Math.floor((c + 0.5) * s) % n) + n) % n;
  });
  if (m === 0) return { g: g, block: null, local: null };
  const span = 2 * rb;
  const block = g.map((c) => Math.floor(c / span));
  const local = g.map((c, a) => c - block[a] * span + GHOST);
  return { g: g, block: block, local: local };
}

export function finestLevelAt(p, opts) {
  for (let m = opts.levels - 1; m >= 1; m--) {
    const at = cellAtLevel(p, m, opts.dims, opts.rb);
    if (opts.present(m, at.block)) return m;
  }
  return 0;
}

// --- dynamic manager (extension) --------------------------------------------

// Applies one manager decision to a level's maps: blocks that hold a slot and
// are no longer wanted release it; then the newly wanted blocks, in ascending
// block id, take the free slots in ascending slot order.  Returns the new
// maps, which blocks were coarsened and refined, and `exhausted` (the number
// of wanted blocks that found no slot; the reference treats exhaustion as a
// hard failure, common_d3_manage.wgsl:500-525).
export function manageSlots(blockSlot, slotToBlock, want) {
  const bs = Int32Array.from(blockSlot);
  const sb = Int32Array.from(slotToBlock);
  const coarsened = [], refined = [];
  for (let id = 0; id < bs.length; id++) {
    if (bs[id] >= 0 && !want[id]) {
      sb[bs[id]] = -1;
      bs[id] = -1;
      coarsened.push(id);
    }
  }
  const free = [];
  for (let s = 0; s < sb.length; s++) if (sb[s] < 0) free.push(s);
  let next = 0, exhausted = 0;
  for (let id = 0; id < bs.length; id++) {
    if (!want[id] || bs[id] >= 0) continue;
    if (next >= free.length) { exhausted++; continue; }
    const s = free[next++];
    bs[id] = s;
    sb[s] = id;
    refined.push(id);
  }
  return { blockSlot: bs, slotToBlock: sb, coarsened: coarsened, refined: refined, exhausted: exhausted };
}

// Ascending compaction of the slots in use (amr_active_list.wgsl:41-70):
// out[0] = count, out[1 + k] = k-th used slot.
export function activeSlotList(slotToBlock) {
  const out = [0];
  for (let s = 0; s < slotToBlock.length; s++) if (slotToBlock[s] >= 0) out.push(s);
  out[0] = out.length - 1;
  return Int32Array.from(out);
}
