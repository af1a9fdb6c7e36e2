// The 2D quadtree pool's host model, a drop-in for the reference's amr2d.mjs
// subset that describes pools and work on this engine's D2Q9 quadtree
// (lbm_amr_desc q = 9, nz = 1): same export names, arguments and return
// shapes.
//
//   pool geometry       makePool (amr2d.mjs:220-247), poolAtLevel / nbAtLevel
//                       (:255-273), parentOfBlock / quadrantOfBlock (:275-276),
//                       quadrantOrigin (:294), the cell-centred fine <-> coarse
//                       map fineToCoarseUnit / coarseUnitToFine (:68-99)
//   static sets         refineWhere: ascending block ids take ascending slots
//                       (:369-389), which is the engine's allocator's order
//   the work unit       cellUpdatesPerMacroStep (:558-564): the root's cells
//                       plus each level's active tiles x (2 RB)^2 interior
//                       cells x 2^m substeps, the 2D AMR MLUPS numerator
//   slot bookkeeping    activeSlotList (:578-582), activeSlotsFromFreeCount
//                       (:570-572, the free count in quads), and the
//                       deterministic free-list handout grantAssignment /
//                       releaseAssignment (:1078-1105)
//   rescale factors     dcRescaleCoarseToFine / dcRescaleFineToCoarse
//                       (:136-143): 1/2 (tau_f - 1)/(tau_c - 1) and its inverse
//
// Node 12 syntax (no `??`), like the package's other modules.

export const GHOST = 2;
export const RB_DEFAULT = 8;

// Coarse-unit position of fine cell j of a tile whose first coarse cell is
// centred at `origin` (cell-centred: a fine centre sits a quarter coarse cell
// inside its parent), and back to the nearest fine index.
export function fineToCoarseUnit(j, origin, dx) {
  const h = dx === undefined ? 0.5 : dx;
  return origin - h / 2 + (j - GHOST) * h;
}
export function coarseUnitToFine(p, origin, dx) {
  const h = dx === undefined ? 0.5 : dx;
  return Math.round((p - origin + h / 2) / h) + GHOST;
}

function positiveMultiple(n, name, rb) {
  if (!Number.isInteger(n) || n <= 0) throw new Error(name + '=' + n + ' is not a positive integer');
  if (n % rb !== 0) throw new Error(name + '=' + n + ' is not a multiple of RB=' + rb);
}

export function makePool(opts) {
  const dims = opts.dims;
  const rb = opts.rb === undefined ? RB_DEFAULT : opts.rb;
  const ghost = opts.ghost === undefined ? GHOST : opts.ghost;
  positiveMultiple(dims[0], 'NX', rb);
  positiveMultiple(dims[1], 'NY', rb);
  const nbx = dims[0] / rb, nby = dims[1] / rb;
  const fb = 2 * rb + 2 * ghost;
  return {
    dims: [dims[0], dims[1]], rb: rb, GHOST: ghost, FB: fb,
    nb: [nbx, nby],
    nBlocks: nbx * nby,
    tileCells: fb * fb,
    maxSlots: opts.maxSlots === undefined ? nbx * nby : opts.maxSlots,
    blockId: (bx, by) => by * nbx + bx,
    blockOf: (id) => [id % nbx, Math.floor(id / nbx)],
    cellIndex: (slot, fx, fy) => (slot * fb + fy) * fb + fx,
    isInterior: (j) => j >= ghost && j < ghost + 2 * rb,
  };
}

// Level m's pool: level m tiles level m - 1's global grid, i.e. the root
// domain scaled by 2^(m-1), with the same tile shape.
export function poolAtLevel(pool, m, opts) {
  if (m < 1) throw new Error('level ' + m + ' is not a pool level (level 0 is the dense grid)');
  const s = Math.pow(2, m - 1);
  return makePool({
    dims: [pool.dims[0] * s, pool.dims[1] * s], rb: pool.rb, ghost: pool.GHOST,
    maxSlots: opts && opts.maxSlots !== undefined ? opts.maxSlots : pool.maxSlots,
  });
}
export function nbAtLevel(pool, m) {
  const s = Math.pow(2, m - 1);
  return [pool.nb[0] * s, pool.nb[1] * s];
}

export function parentOfBlock(b) { return [b[0] >> 1, b[1] >> 1]; }
export function quadrantOfBlock(b) { return [b[0] & 1, b[1] & 1]; }
// Where a child's interior starts in its parent's tile-local frame, per axis.
export function quadrantOrigin(rb, q, parentGhost) { return (parentGhost === undefined ? GHOST : parentGhost) + q * rb; }

// Blocks the predicate selects take slots in ascending block order; too many
// for the pool is an error, never a silent truncation.
export function refineWhere(pool, predicate) {
  const blockSlot = new Int32Array(pool.nBlocks).fill(-1);
  const slotToBlock = new Int32Array(pool.maxSlots).fill(-1);
  let used = 0;
  for (let id = 0; id < pool.nBlocks; id++) {
    const b = pool.blockOf(id);
    const lo = [b[0] * pool.rb, b[1] * pool.rb];
    const box = { bx: b[0], by: b[1], id: id, lo: lo, hi: [lo[0] + pool.rb, lo[1] + pool.rb],
      mid: [lo[0] + pool.rb / 2, lo[1] + pool.rb / 2] };
    if (!predicate(box)) continue;
    if (used >= pool.maxSlots) throw new Error('refinement needs more than maxSlots=' + pool.maxSlots + ' tiles');
    blockSlot[id] = used;
    slotToBlock[used] = id;
    used++;
  }
  return { blockSlot: blockSlot, slotToBlock: slotToBlock, activeSlots: used };
}

// Cell updates of one macro-step: every root cell once, and each level m's
// active tiles x their (2 RB)^2 interior cells x the level's 2^m substeps (the
// ring is bookkeeping, not lattice; a tile with children still steps).
export function cellUpdatesPerMacroStep(args) {
  let n = args.rootCells;
  const tile = 2 * args.rb * 2 * args.rb;
  Object.keys(args.activeByLevel).forEach((m) => {
    n += args.activeByLevel[m] * tile * Math.pow(2, Number(m));
  });
  return n;
}

// In-use slots of a level from its free count, which counts quads of four
// slots; a negative count read mid-round clamps to zero.
export function activeSlotsFromFreeCount(maxSlots, freeQuads) {
  return maxSlots - 4 * Math.max(0, freeQuads);
}

// The in-use slots of one level, ascending (what the active-list kernel
// writes after its count).
export function activeSlotList(slotToBlock) {
  const out = [];
  for (let s = 0; s < slotToBlock.length; s++) if (slotToBlock[s] >= 0) out.push(s);
  return out;
}

// Serves candidates from a free-list stack of freeCount entries: candidates
// in ascending id order, each popping the top (its free-list index counts
// down from freeCount - 1); once the stack is empty the rest are refused,
// so a starved pool starves its highest ids, and arrival order never matters.
export function grantAssignment(args) {
  const ids = Array.from(args.candidates).sort((a, b) => a - b);
  const granted = [], refused = [];
  let top = args.freeCount;
  ids.forEach((id) => {
    if (top > 0) {
      top -= 1;
      granted.push({ id: id, freeIndex: top });
    } else {
      refused.push(id);
    }
  });
  return { granted: granted, refused: refused, freeCount: top };
}

// The inverse: releases pushed in ascending id order, the j-th at
// freeCount + j.
export function releaseAssignment(args) {
  const ids = Array.from(args.releases).sort((a, b) => a - b);
  return { writes: ids.map((id, j) => ({ id: id, freeIndex: args.freeCount + j })),
    freeCount: args.freeCount + ids.length };
}

// The post-collision non-equilibrium rescale across a level boundary.
export function dcRescaleCoarseToFine(tauCoarse, tauFine) { return 0.5 * (tauFine - 1) / (tauCoarse - 1); }
export function dcRescaleFineToCoarse(tauCoarse, tauFine) { return 2 * (tauCoarse - 1) / (tauFine - 1); }
