// Tests of the drop-in amr2d.mjs (getting-air_amd/js/amr2d.mjs), restating the
// properties and known answers of the reference's own suite for these
// functions (tools/test-amr2d.js):
//   - grantAssignment / releaseAssignment: invariance under arrival order,
//     pops from the top of the stack, a starved pool starves the highest ids,
//     an empty pool grants nothing, distinct indices, release pushes in id
//     order, a round trip restores the count, refusals consume nothing
//     (:1443-1545)
//   - cellUpdatesPerMacroStep's worked example 156672 and its three mutants
//     (:2058-2075); activeSlotsFromFreeCount (:2077-2082); activeSlotList
//     (:2084-2090)
// plus the pool geometry: makePool's refusals and block ids, refineWhere's
// ascending slots and loud exhaustion, poolAtLevel / nbAtLevel, the quadtree
// parent/quadrant split, the fine <-> coarse-unit map round trip, and the
// Dupuis-Chopard pair being exact inverses.
// Run: node tests/js/test-amr2d.mjs   (exit code 1 on any failure)
import {
  GHOST, RB_DEFAULT, fineToCoarseUnit, coarseUnitToFine, makePool, poolAtLevel, nbAtLevel, parentOfBlock,
  quadrantOfBlock, quadrantOrigin, refineWhere, cellUpdatesPerMacroStep, activeSlotsFromFreeCount, activeSlotList,
  grantAssignment, releaseAssignment, dcRescaleCoarseToFine, dcRescaleFineToCoarse,
} from '../../getting-air_amd/js/amr2d.mjs';

let failures = 0;
function ok(name, fn) {
  try {
    fn();
    console.log('ok   ' + name);
  } catch (e) {
    failures++;
    console.log('FAIL ' + name + ': ' + e.message);
  }
}
function assert(c, msg) { if (!c) throw new Error(msg || 'assertion failed'); }
function same(a, b, msg) {
  const x = JSON.stringify(a), y = JSON.stringify(b);
  if (x !== y) throw new Error((msg || 'differ') + ': ' + x + ' vs ' + y);
}
function throws(fn, re) {
  let e = null;
  try { fn(); } catch (x) { e = x; }
  assert(e && re.test(e.message), 'expected an error matching ' + re + ', got ' + e);
}

ok('grant is invariant under the order candidates arrive in', () => {
  const ref = JSON.stringify(grantAssignment({ candidates: [41, 7, 19, 3, 28], freeCount: 6 }));
  [[3, 7, 19, 28, 41], [41, 28, 19, 7, 3], [19, 3, 41, 28, 7], [7, 41, 3, 19, 28]].forEach((p) => {
    assert(JSON.stringify(grantAssignment({ candidates: p, freeCount: 6 })) === ref, 'order ' + p);
  });
});

ok('grant pops down from the top of the stack, one index each', () => {
  const r = grantAssignment({ candidates: [5, 2, 9], freeCount: 4 });
  same(r.granted.map((g) => g.id), [2, 5, 9]);
  same(r.granted.map((g) => g.freeIndex), [3, 2, 1]);
  same(r.refused, []);
  assert(r.freeCount === 1);
});

ok('a starved pool starves the highest ids and never over-grants', () => {
  const r = grantAssignment({ candidates: [8, 1, 5, 3, 9], freeCount: 2 });
  same(r.granted.map((g) => g.id), [1, 3]);
  same(r.refused, [5, 8, 9]);
  assert(r.freeCount === 0);
  r.granted.forEach((g) => assert(g.freeIndex >= 0 && g.freeIndex < 2));
});

ok('an empty pool grants nothing and is not an error', () => {
  const r = grantAssignment({ candidates: [4, 1], freeCount: 0 });
  same(r.granted, []);
  same(r.refused, [1, 4]);
  assert(r.freeCount === 0);
});

ok('every granted free-list index is distinct', () => {
  const ids = [31, 4, 17, 22, 9, 40, 12];
  const r = grantAssignment({ candidates: ids, freeCount: ids.length });
  assert(new Set(r.granted.map((g) => g.freeIndex)).size === ids.length);
});

ok('release pushes in id order from the top, invariant under arrival order', () => {
  const r = releaseAssignment({ releases: [12, 3, 7], freeCount: 5 });
  same(r.writes.map((w) => w.id), [3, 7, 12]);
  same(r.writes.map((w) => w.freeIndex), [5, 6, 7]);
  assert(r.freeCount === 8);
  [[3, 7, 12], [12, 7, 3], [7, 12, 3]].forEach((p) => same(releaseAssignment({ releases: p, freeCount: 5 }), r));
});

ok('grant then release restores the count and touches the same indices', () => {
  const g = grantAssignment({ candidates: [6, 2, 11], freeCount: 5 });
  const r = releaseAssignment({ releases: g.granted.map((x) => x.id), freeCount: g.freeCount });
  assert(r.freeCount === 5);
  same(g.granted.map((x) => x.freeIndex).sort((a, b) => a - b), r.writes.map((x) => x.freeIndex).sort((a, b) => a - b));
});

ok('a refused candidate consumes nothing', () => {
  const a = grantAssignment({ candidates: [1, 2], freeCount: 2 });
  const b = grantAssignment({ candidates: [1, 2, 99, 98, 97], freeCount: 2 });
  assert(a.freeCount === b.freeCount);
  same(a.granted, b.granted.slice(0, 2));
});

ok('cellUpdatesPerMacroStep: the worked example, and each term is live', () => {
  const rb = 8, activeByLevel = { 1: 16, 2: 48, 3: 48 };
  const got = cellUpdatesPerMacroStep({ rootCells: 32 * 32, rb: rb, activeByLevel: activeByLevel });
  assert(got === 156672, 'got ' + got);
  const ring = 1024 + [16, 48, 48].reduce((s, a, i) => s + a * 20 * 20 * Math.pow(2, i + 1), 0);
  const noSub = 1024 + (16 + 48 + 48) * 256;
  [ring, noSub, got - 1024].forEach((v) => assert(got !== v, 'a mutant gives the same count'));
});

ok('cellUpdatesPerMacroStep is the quadtree bench line\'s numerator', () => {
  // bench.py amr_d2q9_shear_layer_1024: 1024^2 root, RB 8, 1536 tiles
  assert(cellUpdatesPerMacroStep({ rootCells: 1024 * 1024, rb: 8, activeByLevel: { 1: 1536 } }) === 1835008);
});

ok('activeSlotsFromFreeCount reads the free count in quads and clamps', () => {
  assert(activeSlotsFromFreeCount(1164, 1164 / 4) === 0);
  assert(activeSlotsFromFreeCount(1164, 279) === 48);
  assert(activeSlotsFromFreeCount(64, 0) === 64);
  assert(activeSlotsFromFreeCount(64, -1) === 64);
});

ok('activeSlotList is the ascending in-use slots (block 0 included)', () => {
  same(activeSlotList(Int32Array.from([-1, 7, -1, -1, 3, 0, -1, 12, 9, -1, -1, 5])), [1, 4, 5, 7, 8, 11]);
  same(activeSlotList(new Int32Array(8).fill(-1)), []);
});

ok('makePool: geometry, block id round trip, refusals', () => {
  const p = makePool({ dims: [64, 32], rb: 8 });
  assert(p.FB === 20 && p.tileCells === 400 && p.nBlocks === 32 && p.maxSlots === 32 && RB_DEFAULT === 8);
  same(p.nb, [8, 4]);
  for (let id = 0; id < p.nBlocks; id++) {
    const b = p.blockOf(id);
    assert(p.blockId(b[0], b[1]) === id);
  }
  assert(p.cellIndex(3, 5, 7) === 3 * 400 + 7 * 20 + 5);
  assert(p.isInterior(GHOST) && !p.isInterior(GHOST - 1) && !p.isInterior(GHOST + 16) && p.isInterior(GHOST + 15));
  throws(() => makePool({ dims: [60, 32], rb: 8 }), /NX=60 is not a multiple/);
  throws(() => makePool({ dims: [64, 0], rb: 8 }), /NY=0 is not a positive/);
});

ok('refineWhere: ascending ids take ascending slots; exhaustion is loud', () => {
  const p = makePool({ dims: [32, 32], rb: 8, maxSlots: 5 });
  const r = refineWhere(p, (b) => b.bx === b.by);
  same(Array.from(r.slotToBlock), [0, 5, 10, 15, -1]);
  assert(r.activeSlots === 4 && r.blockSlot[0] === 0 && r.blockSlot[5] === 1 && r.blockSlot[15] === 3);
  throws(() => refineWhere(makePool({ dims: [32, 32], rb: 8, maxSlots: 3 }), (b) => b.bx === b.by),
    /more than maxSlots=3/);
});

ok('poolAtLevel scales the domain, nbAtLevel its block counts', () => {
  const p = makePool({ dims: [64, 32], rb: 8, maxSlots: 10 });
  const l3 = poolAtLevel(p, 3);
  same(l3.dims, [256, 128]);
  same(l3.nb, nbAtLevel(p, 3));
  assert(l3.FB === p.FB && l3.maxSlots === 10 && poolAtLevel(p, 2, { maxSlots: 4 }).maxSlots === 4);
  throws(() => poolAtLevel(p, 0), /not a pool level/);
});

ok('a block is twice its parent plus its quadrant; the quadrant places it', () => {
  for (let x = 0; x < 9; x++) {
    for (let y = 0; y < 9; y++) {
      const pa = parentOfBlock([x, y]), q = quadrantOfBlock([x, y]);
      assert(2 * pa[0] + q[0] === x && 2 * pa[1] + q[1] === y);
    }
  }
  assert(quadrantOrigin(8, 0) === GHOST && quadrantOrigin(8, 1) === GHOST + 8 && quadrantOrigin(8, 1, 0) === 8);
});

ok('the fine <-> coarse-unit map is cell-centred and round-trips', () => {
  // fine cell GHOST of a block at coarse origin 4 sits a quarter cell inside
  assert(fineToCoarseUnit(GHOST, 4) === 3.75 && fineToCoarseUnit(GHOST + 1, 4) === 4.25);
  for (let j = 0; j < 20; j++) assert(coarseUnitToFine(fineToCoarseUnit(j, 12), 12) === j);
  for (let j = 0; j < 20; j++) assert(coarseUnitToFine(fineToCoarseUnit(j, 12, 0.25), 12, 0.25) === j);
});

ok('the Dupuis-Chopard rescale pair are exact inverses', () => {
  [[0.6, 0.7], [0.8, 1.1], [1.4, 2.3]].forEach((t) => {
    assert(Math.abs(dcRescaleCoarseToFine(t[0], t[1]) * dcRescaleFineToCoarse(t[0], t[1]) - 1) < 1e-15);
  });
});

if (failures) {
  console.log(failures + ' failure(s)');
  process.exit(1);
}
