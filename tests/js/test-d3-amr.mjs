// Tests of the drop-in d3-amr.mjs (getting-air_amd/js/d3-amr.mjs), restating
// the properties the reference's own suite checks (tools/test-d3-amr.js):
//   - pool geometry, storage ratio, block id round trip (:42-76)
//   - the cell-centred fine <-> coarse-unit map (:78-100)
//   - resolveSource against the GLOBAL-FINE route, which shares no arithmetic
//     with it, for every sampled cell and all 27 offsets (:110-161), and at
//     pool levels 2 and 3 (:267-305); neighbour sources land in the interior
//     (:163-186)
//   - refineWhere determinism and map consistency, loud slot exhaustion
//     (:188-213); refineNearBody / checkGeometryCoverage (:215-530)
//   - poolAtLevel and the octant placement of a child in its parent (:238-357)
//   - check21Balance catching planted violations (:359-440)
//   - cascade21: identity at levels = 2, closing a gap, materialising
//     parents, idempotence, nothing added to a uniform tree, the nested-shell
//     counts the reference records (432 / 1280 / 2432), octet completeness,
//     every forced octet necessary, ring-parent coverage (:534-748, :750-781)
//   - refineHierarchy: bit-identical to refineWhere at levels = 2, balanced
//     and parented at depth, finest level covers the body (:781-880)
//   - cellAtLevel / finestLevelAt and their mutants (:880-975)
// plus manageSlots / activeSlotList, the dynamic manager's host twin.
// Run: node tests/js/test-d3-amr.mjs   (exit code 1 on any failure)
import {
  GHOST, fineToCoarseUnit, coarseUnitToFine, makePool, refineWhere, refineNearBody, nearBodyWant,
  resolveSource, toGlobalFine, fromGlobalFine, storageRatio, check21Balance, checkRingParentCoverage,
  checkGeometryCoverage, cascade21, poolAtLevel, parentOfBlock, octantOfBlock, octantOrigin,
  refineHierarchy, cellAtLevel, finestLevelAt, manageSlots, activeSlotList,
} from '../../getting-air_amd/js/d3-amr.mjs';

let failures = 0;
function ok(name, fn) {
  try {
    const m = fn();
    console.log('ok   ' + name + (m ? '  (' + m + ')' : ''));
  } catch (e) {
    failures++;
    console.log('FAIL ' + name + ': ' + (e && e.stack ? e.stack : e));
  }
}
function assert(c, m) { if (!c) throw new Error(m); }
function eq(a, b, what) {
  const sa = JSON.stringify(a), sb = JSON.stringify(b);
  if (sa !== sb) throw new Error(what + ': got ' + sa + ', want ' + sb);
}
function throwsLike(fn, re, what) {
  let msg = null;
  try { fn(); } catch (e) { msg = e.message; }
  assert(msg !== null && re.test(msg), what + ': expected a throw matching ' + re + ', got ' + msg);
}
function lcg(seed) {
  let s = seed >>> 0;
  return () => { s = (Math.imul(s, 1103515245) + 12345) & 0x7fffffff; return s / 0x7fffffff; };
}
const sphereSdf = (c, R) => (q) => Math.hypot(q[0] - c[0], q[1] - c[1], q[2] - c[2]) - R;
const keyOf = (b) => b[0] + ',' + b[1] + ',' + b[2];
const parse = (k) => k.split(',').map(Number);
const sortedKeys = (sets) => sets.map((s) => (s ? Array.from(s).sort().join('|') : null));
const OFF27 = [];
for (let dz = -1; dz <= 1; dz++) for (let dy = -1; dy <= 1; dy++) for (let dx = -1; dx <= 1; dx++) OFF27.push([dx, dy, dz]);

// --- geometry ---------------------------------------------------------------

ok('makePool: FB = 2 RB + 2 GHOST, block count, tile cells; a non-dividing RB throws', () => {
  for (const rb of [2, 4, 8]) {
    const p = makePool({ dims: [32, 32, 32], rb: rb });
    assert(p.FB === 2 * rb + 4 && p.nBlocks === Math.pow(32 / rb, 3) && p.tileCells === p.FB * p.FB * p.FB, 'rb ' + rb);
    assert(p.maxSlots === p.nBlocks, 'default maxSlots');
  }
  throwsLike(() => makePool({ dims: [32, 30, 32], rb: 4 }), /not a multiple of RB/, 'NY = 30');
});

ok('storageRatio: (FB / RB)^3 = 27 at RB 4 and 15.625 at RB 8', () => {
  assert(storageRatio(makePool({ dims: [32, 32, 32], rb: 4 })) === 27, 'RB 4');
  assert(storageRatio(makePool({ dims: [32, 32, 32], rb: 8 })) === 15.625, 'RB 8');
});

ok('blockId and blockOf are inverse over a non-cubic grid', () => {
  const p = makePool({ dims: [24, 16, 32], rb: 4 });
  const seen = new Set();
  for (let id = 0; id < p.nBlocks; id++) {
    const b = p.blockOf(id);
    assert(b.every((c, a) => c >= 0 && c < p.nb[a]), 'block ' + b + ' outside the grid');
    assert(p.blockId(b[0], b[1], b[2]) === id, 'round trip at ' + id);
    seen.add(keyOf(b));
  }
  assert(seen.size === p.nBlocks, 'blockOf is not injective');
});

ok('fine <-> coarse-unit map: children at c -/+ 1/4, the ring continues the line, round trip', () => {
  const o = 7;
  const want = [[0, o - 1.25], [1, o - 0.75], [GHOST, o - 0.25], [GHOST + 1, o + 0.25], [GHOST + 2, o + 0.75]];
  for (const [j, u] of want) assert(Math.abs(fineToCoarseUnit(j, o) - u) < 1e-12, 'j ' + j + ' -> ' + fineToCoarseUnit(j, o));
  for (let j = 0; j < 14; j++) assert(coarseUnitToFine(fineToCoarseUnit(j, o), o) === j, 'round trip j ' + j);
  for (let c = 0; c < 4; c++) {
    const mean = (fineToCoarseUnit(GHOST + 2 * c, o) + fineToCoarseUnit(GHOST + 2 * c + 1, o)) / 2;
    assert(Math.abs(mean - (o + c)) < 1e-12, 'children of ' + (o + c) + ' centred at ' + mean);
  }
});

ok('toGlobalFine / fromGlobalFine round trip from interior cells of several blocks', () => {
  const p = makePool({ dims: [32, 24, 16], rb: 4 });
  let n = 0;
  for (const b of [[0, 0, 0], [7, 5, 3], [2, 4, 1], [6, 0, 2]]) {
    for (let k = 0; k < 20; k++) {
      const l = [GHOST + (k * 3) % 8, GHOST + (k * 5) % 8, GHOST + (k * 7) % 8];
      const back = fromGlobalFine(p, toGlobalFine(p, b, l));
      eq(back.block, b, 'block of ' + l);
      eq(back.local, l, 'local of ' + l);
      n++;
    }
  }
  return n + ' cells';
});

// resolveSource against the global route: where does the source cell sit in
// global fine coordinates, and which block (one division, periodic wrap)
// owns it?  Returns the counts of the three outcomes.
function resolveAgainstGlobal(p, blockSlot, cellStep, offsets) {
  const span = 2 * p.rb;
  let checked = 0, own = 0, neigh = 0, none = 0;
  for (let id = 0; id < p.nBlocks; id++) {
    if (blockSlot[id] < 0) continue;
    const b = p.blockOf(id);
    for (let fz = GHOST; fz < GHOST + span; fz += cellStep) {
      for (let fy = GHOST; fy < GHOST + span; fy += cellStep) {
        for (let fx = GHOST; fx < GHOST + span; fx += cellStep) {
          for (const d of offsets) {
            const src = [fx - d[0], fy - d[1], fz - d[2]];
            const got = resolveSource(p, blockSlot, b, src);
            const want = fromGlobalFine(p, toGlobalFine(p, b, src));
            const wantSlot = blockSlot[p.blockId(want.block[0], want.block[1], want.block[2])];
            checked++;
            if (want.block[0] === b[0] && want.block[1] === b[1] && want.block[2] === b[2]) {
              assert(got && got.own, 'own tile expected for ' + src + ' in ' + b);
              eq([got.fx, got.fy, got.fz], want.local, 'own local ' + src);
              own++;
            } else if (wantSlot < 0) {
              assert(got === null, 'null expected for ' + src + ' in ' + b + ', got ' + JSON.stringify(got));
              none++;
            } else {
              assert(got && !got.own && got.slot === wantSlot, 'neighbour slot ' + wantSlot + ' expected for ' + src);
              eq([got.fx, got.fy, got.fz], want.local, 'neighbour local ' + src);
              neigh++;
            }
          }
        }
      }
    }
  }
  return { checked: checked, own: own, neigh: neigh, none: none };
}

ok('resolveSource agrees with the global-fine route for sampled cells and all 27 offsets (ragged refinement)', () => {
  const p = makePool({ dims: [16, 16, 16], rb: 4 });
  const { blockSlot } = refineWhere(p, (b) => (b.bx + b.by + b.bz) % 3 !== 2);
  const r = resolveAgainstGlobal(p, blockSlot, 3, OFF27);
  assert(r.checked > 5000 && r.neigh > 100 && r.none > 100, JSON.stringify(r));
  return JSON.stringify(r);
});

ok('resolveSource agrees with the global route at pool levels 2 and 3, every interior cell', () => {
  const base = makePool({ dims: [8, 8, 8], rb: 2 });
  const out = [];
  for (const m of [2, 3]) {
    const p = poolAtLevel(base, m);
    const { blockSlot } = refineWhere(p, (b) => (b.bx + b.by + b.bz) % 3 !== 2);
    const r = resolveAgainstGlobal(p, blockSlot, 1, [[1, 0, 0], [-1, 1, 0], [0, -1, 1], [1, 1, 1], [-1, -1, -1]]);
    assert(r.checked > 500 && r.neigh > 20 && r.none > 20, 'L' + m + ' ' + JSON.stringify(r));
    out.push('L' + m + ' ' + r.checked);
  }
  return out.join(', ');
});

ok('a source resolved into a neighbour lands in that tile interior, never its ring', () => {
  const p = makePool({ dims: [16, 16, 16], rb: 4 });
  const { blockSlot } = refineWhere(p, () => true);
  for (const b of [[1, 1, 1], [0, 2, 3], [3, 3, 0]]) {
    for (let fz = GHOST; fz < GHOST + 8; fz++) for (let fy = GHOST; fy < GHOST + 8; fy++) for (let fx = GHOST; fx < GHOST + 8; fx++) {
      for (const d of OFF27) {
        const r = resolveSource(p, blockSlot, b, [fx - d[0], fy - d[1], fz - d[2]]);
        assert(r !== null, 'fully refined: every source resolves');
        assert(p.isInterior(r.fx) && p.isInterior(r.fy) && p.isInterior(r.fz), 'ring cell ' + [r.fx, r.fy, r.fz]);
      }
    }
  }
});

// --- slot assignment --------------------------------------------------------

ok('refineWhere: deterministic, ascending block ids, blockSlot and slotToBlock consistent', () => {
  const p = makePool({ dims: [16, 16, 16], rb: 4 });
  const pred = (b) => Math.hypot(b.mid[0] - 8, b.mid[1] - 8, b.mid[2] - 8) < 5;
  const a = refineWhere(p, pred), b = refineWhere(p, pred);
  eq(Array.from(a.blockSlot), Array.from(b.blockSlot), 'determinism');
  assert(a.activeSlots > 0 && a.activeSlots < p.nBlocks, 'selected ' + a.activeSlots);
  let prev = -1;
  for (let s = 0; s < p.maxSlots; s++) {
    if (s < a.activeSlots) {
      const id = a.slotToBlock[s];
      assert(id > prev && a.blockSlot[id] === s, 'slot ' + s + ' -> block ' + id);
      prev = id;
    } else {
      assert(a.slotToBlock[s] === -1, 'slot ' + s + ' above activeSlots is free');
    }
  }
  for (let id = 0; id < p.nBlocks; id++) {
    if (a.blockSlot[id] >= 0) assert(a.slotToBlock[a.blockSlot[id]] === id, 'block ' + id);
  }
});

ok('refineWhere throws when the refinement needs more than maxSlots', () => {
  const p = makePool({ dims: [16, 16, 16], rb: 4, maxSlots: 3 });
  throwsLike(() => refineWhere(p, () => true), /more than maxSlots/, 'exhaustion');
});

ok('refineNearBody refines the surface blocks and nothing beyond the margin; coverage check passes', () => {
  const p = makePool({ dims: [32, 32, 32], rb: 4 });
  const sdf = sphereSdf([16, 16, 16], 6), margin = 2;
  const r = refineNearBody(p, sdf, margin);
  assert(r.activeSlots > 0, 'nothing refined');
  for (let id = 0; id < p.nBlocks; id++) {
    const b = p.blockOf(id);
    const mid = b.map((c) => c * p.rb + p.rb / 2);
    if (sdf(mid) > margin + p.rb * Math.sqrt(3)) assert(r.blockSlot[id] === -1, 'far block ' + id + ' refined');
  }
  const cov = checkGeometryCoverage(p, r.blockSlot, sdf, margin);
  assert(cov.required > 0 && cov.violations.length === 0, JSON.stringify(cov.violations.slice(0, 2)));
  return r.activeSlots + ' tiles';
});

ok('checkGeometryCoverage catches the block under the body centre being unrefined', () => {
  const p = makePool({ dims: [32, 32, 32], rb: 4 });
  const c = [16, 16, 16], sdf = sphereSdf(c, 5), margin = 2;
  const { blockSlot } = refineNearBody(p, sdf, margin);
  const hole = p.blockId(4, 4, 4);
  assert(blockSlot[hole] >= 0, 'centre block refined at the start');
  blockSlot[hole] = -1;
  const r = checkGeometryCoverage(p, blockSlot, sdf, margin);
  assert(r.violations.length > 0 && r.violations.every((v) => v.block === hole), JSON.stringify(r.violations.slice(0, 2)));
});

ok('nearBodyWant is conservative: a body smaller than a block and rotated thin plates are never missed', () => {
  const p = makePool({ dims: [32, 32, 32], rb: 4 });
  const small = sphereSdf([1, 1, 1], 0.6);
  const r = refineNearBody(p, small, 0.1);
  assert(r.activeSlots > 0 && checkGeometryCoverage(p, r.blockSlot, small, 0.1).nViolations === 0, 'small body');
  const rnd = lcg(20260911);
  const half = [7, 3.5, 1], margin = 1.5;
  for (let t = 0; t < 12; t++) {
    const q = [rnd() - 0.5, rnd() - 0.5, rnd() - 0.5, rnd() - 0.5];
    const n = Math.hypot(q[0], q[1], q[2], q[3]);
    const [w, x, y, z] = q.map((v) => v / n);
    const R = [[1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)],
      [2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)],
      [2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)]];
    const ctr = [16 + (rnd() - 0.5) * 4, 16 + (rnd() - 0.5) * 4, 16 + (rnd() - 0.5) * 4];
    const box = (pt) => {
      const d = [pt[0] - ctr[0], pt[1] - ctr[1], pt[2] - ctr[2]];
      const g = R.map((row, i) => Math.abs(d[0] * row[0] + d[1] * row[1] + d[2] * row[2]) - half[i]);
      return Math.hypot(Math.max(g[0], 0), Math.max(g[1], 0), Math.max(g[2], 0)) + Math.min(Math.max(g[0], g[1], g[2]), 0);
    };
    const rr = refineNearBody(p, box, margin);
    assert(checkGeometryCoverage(p, rr.blockSlot, box, margin).nViolations === 0, 'plate pose ' + t);
  }
  // the predicate is the one refineNearBody applies
  const want = nearBodyWant(small, 0.1);
  assert(want({ lo: [0, 0, 0], hi: [4, 4, 4] }) === true && want({ lo: [16, 16, 16], hi: [20, 20, 20] }) === false, 'predicate');
});

// --- deeper levels ----------------------------------------------------------

ok('poolAtLevel: level 1 is the pool, deeper levels double the block grid with the same tile; level 0 throws', () => {
  const p = makePool({ dims: [16, 8, 16], rb: 4 });
  const l1 = poolAtLevel(p, 1);
  eq(l1.nb, p.nb, 'level 1 nb');
  for (let m = 2; m <= 4; m++) {
    const lm = poolAtLevel(p, m);
    assert(lm.FB === p.FB && lm.rb === p.rb, 'tile at level ' + m);
    eq(lm.nb, p.nb.map((n) => n * Math.pow(2, m - 1)), 'nb at level ' + m);
  }
  throwsLike(() => poolAtLevel(p, 0), /no pool/, 'level 0');
});

ok('octant bits place a child cell inside its parent tile (global route at levels 2 and 3)', () => {
  const base = makePool({ dims: [8, 8, 8], rb: 2 });
  let n = 0;
  for (const m of [2, 3]) {
    const child = poolAtLevel(base, m), par = poolAtLevel(base, m - 1);
    const span = 2 * child.rb;
    for (let id = 0; id < child.nBlocks; id += 3) {
      const b = child.blockOf(id), pb = parentOfBlock(b), q = octantOfBlock(b);
      eq(pb, b.map((x) => Math.floor(x / 2)), 'parent of ' + b);
      for (const j of [0, 1, GHOST, GHOST + 1, GHOST + span - 1, GHOST + span, child.FB - 1]) {
        for (let a = 0; a < 3; a++) {
          const gChild = b[a] * span + (j - GHOST);
          const nPar = 2 * par.dims[a];
          const gPar = ((Math.floor(gChild / 2) % nPar) + nPar) % nPar;
          const pBlock = Math.floor(gPar / (2 * par.rb));
          if (j >= GHOST && j < GHOST + span) {
            assert(pBlock === pb[a], 'L' + m + ' parent block axis ' + a);
            assert(gPar - pBlock * 2 * par.rb + GHOST === octantOrigin(child.rb, q[a]) + ((j - GHOST) >> 1), 'L' + m + ' parent local');
          }
          // continuous position: child centre (g + 1/2) dx_c, parent-local u at
          // (pb 2RB + u - GHOST + 1/2) dx_p with dx_c = dx_p / 2
          const u = fineToCoarseUnit(j, octantOrigin(child.rb, q[a]));
          const uWant = GHOST - 0.5 + (gChild + 0.5) / 2 - pb[a] * 2 * par.rb;
          assert(Math.abs(u - uWant) < 1e-12, 'L' + m + ' continuous coordinate ' + u + ' vs ' + uWant);
          n++;
        }
      }
    }
  }
  return n + ' placements';
});

// --- 2:1 balance ------------------------------------------------------------

const nb8 = (m) => [8 * Math.pow(2, m - 1), 8 * Math.pow(2, m - 1), 8 * Math.pow(2, m - 1)];
const nb4 = (m) => [4 * Math.pow(2, m - 1), 4 * Math.pow(2, m - 1), 4 * Math.pow(2, m - 1)];
const onX = (r, bx) => r.violations.filter((v) => v.block[0] === bx && v.axis === 0);

ok('check21Balance: a single refined level cannot violate', () => {
  const p = makePool({ dims: [32, 32, 32], rb: 4 });
  const { blockSlot } = refineWhere(p, (b) => b.bx < 4);
  const lv1 = new Set();
  for (let id = 0; id < p.nBlocks; id++) if (blockSlot[id] >= 0) lv1.add(keyOf(p.blockOf(id)));
  const r = check21Balance([null, lv1], () => p.nb, { levels: 2 });
  assert(r.violations.length === 0 && r.levels === 2, JSON.stringify(r.violations[0]));
});

ok('check21Balance catches a level-3 leaf whose -x face resolves at level 1', () => {
  const lv1 = new Set(['3,0,0', '2,0,0']);
  const good = check21Balance([null, lv1, new Set(['6,0,0', '5,0,0']), new Set(['12,0,0'])], nb8, { levels: 4 });
  assert(onX(good, 12).length === 0, 'legal next to level 2: ' + JSON.stringify(onX(good, 12)));
  const bad = check21Balance([null, lv1, new Set(['6,0,0']), new Set(['12,0,0'])], nb8, { levels: 4 });
  assert(onX(bad, 12).some((v) => v.dir === -1 && v.nDepth === 1), JSON.stringify(bad.violations));
});

ok('check21Balance reads the shared face of a neighbour, not its deepest corner', () => {
  const lv1 = new Set(['6,0,0', '5,0,0']), lv2 = new Set(['10,0,0', '11,0,0']);
  const far = check21Balance([null, lv1, lv2, new Set(['20,0,0'])], nb8, { levels: 4 });
  assert(onX(far, 6).length === 0, 'far-side level 3 is legal');
  const near = check21Balance([null, lv1, lv2, new Set(['22,0,0'])], nb8, { levels: 4 });
  assert(near.violations.some((v) => v.block[0] === 6 && v.axis === 0 && v.dir === -1 && v.nDepth === 3), 'near side');
});

ok('check21Balance wraps periodically', () => {
  const s = new Set(['0,0,0']);
  const r = check21Balance([null, s, s, s], nb8, { levels: 4 });
  assert(r.violations.some((v) => v.block[0] === 0 && v.axis === 0 && v.dir === -1 && v.nDepth === 0), 'wrap');
});

// --- the cascade ------------------------------------------------------------

function orphans(sets, levels) {
  const out = [];
  for (let m = 2; m < levels; m++) {
    (sets[m] || new Set()).forEach((k) => {
      if (!sets[m - 1].has(keyOf(parentOfBlock(parse(k))))) out.push(m + ':' + k);
    });
  }
  return out;
}
function octetOf(b) {
  const p = parentOfBlock(b), s = new Set();
  for (let k = 0; k < 8; k++) s.add(keyOf([2 * p[0] + (k & 1), 2 * p[1] + ((k >> 1) & 1), 2 * p[2] + ((k >> 2) & 1)]));
  return s;
}

ok('cascade21 is the identity at levels = 2 (20 random level-1 sets)', () => {
  const p = makePool({ dims: [32, 32, 32], rb: 4 });
  const rnd = lcg(12345);
  for (let t = 0; t < 20; t++) {
    const lv1 = new Set();
    for (let id = 0; id < p.nBlocks; id++) if (rnd() < 0.3) lv1.add(keyOf(p.blockOf(id)));
    const r = cascade21([null, lv1], () => p.nb, { levels: 2 });
    assert(r.forced.length === 0, 'trial ' + t + ' forced ' + JSON.stringify(r.forced[0]));
    eq(sortedKeys(r.sets), sortedKeys([null, lv1]), 'trial ' + t);
  }
});

ok('cascade21 closes a 3-vs-1 gap; check21Balance then finds nothing, no orphans, wants kept', () => {
  const start = [null, new Set(['3,0,0']), new Set(['6,0,0']), new Set(['12,0,0'])];
  assert(check21Balance(start, nb4, { levels: 4 }).violations.length > 0, 'premise');
  const c = cascade21(start, nb4, { levels: 4 });
  assert(c.forced.length > 0, 'nothing forced');
  const after = check21Balance(c.sets, nb4, { levels: 4 });
  assert(after.violations.length === 0, JSON.stringify(after.violations[0]));
  assert(orphans(c.sets, 4).length === 0 && c.sets[3].has('12,0,0'), 'orphans / dropped want');
});

ok('cascade21 materialises the parents of a level-3 want with no ancestors', () => {
  const c = cascade21([null, new Set(), new Set(), new Set(['12,4,4'])], nb4, { levels: 4 });
  assert(c.sets[2].has('6,2,2') && c.sets[1].has('3,1,1'), 'parents');
  assert(orphans(c.sets, 4).length === 0 && check21Balance(c.sets, nb4, { levels: 4 }).violations.length === 0, 'tree');
});

ok('cascade21 is idempotent', () => {
  const once = cascade21([null, new Set(['3,0,0']), new Set(['6,0,0']), new Set(['12,0,0', '13,5,2'])], nb4, { levels: 4 });
  const twice = cascade21(once.sets, nb4, { levels: 4 });
  assert(twice.forced.length === 0, 'second sweep forced ' + JSON.stringify(twice.forced[0]));
  eq(sortedKeys(twice.sets), sortedKeys(once.sets), 'sets');
});

ok('cascade21 adds nothing to a uniformly refined tree', () => {
  const sets = [null];
  for (let m = 1; m <= 3; m++) {
    const n = nb4(m)[0], s = new Set();
    for (let z = 0; z < n; z++) for (let y = 0; y < n; y++) for (let x = 0; x < n; x++) s.add(keyOf([x, y, z]));
    sets.push(s);
  }
  assert(cascade21(sets, nb4, { levels: 4 }).forced.length === 0, 'forced on a uniform tree');
});

ok('cascade21 on nested sphere shells gives the reference\'s recorded counts 432 / 1280 / 2432', () => {
  const nbAt = (m) => [8 * Math.pow(2, m - 1), 8 * Math.pow(2, m - 1), 8 * Math.pow(2, m - 1)];
  const shell = (m, band) => {
    const n = nbAt(m)[0], w = 32 / n, out = new Set();
    for (let z = 0; z < n; z++) for (let y = 0; y < n; y++) for (let x = 0; x < n; x++) {
      const d = Math.hypot((x + 0.5) * w - 16, (y + 0.5) * w - 16, (z + 0.5) * w - 16) - 5;
      if (Math.abs(d) <= band) out.add(keyOf([x, y, z]));
    }
    return out;
  };
  const sets = [null, shell(1, 6), shell(2, 3), shell(3, 2.5)];
  assert(check21Balance(sets, nbAt, { levels: 4 }).violations.length > 0, 'premise: unbalanced shells');
  const c = cascade21(sets, nbAt, { levels: 4 });
  assert(check21Balance(c.sets, nbAt, { levels: 4 }).violations.length === 0, 'balanced');
  assert(orphans(c.sets, 4).length === 0, 'orphans');
  for (let m = 1; m < 4; m++) sets[m].forEach((k) => assert(c.sets[m].has(k), 'dropped ' + k + ' at ' + m));
  eq(c.counts.slice(1), [432, 1280, 2432], 'counts (tools/test-d3-amr.js:664-665)');
  assert(checkRingParentCoverage(c.sets, nbAt, { levels: 4 }).violations.length === 0, 'ring parents');
  return c.counts.slice(1).join(' / ');
});

ok('cascade21 output is octet-complete', () => {
  const c = cascade21([null, new Set(), new Set(), new Set(['12,4,4'])], nb4, { levels: 4 });
  for (let m = 2; m < 4; m++) {
    c.sets[m].forEach((k) => {
      octetOf(parse(k)).forEach((sib) => assert(c.sets[m].has(sib), 'level ' + m + ' ' + k + ' without sibling ' + sib));
    });
  }
});

ok('every octet cascade21 forces is necessary (removing it breaks balance, parenting or ring coverage)', () => {
  const c = cascade21([null, new Set(), new Set(), octetOf([12, 4, 4])], nb4, { levels: 4 });
  const groups = new Map();
  for (const f of c.forced) {
    const g = f.level + ':' + (f.octet ? f.octet.join(',') : f.block.join(','));
    if (!groups.has(g)) groups.set(g, []);
    groups.get(g).push(f);
  }
  assert(groups.size > 3, 'only ' + groups.size + ' groups');
  groups.forEach((members, g) => {
    const cut = c.sets.map((s) => (s ? new Set(s) : null));
    for (const f of members) cut[f.level].delete(f.block.join(','));
    const broken = check21Balance(cut, nb4, { levels: 4 }).violations.length + orphans(cut, 4).length +
      checkRingParentCoverage(cut, nb4, { levels: 4 }).violations.length;
    assert(broken > 0, 'removing ' + g + ' leaves a legal tree');
  });
  return groups.size + ' octets';
});

ok('ring-parent coverage catches a missing corner parent that 2:1 balance cannot see', () => {
  const child = new Set(['2,2,2']);
  const faceOnly = new Set(['1,1,1', '0,1,1', '1,0,1', '1,1,0']);
  const A = [null, faceOnly, child];
  assert(checkRingParentCoverage(A, nb4, { levels: 3 }).violations.some((v) => v.missingParent === '0,0,0'), 'corner');
  const B = [null, new Set(Array.from(faceOnly).concat(['0,0,0', '0,0,1', '0,1,0', '1,0,0'])), child];
  assert(checkRingParentCoverage(B, nb4, { levels: 3 }).violations.length === 0, 'eight parents suffice');
  for (const t of [A, B]) assert(check21Balance(t, nb4, { levels: 3 }).violations.filter((v) => v.level === 2).length === 0, 'balance silent');
});

// --- refinement hierarchy ---------------------------------------------------

const bodyPredicate = (sdf, margin) => (box) => {
  let best = Infinity;
  for (let i = 0; i < 8; i++) best = Math.min(best, sdf([i & 1 ? box.hi[0] : box.lo[0], i & 2 ? box.hi[1] : box.lo[1], i & 4 ? box.hi[2] : box.lo[2]]));
  return Math.min(best, sdf([(box.lo[0] + box.hi[0]) / 2, (box.lo[1] + box.hi[1]) / 2, (box.lo[2] + box.hi[2]) / 2])) <= margin;
};

ok('refineHierarchy at levels = 2 is bit-identical to refineWhere / refineNearBody', () => {
  const p = makePool({ dims: [32, 32, 32], rb: 4 });
  const cases = [(b) => b.mid.every((c) => c >= 8 && c < 24), () => true, (b) => (b.bx + b.by + b.bz) % 3 !== 2];
  cases.forEach((pred, k) => {
    const flat = refineWhere(p, pred), h = refineHierarchy(p, { levels: 2, want: pred });
    eq(Array.from(h.byLevel[1].blockSlot), Array.from(flat.blockSlot), 'case ' + k + ' blockSlot');
    eq(Array.from(h.byLevel[1].slotToBlock), Array.from(flat.slotToBlock), 'case ' + k + ' slotToBlock');
    assert(h.byLevel[1].activeSlots === flat.activeSlots, 'case ' + k + ' activeSlots');
  });
  const sdf = sphereSdf([16, 16, 16], 5);
  const flat = refineNearBody(p, sdf, 2), h = refineHierarchy(p, { levels: 2, want: bodyPredicate(sdf, 2) });
  eq(Array.from(h.byLevel[1].blockSlot), Array.from(flat.blockSlot), 'body');
  assert(flat.activeSlots > 0, 'body refines');
});

ok('refineHierarchy at levels 3 and 4: balanced, ring-covered, every level populated and parented, shells nested', () => {
  const p = makePool({ dims: [32, 32, 32], rb: 4 });
  const want = bodyPredicate(sphereSdf([16, 16, 16], 5), 2);
  const out = [];
  for (const levels of [3, 4]) {
    const h = refineHierarchy(p, { levels: levels, want: want });
    const nbAt = (m) => poolAtLevel(p, m).nb;
    assert(check21Balance(h.sets, nbAt, { levels: levels }).violations.length === 0, 'levels ' + levels + ' balance');
    const ring = checkRingParentCoverage(h.sets, nbAt, { levels: levels });
    assert(ring.required > 0 && ring.violations.length === 0, 'levels ' + levels + ' ring ' + JSON.stringify(ring.violations[0]));
    for (let m = 1; m < levels; m++) {
      assert(h.byLevel[m].activeSlots > 0, 'level ' + m + ' empty');
      if (m >= 2) h.sets[m].forEach((k) => assert(h.sets[m - 1].has(keyOf(parentOfBlock(parse(k)))), 'orphan ' + k));
    }
    for (let m = 2; m < levels; m++) {
      const wf = p.rb / Math.pow(2, m - 1), wc = p.rb / Math.pow(2, m - 2);
      let xf = -Infinity, xc = -Infinity;
      h.sets[m].forEach((k) => { xf = Math.max(xf, (parse(k)[0] + 1) * wf); });
      h.sets[m - 1].forEach((k) => { xc = Math.max(xc, (parse(k)[0] + 1) * wc); });
      assert(xc >= xf, 'level ' + (m - 1) + ' does not enclose level ' + m);
    }
    out.push('levels ' + levels + ': ' + h.byLevel.slice(1).map((l) => l.activeSlots).join('/'));
  }
  return out.join(', ');
});

ok('refineHierarchy covers the body at the finest level, cell by cell', () => {
  const p = makePool({ dims: [16, 16, 16], rb: 4 });
  const sdf = sphereSdf([8, 8, 8], 3), margin = 1.5, levels = 3;
  const h = refineHierarchy(p, { levels: levels, want: bodyPredicate(sdf, margin) });
  const fine = h.byLevel[levels - 1], s = Math.pow(2, levels - 2);
  const r = checkGeometryCoverage(fine.pool, fine.blockSlot, (q) => sdf(q.map((v) => v / s)) * s, margin * s);
  assert(r.required > 0 && r.violations.length === 0, JSON.stringify(r.violations[0]));
});

// --- viewer frame -----------------------------------------------------------

ok('cellAtLevel agrees with walking the per-rung affine map L = u/2 - 1/4, and stays in the interior', () => {
  const dims = [16, 16, 16], rb = 4;
  for (const m of [1, 2, 3]) {
    for (const g of [0, 1, 2, 7, 8, 15, 16, 31, 63, 100]) {
      if (g >= dims[0] * Math.pow(2, m)) continue;
      let c = g;
      for (let k = 0; k < m; k++) c = 0.5 * c - 0.25;
      const r = cellAtLevel([c, c, c], m, dims, rb);
      assert(r.g[0] === g, 'level ' + m + ' cell ' + g + ' -> ' + r.g[0]);
      assert(r.local[0] >= GHOST && r.local[0] < GHOST + 2 * rb, 'ring reached');
    }
  }
});

ok('cellAtLevel at m = 0 is the dense cell and wraps against the level\'s own dims', () => {
  const dims = [16, 8, 4], rb = 4;
  eq(cellAtLevel([0.49, 0, 0], 0, dims, rb).g, [0, 0, 0], '0.49');
  eq(cellAtLevel([0.51, 0, 0], 0, dims, rb).g, [1, 0, 0], '0.51');
  eq(cellAtLevel([-0.6, -0.6, -0.6], 0, dims, rb).g, [15, 7, 3], 'lower wrap');
  eq(cellAtLevel([15.6, 0, 0], 0, dims, rb).g, [0, 0, 0], 'upper wrap');
  assert(cellAtLevel([15.6, 0, 0], 2, dims, rb).g[0] === 0 && cellAtLevel([-0.6, 0, 0], 1, dims, rb).g[0] === 31, 'level wrap');
});

ok('finestLevelAt picks the deepest covering level', () => {
  const o = { levels: 3, dims: [16, 16, 16], rb: 4, present: (m, b) => m === 1 || (b[0] === 0 && b[1] === 0 && b[2] === 0) };
  assert(finestLevelAt([0, 0, 0], o) === 2 && finestLevelAt([1.4, 0, 0], o) === 2, 'inside level-2 block 0');
  assert(finestLevelAt([1.6, 0, 0], o) === 1 && finestLevelAt([12, 12, 12], o) === 1, 'fallback to level 1');
  assert(finestLevelAt([1.4, 0, 0], { levels: 3, dims: [16, 16, 16], rb: 4, present: () => false }) === 0, 'dense');
});

ok('the viewer-frame check sees a dropped half cell, rounding and the parent scale', () => {
  const dims = [16, 16, 16], rb = 4, n = (m) => 16 * Math.pow(2, m);
  const wrap = (v, m) => ((v % n(m)) + n(m)) % n(m);
  const mutants = [(x, m) => Math.floor(x * Math.pow(2, m)), (x, m) => Math.round((x + 0.5) * Math.pow(2, m)),
    (x, m) => Math.floor((x + 0.5) * Math.pow(2, m - 1))];
  mutants.forEach((mut, k) => {
    let differs = false;
    for (const m of [1, 2]) for (let i = 0; i < 40; i++) if (cellAtLevel([0.37 * i, 0, 0], m, dims, rb).g[0] !== wrap(mut(0.37 * i, m), m)) differs = true;
    assert(differs, 'mutant ' + k + ' not seen');
  });
});

// --- dynamic manager twin ---------------------------------------------------

ok('manageSlots keeps survivors in place, frees unwanted slots, fills new blocks lowest slot first', () => {
  const bs = Int32Array.from([-1, 0, -1, 1, 2, -1]);
  const sb = Int32Array.from([1, 3, 4, -1]);
  const r = manageSlots(bs, sb, [1, 0, 1, 1, 0, 1]);
  eq(Array.from(r.blockSlot), [0, -1, 2, 1, -1, 3], 'blockSlot');
  eq(Array.from(r.slotToBlock), [0, 3, 2, 5], 'slotToBlock');
  eq(r.coarsened, [1, 4], 'coarsened');
  eq(r.refined, [0, 2, 5], 'refined');
  assert(r.exhausted === 0, 'exhausted');
  const full = manageSlots(Int32Array.from([0, -1, -1]), Int32Array.from([0]), [1, 1, 1]);
  assert(full.exhausted === 2, 'exhaustion counted');
  eq(Array.from(activeSlotList(r.slotToBlock)), [4, 0, 1, 2, 3], 'active list');
  eq(Array.from(activeSlotList(Int32Array.from([-1, 7, -1, 2]))), [2, 1, 3], 'active list ascending');
});

process.exitCode = failures ? 1 : 0;
