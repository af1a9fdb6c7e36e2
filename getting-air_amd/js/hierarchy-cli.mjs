// Prints refineHierarchy's block maps for a pinned sphere as JSON, for hosts
// outside Node (pylbm.scenarios.sphere_hierarchy drives it):
//
//   node hierarchy-cli.mjs '{"dims": [nx, ny, nz], "rb": 4, "levels": 3,
//     "centre": [x, y, z], "radius": r, "margin": 2}'
//
// -> {"levels": L, "byLevel": [null, {"blockSlot": [...], "slotToBlock":
// [...], "activeSlots": n}, ...]}, the finest level wanting every block
// within `margin` L0 cells of the sphere (nearBodyWant) and cascade21 closing
// the coarser levels (refineHierarchy, d3-amr.mjs:603-651).  Slot arrays are
// trimmed to the slots in use.
import { makePool, nearBodyWant, refineHierarchy } from './d3-amr.mjs';

const a = JSON.parse(process.argv[2]);
const pool = makePool({ dims: a.dims, rb: a.rb });
const c = a.centre;
const sdf = (p) => {
  const dx = p[0] - c[0], dy = p[1] - c[1], dz = p[2] - c[2];
  return Math.sqrt(dx * dx + dy * dy + dz * dz) - a.radius;
};
const h = refineHierarchy(pool, { levels: a.levels, want: nearBodyWant(sdf, a.margin) });
const byLevel = [null];
for (let m = 1; m < a.levels; m++) {
  const lv = h.byLevel[m];
  byLevel.push({
    blockSlot: Array.from(lv.blockSlot),
    slotToBlock: Array.from(lv.slotToBlock.subarray(0, lv.activeSlots)),
    activeSlots: lv.activeSlots,
  });
}
process.stdout.write(JSON.stringify({ levels: a.levels, byLevel: byLevel }));
