'use strict';
// Drives the MI355X step from JavaScript as World does (src/lib/world.ts:847, :1861-1995): an
// empty GpuSolver3D filled with addBodies/appendJoints/appendSprings/appendWorldJoint, aux and
// lifecycle edits, then steps; writes the resulting records for the caller to compare.
// argv: module path, input JSON path, output path.
const fs = require('fs');
const g = require(process.argv[2]);
const input = JSON.parse(fs.readFileSync(process.argv[3], 'utf8'));
const f32 = (a) => Float32Array.from(a);

async function main() {
  const s = new g.GpuSolver3D(null, null, null, { params: input.params, capacity: { bodies: 64, joints: 2 } });
  const first = s.addBodies(f32(input.pyramid));
  const base = s.addBodies(f32(input.chain));
  const slots = s.appendJoints(input.specs);
  const spring = s.appendSprings(input.springs);
  const drag = s.appendWorldJoint(input.drag.b, input.drag.world, input.drag.body);
  s.setWorldAnchor(drag, input.drag.moveTo);
  s.setBodyAux(input.aux.indices, input.aux.scales, input.aux.linear, input.aux.angular);
  s.setLifecycle(input.disable, 2);
  s.steps(input.steps, 1);
  s.params.gravity = input.gravity2;  // a params edit between steps
  s.steps(input.steps, 1);
  const bodies = await s.readBodies();
  const joints = await s.readJoints();
  const counters = await s.readCounters();
  const contacts = await s.readContactList();
  s.syncParams();
  const colorCap = s.params.colorCap;
  fs.writeFileSync(process.argv[4], Buffer.concat([Buffer.from(bodies.buffer), Buffer.from(joints.buffer)]));
  // fixed storage: a capacity refusal arrives as the reference's StorageCapacityError
  const fixed = new g.GpuSolver3D(f32(input.pyramid), null, null, { params: Object.assign({}, input.params, { flags: g.FLAG_DEFAULTS | g.FLAG_FIXED_STORAGE }), capacity: { joints: 1 } });
  fixed.appendJoints([input.specs[4]].map((j) => Object.assign({}, j, { a: 1, b: 2 })));
  let refused = 'none';
  try {
    fixed.appendJoints([Object.assign({}, input.specs[4], { a: 1, b: 2 })]);
  } catch (e) {
    refused = e instanceof g.StorageCapacityError ? 'StorageCapacityError' : e.constructor.name;
  }
  fixed.destroy();
  s.destroy();
  let threw = 'none';
  try { s.step(); } catch (e) { threw = e.message; }
  console.log(JSON.stringify({ first, base, slots, spring, drag, counters, contacts: contacts.length, refused, afterDestroy: threw,
    colorCap, jointCount: joints.length / 40 }));
}
main().catch((e) => { console.error(e); process.exit(1); });
