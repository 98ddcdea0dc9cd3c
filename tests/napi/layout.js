'use strict';
// Prints the bytes the JavaScript GpuSolver3D packs for params, a joint spec and a spring spec,
// and what its constructor throws on invalid params (no GPU call: the C ABI refuses first).
const g = require(process.argv[2]);
const out = { exports: Object.keys(g.addon).sort(), sizes: g.addon.structSizes() };
out.params = Buffer.from(g.packParams(g.defaultParams({ dt: 1 / 240, colorCap: 24, flags: g.FLAG_DEFAULTS | g.FLAG_ADAPT_COLORS }))).toString('hex');
const methods = Object.getOwnPropertyNames(g.GpuSolver3D.prototype).sort();
out.methods = methods;
try {
  new g.GpuSolver3D(new Float32Array(40), null, null, { params: { colorRounds: 3 } });
  out.create = 'no error';
} catch (e) {
  out.create = `${e.constructor.name}: ${e.message}`;
}
try {
  g.addon.step({}, false);
  out.handle = 'no error';
} catch (e) {
  out.handle = `${e.constructor.name}: ${e.message}`;
}
// the spec packers, through a stand-in addon call that hands back the buffer it was given
const captured = {};
const proto = g.GpuSolver3D.prototype;
const fake = Object.create(proto);
fake.live = () => 0;
const realAppend = g.addon.appendJoints, realSprings = g.addon.appendSprings;
g.addon.appendJoints = (h, buf) => { captured.joint = Buffer.from(buf).toString('hex'); };
g.addon.appendSprings = (h, buf) => { captured.spring = Buffer.from(buf).toString('hex'); };
proto.appendJoints.call(fake, [{ a: 3, b: 7, rA: [0.5, 0, -1], rB: [0, 0.25, 0], yield: 40, breakBend: 0.5, rest: [0, 0, 0.1, 1], linear: true }], Infinity, false);
proto.appendSprings.call(fake, [{ a: -1, b: 2, rA: [1, 2, 3], rB: [4, 5, 6], stiffness: 50, rest: 1.5 }]);
g.addon.appendJoints = realAppend;
g.addon.appendSprings = realSprings;
out.joint = captured.joint;
out.spring = captured.spring;
console.log(JSON.stringify(out));
