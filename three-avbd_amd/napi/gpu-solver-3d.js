'use strict';
/**
 * GpuSolver3D over the MI355X step (libavbd_mi355x.so through avbd_node.node): the reference's
 * WebGPU solver class (src/avbd3d/gpu/solver.ts:504) with the same method names, argument meaning
 * and errors, for the host code that drives it (World, src/lib/world.ts:847, :1736-1995). Records
 * keep the reference's layouts (src/avbd3d/gpu/layout.ts), so readers of body records are
 * unchanged. Reads resolve at once: the library's reads synchronise its stream, where the
 * reference's resolve when WebGPU's mapAsync does (solver.ts:2418-2430).
 */
const path = require('path');

const addon = require(process.env.AVBD_NODE_ADDON || path.join(__dirname, 'avbd_node.node'));
const SIZES = addon.structSizes();

// layout.ts:49-58, :84-110
const BODY_FLOATS = 40, JOINT_FLOATS = 40, MANIFOLD_WORDS = 8, CONTACT_WORDS = 16;
const B_POS = 0, B_ROT = 4, B_INITIAL_POS = 8, B_INITIAL_ROT = 12, B_SIZE = 16, B_MOMENT = 20;
const B_INERTIAL_POS = 24, B_INERTIAL_ROT = 28, B_VEL = 32, B_ANGVEL = 36;
const J_PEN_LIN = 0, J_PEN_ANG = 4, J_LAM_LIN = 8, J_LAM_ANG = 12, J_C0_LIN = 16, J_C0_ANG = 20;
const J_RA = 24, J_RB = 28, J_REST = 32, J_REST_START = 36, J_YIELD = 23, J_BREAK_BEND = 31;
const T_NONE = 0, T_JOINT = 1, T_SPRING = 2;
// GpuParams3D.flags (layout.ts:156-166) and this library's host flags (include/avbd.h)
const FLAG_MATCH_NEAREST = 1, FLAG_FACE_BIAS = 2, FLAG_REUSE_CONTACTS = 4, FLAG_START_AT_REST = 8, FLAG_MASS_PENALTY = 16;
const FLAG_DEFAULTS = 31, FLAG_ADAPT_COLORS = 128, FLAG_FIXED_STORAGE = 256;
const FIELDS = { position: 1, rotation: 2, velocity: 4, angularVelocity: 8 };

/** StorageCapacityError (src/lib/storage-budget.ts): a capacity refused before anything changed. */
class StorageCapacityError extends Error {}

function rethrow(e) {
  if (e && e.code === 'StorageCapacityError') {
    const err = new StorageCapacityError(e.message);
    err.code = e.code;
    throw err;
  }
  throw e;
}
function call(fn, ...args) {
  try {
    return fn(...args);
  } catch (e) {
    return rethrow(e);
  }
}

/** gpuParams3D's defaults (solver.ts:213-226) with z up. */
function defaultParams(over = {}) {
  return Object.assign({
    dt: 1 / 60, gravity: -9.81, betaLin: 1e4, betaAng: 100, gamma: 0.999, alpha: 0.99, contactMargin: 0.01,
    up: [0, 0, 1], flags: FLAG_DEFAULTS, iterations: 10, colorCap: 64, colorRounds: 16,
    primalLanes: [2 ** 31, 2 ** 15, 2 ** 12, 2 ** 11],
  }, over);
}
/** avbd_params as bytes (include/avbd.h: 10 f32, then 8 u32). */
function packParams(p) {
  const buf = new ArrayBuffer(SIZES.params);
  const f = new Float32Array(buf), u = new Uint32Array(buf);
  f.set([p.dt, p.gravity, p.betaLin, p.betaAng, p.gamma, p.alpha, p.contactMargin, p.up[0], p.up[1], p.up[2]], 0);
  u.set([p.flags >>> 0, p.iterations, p.colorCap, p.colorRounds], 10);
  u.set(p.primalLanes.map((x) => Math.min(x, 2 ** 32 - 1) >>> 0), 14);
  return buf;
}
function unpackParams(buf) {
  const f = new Float32Array(buf), u = new Uint32Array(buf);
  return {
    dt: f[0], gravity: f[1], betaLin: f[2], betaAng: f[3], gamma: f[4], alpha: f[5], contactMargin: f[6],
    up: [f[7], f[8], f[9]], flags: u[10], iterations: u[11], colorCap: u[12], colorRounds: u[13],
    primalLanes: [u[14], u[15], u[16], u[17]],
  };
}
const vec3 = (v) => [v[0], v[1], v[2]];

class GpuSolver3D {
  /**
   * new GpuSolver3D(...) (solver.ts:722): `bodies` and `joints` as records (Float32Array, 40
   * floats each), `info` (Int32Array, 4 per joint: type, a, b, 0), options.params (defaults
   * above), options.capacity {bodies, joints, pairs, manifolds, contacts, hullWords}.
   */
  constructor(bodies, joints = null, info = null, options = {}) {
    this.params = defaultParams(options.params || {});
    const c = options.capacity || {};
    const cap = new Uint32Array([c.bodies || 0, c.joints || 0, c.pairs || 0, c.manifolds || 0, c.contacts || 0, c.hullWords || 0]);
    this.handle = call(addon.create, packParams(this.params), bodies || new Float32Array(0), joints, info, cap);
    this.destroyed = false;
    this.sent = unpackParams(packParams(this.params));
  }
  get bodyCount() { return call(addon.getSizes, this.live())[0]; }
  get jointCount() { return call(addon.getSizes, this.live())[1]; }
  live() {
    if (this.destroyed) throw new Error('Solver is destroyed');
    return this.handle;
  }
  /**
   * The params the next step runs with: fields this object changed since they were last sent
   * win, the rest are the library's (its adaptive colouring moves colorCap and colorRounds).
   */
  syncParams() {
    const h = this.live();
    const buf = new ArrayBuffer(SIZES.params);
    call(addon.getParams, h, buf);
    const cur = unpackParams(buf), sent = this.sent || cur, out = {};
    for (const k of Object.keys(cur)) out[k] = JSON.stringify(this.params[k]) !== JSON.stringify(sent[k]) ? this.params[k] : cur[k];
    call(addon.setParams, h, packParams(out));
    call(addon.getParams, h, buf);
    this.params = unpackParams(buf);
    this.sent = unpackParams(buf);
  }
  /** GpuSolver3D.step (solver.ts:2112): one substep with the current params. */
  step(options = {}) {
    this.syncParams();
    call(addon.step, this.handle, !!options.reuseManifolds);
  }
  /** `steps` steps of `substeps` substeps (the first detects, the rest reuse manifolds). */
  steps(steps, substeps = 1) {
    this.syncParams();
    call(addon.steps, this.handle, steps, substeps);
  }
  readBodies() {
    const out = new Float32Array(this.bodyCount * BODY_FLOATS);
    if (out.length) call(addon.readBodies, this.live(), out, 0);
    return Promise.resolve(out);
  }
  readJoints() {
    const out = new Float32Array(this.jointCount * JOINT_FLOATS);
    if (out.length) call(addon.readJoints, this.live(), out, 0);
    return Promise.resolve(out);
  }
  jointInfo() {
    const out = new Int32Array(this.jointCount * 4);
    if (out.length) call(addon.readJointInfo, this.live(), out);
    return out;
  }
  readCounters() {
    const u = new Uint32Array(SIZES.counters / 4);
    call(addon.readCounters, this.live(), u);
    return Promise.resolve({ pairs: u[0], contacts: u[1], manifolds: u[2], overflow: u[3], clashes: u[4], colors: u[5], colorCap: u[6], step: u[7] });
  }
  /** readContactList (solver.ts:2311-2346): one entry per contact point. */
  async readContactList() {
    const c = await this.readCounters();
    const man = new Uint32Array(Math.max(c.manifolds, 1) * MANIFOLD_WORDS);
    const con = new Uint32Array(Math.max(c.contacts, 1) * CONTACT_WORDS);
    const nm = call(addon.readContacts, this.live(), man, con);
    const mf = new Float32Array(man.buffer), cf = new Float32Array(con.buffer);
    const out = [];
    for (let m = 0; m < nm; m++) {
      const o = m * MANIFOLD_WORDS, first = man[o + 2], count = man[o + 3] & 15;
      for (let k = first; k < first + count; k++) {
        const q = k * CONTACT_WORDS;
        out.push({ a: man[o], b: man[o + 1], normal: [mf[o + 4], mf[o + 5], mf[o + 6]], rA: [cf[q], cf[q + 1], cf[q + 2]],
          rB: [cf[q + 4], cf[q + 5], cf[q + 6]], lambda: [cf[q + 12], cf[q + 13], cf[q + 14]], feature: con[q + 3] & 0x7fffffff });
      }
    }
    return out;
  }
  async readPairs() {
    const c = await this.readCounters();
    const out = new Uint32Array(Math.max(c.pairs, 1) * 2);
    const n = call(addon.readPairs, this.live(), out);
    return out.subarray(0, 2 * n);
  }
  /** rewriteBodies (solver.ts:1774) with records: consecutive runs in one write each. */
  rewriteBodies(indices, records) {
    const h = this.live();
    for (let k = 0; k < indices.length;) {
      let e = k + 1;
      while (e < indices.length && indices[e] === indices[e - 1] + 1) e++;
      call(addon.rewriteBodies, h, records.subarray(k * BODY_FLOATS, e * BODY_FLOATS), indices[k]);
      k = e;
    }
  }
  /** addBodies (solver.ts:1544): records; the first one's index, or -1 if none fit. */
  addBodies(records) { return call(addon.addBodies, this.live(), records); }
  addBody(record) { return this.addBodies(record); }
  /** patchBody (solver.ts:1103): `fields` a Set of position/rotation/velocity/angularVelocity. */
  patchBody(index, fields, record, fixed = false, marker = true) {
    let mask = 0;
    for (const f of fields) mask |= FIELDS[f] || 0;
    call(addon.patchBody, this.live(), index, mask, record, fixed, marker);
  }
  rewriteFixed(indices, positions, size, friction, rotations) {
    call(addon.rewriteFixed, this.live(), Uint32Array.from(indices), Float32Array.from(positions), Float32Array.from(size), friction,
      rotations ? Float32Array.from(rotations) : null);
  }
  /** appendJoints (solver.ts:1637): JointSpec objects; returns their slots. */
  appendJoints(joints, fracture = Infinity, linear = false) {
    const stride = SIZES.jointSpec;
    const buf = new ArrayBuffer(Math.max(1, joints.length) * stride);
    joints.forEach((j, k) => {
      const f = new Float32Array(buf, k * stride, stride / 4), i = new Int32Array(buf, k * stride, stride / 4);
      i[0] = j.a; i[1] = j.b;
      f.set(vec3(j.rA), 2); f.set(vec3(j.rB), 5);
      f[8] = j.angular === undefined ? NaN : j.angular;
      f[9] = j.fracture === undefined ? NaN : j.fracture;
      i[10] = j.linear === undefined ? -1 : (j.linear ? 1 : 0);
      f.set(j.rest ? [j.rest[0], j.rest[1], j.rest[2], j.rest[3]] : [0, 0, 0, 0], 11);
      f[15] = j.yield === undefined ? NaN : j.yield;
      f[16] = j.breakBend === undefined ? NaN : j.breakBend;
    });
    const slots = new Uint32Array(Math.max(1, joints.length));
    call(addon.appendJoints, this.live(), buf, joints.length, fracture, linear, slots);
    return Array.from(slots.subarray(0, joints.length));
  }
  appendJoint(a, b, rA, rB, stiffnessLin, stiffnessAng) {
    return call(addon.appendJoint, this.live(), a, b, Float32Array.from(vec3(rA)), Float32Array.from(vec3(rB)), stiffnessLin, stiffnessAng);
  }
  appendWorldJoint(b, anchorWorld, anchorBody, policy = 'legacy') {
    if (policy !== 'legacy') throw new Error(`appendWorldJoint: policy ${policy} is not supported (legacy only)`);
    return call(addon.appendWorldJoint, this.live(), b, Float32Array.from(vec3(anchorWorld)), Float32Array.from(vec3(anchorBody)));
  }
  /** appendSprings (solver.ts:1685): {a, b, rA, rB, stiffness, rest}; returns their slots. */
  appendSprings(springs) {
    const stride = SIZES.springSpec;
    const buf = new ArrayBuffer(Math.max(1, springs.length) * stride);
    springs.forEach((s, k) => {
      const f = new Float32Array(buf, k * stride, stride / 4), i = new Int32Array(buf, k * stride, stride / 4);
      i[0] = s.a; i[1] = s.b;
      f.set(vec3(s.rA), 2); f.set(vec3(s.rB), 5);
      f[8] = s.stiffness; f[9] = s.rest;
    });
    const slots = new Uint32Array(Math.max(1, springs.length));
    call(addon.appendSprings, this.live(), buf, springs.length, slots);
    return Array.from(slots.subarray(0, springs.length));
  }
  releaseJoints(slots) { call(addon.releaseJoints, this.live(), Uint32Array.from(slots)); }
  setWorldAnchor(slot, p) { call(addon.setWorldAnchor, this.live(), slot, Float32Array.from(vec3(p))); }
  disableConstraint(slot) { call(addon.disableConstraint, this.live(), slot); }
  setBodyAux(indices, scales, linear, angular) {
    call(addon.setBodyAux, this.live(), Uint32Array.from(indices), Float32Array.from(scales),
      linear ? Float32Array.from(linear) : null, angular ? Float32Array.from(angular) : null);
  }
  setLifecycle(index, bits) { call(addon.setLifecycle, this.live(), index, bits); }
  setFilters(indices, groups, collidesWith) {
    // the ABI takes every body's pair of masks: unchanged bodies keep the defaults (all groups)
    const n = this.bodyCount;
    const g = new Uint32Array(n).fill(0xffffffff), w = new Uint32Array(n).fill(0xffffffff);
    if (this.filters) { g.set(this.filters.g.subarray(0, n)); w.set(this.filters.w.subarray(0, n)); }
    for (let k = 0; k < indices.length; k++) { g[indices[k]] = groups[k] >>> 0; w[indices[k]] = collidesWith[k] >>> 0; }
    this.filters = { g, w };
    call(addon.setFilters, this.live(), g, w);
  }
  /** adapt (solver.ts:2362): the colour cap and storage from a fresh counters read. */
  adapt() { call(addon.adapt, this.live()); }
  destroy() {
    if (this.destroyed) return;
    addon.destroy(this.handle);
    this.destroyed = true;
  }
}

module.exports = {
  GpuSolver3D, StorageCapacityError, defaultParams, packParams, unpackParams, addon,
  BODY_FLOATS, JOINT_FLOATS, MANIFOLD_WORDS, CONTACT_WORDS,
  B_POS, B_ROT, B_INITIAL_POS, B_INITIAL_ROT, B_SIZE, B_MOMENT, B_INERTIAL_POS, B_INERTIAL_ROT, B_VEL, B_ANGVEL,
  J_PEN_LIN, J_PEN_ANG, J_LAM_LIN, J_LAM_ANG, J_C0_LIN, J_C0_ANG, J_RA, J_RB, J_REST, J_REST_START, J_YIELD, J_BREAK_BEND,
  T_NONE, T_JOINT, T_SPRING, FLAG_MATCH_NEAREST, FLAG_FACE_BIAS, FLAG_REUSE_CONTACTS, FLAG_START_AT_REST,
  FLAG_MASS_PENALTY, FLAG_DEFAULTS, FLAG_ADAPT_COLORS, FLAG_FIXED_STORAGE,
};
