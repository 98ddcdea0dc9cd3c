// TEST INFRASTRUCTURE ONLY — the oracle, built into oracle/liboracle.so and loaded by tests/,
// the smoke check and the bench's CPU-baseline leg (as the checker, never as the thing measured
// or shipped). It is a CPU restatement of the reference's algorithms:
//   - ref3d.hpp: the reference's f64 CPU solver (src/avbd3d/ref/*.ts), pinned by the 14 golden
//     trajectories of tests/golden/oracle3d at 1e-9;
//   - gpu3d.hpp: the WebGPU step's semantics (src/avbd3d/gpu/wgsl-*.ts,
//     src/avbd2d/gpu/wgsl-topology.ts), templated on the float type, pinned against ref3d.hpp by
//     the reference's seeded single-step protocol (tests-gpu/avbd3d-gpu.gpu.test.ts:35-99) and
//     by the known-answer checks listed in each test.
// Build: make -C oracle (g++ -O2 -ffp-contract=off, no fast-math).
#include <cstring>
#include "ref3d.hpp"
#include "scenes.hpp"
#include "gpu3d.hpp"

using namespace refo;

extern "C" {

// ---- FP64 reference solver -----------------------------------------------------------------
void *oref_create(void) { return new Solver(); }
void oref_destroy(void *h) { delete (Solver *)h; }
void oref_build_scene(void *h, int id) { buildScene(*(Solver *)h, id); }
void oref_set_params(void *h, double dt, double gravity, int iterations, double alpha, double betaLin, double betaAng, double gamma) {
  Solver &s = *(Solver *)h;
  s.dt = dt; s.gravity = gravity; s.iterations = iterations; s.alpha = alpha; s.betaLin = betaLin; s.betaAng = betaAng; s.gamma = gamma;
}
int oref_add_body(void *h, const double *size, double density, double friction, const double *pos, const double *rot, const double *vel, const double *angvel) {
  Solver &s = *(Solver *)h;
  Rigid *b = s.addBody(v3(size[0], size[1], size[2]), density, friction, v3(pos[0], pos[1], pos[2]), vel ? v3(vel[0], vel[1], vel[2]) : V3());
  if (rot) b->positionAng = q4(rot[0], rot[1], rot[2], rot[3]);
  if (angvel) b->velocityAng = v3(angvel[0], angvel[1], angvel[2]);
  return b->index;
}
static Rigid *bodyAt(Solver &s, int i) { return i < 0 ? nullptr : s.bodies[i]; }
int oref_add_joint(void *h, int a, int b, const double *rA, const double *rB, double stiffLin, double stiffAng, double fracture) {
  Solver &s = *(Solver *)h;
  s.addJoint(bodyAt(s, a), s.bodies[b], v3(rA[0], rA[1], rA[2]), v3(rB[0], rB[1], rB[2]), stiffLin, stiffAng, fracture);
  return (int)s.forces.size() - 1;
}
int oref_add_spring(void *h, int a, int b, const double *rA, const double *rB, double stiffness, double rest) {
  Solver &s = *(Solver *)h;
  s.addSpring(s.bodies[a], s.bodies[b], v3(rA[0], rA[1], rA[2]), v3(rB[0], rB[1], rB[2]), stiffness, rest);
  return (int)s.forces.size() - 1;
}
int oref_add_ignore(void *h, int a, int b) {
  Solver &s = *(Solver *)h;
  s.addIgnore(s.bodies[a], s.bodies[b]);
  return (int)s.forces.size() - 1;
}
void oref_step(void *h, int n) { for (int i = 0; i < n; i++) ((Solver *)h)->step(); }
int oref_body_count(void *h) { return (int)((Solver *)h)->bodies.size(); }
int oref_force_count(void *h) { return (int)((Solver *)h)->forces.size(); }
/** Poses: per body position xyz, orientation xyzw (the golden fixture's layout). */
void oref_poses(void *h, double *out) {
  Solver &s = *(Solver *)h;
  for (size_t i = 0; i < s.bodies.size(); i++) {
    const Rigid &b = *s.bodies[i];
    double *o = out + 7 * i;
    o[0] = b.positionLin.x; o[1] = b.positionLin.y; o[2] = b.positionLin.z;
    o[3] = b.positionAng.x; o[4] = b.positionAng.y; o[5] = b.positionAng.z; o[6] = b.positionAng.w;
  }
}
/** Full body state as 40-float records (layout.ts:8-58), f64; the up-velocity history is prevVel.z. */
void oref_body_records(void *h, double *out) {
  Solver &s = *(Solver *)h;
  for (size_t i = 0; i < s.bodies.size(); i++) {
    const Rigid &b = *s.bodies[i];
    double *o = out + 40 * i;
    for (int k = 0; k < 40; k++) o[k] = 0;
    for (int k = 0; k < 3; k++) {
      o[k] = b.positionLin[k]; o[8 + k] = b.initialLin[k]; o[16 + k] = b.size[k]; o[20 + k] = b.moment[k];
      o[24 + k] = b.inertialLin[k]; o[32 + k] = b.velocityLin[k]; o[36 + k] = b.velocityAng[k];
    }
    for (int k = 0; k < 4; k++) { o[4 + k] = b.positionAng[k]; o[12 + k] = b.initialAng[k]; o[28 + k] = b.inertialAng[k]; }
    o[3] = b.friction; o[19] = b.mass; o[23] = b.radius; o[35] = b.prevVelocityLin.z;
  }
}
/** Forces in list order: kind, bodyA (-1 world), bodyB, contact count. */
void oref_forces(void *h, int *out) {
  Solver &s = *(Solver *)h;
  for (size_t i = 0; i < s.forces.size(); i++) {
    const Force &f = *s.forces[i];
    out[4 * i] = f.kind; out[4 * i + 1] = f.bodyA ? f.bodyA->index : -1; out[4 * i + 2] = f.bodyB->index; out[4 * i + 3] = (int)f.contacts.size();
  }
}
/**
 * Joint and spring records in list order (layout.ts:84-110 + info): 40 f64 each plus
 * (type 1 joint / 2 spring, a, b, 0). Infinity stays Infinity here; the caller maps it to BIG.
 */
int oref_joint_records(void *h, double *rec, int *info) {
  Solver &s = *(Solver *)h;
  int n = 0;
  for (const Force *fp : s.forces) {
    const Force &f = *fp;
    if (f.kind != F_JOINT && f.kind != F_SPRING) continue;
    double *o = rec + 40 * n;
    for (int k = 0; k < 40; k++) o[k] = 0;
    if (f.kind == F_JOINT) {
      for (int k = 0; k < 3; k++) {
        o[k] = f.penaltyLin[k]; o[4 + k] = f.penaltyAng[k]; o[8 + k] = f.lambdaLin[k]; o[12 + k] = f.lambdaAng[k];
        o[16 + k] = f.C0Lin[k]; o[20 + k] = f.C0Ang[k]; o[24 + k] = f.rA[k]; o[28 + k] = f.rB[k];
      }
      o[3] = f.broken ? 0 : f.stiffnessLin; o[7] = f.broken ? 0 : f.stiffnessAng;
      o[11] = f.fractureLinear ? -f.fracture : f.fracture; o[15] = f.torqueArm; o[23] = INF; o[31] = INF;
      o[35] = 1; o[39] = 1;
    } else {
      o[3] = f.stiffness;
      for (int k = 0; k < 3; k++) { o[24 + k] = f.rA[k]; o[28 + k] = f.rB[k]; }
      o[27] = f.rest; o[35] = 1; o[39] = 1;
    }
    info[4 * n] = f.kind == F_JOINT ? 1 : 2; info[4 * n + 1] = f.bodyA ? f.bodyA->index : -1; info[4 * n + 2] = f.bodyB->index; info[4 * n + 3] = 0;
    n++;
  }
  return n;
}
/**
 * Manifolds in list order as the GPU's previous-step records (solver.ts:1320-1359 seedFrom):
 * manifold (a, b, first, count; basis row 0 xyz, friction), 8 words; contacts 16 words
 * (rA, key|STICK; rB, C0.x; pen, C0.y; lam, C0.z) as f64 with the key as a number.
 */
int oref_manifold_records(void *h, double *man, double *con, int *counts) {
  Solver &s = *(Solver *)h;
  int m = 0, k = 0;
  for (const Force *fp : s.forces) {
    const Force &f = *fp;
    if (f.kind != F_MANIFOLD) continue;
    double *o = man + 8 * m;
    o[0] = f.bodyA->index; o[1] = f.bodyB->index; o[2] = k; o[3] = (double)f.contacts.size();
    o[4] = f.basis[0]; o[5] = f.basis[1]; o[6] = f.basis[2]; o[7] = f.friction;
    for (const Contact &c : f.contacts) {
      double *q = con + 16 * k;
      for (int r = 0; r < 3; r++) { q[r] = c.rA[r]; q[4 + r] = c.rB[r]; q[8 + r] = c.penalty[r]; q[12 + r] = c.lambda[r]; }
      q[3] = (double)(c.feature | (c.stick ? 0x80000000u : 0u));
      q[7] = c.C0.x; q[11] = c.C0.y; q[15] = c.C0.z;
      k++;
    }
    m++;
  }
  counts[0] = m; counts[1] = k;
  return m;
}
/** Box-box narrowphase of bodies a, b of the solver (collide.ts:312): basis 9, contacts (key, rA, rB). */
int oref_collide(void *h, int a, int b, double *basis, double *out) {
  Solver &s = *(Solver *)h;
  M3 B;
  std::vector<RawContact> c = collide(*s.bodies[a], *s.bodies[b], B);
  for (int i = 0; i < 9; i++) basis[i] = B[i];
  for (size_t i = 0; i < c.size(); i++) {
    out[7 * i] = c[i].feature;
    for (int r = 0; r < 3; r++) { out[7 * i + 1 + r] = c[i].rA[r]; out[7 * i + 4 + r] = c[i].rB[r]; }
  }
  return (int)c.size();
}
/** solve6 (math.ts:244) on a dense row-major 6x6 (lower triangle read). */
void oref_solve6(const double *A, const double *b, double *x) {
  M3 lin, ang, cr;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) { lin[i * 3 + j] = A[i * 6 + j]; ang[i * 3 + j] = A[(3 + i) * 6 + 3 + j]; cr[i * 3 + j] = A[(3 + i) * 6 + j]; }
  V3 xl, xa;
  solve6(lin, ang, cr, v3(b[0], b[1], b[2]), v3(b[3], b[4], b[5]), xl, xa);
  for (int i = 0; i < 3; i++) { x[i] = xl[i]; x[3 + i] = xa[i]; }
}

}  // extern "C"
