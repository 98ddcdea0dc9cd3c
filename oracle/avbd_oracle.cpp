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
#include <string>
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

// ---- the GPU step's semantics (gpu3d.hpp), f32 or f64 ----------------------------------------
#include "hull3d.hpp"
namespace {
struct ModelHandle { int prec; gpuo::Model<float> *f; gpuo::Model<double> *d; };
template <class T> gpuo::Model<T> *as(void *h);
template <> gpuo::Model<float> *as<float>(void *h) { return ((ModelHandle *)h)->f; }
template <> gpuo::Model<double> *as<double>(void *h) { return ((ModelHandle *)h)->d; }
#define DISPATCH(h, call) do { if (((ModelHandle *)(h))->prec == 32) { auto *M = as<float>(h); call; } else { auto *M = as<double>(h); call; } } while (0)

template <class T> void setBodies(gpuo::Model<T> *M, const double *r, int n) {
  M->bodies.resize(n);
  for (int i = 0; i < n; i++) {
    const double *o = r + 40 * i;
    gpuo::Body<T> &b = M->bodies[i];
    auto v = [&](int k) { gpuo::V3<T> x; x.x = (T)o[k]; x.y = (T)o[k + 1]; x.z = (T)o[k + 2]; return x; };
    auto q = [&](int k) { gpuo::Q4<T> x; x.x = (T)o[k]; x.y = (T)o[k + 1]; x.z = (T)o[k + 2]; x.w = (T)o[k + 3]; return x; };
    b.pos = v(0); b.friction = (T)o[3]; b.rot = q(4); b.ipos = v(8); b.lifecycle = (uint32_t)o[11]; b.irot = q(12);
    b.size = v(16); b.mass = (T)o[19]; b.moment = v(20); b.radius = (T)o[23]; b.inpos = v(24); b.stamp = (uint32_t)o[27];
    b.inrot = q(28); b.vel = v(32); b.prevUp = (T)o[35]; b.angvel = v(36);
    double sh = o[39];
    if (sh >= 3) { b.shape = gpuo::SHAPE_HULL; b.hull = (uint32_t)(sh - 3 + 0.5); } else { b.shape = (int)sh; b.hull = 0; }
  }
  M->filterGroups.assign(n, 0xffffffffu); M->filterMask.assign(n, 0xffffffffu);
  M->refPos.resize(n); M->refRot.resize(n); M->refSet.assign(n, false);
  M->colorA.assign(n, gpuo::NO_COLOR); M->colorB.assign(n, gpuo::NO_COLOR);
}
template <class T> void getBodies(gpuo::Model<T> *M, double *r) {
  for (size_t i = 0; i < M->bodies.size(); i++) {
    double *o = r + 40 * i;
    const gpuo::Body<T> &b = M->bodies[i];
    auto v = [&](int k, const gpuo::V3<T> &x) { o[k] = x.x; o[k + 1] = x.y; o[k + 2] = x.z; };
    auto q = [&](int k, const gpuo::Q4<T> &x) { o[k] = x.x; o[k + 1] = x.y; o[k + 2] = x.z; o[k + 3] = x.w; };
    v(0, b.pos); o[3] = b.friction; q(4, b.rot); v(8, b.ipos); o[11] = b.lifecycle; q(12, b.irot);
    v(16, b.size); o[19] = b.mass; v(20, b.moment); o[23] = b.radius; v(24, b.inpos); o[27] = b.stamp;
    q(28, b.inrot); v(32, b.vel); o[35] = b.prevUp; v(36, b.angvel);
    o[39] = b.shape == gpuo::SHAPE_HULL ? 3.0 + b.hull : (double)b.shape;
  }
}
template <class T> void setJoints(gpuo::Model<T> *M, const double *r, const int *info, int n) {
  M->joints.resize(n);
  M->noCollide.erase(std::remove_if(M->noCollide.begin(), M->noCollide.end(), [](const typename gpuo::Model<T>::NC &e) { return e.slot != gpuo::IGNORE_SLOT; }), M->noCollide.end());
  for (int j = 0; j < n; j++) {
    const double *o = r + 40 * j;
    gpuo::Joint<T> &k = M->joints[j];
    auto v = [&](int q) { gpuo::V3<T> x; x.x = (T)o[q]; x.y = (T)o[q + 1]; x.z = (T)o[q + 2]; return x; };
    auto fin = [](double x) { return (T)(x > gpuo::BIGD ? gpuo::BIGD : (x < -gpuo::BIGD ? -gpuo::BIGD : x)); };
    k.penLin = v(0); k.stiffLin = fin(o[3]); k.penAng = v(4); k.stiffAng = fin(o[7]); k.lamLin = v(8); k.fracture = fin(o[11]);
    k.lamAng = v(12); k.torqueArm = (T)o[15]; k.c0Lin = v(16); k.pad0 = 0; k.c0Ang = v(20); k.yield = fin(o[23]);
    k.rA = v(24); k.springRest = (T)o[27]; k.rB = v(28); k.breakBend = fin(o[31]);
    k.rest.x = (T)o[32]; k.rest.y = (T)o[33]; k.rest.z = (T)o[34]; k.rest.w = (T)o[35];
    k.restStart.x = (T)o[36]; k.restStart.y = (T)o[37]; k.restStart.z = (T)o[38]; k.restStart.w = (T)o[39];
    k.type = info[4 * j]; k.a = info[4 * j + 1]; k.b = info[4 * j + 2];
    if (k.a >= 0) M->noCollide.push_back({(uint32_t)std::max(k.a, k.b), (uint32_t)std::min(k.a, k.b), (uint32_t)j});
  }
  std::stable_sort(M->noCollide.begin(), M->noCollide.end(), [](const typename gpuo::Model<T>::NC &x, const typename gpuo::Model<T>::NC &y) { return x.hi < y.hi || (x.hi == y.hi && x.lo < y.lo); });
}
template <class T> void getJoints(gpuo::Model<T> *M, double *r) {
  for (size_t j = 0; j < M->joints.size(); j++) {
    double *o = r + 40 * j;
    const gpuo::Joint<T> &k = M->joints[j];
    auto v = [&](int q, const gpuo::V3<T> &x) { o[q] = x.x; o[q + 1] = x.y; o[q + 2] = x.z; };
    v(0, k.penLin); o[3] = k.stiffLin; v(4, k.penAng); o[7] = k.stiffAng; v(8, k.lamLin); o[11] = k.fracture;
    v(12, k.lamAng); o[15] = k.torqueArm; v(16, k.c0Lin); o[19] = 0; v(20, k.c0Ang); o[23] = k.yield;
    v(24, k.rA); o[27] = k.springRest; v(28, k.rB); o[31] = k.breakBend;
    o[32] = k.rest.x; o[33] = k.rest.y; o[34] = k.rest.z; o[35] = k.rest.w;
    o[36] = k.restStart.x; o[37] = k.restStart.y; o[38] = k.restStart.z; o[39] = k.restStart.w;
  }
}
template <class T> void setContacts(std::vector<gpuo::ManifoldT<T>> &ms, std::vector<gpuo::Contact<T>> &cs, const double *man, int nm, const double *con, int nc) {
  ms.resize(nm); cs.resize(nc);
  for (int m = 0; m < nm; m++) {
    const double *o = man + 8 * m;
    gpuo::ManifoldT<T> &x = ms[m];
    x.a = (uint32_t)o[0]; x.b = (uint32_t)o[1]; x.first = (uint32_t)o[2]; x.countStep = (uint32_t)o[3];
    x.n.x = (T)o[4]; x.n.y = (T)o[5]; x.n.z = (T)o[6]; x.friction = (T)o[7];
  }
  for (int c = 0; c < nc; c++) {
    const double *o = con + 16 * c;
    gpuo::Contact<T> &k = cs[c];
    for (int r = 0; r < 3; r++) { k.rA[r] = (T)o[r]; k.rB[r] = (T)o[4 + r]; k.pen[r] = (T)o[8 + r]; k.lam[r] = (T)o[12 + r]; }
    k.key = (uint32_t)o[3]; k.c0x = (T)o[7]; k.c0y = (T)o[11]; k.c0z = (T)o[15];
  }
}
template <class T> void getContacts(const std::vector<gpuo::ManifoldT<T>> &ms, const std::vector<gpuo::Contact<T>> &cs, double *man, double *con) {
  for (size_t m = 0; m < ms.size(); m++) {
    double *o = man + 8 * m;
    const gpuo::ManifoldT<T> &x = ms[m];
    o[0] = x.a; o[1] = x.b; o[2] = x.first; o[3] = x.countStep; o[4] = x.n.x; o[5] = x.n.y; o[6] = x.n.z; o[7] = x.friction;
  }
  for (size_t c = 0; c < cs.size(); c++) {
    double *o = con + 16 * c;
    const gpuo::Contact<T> &k = cs[c];
    for (int r = 0; r < 3; r++) { o[r] = k.rA[r]; o[4 + r] = k.rB[r]; o[8 + r] = k.pen[r]; o[12 + r] = k.lam[r]; }
    o[3] = k.key; o[7] = k.c0x; o[11] = k.c0y; o[15] = k.c0z;
  }
}
}  // namespace

// The hull SAT queries of bodies a and b (hull3d.hpp collidePoly's three queries), for tests that
// check them against a brute force: faSep, faFace, fbSep, fbFace, edgeSep, edgeValid, ea, eb.
template <class T> void hullQuery(gpuo::Model<T> *M, int a, int b, double *out) {
  gpuo::HullOps<T> H;
  H.w = M->hulls.data();
  auto A = H.polyOf(M->bodies[a]), B = H.polyOf(M->bodies[b]);
  T fa, fb;
  uint32_t faf, fbf;
  H.queryFaces(A, B, fa, faf);
  H.queryFaces(B, A, fb, fbf);
  auto e = H.queryEdges(A, B);
  out[0] = fa; out[1] = faf; out[2] = fb; out[3] = fbf; out[4] = e.sep; out[5] = e.valid ? 1 : 0; out[6] = e.ea; out[7] = e.eb;
}
// finishBody's floored LDLT on a given 6x6 system (row-major, symmetric) and right-hand side:
// x = A^-1 rhs, the pivots D and each pivot S before its floor, in f32 (prec 32) or f64.
template <class T> void solve6(const double *A, const double *r, double *x, double *D, double *S) {
  typename gpuo::Model<T>::Acc acc;
  const int I[6] = {0, 1, 1, 2, 2, 2}, J[6] = {0, 0, 1, 0, 1, 2};
  for (int e = 0; e < 6; e++) { acc.lin[e] = (T)A[6 * I[e] + J[e]]; acc.ang[e] = (T)A[6 * (3 + I[e]) + 3 + J[e]]; }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) acc.cr[3 * i + j] = (T)A[6 * (3 + i) + j];
  for (int i = 0; i < 3; i++) { acc.rLin[i] = (T)r[i]; acc.rAng[i] = (T)r[3 + i]; }
  T xx[6], dd[6], ss[6];
  gpuo::Model<T>::solve6Floor(acc, xx, dd, ss);
  for (int k = 0; k < 6; k++) { x[k] = xx[k]; D[k] = dd[k]; S[k] = ss[k]; }
}
extern "C" {
void *omodel_create(int prec) {
  ModelHandle *h = new ModelHandle{prec, nullptr, nullptr};
  if (prec == 32) h->f = new gpuo::Model<float>(); else h->d = new gpuo::Model<double>();
  return h;
}
void omodel_destroy(void *h) { ModelHandle *m = (ModelHandle *)h; delete m->f; delete m->d; delete m; }
void omodel_set_params(void *h, const gpuo::Params *p) { DISPATCH(h, M->p = *p); }
void omodel_get_params(void *h, gpuo::Params *p) { DISPATCH(h, *p = M->p); }
void omodel_set_bodies(void *h, const double *rec, int n) { DISPATCH(h, setBodies(M, rec, n)); }
void omodel_get_bodies(void *h, double *rec) { DISPATCH(h, getBodies(M, rec)); }
void omodel_set_joints(void *h, const double *rec, const int *info, int n) { DISPATCH(h, setJoints(M, rec, info, n)); }
void omodel_get_joints(void *h, double *rec) { DISPATCH(h, getJoints(M, rec)); }
void omodel_add_ignore(void *h, int a, int b) {
  DISPATCH(h, {
    M->noCollide.push_back({(uint32_t)std::max(a, b), (uint32_t)std::min(a, b), gpuo::IGNORE_SLOT});
    std::stable_sort(M->noCollide.begin(), M->noCollide.end(), [](const auto &x, const auto &y) { return x.hi < y.hi || (x.hi == y.hi && x.lo < y.lo); });
  });
}
void omodel_set_hulls(void *h, const uint32_t *words, int n) { DISPATCH(h, M->hulls.assign(words, words + n)); }
void omodel_set_filters(void *h, const uint32_t *groups, const uint32_t *mask) {
  DISPATCH(h, { for (size_t i = 0; i < M->bodies.size(); i++) { M->filterGroups[i] = groups[i]; M->filterMask[i] = mask[i]; } });
}
/** Last step's manifolds and contacts (the warm-start source), as seedFrom writes them. */
void omodel_set_prev(void *h, const double *man, int nm, const double *con, int nc) { DISPATCH(h, setContacts(M->manifolds, M->contacts, man, nm, con, nc)); }
void omodel_set_fixed_colors(void *h, const uint32_t *colors) {
  DISPATCH(h, { M->fixedColors = colors != nullptr; if (colors) M->fixed.assign(colors, colors + M->bodies.size()); });
}
void omodel_set_color_state(void *h, const uint32_t *state) { DISPATCH(h, M->colorA.assign(state, state + M->bodies.size())); }
void omodel_set_step_count(void *h, uint32_t s) { DISPATCH(h, M->stepCount = s); }
void omodel_set_threads(void *h, int n) { DISPATCH(h, M->threads = n < 1 ? 1 : n); }
void omodel_hull_query(void *h, int a, int b, double *out) { DISPATCH(h, hullQuery(M, a, b, out)); }
void omodel_solve6(int prec, const double *A, const double *r, double *x, double *D, double *S) {
  if (prec == 32) solve6<float>(A, r, x, D, S); else solve6<double>(A, r, x, D, S);
}
// the aux records (4 per body: gravity scale, linear and angular decay factors, damped 0/1)
void omodel_set_aux(void *h, const double *r) {
  DISPATCH(h, for (size_t i = 0; i < M->bodies.size(); i++) {
    const double *o = r + 4 * i;
    M->bodies[i].auxScale = o[0]; M->bodies[i].auxLin = o[1]; M->bodies[i].auxAng = o[2]; M->bodies[i].damped = o[3] != 0.0;
  });
}
// updateRefs' reference poses (8 per body: pos xyz, 0, rot xyzw); a zero rotation is "none yet"
void omodel_set_refs(void *h, const double *r) {
  DISPATCH(h, for (size_t i = 0; i < M->bodies.size(); i++) {
    const double *o = r + 8 * i;
    M->refPos[i].x = o[0]; M->refPos[i].y = o[1]; M->refPos[i].z = o[2];
    M->refRot[i].x = o[4]; M->refRot[i].y = o[5]; M->refRot[i].z = o[6]; M->refRot[i].w = o[7];
    M->refSet[i] = (o[4] * o[4] + o[5] * o[5] + o[6] * o[6] + o[7] * o[7]) != 0.0;
  });
}
void omodel_step(void *h, int n, int reuseManifolds) { DISPATCH(h, for (int i = 0; i < n; i++) M->step(reuseManifolds != 0)); }
/** The collision phase of one step only (pairs, manifolds, contacts), then adjacency and colour. */
void omodel_collide(void *h, int reuseManifolds) {
  DISPATCH(h, { M->stepCount++; M->p.step = M->stepCount; M->collision(reuseManifolds != 0); M->adjacency(); M->colour(); });
}
void omodel_solve(void *h) { DISPATCH(h, M->solve()); }
/** counts: pairs, manifolds, contacts, numColors, clashes, adjacency entries, colour cap used */
void omodel_counts(void *h, uint32_t *out) {
  DISPATCH(h, { out[0] = M->pairs.size(); out[1] = M->manifolds.size(); out[2] = M->contacts.size(); out[3] = M->numColors; out[4] = M->clashes; out[5] = M->adjList.size(); out[6] = M->colorCapUsed; });
}
void omodel_get_pairs(void *h, uint32_t *out) { DISPATCH(h, for (size_t i = 0; i < M->pairs.size(); i++) { out[2 * i] = M->pairs[i].first; out[2 * i + 1] = M->pairs[i].second; }); }
void omodel_get_contacts(void *h, double *man, double *con) { DISPATCH(h, getContacts(M->manifolds, M->contacts, man, con)); }
void omodel_get_colors(void *h, uint32_t *state) { DISPATCH(h, for (size_t i = 0; i < M->bodies.size(); i++) state[i] = M->colorA[i]); }
void omodel_get_adjacency(void *h, uint32_t *start, uint32_t *list) {
  DISPATCH(h, { for (size_t i = 0; i < M->adjStart.size(); i++) start[i] = M->adjStart[i]; for (size_t i = 0; i < M->adjList.size(); i++) list[i] = M->adjList[i]; });
}
}  // extern "C"

// ---- adjoint (adjoint3d.hpp) ----
#include "adjoint3d.hpp"
namespace {
struct TapeOwner { gpuo::Tape tape; };
std::string g_adjError;
}
extern "C" {
/** Starts recording a tape on the model (replacing any earlier one); stop with on = 0. */
void omodel_tape(void *h, int on) {
  ModelHandle *m = (ModelHandle *)h;
  static std::unordered_map<void *, TapeOwner *> owners;
  TapeOwner *&o = owners[h];
  if (on) { delete o; o = new TapeOwner(); }
  gpuo::Tape *t = on ? &o->tape : nullptr;
  if (m->prec == 32) m->f->tape = t; else m->d->tape = t;
  if (!on) return;
}
/** Frozen replay over the model's recorded tape, from the body records the tape started at. */
void *ofrozen_create(void *h, const double *records, int n) {
  ModelHandle *m = (ModelHandle *)h;
  const gpuo::Tape *t = m->prec == 32 ? m->f->tape : m->d->tape;
  const gpuo::Params &p = m->prec == 32 ? m->f->p : m->d->p;
  uint32_t joints = m->prec == 32 ? (uint32_t)m->f->joints.size() : (uint32_t)m->d->joints.size();
  if (!t) { g_adjError = "no tape recorded"; return nullptr; }
  try {
    return new gpuo::Frozen(p, *t, records, (size_t)n, joints);
  } catch (const std::exception &e) { g_adjError = e.what(); return nullptr; }
}
const char *ofrozen_error(void) { return g_adjError.c_str(); }
void ofrozen_destroy(void *f) { delete (gpuo::Frozen *)f; }
int ofrozen_steps(void *f) { return (int)((gpuo::Frozen *)f)->tape.steps.size(); }
/** run: state (16 per body) and friction (per body, or NULL) -> final state (16 per body). */
void ofrozen_run(void *f, const double *state, const double *mu, double *out) { ((gpuo::Frozen *)f)->run(state, mu, out); }
/** Linearise at the recorded model's own states (friction per body, or NULL). */
int ofrozen_load_recorded(void *f, const double *mu) {
  try { ((gpuo::Frozen *)f)->loadRecorded(mu); return 0; } catch (const std::exception &e) { g_adjError = e.what(); return -1; }
}
/** Reverse over the last run: adjoint of the final state -> adjoints of the initial state and friction. */
void ofrozen_backward(void *f, const double *dFinal, double *dInit, double *dMu) { ((gpuo::Frozen *)f)->backward(dFinal, dInit, dMu); }
/** Tape statistics: steps, contacts in total, floored pivots, sliding / clamped evaluations. */
void ofrozen_stats(void *f, double *out) {
  const gpuo::Tape &t = ((gpuo::Frozen *)f)->tape;
  double contacts = 0, floors = 0, slide = 0, clamp = 0;
  for (const gpuo::TapeStep &s : t.steps) {
    contacts += s.contacts.size();
    for (const auto &d : s.delta) for (double x : d) floors += x != 0;
    for (const auto &b : s.branch) for (uint8_t x : b) { slide += (x & 2) != 0; clamp += (x & 1) != 0; }
  }
  out[0] = t.steps.size(); out[1] = contacts; out[2] = floors; out[3] = slide; out[4] = clamp;
}
/** Central differences (h = 1e-6) of the quaternion helpers against their reverses at the given
 *  point; returns the largest absolute difference. */
double ofrozen_check_helpers(const double *q, const double *b, const double *w, const double *g) {
  typedef gpuo::Frozen F; typedef gpuo::Ops<double> O;
  gpuo::Q4<double> Q = O::q(q[0], q[1], q[2], q[3]), B = O::q(b[0], b[1], b[2], b[3]);
  gpuo::V3<double> W = O::v(w[0], w[1], w[2]), G3 = O::v(g[0], g[1], g[2]);
  gpuo::Q4<double> G4 = O::q(g[0], g[1], g[2], g[3]);
  double h = 1e-6, worst = 0;
  auto upd = [&](double a, double fd) { worst = std::max(worst, std::fabs(a - fd)); };
  // qadd
  { gpuo::Q4<double> qb = O::q(0, 0, 0, 0); gpuo::V3<double> wb = O::v(0, 0, 0); F::qaddRev(Q, W, G4, qb, wb);
    for (int k = 0; k < 4; k++) {
      gpuo::Q4<double> qp = Q, qm = Q; (&qp.x)[k] += h; (&qm.x)[k] -= h;
      upd((&qb.x)[k], (O::dot4(O::qadd(qp, W), G4) - O::dot4(O::qadd(qm, W), G4)) / (2 * h));
    }
    for (int k = 0; k < 3; k++) {
      gpuo::V3<double> wp = W, wm = W; wp[k] += h; wm[k] -= h;
      upd(wb[k], (O::dot4(O::qadd(Q, wp), G4) - O::dot4(O::qadd(Q, wm), G4)) / (2 * h));
    } }
  // qsub
  { gpuo::Q4<double> ab = O::q(0, 0, 0, 0), bb = ab; F::qsubRev(Q, B, G3, ab, bb);
    for (int k = 0; k < 4; k++) {
      gpuo::Q4<double> p1 = Q, m1 = Q; (&p1.x)[k] += h; (&m1.x)[k] -= h;
      upd((&ab.x)[k], (O::dot(O::qsub(p1, B), G3) - O::dot(O::qsub(m1, B), G3)) / (2 * h));
      gpuo::Q4<double> p2 = B, m2 = B; (&p2.x)[k] += h; (&m2.x)[k] -= h;
      upd((&bb.x)[k], (O::dot(O::qsub(Q, p2), G3) - O::dot(O::qsub(Q, m2), G3)) / (2 * h));
    } }
  // qrotate
  { gpuo::Q4<double> qb = O::q(0, 0, 0, 0); F::qrotateRev(Q, W, G3, qb);
    for (int k = 0; k < 4; k++) {
      gpuo::Q4<double> p1 = Q, m1 = Q; (&p1.x)[k] += h; (&m1.x)[k] -= h;
      upd((&qb.x)[k], (O::dot(O::qrotate(p1, W), G3) - O::dot(O::qrotate(m1, W), G3)) / (2 * h));
    } }
  return worst;
}
}  // extern "C"

// ---- SDF collision backend (sdf3d.hpp) ----
#include "sdf3d.hpp"
namespace {
/** A field as the tests describe it (the same field order as the library's avbd_sdf_field). */
struct OField {
  uint32_t kind, convex;
  float lo[3], hi[3], com[3], prm[8];
  uint32_t dims[3];
  float origin[3], cell;
  const float *grid, *seeds;
  uint32_t seedCount;
  const float *supports;
  uint32_t supportCount;
};
/** CollisionParams (ref/sdf/manifold.ts:13-28), the library's avbd_sdf_params field order. */
struct OPrm {
  uint32_t optimizer, cacheSize, randomSamples, warmStart, curvature, skipConvex;
  float epsilon, alpha, replacement, tol;
  uint32_t maxIt;
  float stepSize, margin, allowedPen;
  uint32_t signedSep, meshVertex, refreshDeeper, sgdStopOnly;
};
sdfo::Field toField(const OField &o) {
  sdfo::Field f;
  f.kind = o.kind; f.convex = o.convex;
  f.lo = sdfo::v3(o.lo[0], o.lo[1], o.lo[2]); f.hi = sdfo::v3(o.hi[0], o.hi[1], o.hi[2]);
  f.com = sdfo::v3(o.com[0], o.com[1], o.com[2]);
  for (int k = 0; k < 8; k++) f.prm[k] = o.prm[k];
  f.warpType = (uint32_t)o.prm[4];
  for (int k = 0; k < 3; k++) f.dims[k] = o.dims[k];
  f.origin = sdfo::v3(o.origin[0], o.origin[1], o.origin[2]); f.cell = o.cell;
  if (o.kind == sdfo::K_VOXEL && o.grid) f.grid.assign(o.grid, o.grid + 4ull * o.dims[0] * o.dims[1] * o.dims[2]);
  for (uint32_t i = 0; i < o.seedCount; i++) f.seeds.push_back(sdfo::v3(o.seeds[3 * i], o.seeds[3 * i + 1], o.seeds[3 * i + 2]));
  for (uint32_t i = 0; i < o.supportCount; i++) f.supports.push_back(sdfo::v3(o.supports[3 * i], o.supports[3 * i + 1], o.supports[3 * i + 2]));
  return f;
}
sdfo::Prm toPrm(const OPrm &o, uint32_t maxPoints) {
  sdfo::Prm p;
  p.cacheSize = o.cacheSize; p.randomSamples = o.randomSamples; p.epsilon = o.epsilon; p.alpha = o.alpha; p.tol = o.tol;
  p.margin = o.margin; p.replacement = o.replacement; p.signedSep = o.signedSep != 0; p.meshVertex = o.meshVertex != 0;
  p.maxIt = o.maxIt; p.warmStart = o.warmStart != 0; p.curvature = o.curvature != 0; p.skipConvex = o.skipConvex != 0;
  p.optimizer = o.optimizer; p.stepSize = o.optimizer == 2 ? o.stepSize : 0.0f; p.sgdStopOnly = o.optimizer == 2 && o.sgdStopOnly != 0;
  p.maxPoints = maxPoints;
  return p;
}
/** Result records in the library's avbd_sdf_read_results layout: identity (a, b, key, count),
 *  stats (queries, iterations, id counter, rolling), then maxPoints points of 28 words. */
void putResult(const sdfo::Result &r, uint32_t maxPoints, uint32_t *o) {
  std::memset(o, 0, 4ull * (8 + 28 * maxPoints));
  o[0] = r.a; o[1] = r.b; o[2] = r.key; o[3] = r.count; o[4] = r.queries; o[5] = r.iters; o[6] = r.idc; o[7] = r.rolling;
  for (uint32_t i = 0; i < r.count; i++) {
    const sdfo::Point &p = r.pts[i];
    float *f = (float *)(o + 8 + 28 * i);
    uint32_t *u = o + 8 + 28 * i;
    f[0] = p.x.x; f[1] = p.x.y; f[2] = p.x.z; f[3] = p.depth;
    f[4] = p.n.x; f[5] = p.n.y; f[6] = p.n.z; f[7] = p.sep;
    f[8] = p.rA.x; f[9] = p.rA.y; f[10] = p.rA.z; f[11] = p.soft;
    f[12] = p.rB.x; f[13] = p.rB.y; f[14] = p.rB.z; f[15] = p.rank;
    u[16] = p.active; u[17] = p.cached; u[18] = p.id; u[19] = p.exported;
    for (int k = 0; k < 4; k++) { f[20 + k] = p.eA[k]; f[24 + k] = p.eB[k]; }
  }
}
void getPoint(const float *f, sdfo::Point &p) {
  const uint32_t *u = (const uint32_t *)f;
  p.x = sdfo::v3(f[0], f[1], f[2]); p.depth = f[3];
  p.n = sdfo::v3(f[4], f[5], f[6]); p.sep = f[7];
  p.rA = sdfo::v3(f[8], f[9], f[10]); p.soft = f[11];
  p.rB = sdfo::v3(f[12], f[13], f[14]); p.rank = f[15];
  p.active = u[16]; p.cached = u[17]; p.id = u[18]; p.exported = u[19];
  for (int k = 0; k < 4; k++) { p.eA[k] = f[20 + k]; p.eB[k] = f[24 + k]; }
}
/** A standalone body for the pinning harness: the CPU reference's SdfBody (body.ts), its
 *  orientation normalised, its bounds the field's corners through the pose. */
sdfo::SBody cpuBody(const sdfo::Field *f, const float *pose) {
  sdfo::SBody b;
  b.f = f;
  b.pos = sdfo::v3(pose[0], pose[1], pose[2]);
  float qx = pose[3], qy = pose[4], qz = pose[5], qw = pose[6];
  float norm = std::sqrt(((qx * qx + qy * qy) + qz * qz) + qw * qw);
  b.rot = sdfo::q4(qx / norm, qy / norm, qz / norm, qw / norm);
  b.scale = pose[7];
  b.fixed = pose[8] != 0.0f;
  b.vel = sdfo::v3(pose[9], pose[10], pose[11]);
  b.omega = sdfo::v3(pose[12], pose[13], pose[14]);
  b.anchor = sdfo::q4(0, 0, 0, 1);
  sdfo::bodyBounds(b);
  return b;
}
struct BridgeHandle { gpuo::Model<float> *model; sdfo::Bridge bridge; };
}  // namespace
extern "C" {
/** Installs the SDF bridge on an f32 model (setCollisionBackend): every body gets a field
 *  (body_field), a scale and a field rotation (4 per body, xyzw). */
void *osdf_create(void *h, const OField *fields, uint32_t nf, const uint32_t *bodyField, const float *bodyScale, const float *bodyRot,
                  const OPrm *prm, int automatic, uint32_t maxPairs, uint32_t maxPoints, float cellSize) {
  ModelHandle *m = (ModelHandle *)h;
  if (m->prec != 32) return nullptr;
  BridgeHandle *b = new BridgeHandle();
  b->model = m->f;
  for (uint32_t i = 0; i < nf; i++) b->bridge.fields.push_back(toField(fields[i]));
  size_t n = m->f->bodies.size();
  for (size_t i = 0; i < n; i++) {
    sdfo::BodyConfig c;
    c.field = bodyField[i];
    c.scale = bodyScale ? bodyScale[i] : 1.0f;
    c.rot = bodyRot ? sdfo::q4(bodyRot[4 * i], bodyRot[4 * i + 1], bodyRot[4 * i + 2], bodyRot[4 * i + 3]) : sdfo::q4(0, 0, 0, 1);
    b->bridge.cfg.push_back(c);
  }
  b->bridge.prm = toPrm(*prm, maxPoints);
  b->bridge.allowedPen = prm->allowedPen;
  b->bridge.refreshDeeper = prm->refreshDeeper != 0;
  b->bridge.automatic = automatic != 0;
  b->bridge.maxPairs = maxPairs;
  b->bridge.cellSize = cellSize;
  sdfo::Bridge *br = &b->bridge;
  m->f->collisionHook = [br](gpuo::Model<float> &M) { br->encode(M); };
  m->f->manifolds.clear();  // setCollisionBackend: no later step warm-starts from the last contacts
  m->f->contacts.clear();
  return b;
}
void osdf_destroy(void *b) {
  BridgeHandle *x = (BridgeHandle *)b;
  x->model->collisionHook = nullptr;
  x->model->manifolds.clear();
  x->model->contacts.clear();
  delete x;
}
void osdf_set_pairs(void *b, const uint32_t *pairs, uint32_t n) {
  BridgeHandle *x = (BridgeHandle *)b;
  x->bridge.pairs.clear();
  for (uint32_t i = 0; i < n; i++) x->bridge.pairs.push_back(sdfo::PairSpec{pairs[3 * i], pairs[3 * i + 1], pairs[3 * i + 2]});
}
void osdf_set_params(void *b, const OPrm *prm) {
  BridgeHandle *x = (BridgeHandle *)b;
  uint32_t mp = x->bridge.prm.maxPoints;
  x->bridge.prm = toPrm(*prm, mp);
  x->bridge.allowedPen = prm->allowedPen;
  x->bridge.refreshDeeper = prm->refreshDeeper != 0;
}
/** The last step's results sorted by (a, b) into out (room for cap records); returns how many,
 *  and the candidates past the capacity in *overflow. */
uint32_t osdf_results(void *b, uint32_t *out, uint32_t cap, uint32_t *overflow) {
  BridgeHandle *x = (BridgeHandle *)b;
  uint32_t mp = x->bridge.prm.maxPoints, n = (uint32_t)x->bridge.last.size();
  for (uint32_t i = 0; i < n && i < cap; i++) putResult(x->bridge.last[i], mp, out + (size_t)i * (8 + 28 * mp));
  if (overflow) *overflow = x->bridge.overflowPairs;
  return n;
}
/** The prepared bodies of the last step (28 floats each: the SdfBody record of wgsl-sdf-bridge.ts). */
void osdf_bodies(void *b, float *out) {
  BridgeHandle *x = (BridgeHandle *)b;
  for (size_t i = 0; i < x->bridge.sb.size(); i++) {
    const sdfo::SBody &s = x->bridge.sb[i];
    float *o = out + 28 * i;
    float v[28] = {s.pos.x, s.pos.y, s.pos.z, s.scale, s.rot.x, s.rot.y, s.rot.z, s.rot.w, s.vel.x, s.vel.y, s.vel.z, s.fixed ? 1.0f : 0.0f,
                   s.omega.x, s.omega.y, s.omega.z, 0, s.lo.x, s.lo.y, s.lo.z, 0, s.hi.x, s.hi.y, s.hi.z, 0, s.anchor.x, s.anchor.y, s.anchor.z, s.anchor.w};
    std::memcpy(o, v, sizeof v);
  }
}

// pinning harness: the pieces of Algorithm 1 on their own
/** phi, gradient xyz per point of a field. */
void osdf_sample(const OField *fo, const float *p, uint32_t n, float *out) {
  sdfo::Field f = toField(*fo);
  for (uint32_t i = 0; i < n; i++) {
    sdfo::Sample s = sdfo::localField(f, sdfo::v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]));
    out[4 * i] = s.phi; out[4 * i + 1] = s.g.x; out[4 * i + 2] = s.g.y; out[4 * i + 3] = s.g.z;
  }
}
/** which: 0 sin, 1 cos, 2 atan2(x[2i], x[2i+1]). */
void osdf_trig(int which, const float *x, uint32_t n, float *out) {
  for (uint32_t i = 0; i < n; i++)
    out[i] = which == 0 ? sdfo::ssin(x[i]) : (which == 1 ? sdfo::scos(x[i]) : sdfo::satan2(x[2 * i], x[2 * i + 1]));
}
/** One optimizer run between body A (pose 15 floats: pos, rot, scale, static, vel, omega) and body B
 *  over the domain lo/hi, with repulsion points rep: kind 0 Nelder-Mead (from seed, initial step
 *  `step`), 1 the ellipsoid method (E row-major 9 floats in, final out), 2 gradient descent (PCG state
 *  in and out), 3 the objective and its gradient at the seed. out: x xyz, value/gradient,
 *  iterations, queries. */
void osdf_optimize(int kind, const OField *fa, const float *poseA, const OField *fb, const float *poseB, const float *domain,
                   const float *rep, uint32_t nrep, const float *seed, const float *step, const OPrm *prm, float *E, uint64_t *pcg,
                   float *out, uint32_t *counts) {
  sdfo::Field FA = toField(*fa), FB = toField(*fb);
  sdfo::SBody A = cpuBody(&FA, poseA), B = cpuBody(&FB, poseB);
  sdfo::Prm P = toPrm(*prm, 20);
  sdfo::PcgRng rng;
  if (pcg) rng.state = *pcg;
  sdfo::PairUpdate<sdfo::PcgRng, true> u(P, A, B, rng);
  u.lo = sdfo::v3(domain[0], domain[1], domain[2]);
  u.hi = sdfo::v3(domain[3], domain[4], domain[5]);
  for (uint32_t i = 0; i < nrep; i++) { u.work.pts[i].x = sdfo::v3(rep[3 * i], rep[3 * i + 1], rep[3 * i + 2]); }
  u.work.count = nrep;
  sdfo::V s = sdfo::v3(seed[0], seed[1], seed[2]), x{};
  if (kind == 0) x = u.nelderMead(s, sdfo::v3(step[0], step[1], step[2]));
  else if (kind == 1) {
    float M[3][3];
    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) M[r][c] = E[3 * r + c];
    x = u.ellipsoidMethod(s, M);
    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) E[3 * r + c] = u.E[r][c];
  } else if (kind == 2) x = u.gradientDescent(s);
  else {
    sdfo::V g; float value;
    u.objectiveGradient(s, g, value);
    out[0] = value; out[1] = g.x; out[2] = g.y; out[3] = g.z;
    counts[0] = 0; counts[1] = u.work.queries;
    return;
  }
  if (pcg) *pcg = rng.state;
  out[0] = x.x; out[1] = x.y; out[2] = x.z;
  out[3] = kind == 0 ? u.objective(x) : 0.0f;
  counts[0] = u.work.iters;
  counts[1] = u.work.queries - (kind == 0 ? 1u : 0u);
}
/** The CPU reference's manifold update (manifold.ts update) on one pair, with its PCG stream:
 *  cache (ncache points of 28 floats, the result point layout) in, the updated points out (room
 *  for 20); stats: pairs, contacts, queries, iterations; returns the point count. */
uint32_t osdf_manifold_cpu(const OField *fa, const float *poseA, const OField *fb, const float *poseB, const OPrm *prm, float dt,
                           const float *cache, uint32_t ncache, uint64_t *pcg, float *out, uint32_t *stats) {
  sdfo::Field FA = toField(*fa), FB = toField(*fb);
  sdfo::SBody A = cpuBody(&FA, poseA), B = cpuBody(&FB, poseB);
  sdfo::Prm P = toPrm(*prm, 20);
  P.dt = dt;
  sdfo::PcgRng rng;
  rng.state = *pcg;
  sdfo::Result prev;
  prev.pts.resize(20);
  prev.count = ncache;
  for (uint32_t i = 0; i < ncache; i++) getPoint(cache + 28 * i, prev.pts[i]);
  sdfo::PairUpdate<sdfo::PcgRng, true> u(P, A, B, rng);
  u.run(0, 0, 0, P.warmStart && ncache ? &prev : nullptr, 0);
  bool valid = !(u.lo.x > u.hi.x || u.lo.y > u.hi.y || u.lo.z > u.hi.z);
  *pcg = rng.state;
  uint32_t active = 0;
  uint32_t buf[8 + 28 * 20];
  putResult(u.work, 20, buf);
  std::memcpy(out, buf + 8, 4ull * 28 * u.work.count);
  for (uint32_t i = 0; i < u.work.count; i++) active += u.work.pts[i].active;
  stats[0] = valid ? 1u : 0u; stats[1] = active; stats[2] = u.work.queries; stats[3] = u.work.iters;
  return u.work.count;
}
}  // extern "C"
