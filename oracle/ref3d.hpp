// TEST INFRASTRUCTURE ONLY — the oracle. Nothing in the product links, imports or runs it;
// tests, the smoke check and the bench's CPU-baseline leg use it as the checker.
//
// FP64 restatement of the reference's CPU AVBD solver (src/avbd3d/ref/*.ts, itself a port of
// avbd-demo3d), operation for operation, so the golden trajectories in
// tests/golden/oracle3d (from the upstream C++ in f64 without FMA contraction) are reproduced
// to round-off. Compile with -ffp-contract=off. Intrusive newest-first lists of the C++ demo
// are arrays in creation order walked in reverse, as in the TS port.
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>
#include <algorithm>

namespace refo {

typedef double R;
static const R INF = std::numeric_limits<R>::infinity();

// --- math.ts ----------------------------------------------------------------------------------
struct V3 { R x = 0, y = 0, z = 0; R &operator[](int i) { return (&x)[i]; } R operator[](int i) const { return (&x)[i]; } };
struct Q4 { R x = 0, y = 0, z = 0, w = 1; R &operator[](int i) { return (&x)[i]; } R operator[](int i) const { return (&x)[i]; } };
struct M3 { R m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}; R &operator[](int i) { return m[i]; } R operator[](int i) const { return m[i]; } };

static inline V3 v3(R x, R y, R z) { V3 v; v.x = x; v.y = y; v.z = z; return v; }
static inline Q4 q4(R x, R y, R z, R w) { Q4 q; q.x = x; q.y = y; q.z = z; q.w = w; return q; }
static inline R rmin(R a, R b) { return a < b ? a : b; }
static inline R rmax(R a, R b) { return a > b ? a : b; }
static inline R clampr(R x, R a, R b) { return rmax(a, rmin(b, x)); }
static inline R sgn(R x) { return x < 0 ? -1 : x > 0 ? 1 : 0; }
// math.ts:22-45
static inline R dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline R lengthSq(const V3 &v) { return dot(v, v); }
static inline R length(const V3 &v) { return std::sqrt(lengthSq(v)); }
static inline V3 add(const V3 &a, const V3 &b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 sub(const V3 &a, const V3 &b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 scale(const V3 &a, R s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 div3(const V3 &a, R s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline V3 neg(const V3 &a) { return v3(-a.x, -a.y, -a.z); }
static inline V3 abs3(const V3 &a) { return v3(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)); }
static inline V3 addScaled(const V3 &a, const V3 &b, R s) { return v3(a.x + b.x * s, a.y + b.y * s, a.z + b.z * s); }
static inline V3 normalize3(const V3 &v) { return div3(v, length(v)); }
static inline V3 cross(const V3 &a, const V3 &b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
// math.ts:57-113
static inline Q4 qmul(const Q4 &a, const Q4 &b) {
  return q4(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
static inline Q4 conj(const Q4 &q) { return q4(-q.x, -q.y, -q.z, q.w); }
static inline R qlengthSq(const Q4 &q) { return (q.x * q.x + q.y * q.y + q.z * q.z) + q.w * q.w; }
static inline Q4 qinverse(const Q4 &q) { R l = qlengthSq(q); return q4(-q.x / l, -q.y / l, -q.z / l, q.w / l); }
static inline Q4 qnormalize(const Q4 &q) { R l = std::sqrt(qlengthSq(q)); return q4(q.x / l, q.y / l, q.z / l, q.w / l); }
static inline V3 qsub(const Q4 &a, const Q4 &b) { Q4 s = qmul(a, qinverse(b)); return v3(s.x * 2, s.y * 2, s.z * 2); }
static inline Q4 qaddv(const Q4 &a, const V3 &b) {
  Q4 s = qmul(q4(b.x, b.y, b.z, 0), a);
  return qnormalize(q4(a.x + s.x * 0.5, a.y + s.y * 0.5, a.z + s.z * 0.5, a.w + s.w * 0.5));
}
static inline V3 rotate(const Q4 &q, const V3 &v) {
  R ux = q.x, uy = q.y, uz = q.z, w = q.w;
  R tx = (uy * v.z - uz * v.y) * 2, ty = (uz * v.x - ux * v.z) * 2, tz = (ux * v.y - uy * v.x) * 2;
  return v3(v.x + tx * w + (uy * tz - uz * ty), v.y + ty * w + (uz * tx - ux * tz), v.z + tz * w + (ux * ty - uy * tx));
}
static inline V3 transform(const V3 &p, const Q4 &q, const V3 &v) { V3 t = rotate(q, v); return add(t, p); }
static inline V3 rotateInv(const Q4 &q, const V3 &v) { return rotate(conj(q), v); }
// math.ts:142-238 (row-major)
static inline M3 diagonal(R a, R b, R c) { M3 o; o[0] = a; o[4] = b; o[8] = c; return o; }
static inline M3 skew(const V3 &r) { M3 o; o[1] = -r.z; o[2] = r.y; o[3] = r.z; o[5] = -r.x; o[6] = -r.y; o[7] = r.x; return o; }
static inline M3 transpose(const M3 &a) { M3 o; o[0] = a[0]; o[1] = a[3]; o[2] = a[6]; o[3] = a[1]; o[4] = a[4]; o[5] = a[7]; o[6] = a[2]; o[7] = a[5]; o[8] = a[8]; return o; }
static inline M3 mul(const M3 &a, const M3 &b) {
  M3 o;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) o[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
  return o;
}
static inline V3 mulv(const M3 &a, const V3 &v) {
  return v3(a[0] * v.x + a[1] * v.y + a[2] * v.z, a[3] * v.x + a[4] * v.y + a[5] * v.z, a[6] * v.x + a[7] * v.y + a[8] * v.z);
}
static inline void addm(M3 &o, const M3 &a) { for (int i = 0; i < 9; i++) o[i] += a[i]; }
static inline M3 scalem(const M3 &a, R s) { M3 o; for (int i = 0; i < 9; i++) o[i] = a[i] * s; return o; }
static inline M3 negm(const M3 &a) { M3 o; for (int i = 0; i < 9; i++) o[i] = -a[i]; return o; }
static inline M3 outer(const V3 &a, const V3 &b) { M3 o; for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) o[i * 3 + j] = b[j] * a[i]; return o; }
static inline M3 diagonalize(const M3 &m) {
  return diagonal(std::sqrt(m[0] * m[0] + m[3] * m[3] + m[6] * m[6]), std::sqrt(m[1] * m[1] + m[4] * m[4] + m[7] * m[7]),
                  std::sqrt(m[2] * m[2] + m[5] * m[5] + m[8] * m[8]));
}
static inline M3 orthonormal(const V3 &n) {
  V3 t1 = std::fabs(n.x) > std::fabs(n.z) ? v3(-n.y, n.x, 0) : v3(0, -n.z, n.y);
  t1 = normalize3(t1);
  V3 t2 = cross(n, t1);
  M3 o;
  o[0] = n.x; o[1] = n.y; o[2] = n.z; o[3] = t1.x; o[4] = t1.y; o[5] = t1.z; o[6] = t2.x; o[7] = t2.y; o[8] = t2.z;
  return o;
}
// math.ts:244-302: unpivoted LDLᵀ of [lin crossᵀ; cross ang], lower triangle only
static inline void solve6(const M3 &aLin, const M3 &aAng, const M3 &aCross, const V3 &bLin, const V3 &bAng, V3 &xLin, V3 &xAng) {
  R A11 = aLin[0], A21 = aLin[3], A22 = aLin[4], A31 = aLin[6], A32 = aLin[7], A33 = aLin[8];
  R A41 = aCross[0], A42 = aCross[1], A43 = aCross[2], A44 = aAng[0];
  R A51 = aCross[3], A52 = aCross[4], A53 = aCross[5], A54 = aAng[3], A55 = aAng[4];
  R A61 = aCross[6], A62 = aCross[7], A63 = aCross[8], A64 = aAng[6], A65 = aAng[7], A66 = aAng[8];
  R L21 = A21 / A11, L31 = A31 / A11, L41 = A41 / A11, L51 = A51 / A11, L61 = A61 / A11, D1 = A11;
  R D2 = A22 - L21 * L21 * D1;
  R L32 = (A32 - L21 * L31 * D1) / D2, L42 = (A42 - L21 * L41 * D1) / D2, L52 = (A52 - L21 * L51 * D1) / D2, L62 = (A62 - L21 * L61 * D1) / D2;
  R D3 = A33 - (L31 * L31 * D1 + L32 * L32 * D2);
  R L43 = (A43 - L31 * L41 * D1 - L32 * L42 * D2) / D3, L53 = (A53 - L31 * L51 * D1 - L32 * L52 * D2) / D3, L63 = (A63 - L31 * L61 * D1 - L32 * L62 * D2) / D3;
  R D4 = A44 - (L41 * L41 * D1 + L42 * L42 * D2 + L43 * L43 * D3);
  R L54 = (A54 - L41 * L51 * D1 - L42 * L52 * D2 - L43 * L53 * D3) / D4, L64 = (A64 - L41 * L61 * D1 - L42 * L62 * D2 - L43 * L63 * D3) / D4;
  R D5 = A55 - (L51 * L51 * D1 + L52 * L52 * D2 + L53 * L53 * D3 + L54 * L54 * D4);
  R L65 = (A65 - L51 * L61 * D1 - L52 * L62 * D2 - L53 * L63 * D3 - L54 * L64 * D4) / D5;
  R D6 = A66 - (L61 * L61 * D1 + L62 * L62 * D2 + L63 * L63 * D3 + L64 * L64 * D4 + L65 * L65 * D5);
  R y1 = bLin.x, y2 = bLin.y - L21 * y1, y3 = bLin.z - L31 * y1 - L32 * y2;
  R y4 = bAng.x - L41 * y1 - L42 * y2 - L43 * y3, y5 = bAng.y - L51 * y1 - L52 * y2 - L53 * y3 - L54 * y4;
  R y6 = bAng.z - L61 * y1 - L62 * y2 - L63 * y3 - L64 * y4 - L65 * y5;
  R z1 = y1 / D1, z2 = y2 / D2, z3 = y3 / D3, z4 = y4 / D4, z5 = y5 / D5, z6 = y6 / D6;
  xAng.z = z6;
  xAng.y = z5 - L65 * xAng.z;
  xAng.x = z4 - L54 * xAng.y - L64 * xAng.z;
  xLin.z = z3 - L43 * xAng.x - L53 * xAng.y - L63 * xAng.z;
  xLin.y = z2 - L32 * xLin.z - L42 * xAng.x - L52 * xAng.y - L62 * xAng.z;
  xLin.x = z1 - L21 * xLin.y - L31 * xLin.z - L41 * xAng.x - L51 * xAng.y - L61 * xAng.z;
}

// --- body.ts ----------------------------------------------------------------------------------
static const R PENALTY_MIN = 1.0, PENALTY_MAX = 10000000000.0, COLLISION_MARGIN = 0.01, STICK_THRESH = 0.00001;

struct Force;
struct Rigid {
  std::vector<Force *> forces;
  V3 positionLin; Q4 positionAng; V3 initialLin; Q4 initialAng; V3 inertialLin; Q4 inertialAng;
  V3 velocityLin, velocityAng, prevVelocityLin, size, moment;
  R mass = 0, friction = 0, radius = 0;
  int index = 0;
  bool constrainedTo(const Rigid *other) const;
};

struct BodySystem { M3 lhsLin, lhsAng, lhsCross; V3 rhsLin, rhsAng; };

enum ForceKind { F_MANIFOLD = 0, F_JOINT = 1, F_SPRING = 2, F_IGNORE = 3 };

struct Contact { uint32_t feature = 0; V3 rA, rB, C0, penalty, lambda; bool stick = false; };

struct Force {
  int kind;
  Rigid *bodyA; // null: attached to the world
  Rigid *bodyB;
  // Manifold (manifold.ts:81)
  std::vector<Contact> contacts; M3 basis; R friction = 0;
  // Joint (forces.ts:82)
  V3 rA, rB, C0Lin, C0Ang, penaltyLin, penaltyAng, lambdaLin, lambdaAng;
  R stiffnessLin = INF, stiffnessAng = 0, fracture = INF, torqueArm = 0;
  bool broken = false, fractureLinear = false;
  // Spring (forces.ts:325)
  R rest = -1, stiffness = 0;
  Force(int k, Rigid *a, Rigid *b) : kind(k), bodyA(a), bodyB(b) {}
};

inline bool Rigid::constrainedTo(const Rigid *other) const {
  for (const Force *f : forces)
    if ((f->bodyA == this && f->bodyB == other) || (f->bodyA == other && f->bodyB == this)) return true;
  return false;
}

// --- collide.ts ---------------------------------------------------------------------------------
struct RawContact { uint32_t feature; V3 rA, rB; };
static const int MAX_CONTACTS = 8, MAX_POLY_VERTS = 16;
static const R SAT_AXIS_EPSILON = 1.0e-6, PLANE_EPSILON = 1.0e-5, CONTACT_MERGE_DIST_SQ = 1.0e-6, FLT_MAXD = 3.4028234663852886e38;
enum { AXIS_FACE_A = 0, AXIS_FACE_B = 1, AXIS_EDGE = 2 };
struct OBB { V3 center, half, axis[3]; };
struct SatAxis { int type, indexA, indexB; R separation; V3 normalAB; bool valid; };

static inline OBB makeOBB(const Rigid &b) {
  OBB o;
  o.center = b.positionLin;
  o.half = scale(b.size, 0.5);
  o.axis[0] = rotate(b.positionAng, v3(1, 0, 0));
  o.axis[1] = rotate(b.positionAng, v3(0, 1, 0));
  o.axis[2] = rotate(b.positionAng, v3(0, 0, 1));
  return o;
}
static inline R absDot(const V3 &a, const V3 &b) { return std::fabs(dot(a, b)); }
static inline V3 supportPoint(const OBB &box, const V3 &dir) {
  R sx = dot(dir, box.axis[0]) >= 0 ? 1 : -1, sy = dot(dir, box.axis[1]) >= 0 ? 1 : -1, sz = dot(dir, box.axis[2]) >= 0 ? 1 : -1;
  V3 p = addScaled(box.center, box.axis[0], box.half.x * sx);
  p = addScaled(p, box.axis[1], box.half.y * sy);
  return addScaled(p, box.axis[2], box.half.z * sz);
}
static inline void faceAxes(const OBB &box, int k, V3 &u, V3 &v, R &eu, R &ev) {
  if (k == 0) { u = box.axis[1]; v = box.axis[2]; eu = box.half.y; ev = box.half.z; }
  else if (k == 1) { u = box.axis[0]; v = box.axis[2]; eu = box.half.x; ev = box.half.z; }
  else { u = box.axis[0]; v = box.axis[1]; eu = box.half.x; ev = box.half.y; }
}
static std::vector<V3> clipPolygonAgainstPlane(const std::vector<V3> &in, const V3 &n, R offset) {
  std::vector<V3> out;
  if (in.empty()) return out;
  V3 a = in.back();
  R da = dot(n, a) - offset;
  for (const V3 &b : in) {
    R db = dot(n, b) - offset;
    bool aIn = da <= PLANE_EPSILON, bIn = db <= PLANE_EPSILON;
    if (aIn != bIn) {
      R t = 0, denom = da - db;
      if (std::fabs(denom) > SAT_AXIS_EPSILON) t = clampr(da / denom, 0, 1);
      if ((int)out.size() < MAX_POLY_VERTS) out.push_back(addScaled(a, sub(b, a), t));
    }
    if (bIn && (int)out.size() < MAX_POLY_VERTS) out.push_back(b);
    a = b;
    da = db;
  }
  return out;
}
struct ContactBuilder {
  std::vector<RawContact> contacts; std::vector<V3> midpoints; const Rigid *A, *B;
  void add(const V3 &xA, const V3 &xB, uint32_t feature) {
    V3 mid = scale(refo::add(xA, xB), 0.5);
    for (const V3 &m : midpoints) if (lengthSq(sub(mid, m)) < CONTACT_MERGE_DIST_SQ) return;
    if ((int)contacts.size() >= MAX_CONTACTS) return;
    contacts.push_back({feature, rotateInv(A->positionAng, sub(xA, A->positionLin)), rotateInv(B->positionAng, sub(xB, B->positionLin))});
    midpoints.push_back(mid);
  }
};
static bool testAxis(const OBB &A, const OBB &B, const V3 &delta, const V3 &axis, int type, int ia, int ib, SatAxis &best) {
  R lenSq = lengthSq(axis);
  if (lenSq < SAT_AXIS_EPSILON) return true;
  R invLen = 1 / std::sqrt(lenSq);
  V3 n = scale(axis, invLen);
  if (dot(n, delta) < 0) n = neg(n);
  R distance = std::fabs(dot(delta, n));
  R rA = A.half.x * absDot(n, A.axis[0]) + A.half.y * absDot(n, A.axis[1]) + A.half.z * absDot(n, A.axis[2]);
  R rB = B.half.x * absDot(n, B.axis[0]) + B.half.y * absDot(n, B.axis[1]) + B.half.z * absDot(n, B.axis[2]);
  R separation = distance - (rA + rB);
  if (separation > 0) return false;
  if (!best.valid || separation > best.separation) { best.valid = true; best.type = type; best.indexA = ia; best.indexB = ib; best.separation = separation; best.normalAB = n; }
  return true;
}
static void supportEdge(const OBB &box, int k, const V3 &dir, V3 &e0, V3 &e1) {
  int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
  R s1 = dot(dir, box.axis[k1]) >= 0 ? 1 : -1, s2 = dot(dir, box.axis[k2]) >= 0 ? 1 : -1;
  V3 c = addScaled(box.center, box.axis[k1], box.half[k1] * s1);
  c = addScaled(c, box.axis[k2], box.half[k2] * s2);
  V3 e = scale(box.axis[k], box.half[k]);
  e0 = sub(c, e);
  e1 = add(c, e);
}
static void closestPointsOnSegments(const V3 &p0, const V3 &p1, const V3 &q0, const V3 &q1, V3 &xa, V3 &xb) {
  V3 d1 = sub(p1, p0), d2 = sub(q1, q0), r = sub(p0, q0);
  R a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r), s = 0, t = 0;
  if (a <= SAT_AXIS_EPSILON && e <= SAT_AXIS_EPSILON) { xa = p0; xb = q0; return; }
  if (a <= SAT_AXIS_EPSILON) t = clampr(f / e, 0, 1);
  else {
    R c = dot(d1, r);
    if (e <= SAT_AXIS_EPSILON) s = clampr(-c / a, 0, 1);
    else {
      R b = dot(d1, d2), denom = a * e - b * b;
      if (std::fabs(denom) > SAT_AXIS_EPSILON) s = clampr((b * f - c * e) / denom, 0, 1);
      t = (b * s + f) / e;
      if (t < 0) { t = 0; s = clampr(-c / a, 0, 1); }
      else if (t > 1) { t = 1; s = clampr((b - c) / a, 0, 1); }
    }
  }
  xa = addScaled(p0, d1, s);
  xb = addScaled(q0, d2, t);
}
// collide.ts:246-292
static std::vector<RawContact> buildFaceManifold(const Rigid &bA, const Rigid &bB, const OBB &A, const OBB &B, bool refIsA, int refAxis, const V3 &nAB) {
  const OBB &R_ = refIsA ? A : B;
  const OBB &I = refIsA ? B : A;
  V3 outward = refIsA ? nAB : neg(nAB);
  R sign = dot(outward, R_.axis[refAxis]) >= 0 ? 1 : -1;
  V3 rn = scale(R_.axis[refAxis], sign);
  V3 rc = addScaled(R_.center, rn, R_.half[refAxis]);
  V3 ru, rv; R eu, ev;
  faceAxes(R_, refAxis, ru, rv, eu, ev);
  int inc = 0; R best = -FLT_MAXD;
  for (int i = 0; i < 3; i++) { R d = absDot(I.axis[i], rn); if (d > best) { best = d; inc = i; } }
  R isg = dot(I.axis[inc], rn) > 0 ? -1 : 1;
  V3 fn = scale(I.axis[inc], isg);
  V3 fc = addScaled(I.center, fn, I.half[inc]);
  V3 iu, iv; R ieu, iev;
  faceAxes(I, inc, iu, iv, ieu, iev);
  auto vert = [&](R su, R sv) { V3 p = addScaled(fc, iu, su * ieu); return addScaled(p, iv, sv * iev); };
  std::vector<V3> poly = {vert(1, 1), vert(-1, 1), vert(-1, -1), vert(1, -1)};
  V3 pn[4] = {ru, neg(ru), rv, neg(rv)};
  R ext[4] = {eu, eu, ev, ev};
  for (int k = 0; k < 4; k++) {
    poly = clipPolygonAgainstPlane(poly, pn[k], dot(pn[k], rc) + ext[k]);
    if (poly.empty()) return {};
  }
  ContactBuilder b; b.A = &bA; b.B = &bB;
  uint32_t prefix = ((uint32_t)(refIsA ? AXIS_FACE_A : AXIS_FACE_B) << 24) | ((uint32_t)(refAxis & 0xff) << 16) | ((uint32_t)(inc & 0xff) << 8);
  for (size_t i = 0; i < poly.size() && (int)b.contacts.size() < MAX_CONTACTS; i++) {
    const V3 &p = poly[i];
    R distance = dot(sub(p, rc), rn);
    if (distance > PLANE_EPSILON) continue;
    V3 pRef = addScaled(p, rn, -distance);
    b.add(refIsA ? pRef : p, refIsA ? p : pRef, prefix | (uint32_t)(i & 0xff));
  }
  if (b.contacts.empty()) b.add(supportPoint(A, nAB), supportPoint(B, neg(nAB)), prefix);
  return b.contacts;
}
// collide.ts:312-345
static std::vector<RawContact> collide(const Rigid &bA, const Rigid &bB, M3 &basisOut) {
  OBB A = makeOBB(bA), B = makeOBB(bB);
  V3 delta = sub(B.center, A.center);
  SatAxis face{AXIS_FACE_A, 0, 0, -FLT_MAXD, V3(), false}, edge{AXIS_EDGE, 0, 0, -FLT_MAXD, V3(), false};
  for (int i = 0; i < 3; i++) if (!testAxis(A, B, delta, A.axis[i], AXIS_FACE_A, i, -1, face)) return {};
  for (int i = 0; i < 3; i++) if (!testAxis(A, B, delta, B.axis[i], AXIS_FACE_B, -1, i, face)) return {};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      if (!testAxis(A, B, delta, cross(A.axis[i], B.axis[j]), AXIS_EDGE, i, j, edge)) return {};
  if (!face.valid) return {};
  SatAxis best = face;
  if (edge.valid && 0.95 * edge.separation > face.separation + 0.01) best = edge;
  basisOut = orthonormal(neg(best.normalAB));
  if (best.type == AXIS_EDGE) {
    V3 a0, a1, b0, b1, xa, xb;
    supportEdge(A, best.indexA, best.normalAB, a0, a1);
    supportEdge(B, best.indexB, neg(best.normalAB), b0, b1);
    closestPointsOnSegments(a0, a1, b0, b1, xa, xb);
    ContactBuilder b; b.A = &bA; b.B = &bB;
    uint32_t key = ((uint32_t)AXIS_EDGE << 24) | ((uint32_t)(best.indexA & 0xff) << 8) | (uint32_t)(best.indexB & 0xff);
    b.add(xa, xb, key);
    if (b.contacts.empty()) b.add(supportPoint(A, best.normalAB), supportPoint(B, neg(best.normalAB)), key);
    return b.contacts;
  }
  if (best.type == AXIS_FACE_A) return buildFaceManifold(bA, bB, A, B, true, best.indexA, best.normalAB);
  return buildFaceManifold(bA, bB, A, B, false, best.indexB, best.normalAB);
}

// --- solver.ts ----------------------------------------------------------------------------------
struct Solver {
  R dt = 1.0 / 60, gravity = -10;
  int iterations = 10;
  R alpha = 0.99, betaLin = 10000, betaAng = 100, gamma = 0.999;
  std::vector<Rigid *> bodies;
  std::vector<Force *> forces;
  ~Solver() { clear(); }
  void clear() {
    for (Force *f : forces) delete f;
    for (Rigid *b : bodies) delete b;
    forces.clear();
    bodies.clear();
  }
  // body.ts:34-56
  Rigid *addBody(const V3 &size, R density, R friction, const V3 &pos, const V3 &vel = V3()) {
    Rigid *b = new Rigid();
    b->size = size; b->friction = friction; b->positionLin = pos; b->velocityLin = vel; b->prevVelocityLin = vel;
    b->mass = size.x * size.y * size.z * density;
    b->moment = v3(((size.y * size.y + size.z * size.z) / 12) * b->mass, ((size.x * size.x + size.z * size.z) / 12) * b->mass,
                   ((size.x * size.x + size.y * size.y) / 12) * b->mass);
    b->radius = length(v3(size.x * 0.5, size.y * 0.5, size.z * 0.5));
    b->index = (int)bodies.size();
    bodies.push_back(b);
    return b;
  }
  Force *link(Force *f) {
    forces.push_back(f);
    if (f->bodyA) f->bodyA->forces.push_back(f);
    f->bodyB->forces.push_back(f);
    return f;
  }
  static void unlink(Force *f) {
    Rigid *bs[2] = {f->bodyA, f->bodyB};
    for (Rigid *b : bs) {
      if (!b) continue;
      auto it = std::find(b->forces.begin(), b->forces.end(), f);
      if (it != b->forces.end()) b->forces.erase(it);
    }
  }
  // forces.ts:129-147
  Force *addJoint(Rigid *a, Rigid *b, const V3 &rA, const V3 &rB, R stiffLin = INF, R stiffAng = 0, R fracture = INF) {
    Force *f = new Force(F_JOINT, a, b);
    f->rA = rA; f->rB = rB; f->stiffnessLin = stiffLin; f->stiffnessAng = stiffAng; f->fracture = fracture;
    V3 s = a ? a->size : V3();
    f->torqueArm = lengthSq(v3(s.x + b->size.x, s.y + b->size.y, s.z + b->size.z));
    return link(f);
  }
  // forces.ts:331-342
  Force *addSpring(Rigid *a, Rigid *b, const V3 &rA, const V3 &rB, R stiffness, R rest = -1) {
    Force *f = new Force(F_SPRING, a, b);
    f->rA = rA; f->rB = rB; f->stiffness = stiffness; f->rest = rest;
    if (rest < 0) f->rest = length(sub(transform(a->positionLin, a->positionAng, rA), transform(b->positionLin, b->positionAng, rB)));
    return link(f);
  }
  Force *addIgnore(Rigid *a, Rigid *b) { return link(new Force(F_IGNORE, a, b)); }

  // --- manifold.ts ---
  bool manifoldInit(Force &m) {
    Rigid &A = *m.bodyA, &B = *m.bodyB;
    m.friction = std::sqrt(A.friction * B.friction);
    std::vector<Contact> old = std::move(m.contacts);
    std::vector<RawContact> raw = collide(A, B, m.basis);
    m.contacts.clear();
    for (const RawContact &c : raw) {
      const Contact *prev = nullptr;
      for (const Contact &o : old) if (o.feature == c.feature) { prev = &o; break; }
      Contact k;
      k.feature = c.feature;
      if (!prev) { k.rA = c.rA; k.rB = c.rB; }
      else { k.rA = prev->stick ? prev->rA : c.rA; k.rB = prev->stick ? prev->rB : c.rB; k.penalty = prev->penalty; k.lambda = prev->lambda; k.stick = prev->stick; }
      m.contacts.push_back(k);
    }
    for (Contact &c : m.contacts) {
      V3 xA = transform(A.positionLin, A.positionAng, c.rA), xB = transform(B.positionLin, B.positionAng, c.rB);
      c.C0 = mulv(m.basis, sub(xA, xB));
      c.C0.x += COLLISION_MARGIN;
      for (int i = 0; i < 3; i++) {
        c.lambda[i] = c.lambda[i] * alpha * gamma;
        c.penalty[i] = clampr(c.penalty[i] * gamma, PENALTY_MIN, PENALTY_MAX);
      }
    }
    return !m.contacts.empty();
  }
  struct Eval { M3 jBLin, jAAng, jBAng, K; V3 C, F; R frictionScale, bounds; };
  static M3 angularJacobian(const V3 &r, const M3 &lin) {
    M3 o;
    for (int k = 0; k < 3; k++) { V3 row = cross(r, v3(lin[k * 3], lin[k * 3 + 1], lin[k * 3 + 2])); o[k * 3] = row.x; o[k * 3 + 1] = row.y; o[k * 3 + 2] = row.z; }
    return o;
  }
  void evaluate(const Force &m, const Contact &c, R a, const V3 &dqALin, const V3 &dqAAng, const V3 &dqBLin, const V3 &dqBAng, Eval &e) const {
    V3 rAW = rotate(m.bodyA->positionAng, c.rA), rBW = rotate(m.bodyB->positionAng, c.rB);
    e.jBLin = negm(m.basis);
    e.jAAng = angularJacobian(rAW, m.basis);
    e.jBAng = angularJacobian(rBW, e.jBLin);
    e.K = diagonal(c.penalty.x, c.penalty.y, c.penalty.z);
    e.C = scale(c.C0, 1 - a);
    e.C = add(e.C, mulv(m.basis, dqALin));
    e.C = add(e.C, mulv(e.jBLin, dqBLin));
    e.C = add(e.C, mulv(e.jAAng, dqAAng));
    e.C = add(e.C, mulv(e.jBAng, dqBAng));
    e.F = add(mulv(e.K, e.C), c.lambda);
    e.F.x = rmin(e.F.x, 0);
    e.bounds = std::fabs(e.F.x) * m.friction;
    e.frictionScale = std::sqrt(e.F.y * e.F.y + e.F.z * e.F.z);
    if (e.frictionScale > e.bounds && e.frictionScale > 0) { e.F.y *= e.bounds / e.frictionScale; e.F.z *= e.bounds / e.frictionScale; }
  }
  void manifoldPrimal(Force &m, Rigid &body, R a, BodySystem &sys) {
    V3 dqALin = sub(m.bodyA->positionLin, m.bodyA->initialLin), dqAAng = qsub(m.bodyA->positionAng, m.bodyA->initialAng);
    V3 dqBLin = sub(m.bodyB->positionLin, m.bodyB->initialLin), dqBAng = qsub(m.bodyB->positionAng, m.bodyB->initialAng);
    bool isA = &body == m.bodyA;
    for (Contact &c : m.contacts) {
      Eval e;
      evaluate(m, c, a, dqALin, dqAAng, dqBLin, dqBAng, e);
      const M3 &jLin = isA ? m.basis : e.jBLin;
      const M3 &jAng = isA ? e.jAAng : e.jBAng;
      M3 jLinT = transpose(jLin), jAngT = transpose(jAng), jAngTk = mul(jAngT, e.K);
      addm(sys.lhsLin, mul(mul(jLinT, e.K), jLin));
      addm(sys.lhsAng, mul(jAngTk, jAng));
      addm(sys.lhsCross, mul(jAngTk, jLin));
      sys.rhsLin = add(sys.rhsLin, mulv(jLinT, e.F));
      sys.rhsAng = add(sys.rhsAng, mulv(jAngT, e.F));
    }
  }
  void manifoldDual(Force &m, R a) {
    V3 dqALin = sub(m.bodyA->positionLin, m.bodyA->initialLin), dqAAng = qsub(m.bodyA->positionAng, m.bodyA->initialAng);
    V3 dqBLin = sub(m.bodyB->positionLin, m.bodyB->initialLin), dqBAng = qsub(m.bodyB->positionAng, m.bodyB->initialAng);
    for (Contact &c : m.contacts) {
      Eval e;
      evaluate(m, c, a, dqALin, dqAAng, dqBLin, dqBAng, e);
      c.lambda = e.F;
      if (e.F.x < 0) c.penalty.x = rmin(c.penalty.x + betaLin * std::fabs(e.C.x), PENALTY_MAX);
      if (e.frictionScale <= e.bounds) {
        c.penalty.y = rmin(c.penalty.y + betaLin * std::fabs(e.C.y), PENALTY_MAX);
        c.penalty.z = rmin(c.penalty.z + betaLin * std::fabs(e.C.z), PENALTY_MAX);
        c.stick = std::sqrt(e.C.y * e.C.y + e.C.z * e.C.z) < STICK_THRESH;
      }
    }
  }
  // --- forces.ts Joint ---
  V3 evaluateLin(const Force &j) const {
    V3 pA = j.bodyA ? transform(j.bodyA->positionLin, j.bodyA->positionAng, j.rA) : j.rA;
    return sub(pA, transform(j.bodyB->positionLin, j.bodyB->positionAng, j.rB));
  }
  V3 evaluateAng(const Force &j) const {
    Q4 rotA = j.bodyA ? j.bodyA->positionAng : Q4();
    const Q4 &rotB = j.bodyB->positionAng;
    V3 out = qsub(rotA, rotB);
    if (rotA.x * rotB.x + rotA.y * rotB.y + rotA.z * rotB.z + rotA.w * rotB.w < 0) out = neg(out);
    return scale(out, j.torqueArm);
  }
  static V3 componentMin(const V3 &a, R b) { return v3(rmin(a.x, b), rmin(a.y, b), rmin(a.z, b)); }
  bool jointInit(Force &j) {
    j.C0Lin = evaluateLin(j);
    j.C0Ang = evaluateAng(j);
    for (int i = 0; i < 3; i++) {
      j.lambdaLin[i] = j.lambdaLin[i] * alpha * gamma;
      j.lambdaAng[i] = j.lambdaAng[i] * alpha * gamma;
      j.penaltyLin[i] = clampr(j.penaltyLin[i] * gamma, PENALTY_MIN, PENALTY_MAX);
      j.penaltyAng[i] = clampr(j.penaltyAng[i] * gamma, PENALTY_MIN, PENALTY_MAX);
    }
    j.penaltyLin = componentMin(j.penaltyLin, j.stiffnessLin);
    j.penaltyAng = componentMin(j.penaltyAng, j.stiffnessAng);
    return !j.broken;
  }
  static M3 gsBallSocket(int k, const V3 &v) {
    M3 o = diagonal(-v[k], -v[k], -v[k]);
    o[k] += v.x; o[3 + k] += v.y; o[6 + k] += v.z;
    return o;
  }
  void jointPrimal(Force &j, Rigid &body, R a, BodySystem &sys) {
    bool isA = &body == j.bodyA;
    if (lengthSq(j.penaltyLin) > 0) {
      M3 K = diagonal(j.penaltyLin.x, j.penaltyLin.y, j.penaltyLin.z);
      V3 C = evaluateLin(j);
      if (j.stiffnessLin == INF) C = addScaled(C, j.C0Lin, -a);
      V3 F = add(mulv(K, C), j.lambdaLin);
      R sg = isA ? 1 : -1;
      M3 jLin = diagonal(sg, sg, sg);
      M3 jAng = isA ? skew(neg(rotate(j.bodyA->positionAng, j.rA))) : skew(rotate(j.bodyB->positionAng, j.rB));
      M3 jLinT = transpose(jLin), jAngT = transpose(jAng), jAngTk = mul(jAngT, K);
      addm(sys.lhsLin, mul(mul(jLinT, K), jLin));
      addm(sys.lhsAng, mul(jAngTk, jAng));
      addm(sys.lhsCross, mul(jAngTk, jLin));
      V3 r = isA ? rotate(j.bodyA->positionAng, j.rA) : neg(rotate(j.bodyB->positionAng, j.rB));
      M3 H = scalem(gsBallSocket(0, r), F.x);
      addm(H, scalem(gsBallSocket(1, r), F.y));
      addm(H, scalem(gsBallSocket(2, r), F.z));
      addm(sys.lhsAng, diagonalize(H));
      sys.rhsLin = add(sys.rhsLin, mulv(jLinT, F));
      sys.rhsAng = add(sys.rhsAng, mulv(jAngT, F));
    }
    if (lengthSq(j.penaltyAng) > 0) {
      M3 K = diagonal(j.penaltyAng.x, j.penaltyAng.y, j.penaltyAng.z);
      V3 C = evaluateAng(j);
      if (j.stiffnessAng == INF) C = addScaled(C, j.C0Ang, -a);
      V3 F = add(mulv(K, C), j.lambdaAng);
      R s = (isA ? 1 : -1) * j.torqueArm;
      M3 jAng = diagonal(s, s, s), jAngT = transpose(jAng);
      addm(sys.lhsAng, mul(mul(jAngT, K), jAng));
      sys.rhsAng = add(sys.rhsAng, mulv(jAngT, F));
    }
  }
  void jointDual(Force &j, R a) {
    if (lengthSq(j.penaltyLin) > 0) {
      M3 K = diagonal(j.penaltyLin.x, j.penaltyLin.y, j.penaltyLin.z);
      V3 C = evaluateLin(j);
      if (j.stiffnessLin == INF) { C = addScaled(C, j.C0Lin, -a); j.lambdaLin = add(mulv(K, C), j.lambdaLin); }
      j.penaltyLin = addScaled(j.penaltyLin, abs3(C), betaLin);
      j.penaltyLin = componentMin(j.penaltyLin, rmin(j.stiffnessLin, PENALTY_MAX));
    }
    if (lengthSq(j.penaltyAng) > 0) {
      M3 K = diagonal(j.penaltyAng.x, j.penaltyAng.y, j.penaltyAng.z);
      V3 C = evaluateAng(j);
      if (j.stiffnessAng == INF) { C = addScaled(C, j.C0Ang, -a); j.lambdaAng = add(mulv(K, C), j.lambdaAng); }
      // (yield is Infinity for every joint built here: the demo's joint, forces.ts:283-306)
      j.penaltyAng = addScaled(j.penaltyAng, abs3(C), betaAng);
      j.penaltyAng = componentMin(j.penaltyAng, rmin(j.stiffnessAng, PENALTY_MAX));
    }
    if (lengthSq(j.lambdaAng) > j.fracture * j.fracture || (j.fractureLinear && lengthSq(j.lambdaLin) > j.fracture * j.fracture)) {
      j.penaltyLin = V3(); j.penaltyAng = V3(); j.lambdaLin = V3(); j.lambdaAng = V3();
      j.broken = true;
    }
  }
  // forces.ts:348-378
  void springPrimal(Force &s, Rigid &body, BodySystem &sys) {
    Rigid &A = *s.bodyA, &B = *s.bodyB;
    V3 pA = transform(A.positionLin, A.positionAng, s.rA), pB = transform(B.positionLin, B.positionAng, s.rB);
    V3 d = sub(pA, pB);
    R dLen = length(d);
    if (dLen <= 1.0e-6) return;
    V3 n = div3(d, dLen);
    R c = dLen - s.rest, f = s.stiffness * c;
    V3 jl, ja;
    if (&body == &A) { V3 r = rotate(A.positionAng, s.rA); jl = n; ja = cross(r, n); }
    else { V3 r = rotate(B.positionAng, s.rB); jl = neg(n); ja = neg(cross(r, n)); }
    addm(sys.lhsLin, scalem(outer(jl, jl), s.stiffness));
    addm(sys.lhsAng, scalem(outer(ja, ja), s.stiffness));
    addm(sys.lhsCross, scalem(outer(ja, jl), s.stiffness));
    sys.rhsLin = add(sys.rhsLin, scale(jl, f));
    sys.rhsAng = add(sys.rhsAng, scale(ja, f));
  }
  bool initialize(Force &f) {
    switch (f.kind) {
      case F_MANIFOLD: return manifoldInit(f);
      case F_JOINT: return jointInit(f);
      default: return true;
    }
  }
  void updatePrimal(Force &f, Rigid &body, R a, BodySystem &sys) {
    switch (f.kind) {
      case F_MANIFOLD: manifoldPrimal(f, body, a, sys); break;
      case F_JOINT: jointPrimal(f, body, a, sys); break;
      case F_SPRING: springPrimal(f, body, sys); break;
      default: break;
    }
  }
  void updateDual(Force &f, R a) {
    if (f.kind == F_MANIFOLD) manifoldDual(f, a);
    else if (f.kind == F_JOINT) jointDual(f, a);
  }

  // solver.ts:137-223
  void step() {
    const int n = (int)bodies.size();
    for (int i = n - 1; i >= 0; i--) {
      Rigid *A = bodies[i];
      for (int j = i - 1; j >= 0; j--) {
        Rigid *B = bodies[j];
        V3 dp = sub(A->positionLin, B->positionLin);
        R r = A->radius + B->radius;
        if (dot(dp, dp) <= r * r && !A->constrainedTo(B)) link(new Force(F_MANIFOLD, A, B));
      }
    }
    std::vector<Force *> kept;
    for (Force *f : forces) {
      if (initialize(*f)) kept.push_back(f);
      else { unlink(f); delete f; }
    }
    forces = kept;
    for (int b = n - 1; b >= 0; b--) {
      Rigid &body = *bodies[b];
      body.inertialLin = addScaled(body.positionLin, body.velocityLin, dt);
      if (body.mass > 0) body.inertialLin.z += gravity * (dt * dt);
      body.inertialAng = qaddv(body.positionAng, scale(body.velocityAng, dt));
      V3 accel = div3(sub(body.velocityLin, body.prevVelocityLin), dt);
      R accelExt = accel.z * sgn(gravity);
      R w = clampr(accelExt / std::fabs(gravity), 0, 1);
      if (!std::isfinite(w)) w = 0;
      body.initialLin = body.positionLin;
      body.initialAng = body.positionAng;
      if (body.mass > 0) {
        body.positionLin = addScaled(body.positionLin, body.velocityLin, dt);
        body.positionLin.z += gravity * (w * dt * dt);
        body.positionAng = qaddv(body.positionAng, scale(body.velocityAng, dt));
      }
    }
    for (int it = 0; it < iterations; it++) {
      for (int b = n - 1; b >= 0; b--) {
        Rigid &body = *bodies[b];
        if (body.mass <= 0) continue;
        BodySystem sys;
        R m = body.mass;
        const V3 &I = body.moment;
        sys.lhsLin = diagonal(m / (dt * dt), m / (dt * dt), m / (dt * dt));
        sys.lhsAng = diagonal(I.x / (dt * dt), I.y / (dt * dt), I.z / (dt * dt));
        sys.rhsLin = mulv(sys.lhsLin, sub(body.positionLin, body.inertialLin));
        sys.rhsAng = mulv(sys.lhsAng, qsub(body.positionAng, body.inertialAng));
        for (int f = (int)body.forces.size() - 1; f >= 0; f--) updatePrimal(*body.forces[f], body, alpha, sys);
        V3 dxLin, dxAng;
        solve6(sys.lhsLin, sys.lhsAng, sys.lhsCross, neg(sys.rhsLin), neg(sys.rhsAng), dxLin, dxAng);
        body.positionLin = add(body.positionLin, dxLin);
        body.positionAng = qaddv(body.positionAng, dxAng);
      }
      for (Force *f : forces) updateDual(*f, alpha);
    }
    for (Rigid *bp : bodies) {
      Rigid &body = *bp;
      body.prevVelocityLin = body.velocityLin;
      if (body.mass > 0) {
        body.velocityLin = div3(sub(body.positionLin, body.initialLin), dt);
        body.velocityAng = div3(qsub(body.positionAng, body.initialAng), dt);
      }
    }
  }
};

}  // namespace refo
