// Device-side math, record types and the narrowphase of the MI355X AVBD step.
//
// Every expression is written in the evaluation order of the reference's WGSL (PRELUDE_3D in
// src/avbd3d/gpu/layout.ts:307-340, src/avbd3d/gpu/wgsl-collision.ts, wgsl-solve.ts), and the
// library is compiled with -ffp-contract=off and IEEE division/sqrt, so a step rounds the same
// way on every run and on the CPU restatement used to check it. min/max/clamp are written as
// comparisons (not fminf/fmaxf) to fix their behaviour on signed zeros.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace avbd {

struct V { float x, y, z; };
struct Q { float x, y, z, w; };

__host__ __device__ __forceinline__ V vv(float x, float y, float z) { V r; r.x = x; r.y = y; r.z = z; return r; }
__host__ __device__ __forceinline__ Q qq(float x, float y, float z, float w) { Q r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
__device__ __forceinline__ V xyz(float4 a) { return vv(a.x, a.y, a.z); }
__device__ __forceinline__ Q q4(float4 a) { return qq(a.x, a.y, a.z, a.w); }
__device__ __forceinline__ float4 f4(V a, float w) { return make_float4(a.x, a.y, a.z, w); }
__device__ __forceinline__ float4 f4q(Q a) { return make_float4(a.x, a.y, a.z, a.w); }
__device__ __forceinline__ float comp(V a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
__device__ __forceinline__ void setComp(V &a, int i, float v) { if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v; }

__device__ __forceinline__ V add(V a, V b) { return vv(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V sub(V a, V b) { return vv(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V mul(V a, float s) { return vv(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V mulv(V a, V b) { return vv(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ V divs(V a, float s) { return vv(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ V neg(V a) { return vv(-a.x, -a.y, -a.z); }
__device__ __forceinline__ V vabs(V a) { return vv(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
__device__ __forceinline__ float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length(V a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V normalize(V a) { return divs(a, length(a)); }
__device__ __forceinline__ V cross(V a, V b) { return vv(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
__device__ __forceinline__ float mn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float mx(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return mn(mx(x, lo), hi); }
__device__ __forceinline__ float dot4(Q a, Q b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
__device__ __forceinline__ Q qmul(Q a, Q b) {
  return qq(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
__device__ __forceinline__ Q qconj(Q a) { return qq(-a.x, -a.y, -a.z, a.w); }
/** 2·vec(a·b⁻¹): the small-angle rotation vector taking b to a. */
__device__ __forceinline__ V qsub(Q a, Q b) {
  float d = dot4(b, b);
  Q r = qmul(a, qq(-b.x / d, -b.y / d, -b.z / d, b.w / d));
  return vv(r.x * 2.0f, r.y * 2.0f, r.z * 2.0f);
}
__device__ __forceinline__ Q qnormalize(Q a) { float l = sqrtf(dot4(a, a)); return qq(a.x / l, a.y / l, a.z / l, a.w / l); }
/** normalize(a + (w, 0)·a·½): integrate a rotation vector. */
__device__ __forceinline__ Q qadd(Q a, V w) {
  Q m = qmul(qq(w.x, w.y, w.z, 0.0f), a);
  return qnormalize(qq(a.x + m.x * 0.5f, a.y + m.y * 0.5f, a.z + m.z * 0.5f, a.w + m.w * 0.5f));
}
__device__ __forceinline__ V qrotate(Q r, V a) {
  V u = vv(r.x, r.y, r.z);
  V t = mul(cross(u, a), 2.0f);
  return add(add(a, mul(t, r.w)), cross(u, t));
}
/** Contact basis rows (n, t1, t2). */
__device__ __forceinline__ void orthonormal(V n, V *b) {
  V t1 = fabsf(n.x) > fabsf(n.z) ? vv(-n.y, n.x, 0.0f) : vv(0.0f, -n.z, n.y);
  t1 = normalize(t1);
  b[0] = n; b[1] = t1; b[2] = cross(n, t1);
}
__device__ __forceinline__ uint32_t hash32(uint32_t x) {  // lowbias32
  x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
  return x;
}

// --- records ---------------------------------------------------------------------------------
static const uint32_t STICK_BIT = 0x80000000u, SPHERE_FEATURE = 3u << 24;
/** A contact of a rolling SDF surface (sphere, capsule, cylinder, cone, torus; layout.ts:129-134):
 *  its lever arms use the step-start rotation, as a sphere's do (wgsl-solve.ts:203). */
static const uint32_t SDF_ROLLING_BIT = 0x20000000u;
static const uint32_t FLAG_MATCH_NEAREST = 1, FLAG_FACE_BIAS = 2, FLAG_REUSE_CONTACTS = 4, FLAG_START_AT_REST = 8, FLAG_MASS_PENALTY = 16;
static const float SHAPE_SPHERE = 1.0f, SHAPE_HULL = 3.0f;
static const uint32_t NO_COLOR = 255u, PENDING = 256u, MAX_COLORS = 64u;
static const uint32_t T_NONE = 0, T_JOINT = 1, T_SPRING = 2, IGNORE_SLOT = 0xffffffffu;
static const float BIG = 3.0e38f, HARD = 1e30f, PENALTY_MIN = 1.0f, PENALTY_MAX = 1e10f;
// counter words (layout.ts:136-141 + avbd2d/gpu/layout.ts:45-51)
enum { C_PAIRS = 0, C_CONTACTS = 1, C_PREV_CONTACTS = 2, C_OVERFLOW = 3, C_CLASHES = 4, C_NUM_COLORS = 5,
       C_MANIFOLDS = 6, C_PREV_MANIFOLDS = 7, C_HULL_PAIRS = 8, C_CLASHING = 9,  // + 10
       // since the host last cleared them: overflow bits | 0x10000 on clashes, high-water counts
       C_STICKY = 11, C_MAX_CONTACTS = 12, C_MAX_MANIFOLDS = 13,
       C_PREV_OVERFLOW = 14,  // the last step's overflow bits (k_begin)
       // bit 0: this substep keeps the last adjacency (a reuse that kept every manifold, no joint
       // broken since the adjacency was built); bit 1: the next substep rebuilds it (a joint broke
       // since it was built, or the host asked: a lone step, or storage regrown)
       C_TOPO = 15,
       C_BARRIER = 16,  // cooperative solve: grid-barrier waits given up since the host last cleared it
       C_STEP = 17,     // substeps since creation (the contact stamps' clock), advanced by k_begin
       C_BOX_PAIRS = 18,  // box-box pairs the narrowphase deferred to k_narrow_box
       C_WORDS = 20 };

/** One contact point: 4 float4, the w's carrying the key and C(x-) (layout.ts:123-127). */
struct __align__(16) Contact { float4 rA; float4 rB; float4 pen; float4 lam; };

/** The step's uniform parameters (layout.ts Params, the fields the kernels read). */
struct Params {
  float dt, gravity, betaLin, betaAng, gamma, alpha, margin;
  float upx, upy, upz;
  uint32_t flags, iterations, colorCap, rounds;
  uint32_t lanes1, lanes2, lanes4, lanes8;
};

/** A body stream whose elements are `stride` float4 apart: the four pose streams share one
 *  64-byte block per body, so a partner's pose is one cache line, not four. */
struct Strided4 {
  float4 *p;
  uint32_t stride;
  __host__ __device__ __forceinline__ float4 &operator[](uint32_t i) const { return p[i * stride]; }
};
/** Body streams, one per record vec4 (layout.ts:8-15, 49-58): the pose four interleaved
 *  (pos, rot, x-, q- per body), the rest one float4 array each. */
struct Bodies {
  Strided4 pos;    // xyz, friction
  Strided4 rot;    // quaternion xyzw
  Strided4 ipos;   // x- xyz, lifecycle flags (1 = rewritten slot: no contact history)
  Strided4 irot;
  Strided4 size;    // full widths, mass (0: static)           } a second 64-byte block per body:
  Strided4 moment;  // principal moments, bounding radius      } what a body's own update reads
  Strided4 inpos;   // inertial target, step it last moved from its reference pose (u32 bits)
  Strided4 inrot;
  float4 *vel;     // xyz, previous step's velocity along up
  float4 *angvel;  // xyz, shape code (0 box, 1 sphere, 3 + hull header offset)
};

// --- narrowphase (wgsl-collision.ts:383-683) ---------------------------------------------------
struct Box { V c, h, ax[3]; };
struct Sat { uint32_t kind, ia, ib; float sep; V n; bool valid; };
struct Found { uint32_t feature[8]; V xA[8], xB[8]; uint32_t count; };

__device__ __forceinline__ V axisOf(const Box &b, uint32_t k) { return k == 0 ? b.ax[0] : (k == 1 ? b.ax[1] : b.ax[2]); }
__device__ __forceinline__ float hOf(const Box &b, uint32_t k) { return comp(b.h, (int)k); }

__device__ __forceinline__ Box makeBox(V c, Q q, V size) {
  Box b;
  b.c = c;
  b.h = mul(size, 0.5f);
  b.ax[0] = qrotate(q, vv(1.0f, 0.0f, 0.0f));
  b.ax[1] = qrotate(q, vv(0.0f, 1.0f, 0.0f));
  b.ax[2] = qrotate(q, vv(0.0f, 0.0f, 1.0f));
  return b;
}
__device__ __forceinline__ float projRadius(const Box &b, V n) {
  return b.h.x * fabsf(dot(n, b.ax[0])) + b.h.y * fabsf(dot(n, b.ax[1])) + b.h.z * fabsf(dot(n, b.ax[2]));
}
__device__ __forceinline__ bool testAxis(const Box &A, const Box &B, V delta, V axis, uint32_t kind, uint32_t ia, uint32_t ib, Sat &best) {
  float lenSq = dot(axis, axis);
  if (lenSq < 1.0e-6f) return true;
  V n = mul(axis, 1.0f / sqrtf(lenSq));
  if (dot(n, delta) < 0.0f) n = neg(n);
  float sep = fabsf(dot(delta, n)) - (projRadius(A, n) + projRadius(B, n));
  if (sep > 0.0f) return false;
  if (!best.valid || sep > best.sep) { best.kind = kind; best.ia = ia; best.ib = ib; best.sep = sep; best.n = n; best.valid = true; }
  return true;
}
__device__ __forceinline__ V support(const Box &b, V dir) {
  float sx = dot(dir, b.ax[0]) >= 0.0f ? 1.0f : -1.0f, sy = dot(dir, b.ax[1]) >= 0.0f ? 1.0f : -1.0f, sz = dot(dir, b.ax[2]) >= 0.0f ? 1.0f : -1.0f;
  return add(add(add(b.c, mul(b.ax[0], b.h.x * sx)), mul(b.ax[1], b.h.y * sy)), mul(b.ax[2], b.h.z * sz));
}
__device__ __forceinline__ void addFound(Found &f, V xA, V xB, uint32_t feature) {
  V mid = mul(add(xA, xB), 0.5f);
  for (uint32_t i = 0; i < f.count; i++) {
    V d = sub(mid, mul(add(f.xA[i], f.xB[i]), 0.5f));
    if (dot(d, d) < 1.0e-6f) return;
  }
  if (f.count >= 8) return;
  f.feature[f.count] = feature; f.xA[f.count] = xA; f.xB[f.count] = xB; f.count++;
}
__device__ __forceinline__ uint32_t clipPlane(const V *src, uint32_t n, V *dst, V pn, float offset) {
  if (n == 0) return 0;
  uint32_t count = 0;
  V a = src[n - 1];
  float da = dot(pn, a) - offset;
  for (uint32_t i = 0; i < n; i++) {
    V b = src[i];
    float db = dot(pn, b) - offset;
    bool aIn = da <= 1.0e-5f, bIn = db <= 1.0e-5f;
    if (aIn != bIn) {
      float t = 0.0f, denom = da - db;
      if (fabsf(denom) > 1.0e-6f) t = clampf(da / denom, 0.0f, 1.0f);
      if (count < 8) dst[count++] = add(a, mul(sub(b, a), t));
    }
    if (bIn && count < 8) dst[count++] = b;
    a = b; da = db;
  }
  return count;
}
__device__ __forceinline__ uint32_t faceU(uint32_t k) { return k == 0 ? 1u : 0u; }
__device__ __forceinline__ uint32_t faceV(uint32_t k) { return k != 2 ? 2u : 1u; }

__device__ __noinline__ void faceManifold(const Box &A, const Box &B, bool refIsA, uint32_t refAxis, V nAB, Found &found) {
  found.count = 0;
  const Box &R = refIsA ? A : B;
  const Box &I = refIsA ? B : A;
  V outward = refIsA ? nAB : neg(nAB);
  V rax = axisOf(R, refAxis);
  float rs = dot(outward, rax) >= 0.0f ? 1.0f : -1.0f;
  V rn = mul(rax, rs);
  V rc = add(R.c, mul(rn, hOf(R, refAxis)));
  V ru = axisOf(R, faceU(refAxis)), rv = axisOf(R, faceV(refAxis));
  float eu = hOf(R, faceU(refAxis)), ev = hOf(R, faceV(refAxis));
  uint32_t inc = 0;
  float bestDot = -1.0f;
  for (uint32_t k = 0; k < 3; k++) { float d = fabsf(dot(axisOf(I, k), rn)); if (d > bestDot) { bestDot = d; inc = k; } }
  V iax = axisOf(I, inc);
  float is = dot(iax, rn) > 0.0f ? -1.0f : 1.0f;
  V ic = add(I.c, mul(iax, is * hOf(I, inc)));
  V iu = mul(axisOf(I, faceU(inc)), hOf(I, faceU(inc))), iv = mul(axisOf(I, faceV(inc)), hOf(I, faceV(inc)));
  V poly[8], tmp[8];
  poly[0] = add(add(ic, iu), iv);
  poly[1] = add(sub(ic, iu), iv);
  poly[2] = sub(sub(ic, iu), iv);
  poly[3] = sub(add(ic, iu), iv);
  uint32_t n = 4;
  n = clipPlane(poly, n, tmp, ru, dot(ru, rc) + eu);
  n = clipPlane(tmp, n, poly, neg(ru), dot(neg(ru), rc) + eu);
  n = clipPlane(poly, n, tmp, rv, dot(rv, rc) + ev);
  n = clipPlane(tmp, n, poly, neg(rv), dot(neg(rv), rc) + ev);
  if (n == 0) return;
  uint32_t prefix = ((refIsA ? 0u : 1u) << 24) | (refAxis << 16) | (inc << 8);
  for (uint32_t i = 0; i < n && found.count < 8; i++) {
    V pt = poly[i];
    float dist = dot(sub(pt, rc), rn);
    if (dist > 1.0e-5f) continue;
    V onRef = sub(pt, mul(rn, dist));
    addFound(found, refIsA ? onRef : pt, refIsA ? pt : onRef, prefix | i);
  }
  if (found.count == 0) addFound(found, support(A, nAB), support(B, neg(nAB)), prefix);
}
__device__ __forceinline__ void supportEdge(const Box &b, uint32_t k, V dir, V &e0, V &e1) {
  uint32_t k1 = (k + 1) % 3, k2 = (k + 2) % 3;
  V a1 = axisOf(b, k1), a2 = axisOf(b, k2);
  float s1 = dot(dir, a1) >= 0.0f ? 1.0f : -1.0f, s2 = dot(dir, a2) >= 0.0f ? 1.0f : -1.0f;
  V c = add(add(b.c, mul(a1, hOf(b, k1) * s1)), mul(a2, hOf(b, k2) * s2));
  V e = mul(axisOf(b, k), hOf(b, k));
  e0 = sub(c, e); e1 = add(c, e);
}
__device__ __noinline__ void edgeContact(const Box &A, const Box &B, uint32_t ia, uint32_t ib, V nAB, Found &found) {
  found.count = 0;
  V a0, a1, b0, b1;
  supportEdge(A, ia, nAB, a0, a1);
  supportEdge(B, ib, neg(nAB), b0, b1);
  V p0 = a0, q0 = b0, d1 = sub(a1, p0), d2 = sub(b1, q0), r = sub(p0, q0);
  float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r), s = 0.0f, t = 0.0f;
  const float EPS = 1.0e-6f;
  if (a <= EPS && e <= EPS) {
  } else if (a <= EPS) {
    t = clampf(f / e, 0.0f, 1.0f);
  } else {
    float c = dot(d1, r);
    if (e <= EPS) s = clampf(-c / a, 0.0f, 1.0f);
    else {
      float b = dot(d1, d2), denom = a * e - b * b;
      if (fabsf(denom) > EPS) s = clampf((b * f - c * e) / denom, 0.0f, 1.0f);
      t = (b * s + f) / e;
      if (t < 0.0f) { t = 0.0f; s = clampf(-c / a, 0.0f, 1.0f); }
      else if (t > 1.0f) { t = 1.0f; s = clampf((b - c) / a, 0.0f, 1.0f); }
    }
  }
  uint32_t key = (2u << 24) | (ia << 8) | ib;
  addFound(found, add(p0, mul(d1, s)), add(q0, mul(d2, t)), key);
  if (found.count == 0) addFound(found, support(A, nAB), support(B, neg(nAB)), key);
}
/** Box-box SAT over 3 + 3 face and 9 edge axes; sat.n points from A towards B. */
__device__ __noinline__ void collideBoxes(const Box &A, const Box &B, bool faceBias, Sat &sat, Found &found) {
  found.count = 0;
  V delta = sub(B.c, A.c);
  Sat face, edge;
  face.valid = false; face.sep = 0.0f; face.kind = face.ia = face.ib = 0; face.n = vv(0.0f, 0.0f, 0.0f);
  edge = face;
  for (uint32_t i = 0; i < 3; i++) if (!testAxis(A, B, delta, A.ax[i], 0, i, 0, face)) return;
  for (uint32_t i = 0; i < 3; i++) if (!testAxis(A, B, delta, B.ax[i], 1, 0, i, face)) return;
  for (uint32_t i = 0; i < 3; i++)
    for (uint32_t j = 0; j < 3; j++)
      if (!testAxis(A, B, delta, cross(A.ax[i], B.ax[j]), 2, i, j, edge)) return;
  if (!face.valid) return;
  Sat best = face;
  if (edge.valid) {
    bool edgeWins = 0.95f * edge.sep > face.sep + 0.01f;
    if (faceBias) edgeWins = edge.sep > 0.95f * face.sep + 0.01f;
    if (edgeWins) best = edge;
  }
  sat = best;
  if (best.kind == 2) edgeContact(A, B, best.ia, best.ib, best.n, found);
  else faceManifold(A, B, best.kind == 0, best.kind == 0 ? best.ia : best.ib, best.n, found);
}
__device__ __forceinline__ void sphereSphere(const Box &A, const Box &B, Sat &sat, Found &found) {
  found.count = 0;
  V d = sub(B.c, A.c);
  float dist = length(d);
  if (dist > A.h.x + B.h.x) return;
  V n = vv(0.0f, 0.0f, 1.0f);
  if (dist > 0.0f) n = divs(d, dist);
  sat.n = n;
  addFound(found, add(A.c, mul(n, A.h.x)), sub(B.c, mul(n, B.h.x)), SPHERE_FEATURE);
}
struct SphereHit { bool hit; V onSphere, onBox, n; };
/** Sphere S against box X: closest point of the box, or the shallowest face when inside. */
__device__ __forceinline__ SphereHit sphereBox(const Box &S, const Box &X, Q qX) {
  SphereHit out;
  out.hit = false;
  V pl = qrotate(qconj(qX), sub(S.c, X.c));
  V ql = vv(clampf(pl.x, -X.h.x, X.h.x), clampf(pl.y, -X.h.y, X.h.y), clampf(pl.z, -X.h.z, X.h.z));
  float r = S.h.x;
  V nLocal, onBox = ql;
  if (pl.x != ql.x || pl.y != ql.y || pl.z != ql.z) {
    V d = sub(pl, ql);
    float dist = length(d);
    if (dist > r) return out;
    nLocal = divs(d, dist);
  } else {
    V depth = sub(X.h, vabs(pl));
    int k = 0;
    if (depth.y < comp(depth, k)) k = 1;
    if (depth.z < comp(depth, k)) k = 2;
    float s = comp(pl, k) >= 0.0f ? 1.0f : -1.0f;
    nLocal = vv(0.0f, 0.0f, 0.0f);
    setComp(nLocal, k, s);
    setComp(onBox, k, s * comp(X.h, k));
  }
  V n = qrotate(qX, nLocal);
  out.hit = true;
  out.n = neg(n);
  out.onBox = add(X.c, qrotate(qX, onBox));
  out.onSphere = sub(S.c, mul(n, r));
  return out;
}

}  // namespace avbd
