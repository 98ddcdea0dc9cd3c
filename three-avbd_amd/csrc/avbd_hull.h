// Convex hull narrowphase (src/avbd3d/gpu/wgsl-hull.ts): pairs with a hull on either side.
// SAT after Gregorius: face axes of both shapes by support points, then only the edge pairs
// whose Gauss-map arcs cross; face contact clips the incident face (the most anti-parallel one
// holding the deepest vertex) to the reference face's side planes and keeps up to 8 points,
// deepest first, then spread; edge contact is the closest points of the two edges.
// Hull buffer: at header h, (vertex start, count, face start, count), (edge start, count, index
// start, 0); vertices xyz; faces (normal, offset), (first index, count); indices 4 per vec4;
// edges (a, b, face left, face right). Offsets in vec4 units, f32 stored as bits.
#pragma once
#include "avbd_device.h"

namespace avbd {

static const uint32_t NO_HULL = 0xffffffffu, MAX_CLIP = 64;
__constant__ const uint32_t BOX_FACE_VERTS[24] = {1, 3, 7, 5, 0, 2, 6, 4, 2, 3, 7, 6, 0, 1, 5, 4, 4, 5, 7, 6, 0, 1, 3, 2};
__constant__ const uint4 BOX_EDGES[12] = {{0, 1, 3, 5}, {2, 3, 2, 5}, {4, 5, 3, 4}, {6, 7, 2, 4}, {0, 2, 1, 5}, {1, 3, 0, 5},
                                           {4, 6, 1, 4}, {5, 7, 0, 4}, {0, 4, 1, 3}, {1, 5, 0, 3}, {2, 6, 1, 2}, {3, 7, 0, 2}};

/** A hull's faces staged in LDS for one wavefront: world planes (facePlane's arithmetic, done once),
 *  each face's first index (high 16 bits) and vertex count, and the vertex-index list as bytes.
 *  Every face lookup of a pair then reads LDS instead of a chain of dependent global loads. */
static const uint32_t STAGE_F = 64, STAGE_I = 192;
struct FaceStage { float4 plane[STAGE_F]; uint32_t info[STAGE_F]; uint8_t idx[STAGE_I]; };

/** A convex shape as the hull code sees it: a box (hull == NO_HULL, half extents h) or a hull;
 *  fst, when set, holds its faces staged (same values as the global reads). */
struct Poly { V c; Q q; V h; uint32_t hull, vs, nv, fs, nf, es, ne, is; const FaceStage *fst; };

__device__ __forceinline__ Poly polyOf(const uint4 *hulls, V c, Q q, V h, float shape) {
  Poly P;
  P.c = c; P.q = q; P.h = h;
  P.hull = NO_HULL; P.nv = 8; P.nf = 6; P.ne = 12; P.vs = P.fs = P.es = P.is = 0;
  P.fst = nullptr;
  if (shape >= SHAPE_HULL) {
    uint32_t hh = (uint32_t)(shape - SHAPE_HULL + 0.5f);
    uint4 h0 = hulls[hh], h1 = hulls[hh + 1];
    P.hull = hh;
    P.vs = h0.x; P.nv = h0.y; P.fs = h0.z; P.nf = h0.w;
    P.es = h1.x; P.ne = h1.y; P.is = h1.z;
  }
  return P;
}
__device__ __forceinline__ V vertLocal(const uint4 *hulls, const Poly &P, uint32_t v) {
  if (P.hull == NO_HULL) return vv((v & 1) ? P.h.x : -P.h.x, (v & 2) ? P.h.y : -P.h.y, (v & 4) ? P.h.z : -P.h.z);
  uint4 w = hulls[P.vs + v];
  return vv(__uint_as_float(w.x), __uint_as_float(w.y), __uint_as_float(w.z));
}
__device__ __forceinline__ V vert(const uint4 *hulls, const Poly &P, uint32_t v) { return add(P.c, qrotate(P.q, vertLocal(hulls, P, v))); }
__device__ __forceinline__ V faceNormalLocal(const uint4 *hulls, const Poly &P, uint32_t f) {
  if (P.hull == NO_HULL) { V n = vv(0.0f, 0.0f, 0.0f); setComp(n, (int)(f / 2), (f & 1) ? -1.0f : 1.0f); return n; }
  uint4 w = hulls[P.fs + 2 * f];
  return vv(__uint_as_float(w.x), __uint_as_float(w.y), __uint_as_float(w.z));
}
/** Face f's world plane: outward normal n and offset. */
__device__ __forceinline__ void facePlane(const uint4 *hulls, const Poly &P, uint32_t f, V &n, float &off) {
  if (P.fst) {
    float4 w = P.fst->plane[f];
    n = xyz(w);
    off = w.w;
    return;
  }
  V nl;
  float d;
  if (P.hull == NO_HULL) {
    uint32_t k = f / 2;
    nl = vv(0.0f, 0.0f, 0.0f);
    setComp(nl, (int)k, (f & 1) ? -1.0f : 1.0f);
    d = comp(P.h, (int)k);
  } else {
    uint4 w = hulls[P.fs + 2 * f];
    nl = vv(__uint_as_float(w.x), __uint_as_float(w.y), __uint_as_float(w.z));
    d = __uint_as_float(w.w);
  }
  n = qrotate(P.q, nl);
  off = d + dot(n, P.c);
}
__device__ __forceinline__ uint32_t faceVertCount(const uint4 *hulls, const Poly &P, uint32_t f) {
  if (P.fst) return P.fst->info[f] & 0xffffu;
  return P.hull == NO_HULL ? 4u : hulls[P.fs + 2 * f + 1].y;
}
__device__ __forceinline__ uint32_t faceVert(const uint4 *hulls, const Poly &P, uint32_t f, uint32_t j) {
  if (P.fst) return P.fst->idx[(P.fst->info[f] >> 16) + j];
  if (P.hull == NO_HULL) return BOX_FACE_VERTS[f * 4 + j];
  uint32_t k = hulls[P.fs + 2 * f + 1].x + j;
  uint4 w = hulls[P.is + k / 4];
  uint32_t r = k % 4;
  return r == 0 ? w.x : (r == 1 ? w.y : (r == 2 ? w.z : w.w));
}
__device__ __forceinline__ uint4 edgeOf(const uint4 *hulls, const Poly &P, uint32_t e) { return P.hull == NO_HULL ? BOX_EDGES[e] : hulls[P.es + e]; }
__device__ __forceinline__ V polySupport(const uint4 *hulls, const Poly &P, V dir) {
  V l = qrotate(qconj(P.q), dir);
  if (P.hull == NO_HULL) return add(P.c, qrotate(P.q, vv(l.x >= 0.0f ? P.h.x : -P.h.x, l.y >= 0.0f ? P.h.y : -P.h.y, l.z >= 0.0f ? P.h.z : -P.h.z)));
  float best = -3.4e38f;
  V at = vv(0.0f, 0.0f, 0.0f);
  for (uint32_t v = 0; v < P.nv; v++) {
    V p = vertLocal(hulls, P, v);
    float s = dot(p, l);
    if (s > best) { best = s; at = p; }
  }
  return add(P.c, qrotate(P.q, at));
}
__device__ __forceinline__ void queryFaces(const uint4 *hulls, const Poly &A, const Poly &B, float &sep, uint32_t &face) {
  sep = -3.4e38f;
  face = 0;
  for (uint32_t f = 0; f < A.nf; f++) {
    V n;
    float d;
    facePlane(hulls, A, f, n, d);
    float s = dot(n, polySupport(hulls, B, neg(n))) - d;
    if (s > sep) { sep = s; face = f; }
  }
}
__device__ __forceinline__ bool minkowskiFace(V a, V b, V c, V d) {
  V bxa = cross(b, a), dxc = cross(d, c);
  float cba = dot(c, bxa), dba = dot(d, bxa), adc = dot(a, dxc), bdc = dot(b, dxc);
  return cba * dba < 0.0f && adc * bdc < 0.0f && cba * bdc > 0.0f;
}
struct EdgeQuery { float sep; uint32_t ea, eb; V n; bool valid; };
__device__ __noinline__ EdgeQuery queryEdges(const uint4 *hulls, const Poly &A, const Poly &B) {
  EdgeQuery out;
  out.sep = -3.4e38f; out.ea = out.eb = 0; out.n = vv(0.0f, 0.0f, 0.0f); out.valid = false;
  Q toB = qmul(qconj(B.q), A.q);
  for (uint32_t i = 0; i < A.ne; i++) {
    uint4 ea = edgeOf(hulls, A, i);
    V a = qrotate(toB, faceNormalLocal(hulls, A, ea.z)), b = qrotate(toB, faceNormalLocal(hulls, A, ea.w));
    for (uint32_t j = 0; j < B.ne; j++) {
      uint4 eb = edgeOf(hulls, B, j);
      if (!minkowskiFace(a, b, neg(faceNormalLocal(hulls, B, eb.z)), neg(faceNormalLocal(hulls, B, eb.w)))) continue;
      V pa = vert(hulls, A, ea.x), qa = vert(hulls, A, ea.y), pb = vert(hulls, B, eb.x), qb = vert(hulls, B, eb.y);
      V n = cross(sub(qa, pa), sub(qb, pb));
      float l = length(n);
      if (l < 1e-5f * length(sub(qa, pa)) * length(sub(qb, pb))) continue;
      n = divs(n, l);
      if (dot(n, sub(pa, A.c)) < 0.0f) n = neg(n);
      float s = dot(n, sub(pb, pa));
      if (s > out.sep) { out.sep = s; out.ea = i; out.eb = j; out.n = n; out.valid = true; }
    }
  }
  return out;
}
__device__ __forceinline__ uint32_t clipHull(const V *src, uint32_t n, V *dst, V pn, float offset) {
  if (n == 0) return 0;
  uint32_t count = 0;
  V a = src[n - 1];
  float da = dot(pn, a) - offset;
  for (uint32_t i = 0; i < n; i++) {
    V b = src[i];
    float db = dot(pn, b) - offset;
    if ((da <= 1.0e-5f) != (db <= 1.0e-5f)) {
      float t = 0.0f;
      if (fabsf(da - db) > 1.0e-6f) t = clampf(da / (da - db), 0.0f, 1.0f);
      if (count < MAX_CLIP) dst[count++] = add(a, mul(sub(b, a), t));
    }
    if (db <= 1.0e-5f && count < MAX_CLIP) dst[count++] = b;
    a = b;
    da = db;
  }
  return count;
}
/** The clip and point selection of a face contact, given the incident face. */
__device__ __noinline__ void hullFaceClip(const uint4 *hulls, const Poly &A, const Poly &B, bool refIsA, uint32_t refFace, const Poly &R,
                                          const Poly &I, V rn, float pw, uint32_t inc, Found &found);
__device__ __noinline__ void hullFaceContact(const uint4 *hulls, const Poly &A, const Poly &B, bool refIsA, uint32_t refFace, Found &found) {
  found.count = 0;
  const Poly &R = refIsA ? A : B;
  const Poly &I = refIsA ? B : A;
  V rn;
  float pw;
  facePlane(hulls, R, refFace, rn, pw);
  uint32_t deepest = 0;
  float low = 3.4e38f;
  for (uint32_t v = 0; v < I.nv; v++) { float d = dot(vert(hulls, I, v), rn); if (d < low) { low = d; deepest = v; } }
  uint32_t inc = 0;
  float least = 3.4e38f;
  for (uint32_t f = 0; f < I.nf; f++) {
    bool touches = false;
    uint32_t cnt = faceVertCount(hulls, I, f);
    for (uint32_t j = 0; j < cnt; j++) if (faceVert(hulls, I, f, j) == deepest) { touches = true; break; }
    if (!touches) continue;
    V fn;
    float fd;
    facePlane(hulls, I, f, fn, fd);
    float d = dot(fn, rn);
    if (d < least) { least = d; inc = f; }
  }
  hullFaceClip(hulls, A, B, refIsA, refFace, R, I, rn, pw, inc, found);
}
__device__ __noinline__ void hullFaceClip(const uint4 *hulls, const Poly &A, const Poly &B, bool refIsA, uint32_t refFace, const Poly &R,
                                          const Poly &I, V rn, float pw, uint32_t inc, Found &found) {
  found.count = 0;
  V poly[64], tmp[64];
  uint32_t n = faceVertCount(hulls, I, inc);
  if (n > MAX_CLIP) n = MAX_CLIP;
  for (uint32_t j = 0; j < n; j++) poly[j] = vert(hulls, I, faceVert(hulls, I, inc, j));
  uint32_t k = faceVertCount(hulls, R, refFace);
  V centre = vv(0.0f, 0.0f, 0.0f);
  for (uint32_t j = 0; j < k; j++) centre = add(centre, vert(hulls, R, faceVert(hulls, R, refFace, j)));
  centre = divs(centre, (float)k);
  for (uint32_t j = 0; j < k && n > 0; j++) {
    V v0 = vert(hulls, R, faceVert(hulls, R, refFace, j)), v1 = vert(hulls, R, faceVert(hulls, R, refFace, (j + 1) % k));
    V s = cross(sub(v1, v0), rn);
    float l = length(s);
    if (l < 1.0e-6f) continue;
    s = divs(s, l);
    if (dot(s, sub(centre, v0)) > 0.0f) s = neg(s);
    n = clipHull(poly, n, tmp, s, dot(s, v0));
    for (uint32_t m = 0; m < n; m++) poly[m] = tmp[m];
  }
  float depth[64];
  uint32_t index[64];
  uint32_t keep = 0;
  for (uint32_t m = 0; m < n; m++) {
    float dist = dot(rn, poly[m]) - pw;
    if (dist <= 1.0e-5f) { tmp[keep] = poly[m]; depth[keep] = dist; index[keep] = m; keep++; }
  }
  uint32_t prefix = ((refIsA ? 0u : 1u) << 24) | ((refFace < 255u ? refFace : 255u) << 16) | ((inc < 255u ? inc : 255u) << 8);
  uint64_t chosen = 0;
  uint32_t picks = keep < 8u ? keep : 8u;
  for (uint32_t c = 0; c < picks; c++) {
    uint32_t pick = 0;
    float bestScore = -3.4e38f;
    for (uint32_t m = 0; m < keep; m++) {
      if ((chosen >> m) & 1ull) continue;
      float score = -depth[m];
      if (c > 0) {
        score = 3.4e38f;
        for (uint32_t o = 0; o < keep; o++)
          if ((chosen >> o) & 1ull) { V dd = sub(tmp[m], tmp[o]); score = mn(score, dot(dd, dd)); }
      }
      if (score > bestScore) { bestScore = score; pick = m; }
    }
    chosen |= 1ull << pick;
    V pt = tmp[pick];
    V onRef = sub(pt, mul(rn, depth[pick]));
    uint32_t idx = index[pick] < 255u ? index[pick] : 255u;
    addFound(found, refIsA ? onRef : pt, refIsA ? pt : onRef, prefix | idx);
  }
  if (found.count == 0) {
    V nAB = refIsA ? rn : neg(rn);
    addFound(found, polySupport(hulls, A, nAB), polySupport(hulls, B, neg(nAB)), prefix);
  }
}
__device__ __forceinline__ void hullEdgeContact(const uint4 *hulls, const Poly &A, const Poly &B, const EdgeQuery &q, Found &found) {
  found.count = 0;
  uint4 ea = edgeOf(hulls, A, q.ea), eb = edgeOf(hulls, B, q.eb);
  V p0 = vert(hulls, A, ea.x), q0 = vert(hulls, B, eb.x);
  V d1 = sub(vert(hulls, A, ea.y), p0), d2 = sub(vert(hulls, B, eb.y), q0), r = sub(p0, q0);
  float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r), c = dot(d1, r), b = dot(d1, d2);
  float denom = a * e - b * b;
  const float EPS = 1.0e-6f;
  float s = 0.0f;
  if (fabsf(denom) > EPS * a * e) s = clampf((b * f - c * e) / denom, 0.0f, 1.0f);
  float t = (b * s + f) / mx(e, EPS);
  if (t < 0.0f) { t = 0.0f; s = clampf(-c / mx(a, EPS), 0.0f, 1.0f); }
  else if (t > 1.0f) { t = 1.0f; s = clampf((b - c) / mx(a, EPS), 0.0f, 1.0f); }
  uint32_t ia = q.ea < 255u ? q.ea : 255u, ib = q.eb < 255u ? q.eb : 255u;
  addFound(found, add(p0, mul(d1, s)), add(q0, mul(d2, t)), (2u << 24) | (ia << 8) | ib);
}
/** Two convex shapes, at least one a hull; sat.n from A towards B. */
__device__ __noinline__ void collidePoly(const uint4 *hulls, const Poly &A, const Poly &B, bool faceBias, Sat &sat, Found &found) {
  found.count = 0;
  float faSep, fbSep;
  uint32_t faFace, fbFace;
  queryFaces(hulls, A, B, faSep, faFace);
  if (faSep > 0.0f) return;
  queryFaces(hulls, B, A, fbSep, fbFace);
  if (fbSep > 0.0f) return;
  EdgeQuery eq = queryEdges(hulls, A, B);
  if (eq.valid && eq.sep > 0.0f) return;
  float faceSep = mx(faSep, fbSep);
  bool edgeWins = eq.valid && eq.sep > 0.95f * faceSep + 0.01f;
  if (!faceBias) edgeWins = eq.valid && 0.95f * eq.sep > faceSep + 0.01f;
  if (edgeWins) { sat.n = eq.n; hullEdgeContact(hulls, A, B, eq, found); return; }
  V n;
  float d;
  if (fbSep > faSep) {
    facePlane(hulls, B, fbFace, n, d);
    sat.n = neg(n);
    hullFaceContact(hulls, A, B, false, fbFace, found);
    return;
  }
  facePlane(hulls, A, faFace, n, d);
  sat.n = n;
  hullFaceContact(hulls, A, B, true, faFace, found);
}
__device__ __forceinline__ V segmentPoint(V a, V b, V p) {
  V d = sub(b, a);
  return add(a, mul(d, clampf(dot(sub(p, a), d) / mx(dot(d, d), 1e-20f), 0.0f, 1.0f)));
}
/** Sphere (centre c, radius r) against hull X: n from the sphere towards X. */
__device__ __noinline__ SphereHit sphereHull(const uint4 *hulls, V c, float r, const Poly &X) {
  SphereHit out;
  out.hit = false;
  float deepest = -3.4e38f;
  uint32_t deepFace = 0;
  for (uint32_t f = 0; f < X.nf; f++) {
    V n;
    float d;
    facePlane(hulls, X, f, n, d);
    float s = dot(n, c) - d;
    if (s > deepest) { deepest = s; deepFace = f; }
  }
  if (deepest > r) return out;
  if (deepest <= 0.0f) {
    V n;
    float d;
    facePlane(hulls, X, deepFace, n, d);
    out.hit = true;
    out.n = neg(n);
    out.onBox = sub(c, mul(n, deepest));
    out.onSphere = sub(c, mul(n, r));
    return out;
  }
  V best = vv(0.0f, 0.0f, 0.0f);
  float bestDist = 3.4e38f;
  for (uint32_t f = 0; f < X.nf; f++) {
    V pn;
    float pd;
    facePlane(hulls, X, f, pn, pd);
    float s = dot(pn, c) - pd;
    if (s <= 0.0f) continue;
    V p = sub(c, mul(pn, s));
    uint32_t k = faceVertCount(hulls, X, f);
    V centre = vv(0.0f, 0.0f, 0.0f);
    for (uint32_t j = 0; j < k; j++) centre = add(centre, vert(hulls, X, faceVert(hulls, X, f, j)));
    centre = divs(centre, (float)k);
    bool inside = true;
    V edgeBest = vv(0.0f, 0.0f, 0.0f);
    float edgeDist = 3.4e38f;
    for (uint32_t j = 0; j < k; j++) {
      V v0 = vert(hulls, X, faceVert(hulls, X, f, j)), v1 = vert(hulls, X, faceVert(hulls, X, f, (j + 1) % k));
      V side = cross(sub(v1, v0), pn);
      if (dot(side, sub(centre, v0)) > 0.0f) side = neg(side);
      if (dot(side, sub(p, v0)) > 0.0f) inside = false;
      V q = segmentPoint(v0, v1, c);
      V dq3 = sub(q, c);
      float dq = dot(dq3, dq3);
      if (dq < edgeDist) { edgeDist = dq; edgeBest = q; }
    }
    if (!inside) p = edgeBest;
    float dist = length(sub(p, c));
    if (dist < bestDist) { bestDist = dist; best = p; }
  }
  if (bestDist > r) return out;
  V n = vv(0.0f, 0.0f, 1.0f);
  if (bestDist > 0.0f) n = divs(sub(c, best), bestDist);
  out.hit = true;
  out.n = neg(n);
  out.onBox = best;
  out.onSphere = sub(c, mul(n, r));
  return out;
}

/** The hull branch of the narrowphase (wgsl-collision.ts:781-795); sat.n from A towards B. */
__device__ __noinline__ void collideHullPair(const uint4 *hulls, const Box &A, Q qA, float shA, const Box &B, Q qB, float shB,
                                             bool sphereA, bool sphereB, bool faceBias, Sat &sat, Found &found) {
  found.count = 0;
  if (sphereA || sphereB) {
    if (sphereA) {
      SphereHit h = sphereHull(hulls, A.c, A.h.x, polyOf(hulls, B.c, qB, B.h, shB));
      sat.n = h.n;
      if (h.hit) addFound(found, h.onSphere, h.onBox, SPHERE_FEATURE);
    } else {
      SphereHit h = sphereHull(hulls, B.c, B.h.x, polyOf(hulls, A.c, qA, A.h, shA));
      sat.n = neg(h.n);
      if (h.hit) addFound(found, h.onBox, h.onSphere, SPHERE_FEATURE);
    }
    return;
  }
  collidePoly(hulls, polyOf(hulls, A.c, qA, A.h, shA), polyOf(hulls, B.c, qB, B.h, shB), faceBias, sat, found);
}

// ---- one wavefront per pair --------------------------------------------------------------------
// The same queries with the work items (faces, edge pairs) split over the 64 lanes. Each lane scans
// its items in increasing order keeping the first strict extremum, and the lanes are combined by
// (value, index) with the lower index winning ties: the result is the serial loop's, bit for bit.
struct ArgBest { float v; uint32_t i; };
__device__ __forceinline__ ArgBest waveArgMax(ArgBest a) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    float ov = __shfl_xor(a.v, m);
    uint32_t oi = __shfl_xor(a.i, m);
    if (ov > a.v || (ov == a.v && oi < a.i)) { a.v = ov; a.i = oi; }
  }
  return a;
}
__device__ __forceinline__ ArgBest waveArgMin(ArgBest a) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    float ov = __shfl_xor(a.v, m);
    uint32_t oi = __shfl_xor(a.i, m);
    if (ov < a.v || (ov == a.v && oi < a.i)) { a.v = ov; a.i = oi; }
  }
  return a;
}
__device__ __forceinline__ V shflV(V v, int src) { return vv(__shfl(v.x, src), __shfl(v.y, src), __shfl(v.z, src)); }
/** queryFaces with lane l taking faces l, l + 64, ... */
__device__ __forceinline__ void queryFacesWave(const uint4 *hulls, const Poly &A, const Poly &B, uint32_t lane, float &sep, uint32_t &face) {
  ArgBest best{-3.4e38f, 0xffffffffu};
  for (uint32_t f = lane; f < A.nf; f += 64) {
    V n;
    float d;
    facePlane(hulls, A, f, n, d);
    float s = dot(n, polySupport(hulls, B, neg(n))) - d;
    if (s > best.v) { best.v = s; best.i = f; }
  }
  best = waveArgMax(best);
  sep = best.v;
  face = best.i == 0xffffffffu ? 0u : best.i;
}
/** queryEdges with lane l taking edge pairs k = i·B.ne + j = l, l + 64, ... */
__device__ __forceinline__ EdgeQuery queryEdgesWave(const uint4 *hulls, const Poly &A, const Poly &B, uint32_t lane) {
  Q toB = qmul(qconj(B.q), A.q);
  ArgBest best{-3.4e38f, 0xffffffffu};
  V bestN = vv(0.0f, 0.0f, 0.0f);
  uint32_t total = A.ne * B.ne;
  for (uint32_t k = lane; k < total; k += 64) {
    uint32_t i = k / B.ne, j = k - i * B.ne;
    uint4 ea = edgeOf(hulls, A, i);
    V a = qrotate(toB, faceNormalLocal(hulls, A, ea.z)), b = qrotate(toB, faceNormalLocal(hulls, A, ea.w));
    uint4 eb = edgeOf(hulls, B, j);
    if (!minkowskiFace(a, b, neg(faceNormalLocal(hulls, B, eb.z)), neg(faceNormalLocal(hulls, B, eb.w)))) continue;
    V pa = vert(hulls, A, ea.x), qa = vert(hulls, A, ea.y), pb = vert(hulls, B, eb.x), qb = vert(hulls, B, eb.y);
    V n = cross(sub(qa, pa), sub(qb, pb));
    float l = length(n);
    if (l < 1e-5f * length(sub(qa, pa)) * length(sub(qb, pb))) continue;
    n = divs(n, l);
    if (dot(n, sub(pa, A.c)) < 0.0f) n = neg(n);
    float s = dot(n, sub(pb, pa));
    if (s > best.v) { best.v = s; best.i = k; bestN = n; }
  }
  ArgBest w = waveArgMax(best);
  EdgeQuery out;
  out.valid = w.i != 0xffffffffu;
  out.sep = w.v;
  out.ea = out.valid ? w.i / B.ne : 0u;
  out.eb = out.valid ? w.i - out.ea * B.ne : 0u;
  out.n = out.valid ? shflV(bestN, (int)(w.i & 63u)) : vv(0.0f, 0.0f, 0.0f);
  if (!out.valid) out.sep = -3.4e38f;
  return out;
}
/** Per-wavefront staging of one pair's hull data in LDS: both shapes' vertices in local and world
 *  frames and their edges' vertex indices; per edge of B the two negated face normals of its
 *  Gauss-map arc and their cross product (d x c); a list of edge pairs whose arcs cross. Sized for
 *  hulls of up to 32 vertices (at most 90 edges), the hull builder's limit; a larger shape would
 *  take the unstaged path. */
static const uint32_t STAGE_V = 32, STAGE_E = 96, STAGE_LIST = 256;
struct EdgeStage {
  float4 nB[2 * STAGE_E], xB[STAGE_E];
  uint2 eA[STAGE_E], eB[STAGE_E];
  uint32_t list[STAGE_LIST];
};
/** After the edge query: the face contact's clip buffer and both shapes' faces. */
struct ContactStage {
  float4 clip[64];
  FaceStage f[2];
};
struct HullStage {
  float4 vA[STAGE_V], vB[STAGE_V], wA[STAGE_V], wB[STAGE_V];
  union {
    EdgeStage e;     // SAT edge query
    ContactStage c;  // face contact (the edge data is dead by then), or a sphere's hull faces
  };
};
__device__ __forceinline__ void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ bool stageable(const Poly &A, const Poly &B) {
  return A.nv <= STAGE_V && B.nv <= STAGE_V && A.ne <= STAGE_E && B.ne <= STAGE_E;
}
/** Stage hull P's faces into fs (a face a lane): false, with nothing to use, for a box or a hull
 *  beyond the stage's sizes (wavefront-uniform). The caller syncs the wavefront before and after. */
__device__ __forceinline__ bool stageFaces(const uint4 *hulls, const Poly &P, uint32_t lane, FaceStage &fs) {
  if (P.hull == NO_HULL || P.nf > STAGE_F || P.nv > STAGE_V) return false;
  uint32_t end = 0;
  if (lane < P.nf) {
    uint4 w = hulls[P.fs + 2 * lane + 1];
    V n;
    float off;
    facePlane(hulls, P, lane, n, off);  // P.fst is unset: the global reads
    fs.plane[lane] = f4(n, off);
    fs.info[lane] = (w.x << 16) | (w.y & 0xffffu);
    end = w.x + w.y;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) { uint32_t e = (uint32_t)__shfl_xor((int)end, o); end = e > end ? e : end; }
  if (end > STAGE_I) return false;
  // the index list: word t of the hull's list holds indices 4t..4t+3 (each below STAGE_V)
  uint32_t *words = reinterpret_cast<uint32_t *>(fs.idx);
  for (uint32_t t = lane; 4 * t < end; t += 64) {
    uint4 w = hulls[P.is + t];
    words[t] = w.x | (w.y << 8) | (w.z << 16) | (w.w << 24);
  }
  return true;
}
__device__ __forceinline__ void stagePair(const uint4 *hulls, const Poly &A, const Poly &B, uint32_t lane, HullStage &st) {
  waveSync();  // the previous pair's readers are done
  for (uint32_t v = lane; v < A.nv; v += 64) {
    V l = vertLocal(hulls, A, v);
    st.vA[v] = f4(l, 0.0f);
    st.wA[v] = f4(add(A.c, qrotate(A.q, l)), 0.0f);  // vert()
  }
  for (uint32_t v = lane; v < B.nv; v += 64) {
    V l = vertLocal(hulls, B, v);
    st.vB[v] = f4(l, 0.0f);
    st.wB[v] = f4(add(B.c, qrotate(B.q, l)), 0.0f);
  }
  for (uint32_t e = lane; e < A.ne; e += 64) {
    uint4 ea = edgeOf(hulls, A, e);
    st.e.eA[e] = make_uint2(ea.x, ea.y);
  }
  for (uint32_t e = lane; e < B.ne; e += 64) {
    uint4 eb = edgeOf(hulls, B, e);
    V c = neg(faceNormalLocal(hulls, B, eb.z)), d = neg(faceNormalLocal(hulls, B, eb.w));
    st.e.nB[2 * e] = f4(c, 0.0f);
    st.e.nB[2 * e + 1] = f4(d, 0.0f);
    st.e.xB[e] = f4(cross(d, c), 0.0f);  // minkowskiFace's dxc
    st.e.eB[e] = make_uint2(eb.x, eb.y);
  }
  waveSync();
}
/** polySupport with the shape's local vertices from the stage. */
__device__ __forceinline__ V polySupportStaged(const Poly &P, const float4 *verts, V dir) {
  V l = qrotate(qconj(P.q), dir);
  if (P.hull == NO_HULL) return add(P.c, qrotate(P.q, vv(l.x >= 0.0f ? P.h.x : -P.h.x, l.y >= 0.0f ? P.h.y : -P.h.y, l.z >= 0.0f ? P.h.z : -P.h.z)));
  float best = -3.4e38f;
  V at = vv(0.0f, 0.0f, 0.0f);
  for (uint32_t v = 0; v < P.nv; v++) {
    V p = xyz(verts[v]);
    float s = dot(p, l);
    if (s > best) { best = s; at = p; }
  }
  return add(P.c, qrotate(P.q, at));
}
__device__ __forceinline__ void queryFacesStaged(const uint4 *hulls, const Poly &A, const Poly &B, const float4 *vertsB, uint32_t lane, float &sep,
                                                 uint32_t &face) {
  ArgBest best{-3.4e38f, 0xffffffffu};
  for (uint32_t f = lane; f < A.nf; f += 64) {
    V n;
    float d;
    facePlane(hulls, A, f, n, d);
    float s = dot(n, polySupportStaged(B, vertsB, neg(n))) - d;
    if (s > best.v) { best.v = s; best.i = f; }
  }
  best = waveArgMax(best);
  sep = best.v;
  face = best.i == 0xffffffffu ? 0u : best.i;
}
/** Edge pair k = i·B.ne + j's candidate separation (queryEdges' body past the Gauss-map test),
 *  from the stage; keeps (s, k) in `best` when s is larger, or equal with a lower k. */
__device__ __forceinline__ void edgePairStaged(const Poly &A, const Poly &B, const HullStage &st, uint32_t i, uint32_t j, ArgBest &best,
                                               V &bestN) {
  uint2 ea = st.e.eA[i], eb = st.e.eB[j];
  V pa = xyz(st.wA[ea.x]), qa = xyz(st.wA[ea.y]), pb = xyz(st.wB[eb.x]), qb = xyz(st.wB[eb.y]);
  V n = cross(sub(qa, pa), sub(qb, pb));
  float l = length(n);
  if (l < 1e-5f * length(sub(qa, pa)) * length(sub(qb, pb))) return;
  n = divs(n, l);
  if (dot(n, sub(pa, A.c)) < 0.0f) n = neg(n);
  float s = dot(n, sub(pb, pa));
  uint32_t k = i * B.ne + j;
  if (s > best.v || (s == best.v && k < best.i)) { best.v = s; best.i = k; bestN = n; }
}
/** queryEdges over the wavefront, bit for bit the serial loop: lane l holds A's edges l, l + 64, ...
 *  (arc and arc cross product in registers) and walks B's edges from the stage; the pairs whose
 *  arcs cross (the Gauss-map test, the serial one's arithmetic) are listed in LDS and evaluated a
 *  lane each. The winner is the largest separation, lowest pair index on ties: the serial loop's
 *  first maximum. */
__device__ __forceinline__ EdgeQuery queryEdgesStaged(const uint4 *hulls, const Poly &A, const Poly &B, HullStage &st, uint32_t lane) {
  Q toB = qmul(qconj(B.q), A.q);
  ArgBest best{-3.4e38f, 0xffffffffu};
  V bestN = vv(0.0f, 0.0f, 0.0f), zero = bestN;
  uint64_t below = (1ull << lane) - 1ull;
  uint32_t listed = 0;  // wavefront-uniform
  for (uint32_t i0 = 0; i0 < A.ne; i0 += 64) {
    uint32_t i = i0 + lane;
    bool has = i < A.ne;
    V a = zero, b = zero, bxa = zero;
    if (has) {
      uint4 ea = edgeOf(hulls, A, i);
      a = qrotate(toB, faceNormalLocal(hulls, A, ea.z));
      b = qrotate(toB, faceNormalLocal(hulls, A, ea.w));
      bxa = cross(b, a);
    }
    for (uint32_t j = 0; j < B.ne; j++) {
      V c = xyz(st.e.nB[2 * j]), d = xyz(st.e.nB[2 * j + 1]), dxc = xyz(st.e.xB[j]);
      float cba = dot(c, bxa), dba = dot(d, bxa), adc = dot(a, dxc), bdc = dot(b, dxc);
      bool crosses = has && cba * dba < 0.0f && adc * bdc < 0.0f && cba * bdc > 0.0f;
      uint64_t m = __ballot(crosses);
      if (m == 0) continue;
      if (crosses) st.e.list[listed + (uint32_t)__popcll(m & below)] = (i << 16) | j;
      listed += (uint32_t)__popcll(m);
      if (listed > STAGE_LIST - 64) {  // room for one more round of 64
        waveSync();
        for (uint32_t t = lane; t < listed; t += 64) { uint32_t ij = st.e.list[t]; edgePairStaged(A, B, st, ij >> 16, ij & 0xffffu, best, bestN); }
        waveSync();
        listed = 0;
      }
    }
  }
  waveSync();
  for (uint32_t t = lane; t < listed; t += 64) { uint32_t ij = st.e.list[t]; edgePairStaged(A, B, st, ij >> 16, ij & 0xffffu, best, bestN); }
  ArgBest w = waveArgMax(best);
  EdgeQuery out;
  out.valid = w.i != 0xffffffffu;
  out.sep = out.valid ? w.v : -3.4e38f;
  out.ea = out.valid ? w.i / B.ne : 0u;
  out.eb = out.valid ? w.i - out.ea * B.ne : 0u;
  uint64_t owner = __ballot(out.valid && best.i == w.i);
  out.n = owner ? shflV(bestN, (int)__ffsll((unsigned long long)owner) - 1) : vv(0.0f, 0.0f, 0.0f);
  return out;
}
/** Exclusive prefix and total over the wavefront of per-lane counts in {0, 1, 2}. */
__device__ __forceinline__ uint32_t wavePrefix2(uint32_t lane, uint32_t count, uint32_t &total) {
  uint64_t one = __ballot(count >= 1u), two = __ballot(count == 2u), below = (1ull << lane) - 1ull;
  total = (uint32_t)(__popcll(one) + __popcll(two));
  return (uint32_t)(__popcll(one & below) + __popcll(two & below));
}
/** hullFaceClip over the wavefront, bit for bit the serial one: lane m holds polygon vertex m; a
 *  clip against a side plane has lane i emit edge (i-1, i)'s crossing and then vertex i if inside,
 *  placed by an exclusive prefix through `buf` (64 entries of LDS, this wavefront's) in the
 *  serial order, the first MAX_CLIP kept. The kept points stay on their lanes (in polygon order,
 *  as the serial arrays), so each pick is a (score, lane) argmax, lowest lane on ties. The contact
 *  is valid in lane 0. Every lane of the wavefront must take part. */
__device__ __noinline__ void hullFaceClipWave(const uint4 *hulls, const Poly &A, const Poly &B, bool refIsA, uint32_t refFace, const Poly &R,
                                                 const Poly &I, V rn, float pw, uint32_t inc, uint32_t lane, float4 *buf, const float4 *worldR,
                                                 const float4 *worldI, Found &found) {
  found.count = 0;
  V zero = vv(0.0f, 0.0f, 0.0f);
  uint32_t n = faceVertCount(hulls, I, inc);
  if (n > MAX_CLIP) n = MAX_CLIP;
  // world vertices from the stage (vert()'s own arithmetic), one face corner per lane
  V p = lane < n ? xyz(worldI[faceVert(hulls, I, inc, lane)]) : zero;
  uint32_t k = faceVertCount(hulls, R, refFace);
  V corner = lane < k ? xyz(worldR[faceVert(hulls, R, refFace, lane)]) : zero;
  V centre = zero;
  for (uint32_t j = 0; j < k; j++) centre = add(centre, shflV(corner, (int)j));  // the serial sum's order
  centre = divs(centre, (float)k);
  for (uint32_t j = 0; j < k && n > 0; j++) {
    V v0 = shflV(corner, (int)j), v1 = shflV(corner, (int)((j + 1) % k));
    V sn = cross(sub(v1, v0), rn);
    float l = length(sn);
    if (l < 1.0e-6f) continue;
    sn = divs(sn, l);
    if (dot(sn, sub(centre, v0)) > 0.0f) sn = neg(sn);
    float off = dot(sn, v0);
    // clipHull: lane i takes edge (src[i-1], src[i]) (src[n-1] before src[0])
    V a = shflV(p, (int)((lane + n - 1) % n));
    float da = dot(sn, a) - off, db = dot(sn, p) - off;
    bool act = lane < n;
    bool cross_ = act && ((da <= 1.0e-5f) != (db <= 1.0e-5f)), inside = act && db <= 1.0e-5f;
    V x = zero;
    if (cross_) {
      float t = 0.0f;
      if (fabsf(da - db) > 1.0e-6f) t = clampf(da / (da - db), 0.0f, 1.0f);
      x = add(a, mul(sub(p, a), t));
    }
    uint32_t total, at = wavePrefix2(lane, (cross_ ? 1u : 0u) + (inside ? 1u : 0u), total);
    waveSync();  // the last round's reads of buf are done
    if (cross_ && at < MAX_CLIP) buf[at] = f4(x, 0.0f);
    if (inside && at + (cross_ ? 1u : 0u) < MAX_CLIP) buf[at + (cross_ ? 1u : 0u)] = f4(p, 0.0f);
    waveSync();
    n = total < MAX_CLIP ? total : MAX_CLIP;
    p = lane < n ? xyz(buf[lane]) : zero;
  }
  float dist = dot(rn, p) - pw;
  bool kept = lane < n && dist <= 1.0e-5f;
  uint64_t keptMask = __ballot(kept);
  uint32_t keep = (uint32_t)__popcll(keptMask);
  uint32_t prefix = ((refIsA ? 0u : 1u) << 24) | ((refFace < 255u ? refFace : 255u) << 16) | ((inc < 255u ? inc : 255u) << 8);
  uint64_t chosen = 0;  // over lanes
  uint32_t picks = keep < 8u ? keep : 8u;
  for (uint32_t c = 0; c < picks; c++) {
    float score = -dist;
    if (c > 0) {
      score = 3.4e38f;
      for (uint64_t left = chosen; left; left &= left - 1ull) {  // chosen points in increasing order
        V dd = sub(p, shflV(p, (int)__ffsll((unsigned long long)left) - 1));
        score = mn(score, dot(dd, dd));
      }
    }
    bool cand = kept && !((chosen >> lane) & 1ull);
    ArgBest best{-3.4e38f, 0xffffffffu};
    if (cand && score > best.v) { best.v = score; best.i = lane; }
    best = waveArgMax(best);
    uint32_t pick = best.i == 0xffffffffu ? (uint32_t)__ffsll((unsigned long long)keptMask) - 1u : best.i;  // the serial loop's pick 0
    chosen |= 1ull << pick;
    V pt = shflV(p, (int)pick);
    V onRef = sub(pt, mul(rn, __shfl(dist, (int)pick)));
    uint32_t idx = pick < 255u ? pick : 255u;
    if (lane == 0) addFound(found, refIsA ? onRef : pt, refIsA ? pt : onRef, prefix | idx);
  }
  if (lane == 0 && found.count == 0) {
    V nAB = refIsA ? rn : neg(rn);
    addFound(found, polySupport(hulls, A, nAB), polySupport(hulls, B, neg(nAB)), prefix);
  }
}
/** hullFaceContact with the deepest-vertex and incident-face searches spread over the lanes
 *  (first minimum, as the serial loops); the clip and selection too when the wavefront has a
 *  stage (its world vertices, and nB as the clip's buffer), else lane 0 clips and selects. */
__device__ __forceinline__ void hullFaceContactWave(const uint4 *hulls, const Poly &A, const Poly &B, bool refIsA, uint32_t refFace, uint32_t lane,
                                                    HullStage *st, Found &found) {
  found.count = 0;
  const Poly &R = refIsA ? A : B;
  const Poly &I = refIsA ? B : A;
  const float4 *worldR = st ? (refIsA ? st->wA : st->wB) : nullptr, *worldI = st ? (refIsA ? st->wB : st->wA) : nullptr;
  V rn;
  float pw;
  facePlane(hulls, R, refFace, rn, pw);
  ArgBest low{3.4e38f, 0xffffffffu};
  for (uint32_t v = lane; v < I.nv; v += 64) {
    float dd = dot(worldI ? xyz(worldI[v]) : vert(hulls, I, v), rn);
    if (dd < low.v) { low.v = dd; low.i = v; }
  }
  low = waveArgMin(low);
  uint32_t deepest = low.i == 0xffffffffu ? 0u : low.i;
  ArgBest least{3.4e38f, 0xffffffffu};
  for (uint32_t f = lane; f < I.nf; f += 64) {
    bool touches = false;
    uint32_t cnt = faceVertCount(hulls, I, f);
    for (uint32_t j = 0; j < cnt; j++) if (faceVert(hulls, I, f, j) == deepest) { touches = true; break; }
    if (!touches) continue;
    V fn;
    float fd;
    facePlane(hulls, I, f, fn, fd);
    float dd = dot(fn, rn);
    if (dd < least.v) { least.v = dd; least.i = f; }
  }
  least = waveArgMin(least);
  uint32_t inc = least.i == 0xffffffffu ? 0u : least.i;
  if (st) {
    hullFaceClipWave(hulls, A, B, refIsA, refFace, R, I, rn, pw, inc, lane, st->c.clip, worldR, worldI, found);
    return;
  }
  if (lane != 0) return;
  hullFaceClip(hulls, A, B, refIsA, refFace, R, I, rn, pw, inc, found);
}
/** collidePoly with the SAT queries spread over the wavefront; lane 0 builds the contact. */
__device__ __forceinline__ void collidePolyWave(const uint4 *hulls, const Poly &A0, const Poly &B0, bool faceBias, uint32_t lane, HullStage *st,
                                                Sat &sat, Found &found) {
  found.count = 0;
  float faSep, fbSep;
  uint32_t faFace, fbFace;
  Poly A = A0, B = B0;
  bool staged = st != nullptr && stageable(A, B);
  if (staged) stagePair(hulls, A, B, lane, *st);
  if (staged) queryFacesStaged(hulls, A, B, st->vB, lane, faSep, faFace);
  else queryFacesWave(hulls, A, B, lane, faSep, faFace);
  if (faSep > 0.0f) return;
  if (staged) queryFacesStaged(hulls, B, A, st->vA, lane, fbSep, fbFace);
  else queryFacesWave(hulls, B, A, lane, fbSep, fbFace);
  if (fbSep > 0.0f) return;
  EdgeQuery eq = staged ? queryEdgesStaged(hulls, A, B, *st, lane) : queryEdgesWave(hulls, A, B, lane);
  if (eq.valid && eq.sep > 0.0f) return;
  float faceSep = mx(faSep, fbSep);
  bool edgeWins = eq.valid && eq.sep > 0.95f * faceSep + 0.01f;
  if (!faceBias) edgeWins = eq.valid && 0.95f * eq.sep > faceSep + 0.01f;
  if (edgeWins) {
    if (lane == 0) { sat.n = eq.n; hullEdgeContact(hulls, A, B, eq, found); }
    return;
  }
  if (staged) {  // both shapes' faces over the dead edge data, for the face contact
    waveSync();
    if (stageFaces(hulls, A, lane, st->c.f[0])) A.fst = &st->c.f[0];
    if (stageFaces(hulls, B, lane, st->c.f[1])) B.fst = &st->c.f[1];
    waveSync();
  }
  V n;
  float d;
  if (fbSep > faSep) {
    facePlane(hulls, B, fbFace, n, d);
    sat.n = neg(n);
    hullFaceContactWave(hulls, A, B, false, fbFace, lane, staged ? st : nullptr, found);
    return;
  }
  facePlane(hulls, A, faFace, n, d);
  sat.n = n;
  hullFaceContactWave(hulls, A, B, true, faFace, lane, staged ? st : nullptr, found);
}
/** sphereHull with the faces spread over the wavefront (the result is valid in every lane). */
__device__ __forceinline__ SphereHit sphereHullWave(const uint4 *hulls, V c, float r, const Poly &X, uint32_t lane, const float4 *world = nullptr) {
  // a corner of face f: from X's staged world vertices when there are (vert()'s own arithmetic)
  auto corner = [&](uint32_t f, uint32_t j) { uint32_t v = faceVert(hulls, X, f, j); return world ? xyz(world[v]) : vert(hulls, X, v); };
  SphereHit out;
  out.hit = false;
  ArgBest deep{-3.4e38f, 0xffffffffu};
  for (uint32_t f = lane; f < X.nf; f += 64) {
    V n;
    float d;
    facePlane(hulls, X, f, n, d);
    float s = dot(n, c) - d;
    if (s > deep.v) { deep.v = s; deep.i = f; }
  }
  deep = waveArgMax(deep);
  float deepest = deep.v;
  uint32_t deepFace = deep.i == 0xffffffffu ? 0u : deep.i;
  if (deepest > r) return out;
  if (deepest <= 0.0f) {
    V n;
    float d;
    facePlane(hulls, X, deepFace, n, d);
    out.hit = true;
    out.n = neg(n);
    out.onBox = sub(c, mul(n, deepest));
    out.onSphere = sub(c, mul(n, r));
    return out;
  }
  ArgBest near{3.4e38f, 0xffffffffu};
  V best = vv(0.0f, 0.0f, 0.0f);
  for (uint32_t f = lane; f < X.nf; f += 64) {
    V pn;
    float pd;
    facePlane(hulls, X, f, pn, pd);
    float s = dot(pn, c) - pd;
    if (s <= 0.0f) continue;
    V p = sub(c, mul(pn, s));
    uint32_t k = faceVertCount(hulls, X, f);
    V centre = vv(0.0f, 0.0f, 0.0f);
    for (uint32_t j = 0; j < k; j++) centre = add(centre, corner(f, j));
    centre = divs(centre, (float)k);
    bool inside = true;
    V edgeBest = vv(0.0f, 0.0f, 0.0f);
    float edgeDist = 3.4e38f;
    for (uint32_t j = 0; j < k; j++) {
      V v0 = corner(f, j), v1 = corner(f, (j + 1) % k);
      V side = cross(sub(v1, v0), pn);
      if (dot(side, sub(centre, v0)) > 0.0f) side = neg(side);
      if (dot(side, sub(p, v0)) > 0.0f) inside = false;
      V q = segmentPoint(v0, v1, c);
      V dq3 = sub(q, c);
      float dq = dot(dq3, dq3);
      if (dq < edgeDist) { edgeDist = dq; edgeBest = q; }
    }
    if (!inside) p = edgeBest;
    float dist = length(sub(p, c));
    if (dist < near.v) { near.v = dist; near.i = f; best = p; }
  }
  ArgBest w = waveArgMin(near);
  float bestDist = w.v;
  if (w.i != 0xffffffffu) best = shflV(best, (int)(w.i & 63u));
  else best = vv(0.0f, 0.0f, 0.0f);
  if (bestDist > r) return out;
  V n = vv(0.0f, 0.0f, 1.0f);
  if (bestDist > 0.0f) n = divs(sub(c, best), bestDist);
  out.hit = true;
  out.n = neg(n);
  out.onBox = best;
  out.onSphere = sub(c, mul(n, r));
  return out;
}
/** collideHullPair for one wavefront; the contact is valid in lane 0. */
__device__ __forceinline__ void collideHullPairWave(const uint4 *hulls, const Box &A, Q qA, float shA, const Box &B, Q qB, float shB,
                                                    bool sphereA, bool sphereB, bool faceBias, uint32_t lane, HullStage *st, Sat &sat, Found &found) {
  found.count = 0;
  if (sphereA || sphereB) {
    Poly X = sphereA ? polyOf(hulls, B.c, qB, B.h, shB) : polyOf(hulls, A.c, qA, A.h, shA);
    const float4 *world = nullptr;
    if (st && X.nv <= STAGE_V) {  // the hull's world vertices and faces into the stage, once for all its faces
      waveSync();                  // the previous pair's readers are done
      for (uint32_t v = lane; v < X.nv; v += 64) st->wA[v] = f4(vert(hulls, X, v), 0.0f);
      bool faces = stageFaces(hulls, X, lane, st->c.f[0]);
      waveSync();
      world = st->wA;
      if (faces) X.fst = &st->c.f[0];
    }
    if (sphereA) {
      SphereHit h = sphereHullWave(hulls, A.c, A.h.x, X, lane, world);
      sat.n = h.n;
      if (h.hit) addFound(found, h.onSphere, h.onBox, SPHERE_FEATURE);
    } else {
      SphereHit h = sphereHullWave(hulls, B.c, B.h.x, X, lane, world);
      sat.n = neg(h.n);
      if (h.hit) addFound(found, h.onBox, h.onSphere, SPHERE_FEATURE);
    }
    return;
  }
  collidePolyWave(hulls, polyOf(hulls, A.c, qA, A.h, shA), polyOf(hulls, B.c, qB, B.h, shB), faceBias, lane, st, sat, found);
}

}  // namespace avbd
