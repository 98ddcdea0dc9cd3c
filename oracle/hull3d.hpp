// TEST INFRASTRUCTURE ONLY — the oracle's restatement of the hull narrowphase
// (src/avbd3d/gpu/wgsl-hull.ts): Gregorius SAT (faces by support points, edge pairs whose
// Gauss-map arcs cross), incident-face clipping with deepest-then-spread point selection, edge
// closest points, sphere against hull. Hull buffer format (wgsl-hull.ts:14-20): at header h,
// h: vertex start, count, face start, count; h+1: edge start, count, index start, 0; vertices
// xyz; faces 2 vec4 (normal, offset; first index, count); indices 4 per vec4; edges (a, b, face
// left, face right). Offsets in vec4 units; f32 fields stored as bits.
#pragma once
#include "gpu3d.hpp"

namespace gpuo {

static const uint32_t NO_HULL = 0xffffffffu, MAX_CLIP = 64;
static const uint32_t BOX_FACE_VERTS[24] = {1, 3, 7, 5, 0, 2, 6, 4, 2, 3, 7, 6, 0, 1, 5, 4, 4, 5, 7, 6, 0, 1, 3, 2};
static const uint32_t BOX_EDGES[12][4] = {{0, 1, 3, 5}, {2, 3, 2, 5}, {4, 5, 3, 4}, {6, 7, 2, 4}, {0, 2, 1, 5}, {1, 3, 0, 5},
                                          {4, 6, 1, 4}, {5, 7, 0, 4}, {0, 4, 1, 3}, {1, 5, 0, 3}, {2, 6, 1, 2}, {3, 7, 0, 2}};

template <class T> struct HullOps {
  typedef Ops<T> O; typedef V3<T> V; typedef Q4<T> Q;
  const uint32_t *w;  // hull buffer words
  struct Poly { V c; Q q; V h; uint32_t hull, vs, nv, fs, nf, es, ne, is; };
  static T f(uint32_t bits) { float x; std::memcpy(&x, &bits, 4); return (T)x; }
  Poly polyOf(const Body<T> &b) const {
    Poly P;
    P.c = b.pos; P.q = b.rot; P.h = O::mul(b.size, T(0.5));
    P.hull = NO_HULL; P.nv = 8; P.nf = 6; P.ne = 12; P.vs = P.fs = P.es = P.is = 0;
    if (b.shape == SHAPE_HULL) {
      uint32_t h = b.hull;
      P.hull = h;
      P.vs = w[4 * h]; P.nv = w[4 * h + 1]; P.fs = w[4 * h + 2]; P.nf = w[4 * h + 3];
      P.es = w[4 * h + 4]; P.ne = w[4 * h + 5]; P.is = w[4 * h + 6];
    }
    return P;
  }
  V vertLocal(const Poly &P, uint32_t v) const {
    if (P.hull == NO_HULL) return O::v((v & 1) ? P.h.x : -P.h.x, (v & 2) ? P.h.y : -P.h.y, (v & 4) ? P.h.z : -P.h.z);
    const uint32_t *o = w + 4 * (P.vs + v);
    return O::v(f(o[0]), f(o[1]), f(o[2]));
  }
  V vert(const Poly &P, uint32_t v) const { return O::add(P.c, O::qrotate(P.q, vertLocal(P, v))); }
  void facePlane(const Poly &P, uint32_t fi, V &n, T &off) const {
    V nl; T d;
    if (P.hull == NO_HULL) {
      uint32_t k = fi / 2;
      T s = (fi & 1) ? T(-1) : T(1);
      nl = O::v(T(0), T(0), T(0));
      nl[k] = s;
      d = P.h[k];
    } else {
      const uint32_t *o = w + 4 * (P.fs + 2 * fi);
      nl = O::v(f(o[0]), f(o[1]), f(o[2]));
      d = f(o[3]);
    }
    n = O::qrotate(P.q, nl);
    off = d + O::dot(n, P.c);
  }
  uint32_t faceVertCount(const Poly &P, uint32_t fi) const { return P.hull == NO_HULL ? 4u : w[4 * (P.fs + 2 * fi + 1) + 1]; }
  uint32_t faceVert(const Poly &P, uint32_t fi, uint32_t j) const {
    if (P.hull == NO_HULL) return BOX_FACE_VERTS[fi * 4 + j];
    uint32_t k = w[4 * (P.fs + 2 * fi + 1)] + j;
    return w[4 * (P.is + k / 4) + k % 4];
  }
  void edgeOf(const Poly &P, uint32_t e, uint32_t out[4]) const {
    for (int i = 0; i < 4; i++) out[i] = P.hull == NO_HULL ? BOX_EDGES[e][i] : w[4 * (P.es + e) + i];
  }
  V polySupport(const Poly &P, V dir) const {
    V l = O::qrotate(O::qconj(P.q), dir);
    if (P.hull == NO_HULL) return O::add(P.c, O::qrotate(P.q, O::v(l.x >= T(0) ? P.h.x : -P.h.x, l.y >= T(0) ? P.h.y : -P.h.y, l.z >= T(0) ? P.h.z : -P.h.z)));
    T best = T(-3.4e38);
    V at = O::v(T(0), T(0), T(0));
    for (uint32_t v = 0; v < P.nv; v++) {
      V p = vertLocal(P, v);
      T s = O::dot(p, l);
      if (s > best) { best = s; at = p; }
    }
    return O::add(P.c, O::qrotate(P.q, at));
  }
  void queryFaces(const Poly &A, const Poly &B, T &sep, uint32_t &face) const {
    sep = T(-3.4e38); face = 0;
    for (uint32_t fi = 0; fi < A.nf; fi++) {
      V n; T d;
      facePlane(A, fi, n, d);
      T s = O::dot(n, polySupport(B, O::neg(n))) - d;
      if (s > sep) { sep = s; face = fi; }
    }
  }
  static bool minkowskiFace(V a, V b, V c, V d) {
    V bxa = O::cross(b, a), dxc = O::cross(d, c);
    T cba = O::dot(c, bxa), dba = O::dot(d, bxa), adc = O::dot(a, dxc), bdc = O::dot(b, dxc);
    return cba * dba < T(0) && adc * bdc < T(0) && cba * bdc > T(0);
  }
  V faceNormalLocal(const Poly &P, uint32_t fi) const {
    if (P.hull == NO_HULL) { V n = O::v(T(0), T(0), T(0)); n[fi / 2] = (fi & 1) ? T(-1) : T(1); return n; }
    const uint32_t *o = w + 4 * (P.fs + 2 * fi);
    return O::v(f(o[0]), f(o[1]), f(o[2]));
  }
  struct EdgeQuery { T sep; uint32_t ea, eb; V n; bool valid; };
  EdgeQuery queryEdges(const Poly &A, const Poly &B) const {
    EdgeQuery out; out.sep = T(-3.4e38); out.ea = out.eb = 0; out.n = O::v(T(0), T(0), T(0)); out.valid = false;
    Q toB = O::qmul(O::qconj(B.q), A.q);
    for (uint32_t i = 0; i < A.ne; i++) {
      uint32_t ea[4]; edgeOf(A, i, ea);
      V a = O::qrotate(toB, faceNormalLocal(A, ea[2])), b = O::qrotate(toB, faceNormalLocal(A, ea[3]));
      for (uint32_t j = 0; j < B.ne; j++) {
        uint32_t eb[4]; edgeOf(B, j, eb);
        if (!minkowskiFace(a, b, O::neg(faceNormalLocal(B, eb[2])), O::neg(faceNormalLocal(B, eb[3])))) continue;
        V pa = vert(A, ea[0]), qa = vert(A, ea[1]), pb = vert(B, eb[0]), qb = vert(B, eb[1]);
        V n = O::cross(O::sub(qa, pa), O::sub(qb, pb));
        T l = O::length(n);
        if (l < T(1e-5) * O::length(O::sub(qa, pa)) * O::length(O::sub(qb, pb))) continue;
        n = O::divs(n, l);
        if (O::dot(n, O::sub(pa, A.c)) < T(0)) n = O::neg(n);
        T s = O::dot(n, O::sub(pb, pa));
        if (s > out.sep) { out.sep = s; out.ea = i; out.eb = j; out.n = n; out.valid = true; }
      }
    }
    return out;
  }
  static uint32_t clipHull(const V *src, uint32_t n, V *dst, V pn, T offset) {
    if (n == 0) return 0;
    uint32_t count = 0;
    V a = src[n - 1];
    T da = O::dot(pn, a) - offset;
    for (uint32_t i = 0; i < n; i++) {
      V b = src[i];
      T db = O::dot(pn, b) - offset;
      if ((da <= T(1.0e-5)) != (db <= T(1.0e-5))) {
        T t = T(0);
        if (std::fabs(da - db) > T(1.0e-6)) t = O::clampt(da / (da - db), T(0), T(1));
        if (count < MAX_CLIP) dst[count++] = O::add(a, O::mul(O::sub(b, a), t));
      }
      if (db <= T(1.0e-5) && count < MAX_CLIP) dst[count++] = b;
      a = b; da = db;
    }
    return count;
  }
  typedef typename Model<T>::Found Found;
  Found faceContact(const Poly &A, const Poly &B, bool refIsA, uint32_t refFace) const {
    Found found; found.count = 0;
    const Poly &R = refIsA ? A : B;
    const Poly &I = refIsA ? B : A;
    V rn; T pw;
    facePlane(R, refFace, rn, pw);
    uint32_t deepest = 0;
    T low = T(3.4e38);
    for (uint32_t v = 0; v < I.nv; v++) { T d = O::dot(vert(I, v), rn); if (d < low) { low = d; deepest = v; } }
    uint32_t inc = 0;
    T least = T(3.4e38);
    for (uint32_t fi = 0; fi < I.nf; fi++) {
      bool touches = false;
      for (uint32_t j = 0; j < faceVertCount(I, fi); j++) if (faceVert(I, fi, j) == deepest) { touches = true; break; }
      if (!touches) continue;
      V fn; T fd;
      facePlane(I, fi, fn, fd);
      T d = O::dot(fn, rn);
      if (d < least) { least = d; inc = fi; }
    }
    V poly[64], tmp[64];
    uint32_t n = std::min(faceVertCount(I, inc), MAX_CLIP);
    for (uint32_t j = 0; j < n; j++) poly[j] = vert(I, faceVert(I, inc, j));
    uint32_t k = faceVertCount(R, refFace);
    V centre = O::v(T(0), T(0), T(0));
    for (uint32_t j = 0; j < k; j++) centre = O::add(centre, vert(R, faceVert(R, refFace, j)));
    centre = O::divs(centre, (T)(float)k);
    for (uint32_t j = 0; j < k && n > 0; j++) {
      V v0 = vert(R, faceVert(R, refFace, j)), v1 = vert(R, faceVert(R, refFace, (j + 1) % k));
      V s = O::cross(O::sub(v1, v0), rn);
      T l = O::length(s);
      if (l < T(1.0e-6)) continue;
      s = O::divs(s, l);
      if (O::dot(s, O::sub(centre, v0)) > T(0)) s = O::neg(s);
      n = clipHull(poly, n, tmp, s, O::dot(s, v0));
      for (uint32_t m = 0; m < n; m++) poly[m] = tmp[m];
    }
    T depth[64];
    uint32_t index[64];
    uint32_t keep = 0;
    for (uint32_t m = 0; m < n; m++) {
      T dist = O::dot(rn, poly[m]) - pw;
      if (dist <= T(1.0e-5)) { tmp[keep] = poly[m]; depth[keep] = dist; index[keep] = m; keep++; }
    }
    uint32_t prefix = ((refIsA ? 0u : 1u) << 24) | (std::min(refFace, 255u) << 16) | (std::min(inc, 255u) << 8);
    bool chosen[64] = {false};
    for (uint32_t c = 0; c < std::min(keep, 8u); c++) {
      uint32_t pick = 0;
      T bestScore = T(-3.4e38);
      for (uint32_t m = 0; m < keep; m++) {
        if (chosen[m]) continue;
        T score = -depth[m];
        if (c > 0) {
          score = T(3.4e38);
          for (uint32_t o = 0; o < keep; o++) if (chosen[o]) { V dd = O::sub(tmp[m], tmp[o]); score = O::mn(score, O::dot(dd, dd)); }
        }
        if (score > bestScore) { bestScore = score; pick = m; }
      }
      chosen[pick] = true;
      V pt = tmp[pick];
      V onRef = O::sub(pt, O::mul(rn, depth[pick]));
      Model<T>::addFound(found, refIsA ? onRef : pt, refIsA ? pt : onRef, prefix | std::min(index[pick], 255u));
    }
    if (found.count == 0) {
      V nAB = refIsA ? rn : O::neg(rn);
      Model<T>::addFound(found, polySupport(A, nAB), polySupport(B, O::neg(nAB)), prefix);
    }
    return found;
  }
  Found edgeContact(const Poly &A, const Poly &B, const EdgeQuery &q) const {
    Found found; found.count = 0;
    uint32_t ea[4], eb[4];
    edgeOf(A, q.ea, ea); edgeOf(B, q.eb, eb);
    V p0 = vert(A, ea[0]), q0 = vert(B, eb[0]);
    V d1 = O::sub(vert(A, ea[1]), p0), d2 = O::sub(vert(B, eb[1]), q0), r = O::sub(p0, q0);
    T a = O::dot(d1, d1), e = O::dot(d2, d2), fv = O::dot(d2, r), c = O::dot(d1, r), b = O::dot(d1, d2);
    T denom = a * e - b * b;
    const T EPS = T(1.0e-6);
    T s = T(0);
    if (std::fabs(denom) > EPS * a * e) s = O::clampt((b * fv - c * e) / denom, T(0), T(1));
    T t = (b * s + fv) / O::mx(e, EPS);
    if (t < T(0)) { t = T(0); s = O::clampt(-c / O::mx(a, EPS), T(0), T(1)); }
    else if (t > T(1)) { t = T(1); s = O::clampt((b - c) / O::mx(a, EPS), T(0), T(1)); }
    Model<T>::addFound(found, O::add(p0, O::mul(d1, s)), O::add(q0, O::mul(d2, t)), (2u << 24) | (std::min(q.ea, 255u) << 8) | std::min(q.eb, 255u));
    return found;
  }
  Found collidePoly(const Poly &A, const Poly &B, bool faceBias, typename Model<T>::Sat &sat) const {
    Found none; none.count = 0;
    T faSep, fbSep; uint32_t faFace, fbFace;
    queryFaces(A, B, faSep, faFace);
    if (faSep > T(0)) return none;
    queryFaces(B, A, fbSep, fbFace);
    if (fbSep > T(0)) return none;
    EdgeQuery eq = queryEdges(A, B);
    if (eq.valid && eq.sep > T(0)) return none;
    T faceSep = O::mx(faSep, fbSep);
    bool edgeWins = eq.valid && eq.sep > T(0.95) * faceSep + T(0.01);
    if (!faceBias) edgeWins = eq.valid && T(0.95) * eq.sep > faceSep + T(0.01);
    if (edgeWins) { sat.n = eq.n; return edgeContact(A, B, eq); }
    if (fbSep > faSep) {
      V n; T d;
      facePlane(B, fbFace, n, d);
      sat.n = O::neg(n);
      return faceContact(A, B, false, fbFace);
    }
    V n; T d;
    facePlane(A, faFace, n, d);
    sat.n = n;
    return faceContact(A, B, true, faFace);
  }
  static V segmentPoint(V a, V b, V p) {
    V d = O::sub(b, a);
    return O::add(a, O::mul(d, O::clampt(O::dot(O::sub(p, a), d) / O::mx(O::dot(d, d), T(1e-20)), T(0), T(1))));
  }
  typedef typename Model<T>::SphereHit SphereHit;
  SphereHit sphereHull(V c, T r, const Poly &X) const {
    SphereHit out; out.hit = false;
    T deepest = T(-3.4e38);
    uint32_t deepFace = 0;
    for (uint32_t fi = 0; fi < X.nf; fi++) {
      V n; T d;
      facePlane(X, fi, n, d);
      T s = O::dot(n, c) - d;
      if (s > deepest) { deepest = s; deepFace = fi; }
    }
    if (deepest > r) return out;
    if (deepest <= T(0)) {
      V n; T d;
      facePlane(X, deepFace, n, d);
      out.hit = true;
      out.n = O::neg(n);
      out.onBox = O::sub(c, O::mul(n, deepest));
      out.onSphere = O::sub(c, O::mul(n, r));
      return out;
    }
    V best = O::v(T(0), T(0), T(0));
    T bestDist = T(3.4e38);
    for (uint32_t fi = 0; fi < X.nf; fi++) {
      V pn; T pd;
      facePlane(X, fi, pn, pd);
      T s = O::dot(pn, c) - pd;
      if (s <= T(0)) continue;
      V p = O::sub(c, O::mul(pn, s));
      uint32_t k = faceVertCount(X, fi);
      V centre = O::v(T(0), T(0), T(0));
      for (uint32_t j = 0; j < k; j++) centre = O::add(centre, vert(X, faceVert(X, fi, j)));
      centre = O::divs(centre, (T)(float)k);
      bool inside = true;
      V edgeBest = O::v(T(0), T(0), T(0));
      T edgeDist = T(3.4e38);
      for (uint32_t j = 0; j < k; j++) {
        V v0 = vert(X, faceVert(X, fi, j)), v1 = vert(X, faceVert(X, fi, (j + 1) % k));
        V side = O::cross(O::sub(v1, v0), pn);
        if (O::dot(side, O::sub(centre, v0)) > T(0)) side = O::neg(side);
        if (O::dot(side, O::sub(p, v0)) > T(0)) inside = false;
        V q = segmentPoint(v0, v1, c);
        V dq3 = O::sub(q, c);
        T dq = O::dot(dq3, dq3);
        if (dq < edgeDist) { edgeDist = dq; edgeBest = q; }
      }
      if (!inside) p = edgeBest;
      T dist = O::length(O::sub(p, c));
      if (dist < bestDist) { bestDist = dist; best = p; }
    }
    if (bestDist > r) return out;
    V n = O::v(T(0), T(0), T(1));
    if (bestDist > T(0)) n = O::divs(O::sub(c, best), bestDist);
    out.hit = true;
    out.n = O::neg(n);
    out.onBox = best;
    out.onSphere = O::sub(c, O::mul(n, r));
    return out;
  }
};

/** The hull branch of narrowphase (wgsl-collision.ts:781-795). */
template <class T>
typename Model<T>::Found Model<T>::collideHull(uint32_t a, uint32_t b, bool sphereA, bool sphereB, Sat &sat) const {
  HullOps<T> H;
  H.w = hulls.data();
  Found found; found.count = 0;
  if (sphereA || sphereB) {
    if (sphereA) {
      SphereHit h = H.sphereHull(bodies[a].pos, bodies[a].size.x * T(0.5), H.polyOf(bodies[b]));
      sat.n = h.n;
      if (h.hit) addFound(found, h.onSphere, h.onBox, SPHERE_FEATURE);
    } else {
      SphereHit h = H.sphereHull(bodies[b].pos, bodies[b].size.x * T(0.5), H.polyOf(bodies[a]));
      sat.n = Ops<T>::neg(h.n);
      if (h.hit) addFound(found, h.onBox, h.onSphere, SPHERE_FEATURE);
    }
    return found;
  }
  return H.collidePoly(H.polyOf(bodies[a]), H.polyOf(bodies[b]), (p.flags & FLAG_FACE_BIAS) != 0, sat);
}

}  // namespace gpuo
