// TEST INFRASTRUCTURE ONLY — the oracle's restatement of the WebGPU step's semantics.
//
// The reference's GPU step (GpuSolver3D.step, src/avbd3d/gpu/solver.ts:2112-2293) differs from
// its CPU solver in documented ways (colour-parallel Gauss-Seidel, grid broadphase predicate,
// sphere/hull shapes, pivot floor, warm-start heuristics). This file restates those semantics
// from the WGSL (src/avbd3d/gpu/wgsl-collision.ts, wgsl-solve.ts, wgsl-hull.ts, layout.ts
// PRELUDE_3D, src/avbd2d/gpu/wgsl-topology.ts) on the CPU, templated on the float type: in f32
// it is the bit-exactness reference for pairs, contacts and colourings; in f64 it is the
// finite-difference reference for the adjoint. Every operation is written in the order the
// kernels evaluate it; build with -ffp-contract=off.
//
// What is deliberately canonical here and not in the reference: pairs and manifolds are kept
// sorted by (a, b) (the reference's order depends on atomics; nothing numeric depends on it,
// since adjacency lists are sorted before the solve).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <unordered_map>
#include <functional>
#include <omp.h>

namespace gpuo {

// layout.ts / avbd2d layout.ts constants
static const uint32_t STICK_BIT = 0x80000000u, SPHERE_FEATURE = 3u << 24, SDF_ROLLING_BIT = 0x20000000u;
static const uint32_t FLAG_MATCH_NEAREST = 1, FLAG_FACE_BIAS = 2, FLAG_REUSE_CONTACTS = 4, FLAG_START_AT_REST = 8,
                      FLAG_MASS_PENALTY = 16, FLAG_REUSE_MANIFOLDS = 64;
static const int SHAPE_BOX = 0, SHAPE_SPHERE = 1, SHAPE_HULL = 3;
static const uint32_t NO_COLOR = 255, PENDING = 256, MAX_COLORS = 64;
static const uint32_t T_NONE = 0, T_JOINT = 1, T_SPRING = 2, IGNORE_SLOT = 0xffffffffu;
static const double BIGD = 3.0e38, HARDD = 1e30;

static inline uint32_t hash32(uint32_t x) {  // lowbias32 (avbd2d/gpu/layout.ts:107-115)
  x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
  return x;
}

template <class T> struct V3 { T x, y, z; T &operator[](int i) { return (&x)[i]; } T operator[](int i) const { return (&x)[i]; } };
template <class T> struct Q4 { T x, y, z, w; };

template <class T> struct Ops {
  typedef V3<T> V; typedef Q4<T> Q;
  static V v(T x, T y, T z) { V r; r.x = x; r.y = y; r.z = z; return r; }
  static Q q(T x, T y, T z, T w) { Q r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
  static V add(V a, V b) { return v(a.x + b.x, a.y + b.y, a.z + b.z); }
  static V sub(V a, V b) { return v(a.x - b.x, a.y - b.y, a.z - b.z); }
  static V mul(V a, T s) { return v(a.x * s, a.y * s, a.z * s); }
  static V mulv(V a, V b) { return v(a.x * b.x, a.y * b.y, a.z * b.z); }
  static V divs(V a, T s) { return v(a.x / s, a.y / s, a.z / s); }
  static V neg(V a) { return v(-a.x, -a.y, -a.z); }
  static V vabs(V a) { return v(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)); }
  static T dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
  static T length(V a) { return std::sqrt(dot(a, a)); }
  static V normalize(V a) { return divs(a, length(a)); }
  static V cross(V a, V b) { return v(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
  static T mn(T a, T b) { return a < b ? a : b; }
  static T mx(T a, T b) { return a > b ? a : b; }
  static T clampt(T x, T lo, T hi) { return mn(mx(x, lo), hi); }
  static T dot4(Q a, Q b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
  // PRELUDE_3D helpers (layout.ts:307-340)
  static Q qmul(Q a, Q b) {
    return q(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
             a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
  }
  static Q qconj(Q a) { return q(-a.x, -a.y, -a.z, a.w); }
  static V qsub(Q a, Q b) {
    T d = dot4(b, b);
    Q inv = q(-b.x / d, -b.y / d, -b.z / d, b.w / d);
    Q r = qmul(a, inv);
    return v(r.x * T(2), r.y * T(2), r.z * T(2));
  }
  static Q qnormalize(Q a) { T l = std::sqrt(dot4(a, a)); return q(a.x / l, a.y / l, a.z / l, a.w / l); }
  static Q qadd(Q a, V w) {
    Q m = qmul(q(w.x, w.y, w.z, T(0)), a);
    return qnormalize(q(a.x + m.x * T(0.5), a.y + m.y * T(0.5), a.z + m.z * T(0.5), a.w + m.w * T(0.5)));
  }
  static V qrotate(Q r, V a) {
    V u = v(r.x, r.y, r.z);
    V t = mul(cross(u, a), T(2));
    return add(add(a, mul(t, r.w)), cross(u, t));
  }
  static void orthonormal(V n, V basis[3]) {
    V t1 = std::fabs(n.x) > std::fabs(n.z) ? v(-n.y, n.x, T(0)) : v(T(0), -n.z, n.y);
    t1 = normalize(t1);
    basis[0] = n; basis[1] = t1; basis[2] = cross(n, t1);
  }
};

/** Body record fields (layout.ts:219-230) in the model's float type, plus the bitcast words. */
template <class T> struct Body {
  V3<T> pos; T friction; Q4<T> rot; V3<T> ipos; uint32_t lifecycle; Q4<T> irot;
  V3<T> size; T mass; V3<T> moment; T radius; V3<T> inpos; uint32_t stamp; Q4<T> inrot;
  V3<T> vel; T prevUp; V3<T> angvel; int shape; uint32_t hull;  // hull: header offset (vec4 units) when shape == HULL
  // the aux record (GpuSolver3D.setBodyAux, solver.ts:1814): gravity scale, the per-step velocity
  // decay factors exp(-rate dt) as f32, and whether either rate is nonzero
  T auxScale = T(1), auxLin = T(1), auxAng = T(1); bool damped = false;
  /** lifecycleBits / bodyEnabled / bodyCold (layout.ts:26-32): 1 rewritten slot, 2 disabled, 4 cold. */
  uint32_t bits() const { return lifecycle <= 7u ? lifecycle : 0u; }
  bool enabled() const { return (bits() & 2u) == 0u; }
  bool cold() const { return (bits() & 4u) != 0u; }
};
template <class T> struct Joint {  // layout.ts:232-243 (+ info)
  V3<T> penLin; T stiffLin; V3<T> penAng; T stiffAng; V3<T> lamLin; T fracture; V3<T> lamAng; T torqueArm;
  V3<T> c0Lin; T pad0; V3<T> c0Ang; T yield; V3<T> rA; T springRest; V3<T> rB; T breakBend; Q4<T> rest; Q4<T> restStart;
  int type, a, b;
};
struct Manifold { uint32_t a, b, first, countStep; float nx, ny, nz, friction; };
template <class T> struct ManifoldT { uint32_t a, b, first, countStep; V3<T> n; T friction; };
template <class T> struct Contact { V3<T> rA; uint32_t key; V3<T> rB; T c0x; V3<T> pen; T c0y; V3<T> lam; T c0z; };

struct Params {  // the subset of Params (layout.ts:265-305) the semantics depend on
  double dt, gravity, betaLin, betaAng, gamma, alpha, contactMargin, up[3];
  uint32_t flags, iterations, colorCap, colorRounds, step;
  uint32_t lanesRule[4];  // colour sizes from which 1, 2, 4, 8 lanes per body are used (else 16)
};

/** One step's frozen decisions, recorded in f64 whatever the model's float type (the adjoint's
 *  tape): the contact set as collision produced it, the warm-start source of every contact, the
 *  colouring and adjacency, and per iteration the penalties, the clamp/cone branches of every
 *  contact evaluation and the pivot-floor shifts of every body update. */
struct TapeContact { uint32_t a, b, key; double rA[3], rB[3], n[3], friction, pen0[3], lam0[3], c0[3]; int src; };
struct TapeStep {
  std::vector<TapeContact> contacts;
  std::vector<uint32_t> manifoldFirst;              // contacts of manifold m: [first[m], first[m+1])
  std::vector<uint32_t> adjStart, adjList, bucketStart, bucketBodies;
  uint32_t colorCapUsed = 0, clashes = 0;
  std::vector<std::vector<double>> pen;             // [iteration][contact*3 + row]
  std::vector<std::vector<uint8_t>> branch;         // [iteration][contact*3 + role]: bit0 normal clamped, bit1 sliding
  std::vector<std::vector<double>> delta;           // [iteration][body*6 + pivot]: pivot-floor shift
  std::vector<uint8_t> atRest, wMode;               // per body: initial guess held; 0 w interior, 1 w = 0, 2 w = 1
  // the forward's own states (xyz or xyzw per body / contact), for a reverse sweep at them
  std::vector<double> p0, q0, v0, w0, u0, inpos, inrot;  // at the step's start (after collision)
  std::vector<std::vector<double>> P, R, lam;            // P/R: [0] initial guess, [it+1] after sweep it; lam: [it], [iters] final
};
struct Tape { std::vector<TapeStep> steps; };

template <class T> struct Model {
  typedef Ops<T> O; typedef V3<T> V; typedef Q4<T> Q;
  Tape *tape = nullptr;  // when set, step() appends its frozen decisions
  int curIt = -1;
  Params p;
  std::vector<Body<T>> bodies;
  std::vector<Joint<T>> joints;
  std::vector<uint32_t> filterGroups, filterMask;
  std::vector<uint32_t> hulls;  // hull buffer words (vec4u units * 4), wgsl-hull.ts header format
  struct NC { uint32_t hi, lo, slot; };
  std::vector<NC> noCollide;  // sorted by (hi, lo)
  std::vector<V> refPos; std::vector<Q> refRot; std::vector<uint8_t> refSet;  // bytes: written by parallel workers
  std::vector<std::pair<uint32_t, uint32_t>> pairs;
  std::vector<ManifoldT<T>> manifolds, prevManifolds;
  std::vector<Contact<T>> contacts, prevContacts;
  std::vector<int> contactSrc;  // per contact: the previous step's contact it was warm-started from, or -1
  std::vector<uint32_t> adjStart, adjList;
  std::vector<uint32_t> colorA, colorB;  // colour state ping-pong (wgsl-topology.ts)
  std::vector<uint32_t> bucketStart, bucketBodies;  // colourStart[65], bodies by colour
  uint32_t clashMask[2] = {0, 0};
  uint32_t numColors = 0, clashes = 0, colorCapUsed = 0;
  bool fixedColors = false;
  std::vector<uint32_t> fixed;
  uint32_t stepCount = 0;
  /** Host threads (OpenMP) for the CPU baseline: bodies of a proper colour, contacts, pairs and
   *  colouring rounds are independent, so the result equals the one-thread result bit for bit.
   *  Clashing colours and tape recording run serially. */
  int threads = 1;

  T dtT() const { return T(p.dt); }
  bool dynamic(int i) const { return i >= 0 && bodies[i].mass > T(0); }

  // ---- broadphase (wgsl-collision.ts:150-234): the exact pair predicate ----
  struct Shape { V ax[3]; V h; bool sphere; };
  Shape shapeOf(uint32_t i) const {
    const Body<T> &b = bodies[i];
    Shape s;
    s.h = O::mul(b.size, T(0.5));
    s.sphere = b.shape == SHAPE_SPHERE;
    s.ax[0] = O::qrotate(b.rot, O::v(T(1), T(0), T(0)));
    s.ax[1] = O::qrotate(b.rot, O::v(T(0), T(1), T(0)));
    s.ax[2] = O::qrotate(b.rot, O::v(T(0), T(0), T(1)));
    return s;
  }
  static T extent(const Shape &s, V n) {
    if (s.sphere) return s.h.x;
    return s.h.x * std::fabs(O::dot(n, s.ax[0])) + s.h.y * std::fabs(O::dot(n, s.ax[1])) + s.h.z * std::fabs(O::dot(n, s.ax[2]));
  }
  static V aabbHalf(const Shape &s) {
    if (s.sphere) return O::v(s.h.x, s.h.x, s.h.x);
    return O::add(O::add(O::mul(O::vabs(s.ax[0]), s.h.x), O::mul(O::vabs(s.ax[1]), s.h.y)), O::mul(O::vabs(s.ax[2]), s.h.z));
  }
  static bool faceSeparated(const Shape &A, const Shape &B, V d) {
    for (int k = 0; k < 3; k++) {
      if (std::fabs(O::dot(d, A.ax[k])) - extent(A, A.ax[k]) - extent(B, A.ax[k]) > T(1e-3)) return true;
      if (std::fabs(O::dot(d, B.ax[k])) - extent(A, B.ax[k]) - extent(B, B.ax[k]) > T(1e-3)) return true;
    }
    return false;
  }
  bool ignored(uint32_t hi, uint32_t lo) const {
    auto it = std::lower_bound(noCollide.begin(), noCollide.end(), NC{hi, lo, 0},
                               [](const NC &x, const NC &y) { return x.hi < y.hi || (x.hi == y.hi && x.lo < y.lo); });
    for (; it != noCollide.end() && it->hi == hi && it->lo == lo; ++it) {
      if (it->slot == IGNORE_SLOT) return true;
      const Joint<T> &j = joints[it->slot];
      if (j.stiffLin != T(0) || j.stiffAng != T(0)) return true;
    }
    return false;
  }
  /** testPair(P, j) with P the probe of body i. */
  bool testPair(uint32_t i, uint32_t j) const {
    if (!bodies[i].enabled() || !bodies[j].enabled()) return false;  // findPairs / testPair skip disabled bodies
    if (!dynamic(i) && !dynamic(j)) return false;
    if ((filterGroups[i] & filterMask[j]) == 0 || (filterGroups[j] & filterMask[i]) == 0) return false;
    V d = O::sub(bodies[i].pos, bodies[j].pos);
    T r = bodies[i].radius + bodies[j].radius;
    if (O::dot(d, d) > r * r) return false;
    Shape A = shapeOf(i), B = shapeOf(j);
    V ha = aabbHalf(A), hb = aabbHalf(B);
    V lim = O::add(O::add(ha, hb), O::v(T(1e-4), T(1e-4), T(1e-4)));
    V ad = O::vabs(d);
    if (ad.x > lim.x || ad.y > lim.y || ad.z > lim.z) return false;
    if (faceSeparated(A, B, d)) return false;
    if (ignored(std::max(i, j), std::min(i, j))) return false;
    return true;
  }
  /** The pair set of the predicate over every (i, j), i > j, sorted by (a, b); a grid of
   *  bounding-sphere candidates keeps it fast, and candidates only ever pass the exact test. */
  void broadphase() {
    pairs.clear();
    const uint32_t n = (uint32_t)bodies.size();
    if (n < 2) return;
    std::vector<double> rs(n);
    for (uint32_t i = 0; i < n; i++) rs[i] = (double)bodies[i].radius;
    std::vector<double> sorted = rs;
    std::sort(sorted.begin(), sorted.end());
    double cut = sorted[std::min<size_t>(n - 1, (size_t)(0.99 * (n - 1)))] * 2.0;
    std::vector<uint32_t> large;
    double cell = 0;
    for (uint32_t i = 0; i < n; i++) { if (rs[i] > cut) large.push_back(i); else cell = std::max(cell, 2 * rs[i]); }
    cell = std::max(cell, 1e-3) * (1 + 1e-6);
    std::unordered_map<uint64_t, std::vector<uint32_t>> grid;
    auto key = [](int64_t x, int64_t y, int64_t z) { return (uint64_t)((x & 0x1fffff) | ((y & 0x1fffff) << 21) | ((z & 0x1fffff) << 42)); };
    std::vector<int64_t> cx(n), cy(n), cz(n);
    for (uint32_t i = 0; i < n; i++) {
      if (rs[i] > cut) continue;
      cx[i] = (int64_t)std::floor((double)bodies[i].pos.x / cell); cy[i] = (int64_t)std::floor((double)bodies[i].pos.y / cell); cz[i] = (int64_t)std::floor((double)bodies[i].pos.z / cell);
      grid[key(cx[i], cy[i], cz[i])].push_back(i);
    }
    std::vector<std::pair<uint32_t, uint32_t>> out;
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> part(threads);
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads) if (threads > 1)
    for (uint32_t i = 0; i < n; i++) {
      if (rs[i] > cut) continue;
      auto &mine = part[threads > 1 ? omp_get_thread_num() : 0];
      for (int dz = -1; dz <= 1; dz++) for (int dy = -1; dy <= 1; dy++) for (int dx = -1; dx <= 1; dx++) {
        auto it = grid.find(key(cx[i] + dx, cy[i] + dy, cz[i] + dz));
        if (it == grid.end()) continue;
        for (uint32_t j : it->second) if (j < i && testPair(i, j)) mine.push_back({i, j});
      }
    }
    for (auto &v : part) out.insert(out.end(), v.begin(), v.end());
    for (uint32_t l : large)
      for (uint32_t j = 0; j < n; j++) {
        if (j == l) continue;
        bool jl = rs[j] > cut;
        if (jl && j > l) continue;  // large-large once
        uint32_t a = std::max(l, j), b = std::min(l, j);
        if (testPair(a, b)) out.push_back({a, b});
      }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    pairs = out;
  }

  // ---- narrowphase (wgsl-collision.ts:383-683) ----
  struct Box { V c, h, ax[3]; };
  struct Sat { uint32_t kind, ia, ib; T sep; V n; bool valid; };
  struct Found { uint32_t feature[8]; V xA[8], xB[8]; uint32_t count; };
  Box makeBox(uint32_t i) const {
    Box b;
    b.c = bodies[i].pos;
    b.h = O::mul(bodies[i].size, T(0.5));
    b.ax[0] = O::qrotate(bodies[i].rot, O::v(T(1), T(0), T(0)));
    b.ax[1] = O::qrotate(bodies[i].rot, O::v(T(0), T(1), T(0)));
    b.ax[2] = O::qrotate(bodies[i].rot, O::v(T(0), T(0), T(1)));
    return b;
  }
  static T projRadius(const Box &b, V n) {
    return b.h.x * std::fabs(O::dot(n, b.ax[0])) + b.h.y * std::fabs(O::dot(n, b.ax[1])) + b.h.z * std::fabs(O::dot(n, b.ax[2]));
  }
  static bool testAxis(const Box &A, const Box &B, V delta, V axis, uint32_t kind, uint32_t ia, uint32_t ib, Sat &best) {
    T lenSq = O::dot(axis, axis);
    if (lenSq < T(1.0e-6)) return true;
    V n = O::mul(axis, T(1) / std::sqrt(lenSq));
    if (O::dot(n, delta) < T(0)) n = O::neg(n);
    T sep = std::fabs(O::dot(delta, n)) - (projRadius(A, n) + projRadius(B, n));
    if (sep > T(0)) return false;
    if (!best.valid || sep > best.sep) best = Sat{kind, ia, ib, sep, n, true};
    return true;
  }
  static V support(const Box &b, V dir) {
    T sx = O::dot(dir, b.ax[0]) >= T(0) ? T(1) : T(-1), sy = O::dot(dir, b.ax[1]) >= T(0) ? T(1) : T(-1), sz = O::dot(dir, b.ax[2]) >= T(0) ? T(1) : T(-1);
    return O::add(O::add(O::add(b.c, O::mul(b.ax[0], b.h.x * sx)), O::mul(b.ax[1], b.h.y * sy)), O::mul(b.ax[2], b.h.z * sz));
  }
  static void addFound(Found &f, V xA, V xB, uint32_t feature) {
    V mid = O::mul(O::add(xA, xB), T(0.5));
    for (uint32_t i = 0; i < f.count; i++) {
      V d = O::sub(mid, O::mul(O::add(f.xA[i], f.xB[i]), T(0.5)));
      if (O::dot(d, d) < T(1.0e-6)) return;
    }
    if (f.count >= 8) return;
    f.feature[f.count] = feature; f.xA[f.count] = xA; f.xB[f.count] = xB; f.count++;
  }
  static uint32_t clipPlane(const V *src, uint32_t n, V *dst, V pn, T offset) {
    if (n == 0) return 0;
    uint32_t count = 0;
    V a = src[n - 1];
    T da = O::dot(pn, a) - offset;
    for (uint32_t i = 0; i < n; i++) {
      V b = src[i];
      T db = O::dot(pn, b) - offset;
      bool aIn = da <= T(1.0e-5), bIn = db <= T(1.0e-5);
      if (aIn != bIn) {
        T t = T(0), denom = da - db;
        if (std::fabs(denom) > T(1.0e-6)) t = O::clampt(da / denom, T(0), T(1));
        if (count < 8) dst[count++] = O::add(a, O::mul(O::sub(b, a), t));
      }
      if (bIn && count < 8) dst[count++] = b;
      a = b; da = db;
    }
    return count;
  }
  static uint32_t faceU(uint32_t k) { return k == 0 ? 1 : 0; }
  static uint32_t faceV(uint32_t k) { return k != 2 ? 2 : 1; }
  static Found faceManifold(const Box &A, const Box &B, bool refIsA, uint32_t refAxis, V nAB) {
    Found found; found.count = 0;
    const Box &R = refIsA ? A : B;
    const Box &I = refIsA ? B : A;
    V outward = refIsA ? nAB : O::neg(nAB);
    T rs = O::dot(outward, R.ax[refAxis]) >= T(0) ? T(1) : T(-1);
    V rn = O::mul(R.ax[refAxis], rs);
    V rc = O::add(R.c, O::mul(rn, R.h[refAxis]));
    V ru = R.ax[faceU(refAxis)], rv = R.ax[faceV(refAxis)];
    T eu = R.h[faceU(refAxis)], ev = R.h[faceV(refAxis)];
    uint32_t inc = 0; T bestDot = T(-1);
    for (uint32_t k = 0; k < 3; k++) { T d = std::fabs(O::dot(I.ax[k], rn)); if (d > bestDot) { bestDot = d; inc = k; } }
    T is = O::dot(I.ax[inc], rn) > T(0) ? T(-1) : T(1);
    V ic = O::add(I.c, O::mul(I.ax[inc], is * I.h[inc]));
    V iu = O::mul(I.ax[faceU(inc)], I.h[faceU(inc)]), iv = O::mul(I.ax[faceV(inc)], I.h[faceV(inc)]);
    V poly[8], tmp[8];
    poly[0] = O::add(O::add(ic, iu), iv);
    poly[1] = O::add(O::sub(ic, iu), iv);
    poly[2] = O::sub(O::sub(ic, iu), iv);
    poly[3] = O::sub(O::add(ic, iu), iv);
    uint32_t n = 4;
    n = clipPlane(poly, n, tmp, ru, O::dot(ru, rc) + eu);
    n = clipPlane(tmp, n, poly, O::neg(ru), O::dot(O::neg(ru), rc) + eu);
    n = clipPlane(poly, n, tmp, rv, O::dot(rv, rc) + ev);
    n = clipPlane(tmp, n, poly, O::neg(rv), O::dot(O::neg(rv), rc) + ev);
    if (n == 0) return found;
    uint32_t prefix = ((refIsA ? 0u : 1u) << 24) | (refAxis << 16) | (inc << 8);
    for (uint32_t i = 0; i < n && found.count < 8; i++) {
      V pt = poly[i];
      T dist = O::dot(O::sub(pt, rc), rn);
      if (dist > T(1.0e-5)) continue;
      V onRef = O::sub(pt, O::mul(rn, dist));
      addFound(found, refIsA ? onRef : pt, refIsA ? pt : onRef, prefix | i);
    }
    if (found.count == 0) addFound(found, support(A, nAB), support(B, O::neg(nAB)), prefix);
    return found;
  }
  static void supportEdge(const Box &b, uint32_t k, V dir, V &e0, V &e1) {
    uint32_t k1 = (k + 1) % 3, k2 = (k + 2) % 3;
    T s1 = O::dot(dir, b.ax[k1]) >= T(0) ? T(1) : T(-1), s2 = O::dot(dir, b.ax[k2]) >= T(0) ? T(1) : T(-1);
    V c = O::add(O::add(b.c, O::mul(b.ax[k1], b.h[k1] * s1)), O::mul(b.ax[k2], b.h[k2] * s2));
    V e = O::mul(b.ax[k], b.h[k]);
    e0 = O::sub(c, e); e1 = O::add(c, e);
  }
  static Found edgeContact(const Box &A, const Box &B, uint32_t ia, uint32_t ib, V nAB) {
    Found found; found.count = 0;
    V a0, a1, b0, b1;
    supportEdge(A, ia, nAB, a0, a1);
    supportEdge(B, ib, O::neg(nAB), b0, b1);
    V p0 = a0, q0 = b0, d1 = O::sub(a1, p0), d2 = O::sub(b1, q0), r = O::sub(p0, q0);
    T a = O::dot(d1, d1), e = O::dot(d2, d2), f = O::dot(d2, r), s = T(0), t = T(0);
    const T EPS = T(1.0e-6);
    if (a <= EPS && e <= EPS) {
    } else if (a <= EPS) {
      t = O::clampt(f / e, T(0), T(1));
    } else {
      T c = O::dot(d1, r);
      if (e <= EPS) s = O::clampt(-c / a, T(0), T(1));
      else {
        T b = O::dot(d1, d2), denom = a * e - b * b;
        if (std::fabs(denom) > EPS) s = O::clampt((b * f - c * e) / denom, T(0), T(1));
        t = (b * s + f) / e;
        if (t < T(0)) { t = T(0); s = O::clampt(-c / a, T(0), T(1)); }
        else if (t > T(1)) { t = T(1); s = O::clampt((b - c) / a, T(0), T(1)); }
      }
    }
    uint32_t key = (2u << 24) | (ia << 8) | ib;
    addFound(found, O::add(p0, O::mul(d1, s)), O::add(q0, O::mul(d2, t)), key);
    if (found.count == 0) addFound(found, support(A, nAB), support(B, O::neg(nAB)), key);
    return found;
  }
  Found collideBoxes(const Box &A, const Box &B, Sat &sat) const {
    Found none; none.count = 0;
    V delta = O::sub(B.c, A.c);
    Sat face{0, 0, 0, T(0), O::v(T(0), T(0), T(0)), false}, edge = face;
    for (uint32_t i = 0; i < 3; i++) if (!testAxis(A, B, delta, A.ax[i], 0, i, 0, face)) return none;
    for (uint32_t i = 0; i < 3; i++) if (!testAxis(A, B, delta, B.ax[i], 1, 0, i, face)) return none;
    for (uint32_t i = 0; i < 3; i++)
      for (uint32_t j = 0; j < 3; j++)
        if (!testAxis(A, B, delta, O::cross(A.ax[i], B.ax[j]), 2, i, j, edge)) return none;
    if (!face.valid) return none;
    Sat best = face;
    if (edge.valid) {
      bool edgeWins = T(0.95) * edge.sep > face.sep + T(0.01);
      if (p.flags & FLAG_FACE_BIAS) edgeWins = edge.sep > T(0.95) * face.sep + T(0.01);
      if (edgeWins) best = edge;
    }
    sat = best;
    if (best.kind == 2) return edgeContact(A, B, best.ia, best.ib, best.n);
    return faceManifold(A, B, best.kind == 0, best.kind == 0 ? best.ia : best.ib, best.n);
  }
  static Found sphereSphere(const Box &A, const Box &B, Sat &sat) {
    Found found; found.count = 0;
    V d = O::sub(B.c, A.c);
    T dist = O::length(d);
    if (dist > A.h.x + B.h.x) return found;
    V n = O::v(T(0), T(0), T(1));
    if (dist > T(0)) n = O::divs(d, dist);
    sat.n = n;
    addFound(found, O::add(A.c, O::mul(n, A.h.x)), O::sub(B.c, O::mul(n, B.h.x)), SPHERE_FEATURE);
    return found;
  }
  struct SphereHit { bool hit; V onSphere, onBox, n; };
  static SphereHit sphereBox(const Box &S, const Box &X, Q qX) {
    SphereHit out; out.hit = false;
    V pl = O::qrotate(O::qconj(qX), O::sub(S.c, X.c));
    V ql = O::v(O::clampt(pl.x, -X.h.x, X.h.x), O::clampt(pl.y, -X.h.y, X.h.y), O::clampt(pl.z, -X.h.z, X.h.z));
    T r = S.h.x;
    V nLocal, onBox = ql;
    if (pl.x != ql.x || pl.y != ql.y || pl.z != ql.z) {
      V d = O::sub(pl, ql);
      T dist = O::length(d);
      if (dist > r) return out;
      nLocal = O::divs(d, dist);
    } else {
      V depth = O::sub(X.h, O::vabs(pl));
      int k = 0;
      if (depth.y < depth[k]) k = 1;
      if (depth.z < depth[k]) k = 2;
      T s = pl[k] >= T(0) ? T(1) : T(-1);
      nLocal = O::v(T(0), T(0), T(0));
      nLocal[k] = s;
      onBox[k] = s * X.h[k];
    }
    V n = O::qrotate(qX, nLocal);
    out.hit = true;
    out.n = O::neg(n);
    out.onBox = O::add(X.c, O::qrotate(qX, onBox));
    out.onSphere = O::sub(S.c, O::mul(n, r));
    return out;
  }
  // Hull shapes: the oracle's hull narrowphase (hull3d.hpp) when compiled in.
  Found collideHull(uint32_t a, uint32_t b, bool sphereA, bool sphereB, Sat &sat) const;

  T contactPenaltyMin(uint32_t a, uint32_t b) const {
    if (!(p.flags & FLAG_MASS_PENALTY)) return T(1.0);
    T mA = bodies[a].mass <= T(0) ? T(3.4e38) : bodies[a].mass;
    T mB = bodies[b].mass <= T(0) ? T(3.4e38) : bodies[b].mass;
    T dt = dtT();
    return O::clampt(O::mn(mA, mB) / (dt * dt), T(1.0), T(1e10));
  }
  std::unordered_map<uint64_t, uint32_t> prevIndex;  // (a, b) -> previous manifold (hashInsert/hashFind)
  static uint64_t pkey(uint32_t a, uint32_t b) { return ((uint64_t)a << 32) | b; }
  int findPrev(uint32_t a, uint32_t b) const {
    auto it = prevIndex.find(pkey(a, b));
    return it == prevIndex.end() ? -1 : (int)it->second;
  }
  /** Where a pass appends manifolds and contacts: the model's own lists, or a worker's private
   *  lists that are concatenated in pair order afterwards (so the result equals the serial one). */
  struct Sink { std::vector<ManifoldT<T>> &man; std::vector<Contact<T>> &con; std::vector<int> &src; };
  Sink own() { return Sink{manifolds, contacts, contactSrc}; }
  void copyContacts(const ManifoldT<T> &prev, Sink o) const {
    uint32_t a = prev.a, b = prev.b, count = prev.countStep & 15u;
    ManifoldT<T> m = prev;
    m.first = (uint32_t)o.con.size();
    V pA = bodies[a].pos, pB = bodies[b].pos;
    Q qA = bodies[a].rot, qB = bodies[b].rot;
    V basis[3];
    O::orthonormal(prev.n, basis);
    T alpha = T(p.alpha), gamma = T(p.gamma), margin = T(p.contactMargin), pmin = contactPenaltyMin(a, b);
    for (uint32_t i = 0; i < count; i++) {
      Contact<T> k = prevContacts[prev.first + i];
      V d = O::sub(O::add(O::qrotate(qA, k.rA), pA), O::add(O::qrotate(qB, k.rB), pB));
      k.c0x = O::dot(basis[0], d) + margin;
      k.c0y = O::dot(basis[1], d);
      k.c0z = O::dot(basis[2], d);
      k.lam = O::mul(O::mul(k.lam, alpha), gamma);
      for (int r = 0; r < 3; r++) k.pen[r] = O::clampt(k.pen[r] * gamma, T(1.0), T(1e10));
      k.pen.x = O::mx(k.pen.x, pmin);
      o.con.push_back(k);
      o.src.push_back((int)(prev.first + i));
    }
    o.man.push_back(m);
  }
  bool reuseContacts(uint32_t a, uint32_t b, Sink o) const {
    int pm = findPrev(a, b);
    if (pm < 0) return false;
    const ManifoldT<T> &prev = prevManifolds[pm];
    uint32_t count = prev.countStep & 15u, generated = prev.countStep >> 4;
    uint32_t moved = std::max(bodies[a].stamp, bodies[b].stamp);
    if (count == 0 || generated < moved) return false;
    copyContacts(prev, o);
    return true;
  }
  void narrowphase(uint32_t a, uint32_t b, Sink o) const {
    if ((p.flags & FLAG_REUSE_CONTACTS) && reuseContacts(a, b, o)) return;
    Box A = makeBox(a), B = makeBox(b);
    Sat sat; sat.n = O::v(T(0), T(0), T(0));
    Found found; found.count = 0;
    bool sphereA = bodies[a].shape == SHAPE_SPHERE, sphereB = bodies[b].shape == SHAPE_SPHERE;
    bool hullA = bodies[a].shape == SHAPE_HULL, hullB = bodies[b].shape == SHAPE_HULL;
    if (sphereA && sphereB) found = sphereSphere(A, B, sat);
    else if (hullA || hullB) found = collideHull(a, b, sphereA, sphereB, sat);
    else if (sphereA || sphereB) {
      if (sphereA) {
        SphereHit h = sphereBox(A, B, bodies[b].rot);
        sat.n = h.n;
        if (h.hit) addFound(found, h.onSphere, h.onBox, SPHERE_FEATURE);
      } else {
        SphereHit h = sphereBox(B, A, bodies[a].rot);
        sat.n = O::neg(h.n);
        if (h.hit) addFound(found, h.onBox, h.onSphere, SPHERE_FEATURE);
      }
    } else found = collideBoxes(A, B, sat);
    if (found.count == 0) return;
    ManifoldT<T> m;
    m.a = a; m.b = b; m.first = (uint32_t)o.con.size(); m.countStep = found.count | (p.step << 4);
    V n = O::neg(sat.n);
    m.n = n;
    m.friction = std::sqrt(bodies[a].friction * bodies[b].friction);
    Q qA = bodies[a].rot, qB = bodies[b].rot;
    V basis[3];
    O::orthonormal(n, basis);
    bool anySphere = sphereA || sphereB;
    uint32_t prevFirst = 0, prevCount = 0;
    int pm = findPrev(a, b);
    if (pm >= 0) { prevFirst = prevManifolds[pm].first; prevCount = prevManifolds[pm].countStep & 15u; }
    bool matchNearest = (p.flags & FLAG_MATCH_NEAREST) != 0;
    T minSide = O::mn(O::mn(O::mn(A.h.x, A.h.y), A.h.z), O::mn(O::mn(B.h.x, B.h.y), B.h.z)) * T(2.0);
    T penaltyMin = contactPenaltyMin(a, b);
    T alpha = T(p.alpha), gamma = T(p.gamma), margin = T(p.contactMargin);
    for (uint32_t i = 0; i < found.count; i++) {
      V rA = O::qrotate(O::qconj(qA), O::sub(found.xA[i], A.c));
      V rB = O::qrotate(O::qconj(qB), O::sub(found.xB[i], B.c));
      V pen = O::v(T(0), T(0), T(0)), lam = pen;
      uint32_t stick = 0;
      int j = -1;
      for (uint32_t c = prevFirst; c < prevFirst + prevCount; c++)
        if ((prevContacts[c].key & ~STICK_BIT) == found.feature[i]) { j = (int)c; break; }
      if (j < 0 && matchNearest) {
        T bestDist = T(0.05) * minSide;
        for (uint32_t c = prevFirst; c < prevFirst + prevCount; c++) {
          T dist = O::length(O::sub(prevContacts[c].rA, rA));
          if (dist <= bestDist) { bestDist = dist; j = (int)c; }
        }
      }
      if (j >= 0) {
        const Contact<T> &pc = prevContacts[j];
        pen = pc.pen; lam = pc.lam; stick = pc.key & STICK_BIT;
        if (stick != 0 && !anySphere) { rA = pc.rA; rB = pc.rB; }
      }
      V d = O::sub(O::add(O::qrotate(qA, rA), A.c), O::add(O::qrotate(qB, rB), B.c));
      Contact<T> k;
      k.rA = rA; k.rB = rB;
      k.key = found.feature[i] | stick;
      k.c0x = O::dot(basis[0], d) + margin;
      k.c0y = O::dot(basis[1], d);
      k.c0z = O::dot(basis[2], d);
      k.lam = O::mul(O::mul(lam, alpha), gamma);
      for (int r = 0; r < 3; r++) k.pen[r] = O::clampt(pen[r] * gamma, T(1.0), T(1e10));
      k.pen.x = O::mx(k.pen.x, penaltyMin);
      o.con.push_back(k);
      o.src.push_back(j);
    }
    o.man.push_back(m);
  }

  /** An installed collision backend (GpuCollisionBackend3D, solver.ts:493-502): replaces the
   *  refs, broadphase and narrowphase; it fills pairs, manifolds and contacts itself. */
  std::function<void(Model &)> collisionHook;
  /** beginFrame + updateRefs + (hashInsert, broadphase, narrowphase) or reuseManifolds. */
  void collision(bool reuseManifolds) {
    prevManifolds.swap(manifolds); prevContacts.swap(contacts);
    manifolds.clear(); contacts.clear(); contactSrc.clear();
    if (collisionHook) {
      pairs.clear();
      collisionHook(*this);
      return;
    }
    if (p.flags & FLAG_REUSE_CONTACTS) {  // updateRefs (wgsl-collision.ts:34-48)
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
      for (size_t i = 0; i < bodies.size(); i++) {
        Body<T> &b = bodies[i];
        if (!b.enabled()) continue;
        bool moved = !refSet[i] || O::length(O::sub(b.pos, refPos[i])) > T(0.002) || O::length(O::qsub(b.rot, refRot[i])) > T(0.002);
        if (moved) { refPos[i] = b.pos; refRot[i] = b.rot; refSet[i] = true; b.stamp = p.step; }
      }
    }
    prevIndex.clear();
    if (reuseManifolds) {
      for (const ManifoldT<T> &m : prevManifolds) if ((m.countStep & 15u) > 0) copyContacts(m, own());
      return;
    }
    for (size_t m = 0; m < prevManifolds.size(); m++) {  // hashInsert skips rewritten slots
      const ManifoldT<T> &pm = prevManifolds[m];
      if (bodies[pm.a].lifecycle != 0 || bodies[pm.b].lifecycle != 0) continue;
      prevIndex[pkey(pm.a, pm.b)] = (uint32_t)m;
    }
    broadphase();
    if (threads <= 1) {
      for (auto &pr : pairs) narrowphase(pr.first, pr.second, own());
      return;
    }
    // contiguous pair ranges per worker, appended in range order: the serial order exactly
    const int parts = threads * 8;
    std::vector<std::vector<ManifoldT<T>>> pm(parts);
    std::vector<std::vector<Contact<T>>> pc(parts);
    std::vector<std::vector<int>> ps(parts);
    const size_t np = pairs.size();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int w = 0; w < parts; w++) {
      size_t lo = np * (size_t)w / parts, hi = np * (size_t)(w + 1) / parts;
      Sink o{pm[w], pc[w], ps[w]};
      for (size_t k = lo; k < hi; k++) narrowphase(pairs[k].first, pairs[k].second, o);
    }
    for (int w = 0; w < parts; w++) {
      uint32_t base = (uint32_t)contacts.size();
      for (ManifoldT<T> m : pm[w]) { m.first += base; manifolds.push_back(m); }
      contacts.insert(contacts.end(), pc[w].begin(), pc[w].end());
      contactSrc.insert(contactSrc.end(), ps[w].begin(), ps[w].end());
    }
  }

  // ---- adjacency (wgsl-topology.ts:63-153) ----
  bool jointActive(uint32_t j) const {
    const Joint<T> &k = joints[j];
    return k.type != (int)T_NONE && (k.stiffLin != T(0) || k.stiffAng != T(0));
  }
  void endpoints(uint32_t id, int &e0, int &e1) const {
    uint32_t J = (uint32_t)joints.size();
    if (id < J) { e0 = joints[id].a; e1 = joints[id].b; }
    else { e0 = (int)manifolds[id - J].a; e1 = (int)manifolds[id - J].b; }
  }
  void adjacency() {
    const uint32_t n = (uint32_t)bodies.size(), J = (uint32_t)joints.size();
    std::vector<std::vector<std::pair<std::pair<uint32_t, uint32_t>, uint32_t>>> lists(n);
    auto addId = [&](uint32_t id) {
      int e[2];
      endpoints(id, e[0], e[1]);
      for (int s = 0; s < 2; s++) {
        int x = e[s];
        if (!dynamic(x)) continue;
        std::pair<uint32_t, uint32_t> key;
        if (id < J) key = {0u, id};
        else {
          int other = e[0] == x ? e[1] : e[0];
          const ManifoldT<T> &m = manifolds[id - J];
          uint32_t sec = (m.countStep & 15u) == 0 ? 0u : (contacts[m.first].key & ~STICK_BIT);
          key = {(uint32_t)(other + 2), sec};
        }
        lists[x].push_back({key, id});
      }
    };
    for (uint32_t j = 0; j < J; j++) if (jointActive(j)) addId(j);
    for (uint32_t m = 0; m < manifolds.size(); m++) addId(J + m);
    adjStart.assign(n + 1, 0);
    adjList.clear();
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads) if (threads > 1)
    for (uint32_t b = 0; b < n; b++)
      std::stable_sort(lists[b].begin(), lists[b].end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    for (uint32_t b = 0; b < n; b++) {
      adjStart[b] = (uint32_t)adjList.size();
      for (auto &e : lists[b]) adjList.push_back(e.second);
    }
    adjStart[n] = (uint32_t)adjList.size();
  }

  // ---- colouring (wgsl-topology.ts:155-328) ----
  static uint32_t priority(uint32_t i) { return hash32(i + 0x9e3779b9u); }
  static bool beats(uint32_t i, uint32_t j) { uint32_t pi = priority(i), pj = priority(j); return pi > pj || (pi == pj && i > j); }
  int neighbour(uint32_t b, uint32_t e) const {
    int e0, e1;
    endpoints(adjList[e], e0, e1);
    int other = e0 == (int)b ? e1 : e0;
    return dynamic(other) ? other : -1;
  }
  uint32_t smallestFree(uint32_t b, const std::vector<uint32_t> &src, bool skipPending) const {
    uint32_t lo = 0, hi = 0;
    for (uint32_t e = adjStart[b]; e < adjStart[b + 1]; e++) {
      int j = neighbour(b, e);
      if (j < 0) continue;
      uint32_t s = src[j];
      if (skipPending && (s & PENDING)) continue;
      uint32_t col = s & 0xffu;
      if (col == NO_COLOR) continue;
      if (col < 32) lo |= 1u << col; else hi |= 1u << (col - 32);
    }
    if (~lo != 0) return (uint32_t)__builtin_ctz(~lo);
    if (~hi != 0) return 32u + (uint32_t)__builtin_ctz(~hi);
    return MAX_COLORS;
  }
  void colour() {
    const uint32_t n = (uint32_t)bodies.size();
    colorA.resize(n, NO_COLOR); colorB.resize(n, NO_COLOR);
    uint32_t cap = p.colorCap;
    if (fixedColors) {
      for (uint32_t b = 0; b < n; b++) colorA[b] = fixed[b];
      cap = 1;
      for (uint32_t b = 0; b < n; b++) if (fixed[b] != NO_COLOR) cap = std::max(cap, fixed[b] + 1);
    } else {
#pragma omp parallel for schedule(dynamic, 512) num_threads(threads) if (threads > 1)
      for (uint32_t b = 0; b < n; b++) {  // colorCompact: A -> B
        if (!dynamic(b)) { colorB[b] = NO_COLOR; continue; }
        uint32_t col = colorA[b] & 0xffu;
        if (col != NO_COLOR) {
          bool isMax = true;
          for (uint32_t e = adjStart[b]; e < adjStart[b + 1] && isMax; e++) { int j = neighbour(b, e); if (j >= 0 && !beats(b, j)) isMax = false; }
          if (isMax) col = std::min(col, smallestFree(b, colorA, false));
        }
        colorB[b] = col;
      }
#pragma omp parallel for schedule(dynamic, 512) num_threads(threads) if (threads > 1)
      for (uint32_t b = 0; b < n; b++) {  // colorMark: B -> A
        uint32_t col = colorB[b] & 0xffu;
        if (!dynamic(b)) { colorA[b] = NO_COLOR; continue; }
        bool pending = col == NO_COLOR;
        for (uint32_t e = adjStart[b]; e < adjStart[b + 1] && !pending; e++) {
          int j = neighbour(b, e);
          if (j >= 0 && (colorB[j] & 0xffu) == col && beats(j, b)) pending = true;
        }
        colorA[b] = col | (pending ? PENDING : 0);
      }
      for (uint32_t r = 0; r < p.colorRounds; r++) {  // colorRoundAB / BA
        std::vector<uint32_t> &src = (r % 2 == 0) ? colorA : colorB;
        std::vector<uint32_t> &dst = (r % 2 == 0) ? colorB : colorA;
#pragma omp parallel for schedule(dynamic, 512) num_threads(threads) if (threads > 1)
        for (uint32_t b = 0; b < n; b++) {
          uint32_t s = src[b];
          if (!(s & PENDING)) { dst[b] = s; continue; }
          bool isMax = true;
          for (uint32_t e = adjStart[b]; e < adjStart[b + 1] && isMax; e++) {
            int j = neighbour(b, e);
            if (j >= 0 && (src[j] & PENDING) && !beats(b, j)) isMax = false;
          }
          if (isMax) { uint32_t f = smallestFree(b, src, true); dst[b] = f >= cap ? ((cap - 1) | PENDING) : f; }
          else dst[b] = s;
        }
      }
    }
    // colorCount / colorStarts / colorScatter: stable buckets, clash masks
    colorCapUsed = cap;
    numColors = 0; clashes = 0; clashMask[0] = clashMask[1] = 0;
    std::vector<uint32_t> counts(MAX_COLORS, 0), colOf(n, NO_COLOR);
    for (uint32_t b = 0; b < n; b++) {
      if (!dynamic(b)) continue;
      uint32_t s = colorA[b], col = s & 0xffu;
      if (s & PENDING) clashes++;
      if (col == NO_COLOR) col = cap - 1;
      colorA[b] = col | (s & PENDING);
      numColors = std::max(numColors, col + 1);
      colOf[b] = col;
      counts[col]++;
      if (s & PENDING) clashMask[col >> 5] |= 1u << (col & 31);
    }
    bucketStart.assign(MAX_COLORS + 1, 0);
    for (uint32_t c = 0; c < MAX_COLORS; c++) bucketStart[c + 1] = bucketStart[c] + counts[c];
    bucketBodies.assign(bucketStart[MAX_COLORS], 0);
    std::vector<uint32_t> fill(bucketStart.begin(), bucketStart.end() - 1);
    for (uint32_t b = 0; b < n; b++) if (colOf[b] != NO_COLOR) bucketBodies[fill[colOf[b]]++] = b;
  }

  // ---- solve (wgsl-solve.ts) ----
  struct Acc { T lin[6], ang[6], cr[9], rLin[3], rAng[3]; };  // lin/ang lower triangles: 00 10 11 20 21 22
  static void zero(Acc &a) { std::memset(&a, 0, sizeof(a)); for (int i = 0; i < 6; i++) { a.lin[i] = T(0); a.ang[i] = T(0); } }
  /** addRow: the row's scaled Jacobian (k l, k a) once, then one fused multiply-add per entry. */
  static void addRow(Acc &acc, V l, V a, T k, T f) {
    static const int I[6] = {0, 1, 1, 2, 2, 2}, J[6] = {0, 0, 1, 0, 1, 2};
    T kl[3] = {l[0] * k, l[1] * k, l[2] * k}, ka[3] = {a[0] * k, a[1] * k, a[2] * k};
    for (int e = 0; e < 6; e++) { acc.lin[e] = std::fma(kl[I[e]], l[J[e]], acc.lin[e]); acc.ang[e] = std::fma(ka[I[e]], a[J[e]], acc.ang[e]); }
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) acc.cr[i * 3 + j] = std::fma(ka[i], l[j], acc.cr[i * 3 + j]);
    for (int i = 0; i < 3; i++) { acc.rLin[i] = std::fma(l[i], f, acc.rLin[i]); acc.rAng[i] = std::fma(a[i], f, acc.rAng[i]); }
  }
  /** a . b as a chain of fused multiply-adds (x first). */
  static T fdot(V a, V b) { return std::fma(a.z, b.z, std::fma(a.y, b.y, a.x * b.x)); }
  static void accAdd(Acc &a, const Acc &b) {
    for (int i = 0; i < 6; i++) { a.lin[i] += b.lin[i]; a.ang[i] += b.ang[i]; }
    for (int i = 0; i < 9; i++) a.cr[i] += b.cr[i];
    for (int i = 0; i < 3; i++) { a.rLin[i] += b.rLin[i]; a.rAng[i] += b.rAng[i]; }
  }
  Q rotAOf(int a) const { return a < 0 ? O::q(T(0), T(0), T(0), T(1)) : bodies[a].rot; }
  V anchorA(const Joint<T> &k, int a) const { return a < 0 ? k.rA : O::add(O::qrotate(bodies[a].rot, k.rA), bodies[a].pos); }
  V jointLinC(const Joint<T> &k) const { return O::sub(anchorA(k, k.a), O::add(O::qrotate(bodies[k.b].rot, k.rB), bodies[k.b].pos)); }
  V jointAngC(const Joint<T> &k) const {
    Q qb = bodies[k.b].rot;
    T dd = O::dot4(qb, qb);
    Q inv = O::q(-qb.x / dd, -qb.y / dd, -qb.z / dd, qb.w / dd);
    Q d = O::qmul(O::qmul(rotAOf(k.a), k.rest), inv);
    if (d.w < T(0)) d = O::q(-d.x, -d.y, -d.z, -d.w);
    return O::mul(O::mul(O::v(d.x, d.y, d.z), T(2.0)), k.torqueArm);
  }
  static bool hard(T s) { return s >= T(HARDD); }
  V jointLinRows(const Joint<T> &k, T alpha) const { return O::sub(jointLinC(k), hard(k.stiffLin) ? O::mul(k.c0Lin, alpha) : O::v(T(0), T(0), T(0))); }
  V jointAngRows(const Joint<T> &k, T alpha) const { return O::sub(jointAngC(k), hard(k.stiffAng) ? O::mul(k.c0Ang, alpha) : O::v(T(0), T(0), T(0))); }
  void addJoint(Acc &acc, uint32_t jid, T alpha, uint32_t i) const {
    const Joint<T> &k = joints[jid];
    bool isA = (int)i == k.a;
    T sg = isA ? T(1) : T(-1);
    if (k.type == (int)T_SPRING) {
      V rAW = O::qrotate(bodies[k.a].rot, k.rA), rBW = O::qrotate(bodies[k.b].rot, k.rB);
      V d = O::sub(O::add(rAW, bodies[k.a].pos), O::add(rBW, bodies[k.b].pos));
      T len = O::length(d);
      if (len <= T(1.0e-6)) return;
      V n = O::divs(d, len);
      T st = k.stiffLin;
      V r = isA ? rAW : rBW;
      addRow(acc, O::mul(n, sg), O::mul(O::cross(r, n), sg), st, st * (len - k.springRest));
      return;
    }
    if (O::dot(k.penLin, k.penLin) > T(0)) {
      V F = O::add(O::mulv(k.penLin, jointLinRows(k, alpha)), k.lamLin);
      V r = isA ? O::qrotate(bodies[k.a].rot, k.rA) : O::neg(O::qrotate(bodies[k.b].rot, k.rB));
      V v = O::neg(r);
      addRow(acc, O::v(sg, T(0), T(0)), O::v(T(0), -v.z, v.y), k.penLin.x, F.x);
      addRow(acc, O::v(T(0), sg, T(0)), O::v(v.z, T(0), -v.x), k.penLin.y, F.y);
      addRow(acc, O::v(T(0), T(0), sg), O::v(-v.y, v.x, T(0)), k.penLin.z, F.z);
      T rf = O::dot(r, F);
      T g0 = O::length(O::sub(O::mul(r, F.x), O::v(rf, T(0), T(0))));
      T g1 = O::length(O::sub(O::mul(r, F.y), O::v(T(0), rf, T(0))));
      T g2 = O::length(O::sub(O::mul(r, F.z), O::v(T(0), T(0), rf)));
      acc.ang[0] += g0; acc.ang[1] += T(0); acc.ang[2] += g1; acc.ang[3] += T(0); acc.ang[4] += T(0); acc.ang[5] += g2;
    }
    if (O::dot(k.penAng, k.penAng) > T(0)) {
      V F = O::add(O::mulv(k.penAng, jointAngRows(k, alpha)), k.lamAng);
      T s = sg * k.torqueArm;
      V z = O::v(T(0), T(0), T(0));
      addRow(acc, z, O::v(s, T(0), T(0)), k.penAng.x, F.x);
      addRow(acc, z, O::v(T(0), s, T(0)), k.penAng.y, F.y);
      addRow(acc, z, O::v(T(0), T(0), s), k.penAng.z, F.z);
    }
  }
  struct PairBody { uint32_t index; Q rot; V dLin, dAng; };
  PairBody pairBody(uint32_t i) const { return PairBody{i, bodies[i].rot, O::sub(bodies[i].pos, bodies[i].ipos), O::qsub(bodies[i].rot, bodies[i].irot)}; }
  struct Eval { V rAW, rBW, C, F; T frictionScale, bounds; uint8_t bits; };
  Eval evalContact(const Contact<T> &k, const V basis[3], T friction, const PairBody &A, const PairBody &B, T alpha) const {
    Q rotA = A.rot, rotB = B.rot;
    if ((k.key & ~STICK_BIT) == SPHERE_FEATURE || (k.key & SDF_ROLLING_BIT) != 0u) { rotA = bodies[A.index].irot; rotB = bodies[B.index].irot; }
    Eval e;
    e.rAW = O::qrotate(rotA, k.rA);
    e.rBW = O::qrotate(rotB, k.rB);
    T keep = T(1.0) - alpha;
    T c0[3] = {k.c0x, k.c0y, k.c0z};
    // C_r = c0 (1 - alpha) + l . dxA - l . dxB + (rA x l) . dthA + (rB x -l) . dthB, as
    // c0 (1 - alpha) + l . g with g = (dxA - dxB) + (dthA x rA - dthB x rB) (scalar triple product)
    V g = O::add(O::sub(A.dLin, B.dLin), O::sub(O::cross(A.dAng, e.rAW), O::cross(B.dAng, e.rBW)));
    for (int r = 0; r < 3; r++) e.C[r] = std::fma(c0[r], keep, fdot(basis[r], g));
    V F = O::add(O::mulv(k.pen, e.C), k.lam);
    e.bits = F.x < T(0.0) ? 0 : 1;
    F.x = O::mn(F.x, T(0.0));
    e.bounds = std::fabs(F.x) * friction;
    e.frictionScale = std::sqrt(F.y * F.y + F.z * F.z);
    if (e.frictionScale > e.bounds && e.frictionScale > T(0.0)) {
      T s = e.bounds / e.frictionScale;
      F.y = F.y * s; F.z = F.z * s;
      e.bits |= 2;
    }
    e.F = F;
    return e;
  }
  uint32_t lanesFor(uint32_t count) const {
    if (count >= p.lanesRule[0]) return 1;
    if (count >= p.lanesRule[1]) return 2;
    if (count >= p.lanesRule[2]) return 4;
    if (count >= p.lanesRule[3]) return 8;
    return 16;
  }
  /** accumulate(i, lane, lanes) (wgsl-solve.ts:436-501): the body's walk items (one per joint
   *  entry and per contact point of each manifold entry, in adjacency order), lane taking items
   *  lane, lane + lanes, ... in order. With one lane that is the reference's order. */
  Acc accumulate(uint32_t i, uint32_t lane, uint32_t lanes, T alpha) const {
    Acc acc; zero(acc);
    const Body<T> &bd = bodies[i];
    T dt = dtT(), dt2 = dt * dt;
    T m = bd.mass / dt2;
    V I = O::divs(bd.moment, dt2);
    if (lane == 0) {
      acc.lin[0] = m; acc.lin[2] = m; acc.lin[5] = m;
      acc.ang[0] = I.x; acc.ang[2] = I.y; acc.ang[5] = I.z;
      V dx = O::sub(bd.pos, bd.inpos);
      acc.rLin[0] = m * dx.x; acc.rLin[1] = m * dx.y; acc.rLin[2] = m * dx.z;
      V dq = O::qsub(bd.rot, bd.inrot);
      acc.rAng[0] = I.x * dq.x; acc.rAng[1] = I.y * dq.y; acc.rAng[2] = I.z * dq.z;
    }
    const uint32_t J = (uint32_t)joints.size();
    uint32_t q = 0;  // the walk item's index
    for (uint32_t e = adjStart[i]; e < adjStart[i + 1]; e++) {
      uint32_t id = adjList[e];
      if (id < J) {
        if (q++ % lanes == lane) addJoint(acc, id, alpha, i);
        continue;
      }
      const ManifoldT<T> &mf = manifolds[id - J];
      bool isA = i == mf.a;
      PairBody A = pairBody(mf.a), B = pairBody(mf.b);
      V basis[3];
      uint32_t c = mf.first, cEnd = mf.first + (mf.countStep & 15u);
      for (; c < cEnd; c++) {
        if (q++ % lanes != lane) continue;
        const Contact<T> &k = contacts[c];
        O::orthonormal(mf.n, basis);
        Eval ev = evalContact(k, basis, mf.friction, A, B, alpha);
        if (tape && curIt >= 0) tape->steps.back().branch[curIt][c * 3 + (isA ? 0 : 1)] = ev.bits;
        T sg = isA ? T(1) : T(-1);
        V r = isA ? ev.rAW : ev.rBW;
        V l0 = O::mul(basis[0], sg), l1 = O::mul(basis[1], sg), l2 = O::mul(basis[2], sg);
        addRow(acc, l0, O::cross(r, l0), k.pen.x, ev.F.x);
        addRow(acc, l1, O::cross(r, l1), k.pen.y, ev.F.y);
        addRow(acc, l2, O::cross(r, l2), k.pen.z, ev.F.z);
      }
    }
    return acc;
  }
  /** finishBody's LDLT of the body's 6x6 system (wgsl-solve.ts:504-539): unpivoted, each pivot
   *  D_k floored at 1e-5 A_kk; S receives each pivot before its floor. x = -A^-1 rhs as (v, w). */
  static void solve6Floor(const Acc &acc, T x[6], T Dout[6], T Sout[6]) {
    const T PF = T(1e-5);
    T A11 = acc.lin[0], A21 = acc.lin[1], A22 = acc.lin[2], A31 = acc.lin[3], A32 = acc.lin[4], A33 = acc.lin[5];
    T A41 = acc.cr[0], A42 = acc.cr[1], A43 = acc.cr[2], A44 = acc.ang[0];
    T A51 = acc.cr[3], A52 = acc.cr[4], A53 = acc.cr[5], A54 = acc.ang[1], A55 = acc.ang[2];
    T A61 = acc.cr[6], A62 = acc.cr[7], A63 = acc.cr[8], A64 = acc.ang[3], A65 = acc.ang[4], A66 = acc.ang[5];
    T D1 = A11;
    T L21 = A21 / D1, L31 = A31 / D1, L41 = A41 / D1, L51 = A51 / D1, L61 = A61 / D1;
    T S2 = A22 - L21 * L21 * D1, D2 = O::mx(S2, A22 * PF);
    T L32 = (A32 - L21 * L31 * D1) / D2, L42 = (A42 - L21 * L41 * D1) / D2, L52 = (A52 - L21 * L51 * D1) / D2, L62 = (A62 - L21 * L61 * D1) / D2;
    T S3 = A33 - (L31 * L31 * D1 + L32 * L32 * D2), D3 = O::mx(S3, A33 * PF);
    T L43 = (A43 - L31 * L41 * D1 - L32 * L42 * D2) / D3, L53 = (A53 - L31 * L51 * D1 - L32 * L52 * D2) / D3, L63 = (A63 - L31 * L61 * D1 - L32 * L62 * D2) / D3;
    T S4 = A44 - (L41 * L41 * D1 + L42 * L42 * D2 + L43 * L43 * D3), D4 = O::mx(S4, A44 * PF);
    T L54 = (A54 - L41 * L51 * D1 - L42 * L52 * D2 - L43 * L53 * D3) / D4, L64 = (A64 - L41 * L61 * D1 - L42 * L62 * D2 - L43 * L63 * D3) / D4;
    T S5 = A55 - (L51 * L51 * D1 + L52 * L52 * D2 + L53 * L53 * D3 + L54 * L54 * D4), D5 = O::mx(S5, A55 * PF);
    T L65 = (A65 - L51 * L61 * D1 - L52 * L62 * D2 - L53 * L63 * D3 - L54 * L64 * D4) / D5;
    T S6 = A66 - (L61 * L61 * D1 + L62 * L62 * D2 + L63 * L63 * D3 + L64 * L64 * D4 + L65 * L65 * D5), D6 = O::mx(S6, A66 * PF);
    T y1 = acc.rLin[0], y2 = acc.rLin[1] - L21 * y1, y3 = acc.rLin[2] - L31 * y1 - L32 * y2;
    T y4 = acc.rAng[0] - L41 * y1 - L42 * y2 - L43 * y3, y5 = acc.rAng[1] - L51 * y1 - L52 * y2 - L53 * y3 - L54 * y4;
    T y6 = acc.rAng[2] - L61 * y1 - L62 * y2 - L63 * y3 - L64 * y4 - L65 * y5;
    T w3 = y6 / D6, w2 = y5 / D5 - L65 * w3, w1 = y4 / D4 - L54 * w2 - L64 * w3;
    T v3 = y3 / D3 - L43 * w1 - L53 * w2 - L63 * w3;
    T v2 = y2 / D2 - L32 * v3 - L42 * w1 - L52 * w2 - L62 * w3;
    T v1 = y1 / D1 - L21 * v2 - L31 * v3 - L41 * w1 - L51 * w2 - L61 * w3;
    x[0] = v1; x[1] = v2; x[2] = v3; x[3] = w1; x[4] = w2; x[5] = w3;
    Dout[0] = D1; Dout[1] = D2; Dout[2] = D3; Dout[3] = D4; Dout[4] = D5; Dout[5] = D6;
    Sout[0] = D1; Sout[1] = S2; Sout[2] = S3; Sout[3] = S4; Sout[4] = S5; Sout[5] = S6;
  }
  /** finishBody (wgsl-solve.ts:504-558): LDLᵀ with the 1e-5 pivot floor, then the update. */
  void finishBody(uint32_t i, const Acc &acc) {
    T x[6], D[6], S[6];
    solve6Floor(acc, x, D, S);
    if (tape && curIt >= 0) {
      double *dl = &tape->steps.back().delta[curIt][(size_t)i * 6];
      for (int k = 1; k < 6; k++) dl[k] = (double)D[k] - (double)S[k];
    }
    T v1 = x[0], v2 = x[1], v3 = x[2], w1 = x[3], w2 = x[4], w3 = x[5];
    Body<T> &bd = bodies[i];
    bd.pos = O::sub(bd.pos, O::v(v1, v2, v3));
    bd.rot = O::qadd(bd.rot, O::neg(O::v(w1, w2, w3)));
  }
  /** One colour's primal: chunks of 64/lanes bodies; each chunk reads every pose it needs
   *  before any body of it is written (the reference's workgroup barriers), so a clashing colour
   *  is exact and a proper one is order-free. */
  void primalColor(uint32_t c, T alpha) {
    uint32_t start = bucketStart[c], count = bucketStart[c + 1] - start;
    if (count == 0) return;
    uint32_t lanes = lanesFor(count), per = 64 / lanes;
    if (threads > 1 && !tape && !((clashMask[c >> 5] >> (c & 31)) & 1u)) {
      // a proper colour: no body of it reads another's pose, so every system is formed first
      // and every update applied after, in parallel
      std::vector<Acc> all(count);
      std::vector<char> on(count);
#pragma omp parallel for schedule(dynamic, 32) num_threads(threads)
      for (uint32_t k = 0; k < count; k++) {
        uint32_t i = bucketBodies[start + k];
        on[k] = bodies[i].enabled();
        if (!on[k]) continue;
        Acc part[16];
        for (uint32_t l = 0; l < lanes; l++) part[l] = accumulate(i, l, lanes, alpha);
        for (uint32_t s = 8; s > 0; s >>= 1)
          if (s < lanes) for (uint32_t l = 0; l < s; l++) accAdd(part[l], part[l + s]);
        all[k] = part[0];
      }
#pragma omp parallel for schedule(static) num_threads(threads)
      for (uint32_t k = 0; k < count; k++)
        if (on[k]) finishBody(bucketBodies[start + k], all[k]);
      return;
    }
    std::vector<Acc> accs(per);
    std::vector<char> live(per);
    for (uint32_t chunk = 0; chunk * per < count; chunk++) {
      uint32_t k0 = chunk * per, k1 = std::min(count, k0 + per);
      for (uint32_t k = k0; k < k1; k++) {
        uint32_t i = bucketBodies[start + k];
        live[k - k0] = bodies[i].enabled();  // a disabled body keeps its pose
        if (!live[k - k0]) continue;
        Acc part[16];
        for (uint32_t l = 0; l < lanes; l++) part[l] = accumulate(i, l, lanes, alpha);
        for (uint32_t s = 8; s > 0; s >>= 1)
          if (s < lanes) for (uint32_t l = 0; l < s; l++) accAdd(part[l], part[l + s]);
        accs[k - k0] = part[0];
      }
      for (uint32_t k = k0; k < k1; k++)
        if (live[k - k0]) finishBody(bucketBodies[start + k], accs[k - k0]);
    }
  }
  void dualJoint(uint32_t j, T alpha) {
    Joint<T> &k = joints[j];
    if (k.type != (int)T_JOINT) return;
    if ((k.a >= 0 && !bodies[k.a].enabled()) || !bodies[k.b].enabled()) return;
    Joint<T> k0 = k;
    T BIG = T(BIGD);
    V penLin = k0.penLin, penAng = k0.penAng, lamLin = k0.lamLin, lamAng = k0.lamAng;
    T stLin = k0.stiffLin, stAng = k0.stiffAng;
    if (O::dot(penLin, penLin) > T(0)) {
      V C = jointLinRows(k0, alpha);
      if (hard(stLin)) lamLin = O::add(O::mulv(penLin, C), lamLin);
      T cap = O::mn(stLin, T(1e10));
      V np = O::add(penLin, O::mul(O::vabs(C), T(p.betaLin)));
      penLin = O::v(O::mn(np.x, cap), O::mn(np.y, cap), O::mn(np.z, cap));
    }
    T frac = k0.fracture, limit = std::fabs(frac);
    bool yielded = false;
    if (O::dot(penAng, penAng) > T(0)) {
      V C = jointAngRows(k0, alpha);
      if (hard(stAng)) {
        lamAng = O::add(O::mulv(penAng, C), lamAng);
        // a plastic joint (wgsl-solve.ts:586-599): past yield, short of fracture, the angular
        // force is cut back to yield and the bend becomes the new rest
        T yieldForce = k0.yield;
        if (yieldForce < BIG) {
          T force = O::length(lamAng);
          if (force > yieldForce && !(limit < BIG && force > limit)) {
            lamAng = O::mul(lamAng, yieldForce / force);
            yielded = true;
          }
        }
      }
      if (!yielded) {  // a yielding row is at its bound: no penalty ramp
        T cap = O::mn(stAng, T(1e10));
        V np = O::add(penAng, O::mul(O::vabs(C), T(p.betaAng)));
        penAng = O::v(O::mn(np.x, cap), O::mn(np.y, cap), O::mn(np.z, cap));
      }
    }
    // the rest moves to where the bodies are; bent past breakBend from the rest it started with,
    // the joint breaks (wgsl-solve.ts:606-614)
    Q rest = k0.rest;
    bool bentOff = false;
    if (yielded) {
      rest = O::qnormalize(O::qmul(O::qconj(rotAOf(k0.a)), bodies[k0.b].rot));
      if (k0.breakBend < BIG) {
        Q dq = O::qmul(O::qconj(k0.restStart), rest);
        bentOff = T(2.0) * std::atan2(O::length(O::v(dq.x, dq.y, dq.z)), std::fabs(dq.w)) > k0.breakBend;
      }
    }
    bool linear = frac < T(0) && O::dot(lamLin, lamLin) > limit * limit;
    if ((limit < BIG && (O::dot(lamAng, lamAng) > limit * limit || linear)) || bentOff) {
      penLin = O::v(T(0), T(0), T(0)); penAng = penLin; lamLin = penLin; lamAng = penLin;
      stLin = T(0); stAng = T(0);
    }
    k.penLin = penLin; k.penAng = penAng; k.lamLin = lamLin; k.lamAng = lamAng; k.stiffLin = stLin; k.stiffAng = stAng;
    if (yielded) { k.rest = rest; k.c0Ang = O::v(T(0), T(0), T(0)); }
  }
  void dualManifold(uint32_t m, T alpha) {
    const ManifoldT<T> &mf = manifolds[m];
    PairBody A = pairBody(mf.a), B = pairBody(mf.b);
    V basis[3];
    O::orthonormal(mf.n, basis);
    uint32_t end = mf.first + (mf.countStep & 15u);
    for (uint32_t c = mf.first; c < end; c++) {
      Contact<T> &k = contacts[c];
      Eval e = evalContact(k, basis, mf.friction, A, B, alpha);
      if (tape && curIt >= 0) tape->steps.back().branch[curIt][c * 3 + 2] = e.bits;
      k.lam = e.F;
      V pen = k.pen;
      T BL = T(p.betaLin), PMAX = T(1e10);
      if (e.F.x < T(0)) pen.x = O::mn(pen.x + BL * std::fabs(e.C.x), PMAX);
      if (e.frictionScale <= e.bounds) {
        pen.y = O::mn(pen.y + BL * std::fabs(e.C.y), PMAX);
        pen.z = O::mn(pen.z + BL * std::fabs(e.C.z), PMAX);
        bool stick = std::sqrt(e.C.y * e.C.y + e.C.z * e.C.z) < T(0.00001);
        k.key = (k.key & ~STICK_BIT) | (stick ? STICK_BIT : 0u);
      }
      k.pen = pen;
    }
  }
  void warmStartJoints() {
    T alpha = T(p.alpha), gamma = T(p.gamma), decay = alpha * gamma;
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
    for (size_t jj = 0; jj < joints.size(); jj++) {
      Joint<T> &k = joints[jj];
      if (k.type != (int)T_JOINT) continue;
      if ((k.a >= 0 && !bodies[k.a].enabled()) || !bodies[k.b].enabled()) continue;
      k.c0Lin = jointLinC(k);
      k.c0Ang = jointAngC(k);
      if ((k.a >= 0 && bodies[k.a].cold()) || bodies[k.b].cold()) {  // a cold body's joints restart
        k.lamLin = O::v(T(0), T(0), T(0));
        k.lamAng = O::v(T(0), T(0), T(0));
        T fl = O::mn(T(1.0), k.stiffLin), fa = O::mn(T(1.0), k.stiffAng);
        k.penLin = O::v(fl, fl, fl);
        k.penAng = O::v(fa, fa, fa);
        continue;
      }
      k.lamLin = O::mul(k.lamLin, decay);
      k.lamAng = O::mul(k.lamAng, decay);
      for (int r = 0; r < 3; r++) {
        k.penLin[r] = O::mn(O::clampt(k.penLin[r] * gamma, T(1.0), T(1e10)), k.stiffLin);
        k.penAng[r] = O::mn(O::clampt(k.penAng[r] * gamma, T(1.0), T(1e10)), k.stiffAng);
      }
    }
  }
  void warmStartBodies() {
    T dt = dtT();
    V up = O::v(T(p.up[0]), T(p.up[1]), T(p.up[2]));
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1 && !tape)
    for (size_t bi = 0; bi < bodies.size(); bi++) {
      Body<T> &b = bodies[bi];
      if (!b.enabled()) {  // untouched, its flag kept
        if (tape) { tape->steps.back().atRest.push_back(0); tape->steps.back().wMode.push_back(1); }
        continue;
      }
      bool dyn = b.mass > T(0);
      T g = T(p.gravity) * b.auxScale;  // the aux variant's scaled gravity (wgsl-solve.ts:300)
      if (dyn && b.damped) {  // DAMPING_WGSL (wgsl-solve.ts:14-20): decayed before use, and stored
        b.vel = O::mul(b.vel, b.auxLin);
        b.angvel = O::mul(b.angvel, b.auxAng);
      }
      V inpos = O::add(b.pos, O::mul(b.vel, dt));
      if (dyn) inpos = O::add(inpos, O::mul(up, g * (dt * dt)));
      b.inpos = inpos;
      b.inrot = O::qadd(b.rot, O::mul(b.angvel, dt));
      T w = T(0);
      uint8_t wMode = 1;
      if (std::fabs(g) > T(0)) {
        T sg = g > T(0) ? T(1) : (g < T(0) ? T(-1) : T(0));
        T raw = (O::dot(b.vel, up) - b.prevUp) / dt * sg / std::fabs(g);
        w = O::clampt(raw, T(0), T(1));
        wMode = raw > T(0) ? (raw < T(1) ? 0 : 2) : 1;
      }
      bool slow = O::length(b.vel) + O::length(b.angvel) * b.radius < T(0.1);
      bool atRest = (p.flags & FLAG_START_AT_REST) && slow;
      if (tape) { tape->steps.back().atRest.push_back(atRest ? 1 : 0); tape->steps.back().wMode.push_back(wMode); }
      b.ipos = b.pos; b.lifecycle = 0; b.irot = b.rot;
      if (dyn && !atRest) {
        b.pos = O::add(O::add(b.pos, O::mul(b.vel, dt)), O::mul(up, g * (w * dt * dt)));
        b.rot = O::qadd(b.rot, O::mul(b.angvel, dt));
      }
      bool moving = b.vel.x != T(0) || b.vel.y != T(0) || b.vel.z != T(0) || b.angvel.x != T(0) || b.angvel.y != T(0) || b.angvel.z != T(0);
      if (!dyn && moving) {
        b.pos = O::add(b.pos, O::mul(b.vel, dt));
        b.rot = O::qadd(b.rot, O::mul(b.angvel, dt));
      }
    }
  }
  void updateVelocities() {
    T dt = dtT();
    V up = O::v(T(p.up[0]), T(p.up[1]), T(p.up[2]));
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
    for (size_t bi = 0; bi < bodies.size(); bi++) {
      Body<T> &b = bodies[bi];
      if (!b.enabled()) continue;
      T prevUp = O::dot(b.vel, up);
      if (b.mass > T(0)) {
        b.vel = O::divs(O::sub(b.pos, b.ipos), dt);
        b.angvel = O::divs(O::qsub(b.rot, b.irot), dt);
      }
      b.prevUp = prevUp;
    }
  }
  void solve() {
    T alpha = T(p.alpha);
    warmStartJoints();
    warmStartBodies();
    TapeStep *ts = tape ? &tape->steps.back() : nullptr;
    if (ts) {
      for (const Body<T> &b : bodies) {
        for (int k = 0; k < 3; k++) {
          ts->p0.push_back((double)b.ipos[k]); ts->inpos.push_back((double)b.inpos[k]);
          ts->v0.push_back((double)b.vel[k]); ts->w0.push_back((double)b.angvel[k]);
        }
        const T *q = &b.irot.x, *iq = &b.inrot.x;
        for (int k = 0; k < 4; k++) { ts->q0.push_back((double)q[k]); ts->inrot.push_back((double)iq[k]); }
        ts->u0.push_back((double)b.prevUp);
      }
      recordPoses(*ts);
    }
    for (uint32_t it = 0; it < p.iterations; it++) {
      curIt = (int)it;
      if (ts) {
        ts->lam.emplace_back();
        for (const Contact<T> &k : contacts) for (int r = 0; r < 3; r++) ts->lam.back().push_back((double)k.lam[r]);
        ts->pen.emplace_back(); ts->branch.emplace_back(contacts.size() * 3, 0); ts->delta.emplace_back(bodies.size() * 6, 0.0);
        for (const Contact<T> &k : contacts) for (int r = 0; r < 3; r++) ts->pen.back().push_back((double)k.pen[r]);
      }
      for (uint32_t c = 0; c < colorCapUsed; c++) primalColor(c, alpha);
      if (ts) recordPoses(*ts);
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
      for (uint32_t j = 0; j < joints.size(); j++) dualJoint(j, alpha);
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads) if (threads > 1)
      for (uint32_t m = 0; m < manifolds.size(); m++) dualManifold(m, alpha);
    }
    curIt = -1;
    if (ts) {
      ts->lam.emplace_back();
      for (const Contact<T> &k : contacts) for (int r = 0; r < 3; r++) ts->lam.back().push_back((double)k.lam[r]);
    }
    updateVelocities();
  }
  void recordPoses(TapeStep &ts) const {
    ts.P.emplace_back(); ts.R.emplace_back();
    for (const Body<T> &b : bodies) {
      for (int k = 0; k < 3; k++) ts.P.back().push_back((double)b.pos[k]);
      const T *q = &b.rot.x;
      for (int k = 0; k < 4; k++) ts.R.back().push_back((double)q[k]);
    }
  }
  void recordTopology(TapeStep &ts) const {
    for (size_t m = 0; m < manifolds.size(); m++) {
      const ManifoldT<T> &mf = manifolds[m];
      ts.manifoldFirst.push_back((uint32_t)ts.contacts.size());
      for (uint32_t c = mf.first; c < mf.first + (mf.countStep & 15u); c++) {
        const Contact<T> &k = contacts[c];
        TapeContact t;
        t.a = mf.a; t.b = mf.b; t.key = k.key; t.src = contactSrc[c];
        t.friction = (double)mf.friction;
        for (int r = 0; r < 3; r++) {
          t.rA[r] = (double)k.rA[r]; t.rB[r] = (double)k.rB[r]; t.n[r] = (double)mf.n[r];
          t.pen0[r] = (double)k.pen[r]; t.lam0[r] = (double)k.lam[r];
        }
        t.c0[0] = (double)k.c0x; t.c0[1] = (double)k.c0y; t.c0[2] = (double)k.c0z;
        ts.contacts.push_back(t);
      }
    }
    ts.manifoldFirst.push_back((uint32_t)ts.contacts.size());
    ts.adjStart = adjStart; ts.adjList = adjList; ts.bucketStart = bucketStart; ts.bucketBodies = bucketBodies;
    ts.colorCapUsed = colorCapUsed; ts.clashes = clashes;
  }
  /** One step (GpuSolver3D.step). */
  void step(bool reuseManifolds = false) {
    stepCount++;
    p.step = stepCount;
    if (tape) tape->steps.emplace_back();
    collision(reuseManifolds);
    adjacency();
    colour();
    if (tape) recordTopology(tape->steps.back());
    solve();
  }
};

}  // namespace gpuo
