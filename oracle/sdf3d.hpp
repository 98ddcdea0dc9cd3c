// TEST INFRASTRUCTURE ONLY — the oracle's restatement of the SDF collision backend.
//
// The reference's second producer of the manifold/contact ABI: SdfCollisionBridge
// (src/avbd3d/gpu/sdf-bridge.ts), installed on the solver by setCollisionBackend (solver.ts:2078)
// in place of the built-in broadphase and narrowphase. Restated from its WGSL: the fields
// (wgsl-sdf-fields.ts), Algorithm 1 per pair with its three optimizers and the mesh vertex
// witnesses (wgsl-sdf.ts, wgsl-sdf-supports.ts), the body preparation, pair filter and contact
// export (wgsl-sdf-bridge.ts) and the spatial candidate grid (wgsl-sdf-broadphase.ts). f32 only,
// each operation in the order the kernels evaluate it; build with -ffp-contract=off.
//
// The pair update is templated on its random stream and seed sampling: with the per-pair hash
// stream it is the GPU semantics (the bit-exactness reference for the HIP backend); with the CPU
// reference's PCG stream and its local-bounds seed filter it is the CPU Algorithm 1
// (src/avbd3d/ref/sdf/manifold.ts, optimize.ts), which the reference's golden fixtures
// (tests/fixtures/oracle-sdf) pin: fields, Nelder-Mead, the ellipsoid method, gradient descent and
// whole manifold updates.
//
// sin, cos and atan2 are this file's own (Cody-Waite reduction with fused multiply-adds and
// truncated Taylor series, the usual published method); the HIP backend evaluates the same
// series, so both round alike. They are checked against the C library's to a few ulp.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>
#include <algorithm>
#include <array>
#include <type_traits>

namespace sdfo {

struct V { float x, y, z; };
struct Q { float x, y, z, w; };
static inline V v3(float x, float y, float z) { V r; r.x = x; r.y = y; r.z = z; return r; }
static inline Q q4(float x, float y, float z, float w) { Q r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
static inline float comp(V a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline void setc(V &a, int i, float x) { if (i == 0) a.x = x; else if (i == 1) a.y = x; else a.z = x; }
static inline V add(V a, V b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V sub(V a, V b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V mul(V a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V mulv(V a, V b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V divs(V a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline V neg(V a) { return v3(-a.x, -a.y, -a.z); }
static inline float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float length(V a) { return std::sqrt(dot(a, a)); }
static inline V cross(V a, V b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float mn(float a, float b) { return a < b ? a : b; }
static inline float mx(float a, float b) { return a > b ? a : b; }
static inline float clampf(float x, float lo, float hi) { return mn(mx(x, lo), hi); }
static inline V vmn(V a, V b) { return v3(mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z)); }
static inline V vmx(V a, V b) { return v3(mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z)); }
static inline V vclamp(V a, V lo, V hi) { return vmn(vmx(a, lo), hi); }
static inline V vabs(V a) { return v3(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)); }
static inline float signnz(float x) { return x < 0.0f ? -1.0f : 1.0f; }
/** unit (wgsl-sdf-fields.ts): v / |v|, or +y for |v| <= 1e-12. */
static inline V unit(V a) { float l = length(a); if (l > 1e-12f) return divs(a, l); return v3(0.0f, 1.0f, 0.0f); }
/** rot (wgsl-sdf-fields.ts) = qrotate (layout.ts:329): t = 2 q.xyz x v; v + q.w t + q.xyz x t. */
static inline V rot(Q q, V a) {
  V u = v3(q.x, q.y, q.z);
  V t = mul(cross(u, a), 2.0f);
  return add(add(a, mul(t, q.w)), cross(u, t));
}
static inline V invrot(Q q, V a) { return rot(q4(-q.x, -q.y, -q.z, q.w), a); }
static inline Q qmul(Q a, Q b) {  // layout.ts:307-313
  return q4(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
/** orthonormal (layout.ts:335-340): the contact basis rows (n, t1, t2). */
static inline void orthonormal(V n, V b[3]) {
  V t1 = std::fabs(n.x) > std::fabs(n.z) ? v3(-n.y, n.x, 0.0f) : v3(0.0f, -n.z, n.y);
  t1 = divs(t1, length(t1));
  b[0] = n; b[1] = t1; b[2] = cross(n, t1);
}
static inline uint32_t mix(uint32_t x) {  // the kernels' hash (lowbias32)
  x = (x ^ (x >> 16)) * 0x7feb352du; x = (x ^ (x >> 15)) * 0x846ca68bu; return x ^ (x >> 16);
}

// ---- trigonometry --------------------------------------------------------------------------
// x = k pi/2 + r, |r| <= pi/4, by three fused multiply-adds with pi/2 split in three floats;
// then the Taylor series of sin to r^9 and of cos to r^10 (truncation below 2.6e-9 relative).
static const float PIO2_1 = 1.57079637050628662109375f, PIO2_2 = -4.371138828673793e-8f, PIO2_3 = -1.7151245100058819e-15f;
static inline float reduce(float x, int &quadrant) {
  float k = std::rint(x * 0.636619746685028076171875f);
  float r = std::fma(-k, PIO2_1, x);
  r = std::fma(-k, PIO2_2, r);
  r = std::fma(-k, PIO2_3, r);
  quadrant = (int)((long long)k & 3);
  return r;
}
static inline float sinPoly(float r) {
  float z = r * r;
  float p = std::fma(z, 2.75573192239858907e-6f, -1.98412698412698413e-4f);
  p = std::fma(z, p, 8.33333333333333333e-3f);
  p = std::fma(z, p, -1.66666666666666667e-1f);
  return std::fma(r * z, p, r);
}
static inline float cosPoly(float r) {
  float z = r * r;
  float p = std::fma(z, -2.75573192239858907e-7f, 2.48015873015873016e-5f);
  p = std::fma(z, p, -1.38888888888888889e-3f);
  p = std::fma(z, p, 4.16666666666666667e-2f);
  p = std::fma(z, p, -0.5f);
  return std::fma(z, p, 1.0f);
}
static inline float ssin(float x) {
  int q; float r = reduce(x, q);
  float s = sinPoly(r), c = cosPoly(r);
  return q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
}
static inline float scos(float x) {
  int q; float r = reduce(x, q);
  float s = sinPoly(r), c = cosPoly(r);
  return q == 0 ? c : (q == 1 ? -s : (q == 2 ? -c : s));
}
/** atan(u) for |u| <= tan(pi/8) by its Taylor series to u^17. */
static inline float atanSmall(float u) {
  float z = u * u;
  float p = std::fma(z, -0.0588235294117647059f, 0.0666666666666666667f);
  p = std::fma(z, p, -0.0769230769230769231f);
  p = std::fma(z, p, 0.0909090909090909091f);
  p = std::fma(z, p, -0.111111111111111111f);
  p = std::fma(z, p, 0.142857142857142857f);
  p = std::fma(z, p, -0.2f);
  p = std::fma(z, p, 0.333333333333333333f);
  return std::fma(-u * z, p, u);
}
/** atan2 by octant reduction: t = min/max in [0, 1]; atan(t) = pi/4 + atan((t-1)/(t+1)) above
 *  tan(pi/8). atan2(0, 0) = 0. */
static inline float satan2(float y, float x) {
  float ax = std::fabs(x), ay = std::fabs(y);
  float hi = mx(ax, ay), lo = mn(ax, ay);
  if (!(hi > 0.0f)) return 0.0f;
  float t = lo / hi;
  float a = t > 0.414213562373095049f ? 0.785398163397448310f + atanSmall((t - 1.0f) / (t + 1.0f)) : atanSmall(t);
  if (ay > ax) a = 1.57079632679489662f - a;
  if (x < 0.0f) a = 3.14159265358979324f - a;
  if (y < 0.0f) a = -a;
  return a;
}

// ---- fields (wgsl-sdf-fields.ts) -------------------------------------------------------------
enum { K_BOX = 0, K_SPHERE = 1, K_CAPSULE = 2, K_CYLINDER = 3, K_CONE = 4, K_TORUS = 5, K_WARP = 6, K_VOXEL = 7,
       K_SPIKE_PLATE = 8, K_SPIKY_WHEEL = 9 };
struct Field {
  uint32_t kind = 0, convex = 0;
  V lo{}, hi{}, com{};
  float prm[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // the atlas words 20..27 (warp: its type in word 24)
  uint32_t warpType = 0;
  uint32_t dims[3] = {0, 0, 0};
  V origin{};
  float cell = 0;
  std::vector<float> grid;  // 4 per node: phi, gradient xyz; x fastest
  std::vector<V> seeds, supports;
};
struct Sample { float phi; V g; };

static inline V radial(V p, float r) { if (r > 1e-9f) return v3(p.x / r, 0.0f, p.z / r); return v3(1.0f, 0.0f, 0.0f); }
static inline Sample boxField(V p, V h) {
  V q = sub(vabs(p), h);
  V o = vmx(q, v3(0.0f, 0.0f, 0.0f));
  float d = length(o);
  V g = v3(0.0f, 0.0f, 0.0f);
  if (d > 0.0f) {
    g = v3((o.x / d) * signnz(p.x), (o.y / d) * signnz(p.y), (o.z / d) * signnz(p.z));
  } else {
    int axis = 2;
    if (q.x > q.y) { if (q.x > q.z) axis = 0; } else if (q.y > q.z) axis = 1;
    setc(g, axis, signnz(comp(p, axis)));
  }
  return Sample{d + mn(mx(q.x, mx(q.y, q.z)), 0.0f), g};
}
static inline Sample sphereField(V p, float r) {
  float d = length(p);
  V g = v3(0.0f, 1.0f, 0.0f);
  if (d > 1e-9f) g = divs(p, d);
  return Sample{d - r, g};
}
static inline Sample cylinderField(V p, float r, float h) {
  float rl = std::sqrt(p.x * p.x + p.z * p.z);
  float dx = rl - r, dy = std::fabs(p.y) - h;
  float ox = mx(dx, 0.0f), oy = mx(dy, 0.0f);
  float l = std::sqrt(ox * ox + oy * oy);
  V e = radial(p, rl);
  V g = v3(0.0f, signnz(p.y), 0.0f);
  if (dx > dy) g = e;
  if (l > 0.0f) g = add(mul(e, ox / l), v3(0.0f, signnz(p.y) * oy / l, 0.0f));
  return Sample{l + mn(mx(dx, dy), 0.0f), g};
}
static inline Sample coneField(V p, float radius, float h) {
  float qx = std::sqrt(p.x * p.x + p.z * p.z), qy = p.y;
  float cax = qx - mn(qx, qy < 0.0f ? radius : 0.0f), cay = std::fabs(qy) - h;
  float kx = -radius, ky = 2.0f * h;
  float t = clampf((qx * radius + (h - qy) * ky) / (kx * kx + ky * ky), 0.0f, 1.0f);
  float cbx = qx + kx * t, cby = (qy - h) + ky * t;
  float s = (cbx < 0.0f && cay < 0.0f) ? -1.0f : 1.0f;
  float da = cax * cax + cay * cay, db = cbx * cbx + cby * cby;
  float gx = cbx, gy = cby;
  if (da < db) { gx = cax; gy = cay * signnz(qy); }
  float l = std::sqrt(gx * gx + gy * gy);
  if (l > 1e-9f) { float f = s / l; gx = gx * f; gy = gy * f; }
  else if (da < db) { gx = 0.0f; gy = signnz(qy); }
  else { float nl = std::sqrt(ky * ky + radius * radius); gx = ky / nl; gy = radius / nl; }
  return Sample{s * std::sqrt(mn(da, db)), add(mul(radial(p, qx), gx), v3(0.0f, gy, 0.0f))};
}
static inline Sample torusField(V p, float major, float minor) {
  float r = std::sqrt(p.x * p.x + p.z * p.z);
  float qx = r - major, qy = p.y;
  float l = std::sqrt(qx * qx + qy * qy);
  V e = radial(p, r), g = e;
  if (l > 1e-9f) g = add(mul(e, qx / l), v3(0.0f, qy / l, 0.0f));
  return Sample{l - minor, g};
}
static inline Sample warpField(V p, V h, float k, uint32_t kind) {
  if (kind == 2u) {
    V s = v3(ssin(p.x * 7.0f), ssin(p.y * 7.0f), ssin(p.z * 7.0f));
    V c = v3(scos(p.x * 7.0f), scos(p.y * 7.0f), scos(p.z * 7.0f));
    Sample b = boxField(p, h);
    float k7 = 7.0f * k;
    V w = v3((c.x * s.y) * s.z, (s.x * c.y) * s.z, (s.x * s.y) * c.z);
    return Sample{b.phi + ((k * s.x) * s.y) * s.z, add(b.g, mul(w, k7))};
  }
  float angle = k * (kind == 0u ? p.x : p.y);
  float c = scos(angle), s = ssin(angle);
  V q = kind == 0u ? v3(c * p.x - s * p.y, s * p.x + c * p.y, p.z) : v3(c * p.x - s * p.z, s * p.x + c * p.z, p.y);
  Sample b = boxField(q, h);
  V g = b.g;
  if (kind == 0u) return Sample{b.phi, v3(dot(v3(c - k * q.y, s + k * q.x, 0.0f), g), dot(v3(-s, c, 0.0f), g), g.z)};
  return Sample{b.phi, v3(dot(v3(c, s, 0.0f), g), dot(v3(-(k * q.y), k * q.x, 1.0f), g), dot(v3(-s, c, 0.0f), g))};
}
static inline float shapeHash01(uint32_t x) {
  x = (x ^ (x >> 16)) * 0x7feb352du; x = (x ^ (x >> 15)) * 0x846ca68bu; x = x ^ (x >> 16);
  return (float)(x & 0xffffffu) / 16777216.0f;
}
/** i32(floor(x)) with x first held to [lo - 1, hi + 1] (in range the same value). */
static inline int floorIndex(float x, int lo, int hi) {
  float f = std::floor(x);
  f = clampf(f, (float)(lo - 1), (float)(hi + 1));
  return (int)f;
}
static inline Sample spikePlateField(V p, const float *a, const float *b) {
  V h = v3(a[0], a[1], a[2]);
  int count = (int)a[3];
  float radius = b[0];
  V pitch = mul(h, 2.0f / a[3]);
  int ci = floorIndex((p.x + h.x) / pitch.x, 0, count - 1), cj = floorIndex((p.z + h.z) / pitch.z, 0, count - 1);
  ci = ci < 0 ? 0 : (ci > count - 1 ? count - 1 : ci);
  cj = cj < 0 ? 0 : (cj > count - 1 ? count - 1 : cj);
  Sample result = boxField(p, h);
  for (int j = std::max(cj - 1, 0); j <= std::min(cj + 1, count - 1); j++)
    for (int i = std::max(ci - 1, 0); i <= std::min(ci + 1, count - 1); i++) {
      float height = b[1] + (b[2] - b[1]) * shapeHash01(17u * ((uint32_t)i + (uint32_t)j * (uint32_t)count) + 11u);
      V center = v3(-h.x + ((float)i + 0.5f) * pitch.x, (h.y + 0.5f * height) - 0.5f * radius, -h.z + ((float)j + 0.5f) * pitch.z);
      Sample c = coneField(sub(p, center), (radius * (height + radius)) / height, 0.5f * (height + radius));
      if (c.phi < result.phi) result = c;
    }
  return result;
}
static inline Sample spikyWheelField(V p, const float *a, float spikes) {
  float radius = a[0], halfWidth = a[1], spikeRadius = a[2], spikeHeight = a[3];
  Sample result = cylinderField(p, radius, halfWidth);
  float step = (2.0f * 3.1415927410125732f) / spikes, ring = halfWidth - spikeRadius;
  int nearest = (int)std::floor(clampf(satan2(p.z, p.x) / step + 0.5f, -1.0e6f, 1.0e6f));
  for (int k = nearest - 1; k <= nearest + 1; k++) {
    float ang = (float)k * step;
    float sa = ssin(ang), ca = scos(ang);
    V side = v3(-sa, 0.0f, ca), dir = v3(ca, 0.0f, sa);
    for (int j = 0; j < 2; j++) {
      float y = j == 1 ? ring : -ring;
      V r = sub(sub(p, mul(dir, radius + spikeHeight)), v3(0.0f, y, 0.0f));
      Sample c = coneField(v3(dot(r, side), dot(r, dir), r.y), spikeRadius, spikeHeight);
      if (c.phi < result.phi) result = Sample{c.phi, add(add(mul(side, c.g.x), mul(dir, c.g.y)), v3(0.0f, c.g.z, 0.0f))};
    }
  }
  return result;
}
static inline Sample voxelField(const Field &f, V p) {
  uint32_t nx = f.dims[0], ny = f.dims[1], nz = f.dims[2];
  V hiV = add(f.origin, mul(v3((float)(nx - 1), (float)(ny - 1), (float)(nz - 1)), f.cell));
  V q = vclamp(p, f.origin, hiV);
  V coord = divs(sub(q, f.origin), f.cell);
  uint32_t bx = std::min((uint32_t)coord.x, nx - 2), by = std::min((uint32_t)coord.y, ny - 2), bz = std::min((uint32_t)coord.z, nz - 2);
  V t = sub(coord, v3((float)bx, (float)by, (float)bz));
  float acc[4] = {0, 0, 0, 0};
  for (uint32_t i = 0; i < 8; i++) {
    uint32_t ox = i & 1u, oy = (i >> 1) & 1u, oz = (i >> 2) & 1u;
    float wx = ox ? t.x : 1.0f - t.x, wy = oy ? t.y : 1.0f - t.y, wz = oz ? t.z : 1.0f - t.z;
    float w = (wx * wy) * wz;
    size_t node = 4ull * ((bx + ox) + (size_t)nx * ((by + oy) + (size_t)ny * (bz + oz)));
    for (int c = 0; c < 4; c++) acc[c] = acc[c] + f.grid[node + c] * w;
  }
  V outside = sub(p, q);
  float d = length(outside);
  V g = v3(acc[1], acc[2], acc[3]);
  if (d > 0.0f) g = unit(add(g, divs(outside, d)));
  return Sample{acc[0] + d, g};
}
/** localField (wgsl-sdf-fields.ts): the field's value and gradient at a field-local point. */
static inline Sample localField(const Field &f, V p) {
  const float *a = f.prm;
  switch (f.kind) {
    case K_BOX: return boxField(p, v3(a[0], a[1], a[2]));
    case K_SPHERE: return sphereField(p, a[0]);
    case K_CAPSULE: return sphereField(v3(p.x, p.y - clampf(p.y, -a[1], a[1]), p.z), a[0]);
    case K_CYLINDER: return cylinderField(p, a[0], a[1]);
    case K_CONE: return coneField(p, a[0], a[1]);
    case K_TORUS: return torusField(p, a[0], a[1]);
    case K_WARP: return warpField(p, v3(a[0], a[1], a[2]), a[3], f.warpType);
    case K_SPIKE_PLATE: return spikePlateField(p, a, a + 4);
    case K_SPIKY_WHEEL: return spikyWheelField(p, a, a[4]);
    default: return voxelField(f, p);
  }
}

// ---- bodies (sdfPrepareWGSL, wgsl-sdf-bridge.ts) ---------------------------------------------
struct SBody {
  V pos; float scale; Q rot; V vel; bool fixed; V omega; const Field *f; V lo, hi; Q anchor;
};
static inline V localPoint(const SBody &b, V p) { return add(divs(invrot(b.rot, sub(p, b.pos)), b.scale), b.f->com); }
static inline V worldPoint(const SBody &b, V p) { return add(b.pos, rot(b.rot, mul(sub(p, b.f->com), b.scale))); }
static inline Sample worldField(const SBody &b, V p) {
  Sample s = localField(*b.f, localPoint(b, p));
  return Sample{s.phi * b.scale, rot(b.rot, s.g)};
}
/** prepare: the field's bounds corners through the pose, min/max. */
static inline void bodyBounds(SBody &b) {
  V lo = v3(3.4e38f, 3.4e38f, 3.4e38f), hi = v3(-3.4e38f, -3.4e38f, -3.4e38f);
  for (uint32_t c = 0; c < 8; c++) {
    V p = v3((c & 1u) ? b.f->hi.x : b.f->lo.x, (c & 2u) ? b.f->hi.y : b.f->lo.y, (c & 4u) ? b.f->hi.z : b.f->lo.z);
    V w = add(b.pos, rot(b.rot, mul(sub(p, b.f->com), b.scale)));
    lo = vmn(lo, w); hi = vmx(hi, w);
  }
  b.lo = lo; b.hi = hi;
}

// ---- the pass parameters (packSdfParams, sdf-layout.ts:100-118) --------------------------------
struct Prm {
  uint32_t step = 0, cacheSize = 4, randomSamples = 1;
  float dt = 0, epsilon = 0.1f, alpha = 0.01f, tol = 0.01f;
  float margin = 0.02f, replacement = 0.01f;
  bool signedSep = false, meshVertex = false;
  uint32_t maxIt = 50; bool warmStart = true, curvature = true, skipConvex = true;
  uint32_t optimizer = 0; float stepSize = 0.1f; bool sgdStopOnly = false;
  uint32_t maxPoints = 5;
};
/** A pair result's point (the Point struct of wgsl-sdf.ts). exported: the contact written for it
 *  plus 1 (0: none). ell: the symmetric ellipsoid e00 e01 e02 e11 | e12 e22. */
struct Point {
  V x; float depth; V n; float sep; V rA; float soft; V rB; float rank;
  uint32_t active, cached, id, exported;
  float eA[4], eB[4];
};
struct Result {
  uint32_t a = 0xffffffffu, b = 0, key = 0, count = 0;
  uint32_t queries = 0, iters = 0, idc = 0, rolling = 0;
  std::vector<Point> pts;
};
static const uint32_t SPHERE_FEATURE = 3u << 24, SDF_ROLLING_BIT = 0x20000000u, STICK_BIT = 0x80000000u;
static const uint32_t WITNESS_ID = 0x40000000u;

/** The kernels' per-pair stream: rng = hash(rng + 0x9e3779b9). */
struct HashRng {
  uint32_t s = 0;
  uint32_t next() { s = mix(s + 0x9e3779b9u); return s; }
  float uniform() { return (float)(next() >> 8) / 16777216.0f; }
};
/** The CPU reference's stream (ref/sdf/math.ts Random: PCG-XSH-RR 64/32). */
struct PcgRng {
  uint64_t state = 0x853c49e6748fea9bull;
  uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ull + 1442695040888963407ull;
    uint32_t x = (uint32_t)(((old >> 18) ^ old) >> 27), r = (uint32_t)(old >> 59);
    return (x >> r) | (x << ((32u - r) & 31u));
  }
  float uniform() { return (float)(next() >> 8) / 16777216.0f; }
};

/** Algorithm 1 for one pair (the narrowphase entry point of wgsl-sdf.ts). Rng: the random stream;
 *  CpuSampling: the CPU reference's seed filter (a seed is considered only inside the volume's
 *  local bounds, manifold.ts stochasticSample) and stream order. */
template <class Rng, bool CpuSampling> struct PairUpdate {
  const Prm &P;
  const SBody &A, &B;
  Rng &rng;
  Result work;
  V lo, hi;
  float E[3][3];  // the ellipsoid method's final matrix (ellipsoidState)

  PairUpdate(const Prm &prm, const SBody &a, const SBody &b, Rng &r) : P(prm), A(a), B(b), rng(r) { work.pts.resize(P.maxPoints); }

  float smoothMax(float a, float b) const { return 0.5f * ((a + b) + std::sqrt((a - b) * (a - b) + P.epsilon)); }
  float objective(V p) {
    work.queries++;
    float value = smoothMax(worldField(A, p).phi, worldField(B, p).phi);
    for (uint32_t i = 0; i < work.count; i++) value = value - P.alpha * length(sub(p, work.pts[i].x));
    return value;
  }
  /** xyz gradient, w value; one query. */
  void objectiveGradient(V p, V &grad, float &value) {
    work.queries++;
    Sample a = worldField(A, p), b = worldField(B, p);
    float root = std::sqrt((a.phi - b.phi) * (a.phi - b.phi) + P.epsilon);
    value = 0.5f * ((a.phi + b.phi) + root);
    float t = 0.0f;
    if (root > 0.0f) t = (a.phi - b.phi) / root;
    grad = add(mul(a.g, 0.5f * (1.0f + t)), mul(b.g, 0.5f * (1.0f - t)));
    for (uint32_t i = 0; i < work.count; i++) {
      V dx = sub(p, work.pts[i].x);
      float len = length(dx);
      value = value - P.alpha * len;
      if (len > 1e-9f) grad = sub(grad, mul(dx, P.alpha / len));
    }
  }
  /** The deep-cut ellipsoid method (optimize.cpp ellipsoidMethod). */
  V ellipsoidMethod(V seed, float Ein[3][3]) {
    V x = seed;
    float M[3][3];
    std::memcpy(M, Ein, sizeof M);
    V best = vclamp(x, lo, hi);
    float bestValue = 3.4e38f;
    uint32_t it = 0;
    for (; it < P.maxIt; it++) {
      V grad = v3(0.0f, 0.0f, 0.0f);
      float value = 0.0f, violation = 0.0f;
      for (int i = 0; i < 3; i++) {
        float below = comp(lo, i) - comp(x, i), above = comp(x, i) - comp(hi, i);
        if (mx(below, above) > violation) { violation = mx(below, above); grad = v3(0.0f, 0.0f, 0.0f); setc(grad, i, above > below ? 1.0f : -1.0f); }
      }
      if (violation == 0.0f) {
        objectiveGradient(x, grad, value);
        if (value < bestValue) { bestValue = value; best = x; }
        if (length(grad) < P.tol) break;
      }
      V Eg = v3((M[0][0] * grad.x + M[0][1] * grad.y) + M[0][2] * grad.z, (M[1][0] * grad.x + M[1][1] * grad.y) + M[1][2] * grad.z,
                (M[2][0] * grad.x + M[2][1] * grad.y) + M[2][2] * grad.z);
      float gEg = dot(grad, Eg);
      if (!(gEg > 1e-12f)) break;
      float s = std::sqrt(gEg);
      float a = mn((violation > 0.0f ? violation : value - bestValue) / s, 0.5f);
      V step = mul(Eg, (1.0f + 3.0f * a) / (4.0f * s));
      x = sub(x, step);
      float cut = 2.0f * (1.0f + 3.0f * a) / ((4.0f * (1.0f + a)) * gEg);
      float shrink = 9.0f * (1.0f - a * a) / 8.0f;
      float N[3][3];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) N[r][c] = (M[r][c] - (comp(Eg, r) * comp(Eg, c)) * cut) * shrink;
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) M[r][c] = (N[r][c] + N[c][r]) * 0.5f;
      if (length(step) < P.tol) break;
    }
    work.iters += it;
    std::memcpy(E, M, sizeof M);
    return best;
  }
  V onSphere() {
    float z = -1.0f + 2.0f * rng.uniform();
    float angle = 6.2831855f * rng.uniform();
    float r = std::sqrt(mx(0.0f, 1.0f - z * z));
    return v3(r * scos(angle), r * ssin(angle), z);
  }
  /** Stochastic projected gradient descent (optimize.cpp gradientDescent). */
  V gradientDescent(V seed) {
    V x = vclamp(seed, lo, hi), best = x;
    float bestValue = 3.4e38f;
    uint32_t it = 0;
    for (; it < P.maxIt; it++) {
      V g; float value;
      objectiveGradient(x, g, value);
      if (value < bestValue) { bestValue = value; best = x; }
      float gl = length(g);
      if (gl < P.tol) break;
      float decay = 1.0f / std::sqrt((float)(it + 1u));
      V noise = mul(onSphere(), (0.25f * decay) * mx(gl, 1.0f));
      V next = vclamp(sub(x, mul(add(g, noise), P.stepSize * decay)), lo, hi);
      bool converged = !P.sgdStopOnly && length(sub(next, x)) < P.tol;
      x = next;
      if (converged) break;
    }
    work.iters += it;
    return best;
  }
  /** Domain-clamped Nelder-Mead (optimize.cpp; the kernel's optimize). */
  V nelderMead(V seed, V step) {
    V x[4]; float f[4];
    x[0] = vclamp(seed, lo, hi);
    for (int i = 0; i < 3; i++) {
      x[i + 1] = x[0];
      float si = comp(step, i);
      setc(x[i + 1], i, comp(x[0], i) + ((comp(x[0], i) + si <= comp(hi, i)) ? si : -si));
      x[i + 1] = vclamp(x[i + 1], lo, hi);
    }
    for (int i = 0; i < 4; i++) f[i] = objective(x[i]);
    uint32_t iterations = 0;
    for (; iterations < P.maxIt; iterations++) {
      for (int i = 1; i < 4; i++)
        for (int j = i; j > 0 && f[j] < f[j - 1]; j--) { std::swap(f[j], f[j - 1]); std::swap(x[j], x[j - 1]); }
      float extent = mx(length(sub(x[1], x[0])), mx(length(sub(x[2], x[0])), length(sub(x[3], x[0]))));
      if (extent < P.tol || f[3] - f[0] < 0.05f * P.tol) break;
      V centroid = divs(add(add(x[0], x[1]), x[2]), 3.0f);
      V reflected = vclamp(add(centroid, sub(centroid, x[3])), lo, hi);
      float fr = objective(reflected);
      if (fr < f[0]) {
        V expanded = vclamp(add(centroid, mul(sub(centroid, x[3]), 2.0f)), lo, hi);
        float fe = objective(expanded);
        x[3] = fe < fr ? expanded : reflected;
        f[3] = mn(fe, fr);
      } else if (fr < f[2]) {
        x[3] = reflected; f[3] = fr;
      } else {
        V contracted = add(centroid, mul(sub(fr < f[3] ? reflected : x[3], centroid), 0.5f));
        float fc = objective(contracted);
        if (fc < mn(fr, f[3])) { x[3] = contracted; f[3] = fc; }
        else for (int i = 1; i < 4; i++) { x[i] = add(x[0], mul(sub(x[i], x[0]), 0.5f)); f[i] = objective(x[i]); }
      }
    }
    work.iters += iterations;
    int best = 0;
    for (int i = 1; i < 4; i++) if (f[i] < f[best]) best = i;
    return x[best];
  }
  bool inside(V p) const { return p.x >= lo.x && p.y >= lo.y && p.z >= lo.z && p.x <= hi.x && p.y <= hi.y && p.z <= hi.z; }
  static bool within(V p, V l, V h) { return p.x >= l.x && p.y >= l.y && p.z >= l.z && p.x <= h.x && p.y <= h.y && p.z <= h.z; }
  /** seedPoint: a uniform point of the volume, or (curvature sampling) a uniformly chosen seed of
   *  either field that lies in it, by reservoir sampling. */
  V seedPoint() {
    float rx = rng.uniform(), ry = rng.uniform(), rz = rng.uniform();
    V p = add(lo, mulv(sub(hi, lo), v3(rx, ry, rz)));
    uint32_t n = 0;
    if (!P.curvature) return p;
    for (int side = 0; side < 2; side++) {
      const SBody &body = side == 0 ? A : B;
      V llo{}, lhi{};
      if (CpuSampling) {  // localBounds of the volume (body.ts localBounds)
        llo = v3(3.4e38f, 3.4e38f, 3.4e38f); lhi = v3(-3.4e38f, -3.4e38f, -3.4e38f);
        for (uint32_t c = 0; c < 8; c++) {
          V q = localPoint(body, v3((c & 1u) ? hi.x : lo.x, (c & 2u) ? hi.y : lo.y, (c & 4u) ? hi.z : lo.z));
          llo = vmn(llo, q); lhi = vmx(lhi, q);
        }
      }
      for (const V &s : body.f->seeds) {
        if (CpuSampling && !within(s, llo, lhi)) continue;
        V cand = worldPoint(body, s);
        if (inside(cand)) { n++; if (rng.next() % n == 0u) p = cand; }
      }
    }
    return p;
  }
  static V velocity(const SBody &b, V p) { return add(b.vel, cross(b.omega, sub(p, b.pos))); }
  void addPoint(const Point &old, V seed, bool cached) {
    float eA[4] = {1, 0, 0, 1}, eB[4] = {0, 1, 0, 0};
    V p;
    if (P.optimizer == 1u) {
      V extent = add(mul(sub(hi, lo), 0.5f), vabs(sub(seed, mul(add(lo, hi), 0.5f))));
      V d = mul(mulv(extent, extent), 3.0f);
      float M[3][3] = {{d.x, 0, 0}, {0, d.y, 0}, {0, 0, d.z}};
      if (cached) {
        float C[3][3] = {{old.eA[0], old.eA[1], old.eA[2]}, {old.eA[1], old.eA[3], old.eB[0]}, {old.eA[2], old.eB[0], old.eB[1]}};
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) M[r][c] = C[r][c] * 0.75f + M[r][c] * 0.25f;
      }
      p = ellipsoidMethod(seed, M);
      eA[0] = E[0][0]; eA[1] = E[0][1]; eA[2] = E[0][2]; eA[3] = E[1][1];
      eB[0] = E[1][2]; eB[1] = E[2][2]; eB[2] = 0; eB[3] = 0;
    } else if (P.optimizer == 2u) {
      p = gradientDescent(seed);
    } else {
      float h = 0.5f * P.tol;
      V step = vmx(mul(sub(hi, lo), 0.1f), v3(h, h, h));
      if (cached) { float e = mx(length(sub(seed, old.x)), 5.0f * P.tol); step = vmn(step, v3(e, e, e)); }
      p = nelderMead(seed, step);
    }
    Sample a = worldField(A, p), b = worldField(B, p);
    work.queries++;
    V na = unit(a.g), nb = neg(unit(b.g));
    V normal = a.phi < b.phi ? na : nb;
    if (cached) normal = dot(na, old.n) > dot(nb, old.n) ? na : nb;
    V offA = mul(a.g, a.phi / mx(dot(a.g, a.g), 1e-8f)), offB = mul(b.g, b.phi / mx(dot(b.g, b.g), 1e-8f));
    float sep = a.phi + b.phi;
    if (a.phi > 0.0f || b.phi > 0.0f) {
      V gap = sub(offA, offB);
      sep = length(gap);
      if (P.signedSep && (a.phi <= 0.0f || b.phi <= 0.0f)) {
        bool fixedA = A.fixed && A.f->kind == K_BOX, fixedB = B.fixed && B.f->kind == K_BOX;
        if (fixedA || fixedB) {
          normal = fixedA ? na : nb;
          sep = dot(gap, normal);
          V moving = fixedA ? b.g : a.g;
          if (dot(moving, moving) < 0.25f) sep = a.phi + b.phi;
        }
      }
      if (a.phi > 0.0f && b.phi > 0.0f && sep > 1e-6f && !(cached && dot(gap, old.n) < 0.0f)) normal = divs(gap, sep);
    }
    uint32_t id = old.id;
    if (!cached) {
      work.idc++;
      if (work.idc == SPHERE_FEATURE) work.idc++;
      if (work.idc == 0u || work.idc >= SDF_ROLLING_BIT) work.idc = 1u;
      id = work.idc;
    }
    Point q;
    q.x = p; q.depth = mn(a.phi, b.phi);
    q.n = normal; q.sep = sep;
    q.rA = rot(A.anchor, invrot(A.rot, sub(sub(p, offA), A.pos))); q.soft = smoothMax(a.phi, b.phi);
    q.rB = rot(B.anchor, invrot(B.rot, sub(sub(p, offB), B.pos))); q.rank = 0.0f;
    q.active = sep < P.margin ? 1u : 0u; q.cached = cached ? 1u : 0u; q.id = id; q.exported = 0;
    std::memcpy(q.eA, eA, sizeof eA); std::memcpy(q.eB, eB, sizeof eB);
    work.pts[work.count] = q;
    work.count++;
  }
  bool cachedIn(const Result *prev, uint32_t id) const {
    if (!prev) return false;
    for (uint32_t j = 0; j < prev->count; j++) if (prev->pts[j].id == id) return true;
    return false;
  }
  void witnessPoint(const Result *prev, V p, float sep, V normal, V a, V b, uint32_t id) {
    uint32_t slot = std::min(work.count, P.maxPoints - 1u);
    Point q;
    q.x = p; q.depth = mn(0.0f, sep); q.n = normal; q.sep = sep;
    q.rA = rot(A.anchor, invrot(A.rot, sub(a, A.pos))); q.soft = 0.0f;
    q.rB = rot(B.anchor, invrot(B.rot, sub(b, B.pos))); q.rank = 0.0f;
    q.active = 1; q.cached = cachedIn(prev, id) ? 1u : 0u; q.id = id; q.exported = 0;
    float eA[4] = {1, 0, 0, 1}, eB[4] = {0, 1, 0, 0};
    std::memcpy(q.eA, eA, sizeof eA); std::memcpy(q.eB, eB, sizeof eB);
    work.pts[slot] = q;
    work.count = slot + 1;
  }
  /** meshPlaneContacts (wgsl-sdf-supports.ts): the mesh's lowest support against a static box top
   *  face that covers the moving AABB. */
  bool meshPlane(const Result *prev) {
    if (!P.meshVertex) return false;
    bool fixedA = A.fixed && A.f->kind == K_BOX, fixedB = B.fixed && B.f->kind == K_BOX;
    if (!fixedA && !fixedB) return false;
    const SBody &mesh = fixedA ? B : A, &plane = fixedA ? A : B;
    uint32_t count = (uint32_t)mesh.f->supports.size();
    if (mesh.fixed || count == 0) return false;
    V half = v3(plane.f->prm[0], plane.f->prm[1], plane.f->prm[2]);
    for (uint32_t i = 0; i < 8; i++) {
      V corner = v3((i & 1u) ? mesh.hi.x : mesh.lo.x, (i & 2u) ? mesh.hi.y : mesh.lo.y, (i & 4u) ? mesh.hi.z : mesh.lo.z);
      V p = localPoint(plane, corner);
      if (std::fabs(p.x) >= half.x || std::fabs(p.z) >= half.z || p.y <= -half.y) return false;
    }
    V up = rot(plane.rot, v3(0.0f, 1.0f, 0.0f));
    V normal = fixedA ? up : neg(up);
    float lowest = 1e30f; uint32_t best = 0;
    for (uint32_t i = 0; i < count; i++) {
      V p = worldPoint(mesh, mesh.f->supports[i]);
      float s = (localPoint(plane, p).y - half.y) * plane.scale;
      if (s < lowest) { lowest = s; best = i; }
    }
    work.queries += count;
    if (lowest >= mn(P.margin, 0.5f * P.tol)) return false;
    V p = worldPoint(mesh, mesh.f->supports[best]);
    V projected = sub(p, mul(up, lowest));
    V a = fixedA ? projected : p, b = fixedA ? p : projected;
    witnessPoint(prev, p, lowest, normal, a, b, WITNESS_ID + best);
    return true;
  }
  void vertexScan(const SBody &mesh, const SBody &other, uint32_t side, float &bw, V &bp, uint32_t &id) {
    uint32_t count = std::min((uint32_t)mesh.f->supports.size(), 0x08000000u);
    for (uint32_t i = 0; i < count; i++) {
      V p = worldPoint(mesh, mesh.f->supports[i]);
      if (!inside(p)) continue;
      Sample s = worldField(other, p);
      work.queries++;
      if (s.phi < bw && dot(s.g, s.g) >= 0.25f) { bw = s.phi; bp = p; id = WITNESS_ID | (side << 27) | i; }
    }
  }
  /** meshMeshContacts: the deepest source vertex of either mesh inside the other's field. */
  bool meshMesh(const Result *prev) {
    if (!P.meshVertex || A.fixed || B.fixed) return false;
    float bw = 1e30f; V bp = v3(0, 0, 0); uint32_t id = 0;
    vertexScan(A, B, 0u, bw, bp, id);
    vertexScan(B, A, 1u, bw, bp, id);
    if (id == 0u || bw >= mn(P.margin, 0.5f * P.tol)) return false;
    V p = bp, a = p, b = p, normal;
    bool fromB = ((id >> 27) & 1u) != 0u;
    if (fromB) { V n = unit(worldField(A, p).g); normal = n; a = sub(p, mul(n, bw)); }
    else { V n = unit(worldField(B, p).g); normal = neg(n); b = sub(p, mul(n, bw)); }
    witnessPoint(prev, p, bw, normal, a, b, id);
    return true;
  }
  /** The whole pair: `prev` the pair's previous result (null: none or warm start off). */
  void run(uint32_t a, uint32_t b, uint32_t key, const Result *prev, uint32_t seedWord) {
    work.a = a; work.b = b; work.key = key; work.count = 0;
    work.queries = work.iters = work.idc = 0;
    bool rollA = A.f->kind >= 1u && A.f->kind <= 5u, rollB = B.f->kind >= 1u && B.f->kind <= 5u;
    work.rolling = (rollA || rollB) ? 1u : 0u;
    float m = 0.5f * P.margin;
    V mm = v3(m, m, m);
    lo = vmx(sub(A.lo, mm), sub(B.lo, mm));
    hi = vmn(add(A.hi, mm), add(B.hi, mm));
    if (lo.x > hi.x || lo.y > hi.y || lo.z > hi.z || (A.fixed && B.fixed)) return;
    (void)seedWord;
    uint32_t cached = 0, active = 0;
    if (prev) {
      work.idc = prev->idc;
      for (uint32_t i = 0; i < std::min(prev->count, P.maxPoints - 1u); i++) {
        const Point &pt = prev->pts[i];
        if (pt.id >= WITNESS_ID) continue;
        if (cached >= P.cacheSize) break;
        cached++; active += pt.active;
        V va = velocity(A, pt.x), vb = velocity(B, pt.x);
        V vel = mul(add(va, vb), 0.5f);
        if (A.fixed) vel = vb; else if (B.fixed) vel = va;
        addPoint(pt, add(pt.x, mul(vel, P.dt)), true);
      }
    }
    bool skip = P.skipConvex && P.optimizer != 2u && active > 0 && active == P.cacheSize && A.f->convex && B.f->convex;
    uint32_t samples = (skip ? 0u : P.randomSamples) + P.cacheSize - cached;
    Point empty;
    std::memset(&empty, 0, sizeof empty);
    for (uint32_t i = 0; i < samples; i++) { V s = seedPoint(); addPoint(empty, s, false); }
    uint32_t unique = 0;
    for (uint32_t i = 0; i < work.count; i++) {
      Point q = work.pts[i];
      bool dup = false;
      for (uint32_t j = 0; j < unique; j++) { V d = sub(q.x, work.pts[j].x); if (dot(d, d) < P.tol * P.tol) dup = true; }
      if (!dup) { work.pts[unique] = q; unique++; }
    }
    work.count = unique;
    for (uint32_t i = 0; i < unique; i++) {
      float rank = work.pts[i].soft + (work.pts[i].cached != 0u ? 0.0f : P.replacement);
      for (uint32_t j = 0; j < unique; j++) rank = rank - P.alpha * length(sub(work.pts[i].x, work.pts[j].x));
      work.pts[i].rank = rank;
    }
    for (uint32_t i = 1; i < unique; i++)
      for (uint32_t j = i; j > 0 && work.pts[j].rank < work.pts[j - 1].rank; j--) std::swap(work.pts[j], work.pts[j - 1]);
    if (!meshPlane(prev)) (void)meshMesh(prev);
  }
};

/** The kernel's stream seed for a pair. */
static inline uint32_t pairSeed(uint32_t a, uint32_t b, uint32_t key, uint32_t step) {
  return mix(key ^ mix(step) ^ mix(a) ^ mix(b + 0x9e3779b9u));
}
static inline uint32_t pairHash(uint32_t a, uint32_t b) { return mix(a ^ mix(b)); }

// ---- the bridge over a gpu3d model -------------------------------------------------------------
struct BodyConfig { uint32_t field; float scale; Q rot; };
struct PairSpec { uint32_t a, b, key; };
/** The solver-side of one step with the bridge installed (encode in sdf-bridge.ts): bodies
 *  prepared from the solver's records, candidates filtered (explicit) or generated (automatic),
 *  Algorithm 1 per pair, and the export into the solver's manifolds and contacts (one manifold per
 *  active point), warm-started from the solved contact the previous result's point was exported
 *  to. Results are kept by pair identity (a, b, key), which is how the kernels map them across
 *  reorderings (explicit: the previous slot of the same key with the same bodies; automatic: the
 *  previous slot of the same bodies). */
struct Bridge {
  std::vector<Field> fields;
  std::vector<BodyConfig> cfg;
  Prm prm;
  float allowedPen = 0.01f;
  bool refreshDeeper = false;
  bool automatic = false;
  uint32_t maxPairs = 1;
  float cellSize = 1.0f;
  std::vector<PairSpec> pairs;                                    // explicit candidates
  std::map<std::tuple<uint32_t, uint32_t, uint32_t>, Result> prev;  // last step's results by identity
  std::vector<Result> last;                                        // last step's results, sorted by (a, b)
  uint32_t stepIndex = 0;
  uint32_t overflowPairs = 0;  // automatic: candidates past the capacity in the last step
  std::vector<SBody> sb;

  template <class Model> void prepare(const Model &m) {
    sb.resize(m.bodies.size());
    for (size_t i = 0; i < m.bodies.size(); i++) {
      const auto &b = m.bodies[i];
      SBody &s = sb[i];
      s.f = &fields[cfg[i].field];
      s.pos = v3(b.pos.x, b.pos.y, b.pos.z);
      s.scale = cfg[i].scale;
      s.rot = qmul(q4(b.rot.x, b.rot.y, b.rot.z, b.rot.w), cfg[i].rot);
      s.vel = v3(b.vel.x, b.vel.y, b.vel.z);
      s.fixed = !(b.mass > 0.0f);
      s.omega = v3(b.angvel.x, b.angvel.y, b.angvel.z);
      s.anchor = cfg[i].rot;
      bodyBounds(s);
    }
  }
  template <class Model> bool passes(const Model &m, uint32_t a, uint32_t b, float margin) const {
    const SBody &A = sb[a], &B = sb[b];
    if (A.fixed && B.fixed) return false;
    if ((m.filterGroups[a] & m.filterMask[b]) == 0u || (m.filterGroups[b] & m.filterMask[a]) == 0u) return false;
    if (A.lo.x > B.hi.x + margin || A.lo.y > B.hi.y + margin || A.lo.z > B.hi.z + margin) return false;
    if (B.lo.x > A.hi.x + margin || B.lo.y > A.hi.y + margin || B.lo.z > A.hi.z + margin) return false;
    if (m.ignored(a, b)) return false;
    return true;
  }
  void cellOf(uint32_t i, int c[3], bool &large) const {
    const SBody &s = sb[i];
    V p = divs(add(mul(s.lo, 0.5f), mul(s.hi, 0.5f)), cellSize);
    for (int k = 0; k < 3; k++) c[k] = (int)clampf(std::floor(comp(p, k)), -1e9f, 1e9f);
    V ext = add(sub(s.hi, s.lo), v3(prm.margin, prm.margin, prm.margin));
    large = ext.x > cellSize || ext.y > cellSize || ext.z > cellSize;
  }
  /** The candidates of the spatial grid (wgsl-sdf-broadphase.ts): a regular body (extent plus
   *  margin within one cell) meets the regular bodies of the 27 cells around its centre cell and
   *  every large body; large bodies meet each other. Then the pair filter with the margin. */
  template <class Model> std::vector<PairSpec> generate(const Model &m) {
    uint32_t n = (uint32_t)sb.size();
    std::vector<std::array<int, 3>> cell(n);
    std::vector<uint8_t> large(n);
    for (uint32_t i = 0; i < n; i++) { bool l; cellOf(i, cell[i].data(), l); large[i] = l; }
    std::vector<PairSpec> out;
    for (uint32_t a = 0; a < n; a++)
      for (uint32_t b = 0; b < a; b++) {
        bool cand = large[a] || large[b];
        if (!cand) {
          cand = true;
          for (int k = 0; k < 3; k++) { long long dlt = (long long)cell[a][k] - cell[b][k]; if (dlt < -1 || dlt > 1) cand = false; }
        }
        if (cand && passes(m, a, b, prm.margin)) out.push_back(PairSpec{a, b, pairHash(a, b)});
      }
    return out;
  }
  /** One step's collision into the model (manifolds, contacts and pairs already cleared). */
  template <class Model> void encode(Model &m) {
    prepare(m);
    std::vector<PairSpec> list;
    overflowPairs = 0;
    if (automatic) {
      list = generate(m);
      if (list.size() > maxPairs) { overflowPairs = (uint32_t)(list.size() - maxPairs); list.resize(maxPairs); }
    } else {
      for (const PairSpec &ps : pairs) if (passes(m, ps.a, ps.b, prm.margin)) list.push_back(ps);
    }
    std::sort(list.begin(), list.end(), [](const PairSpec &x, const PairSpec &y) { return x.a < y.a || (x.a == y.a && x.b < y.b); });
    Prm P = prm;
    P.step = stepIndex;
    P.dt = (float)m.p.dt;
    std::vector<Result> results;
    for (const PairSpec &ps : list) {
      auto it = prev.find(std::make_tuple(ps.a, ps.b, ps.key));
      const Result *pr = (P.warmStart && it != prev.end()) ? &it->second : nullptr;
      HashRng rng;
      rng.s = pairSeed(ps.a, ps.b, ps.key, P.step);
      PairUpdate<HashRng, false> u(P, sb[ps.a], sb[ps.b], rng);
      u.run(ps.a, ps.b, ps.key, pr, 0);
      results.push_back(u.work);
    }
    exportContacts(m, results);
    prev.clear();
    for (const Result &r : results) prev[std::make_tuple(r.a, r.b, r.key)] = r;
    last = results;
    stepIndex++;
  }
  /** exportContacts (sdfExportSource): one manifold per active point. */
  template <class Model> void exportContacts(Model &m, std::vector<Result> &results) {
    float alpha = (float)m.p.alpha, gamma = (float)m.p.gamma, dt = (float)m.p.dt;
    float margin = prm.margin, ap = allowedPen, refresh = refreshDeeper ? prm.tol : 0.0f;
    float ag = alpha * gamma;
    for (Result &r : results) {
      m.pairs.push_back({r.a, r.b});
      uint32_t a = r.a, b = r.b;
      const auto &BA = m.bodies[a];
      const auto &BB = m.bodies[b];
      V pA = v3(BA.pos.x, BA.pos.y, BA.pos.z), pB = v3(BB.pos.x, BB.pos.y, BB.pos.z);
      Q qA = q4(BA.rot.x, BA.rot.y, BA.rot.z, BA.rot.w), qB = q4(BB.rot.x, BB.rot.y, BB.rot.z, BB.rot.w);
      auto pit = prev.find(std::make_tuple(r.a, r.b, r.key));
      const Result *pr = pit != prev.end() ? &pit->second : nullptr;
      for (uint32_t i = 0; i < r.count; i++) {
        Point &pt = r.pts[i];
        pt.exported = 0;
        if (!pt.active) continue;
        V n = neg(pt.n);
        V basis[3];
        orthonormal(n, basis);
        V rA = pt.rA, rB = pt.rB;
        V lam = v3(0, 0, 0), pen = v3(0, 0, 0);
        uint32_t stick = 0;
        if (pt.cached && pr) {
          for (uint32_t j = 0; j < pr->count; j++) {
            const Point &prior = pr->pts[j];
            if (prior.id == pt.id && prior.exported > 0u && prior.exported <= m.prevContacts.size()) {
              const auto &k = m.prevContacts[prior.exported - 1];
              V ob[3];
              orthonormal(neg(prior.n), ob);
              V kl = v3(k.lam.x, k.lam.y, k.lam.z);
              V w = add(add(mul(ob[0], kl.x), mul(ob[1], kl.y)), mul(ob[2], kl.z));
              lam = v3(dot(basis[0], w), dot(basis[1], w), dot(basis[2], w));
              pen = v3(k.pen.x, k.pen.y, k.pen.z);
              stick = r.rolling ? 0u : (k.key & STICK_BIT);
              V krA = v3(k.rA.x, k.rA.y, k.rA.z), krB = v3(k.rB.x, k.rB.y, k.rB.z);
              if (stick != 0u && refresh > 0.0f) {
                V retained = sub(sub(add(pA, rot(qA, krA)), pB), rot(qB, krB));
                float moved = mx(length(sub(rA, krA)), length(sub(rB, krB)));
                if (moved > 3.0f * refresh && pt.sep < dot(n, retained) - refresh) stick = 0u;
              }
              if (stick != 0u) { rA = krA; rB = krB; }
              break;
            }
          }
        }
        V d = sub(sub(add(pA, rot(qA, rA)), pB), rot(qB, rB));
        float sep = stick != 0u ? dot(n, d) : pt.sep;
        float c0x = (sep - margin) + ap, c0y = stick != 0u ? dot(basis[1], d) : 0.0f, c0z = stick != 0u ? dot(basis[2], d) : 0.0f;
        lam = mul(lam, ag);
        pen = v3(clampf(pen.x * gamma, 1.0f, 1e10f), clampf(pen.y * gamma, 1.0f, 1e10f), clampf(pen.z * gamma, 1.0f, 1e10f));
        if (m.p.flags & 16u) {
          float ma = BA.mass <= 0.0f ? 3.4e38f : BA.mass, mb = BB.mass <= 0.0f ? 3.4e38f : BB.mass;
          pen.x = mx(pen.x, clampf(mn(ma, mb) / (dt * dt), 1.0f, 1e10f));
        }
        typename std::remove_reference<decltype(m.contacts[0])>::type k;
        k.rA = {rA.x, rA.y, rA.z};
        k.key = (pt.id & ~STICK_BIT) | stick | (r.rolling ? SDF_ROLLING_BIT : 0u);
        k.rB = {rB.x, rB.y, rB.z}; k.c0x = c0x;
        k.pen = {pen.x, pen.y, pen.z}; k.c0y = c0y;
        k.lam = {lam.x, lam.y, lam.z}; k.c0z = c0z;
        typename std::remove_reference<decltype(m.manifolds[0])>::type man;
        man.a = a; man.b = b; man.first = (uint32_t)m.contacts.size(); man.countStep = 1u | (m.p.step << 4);
        man.n = {n.x, n.y, n.z};
        man.friction = std::sqrt(BA.friction * BB.friction);
        m.contacts.push_back(k);
        m.contactSrc.push_back(-1);
        m.manifolds.push_back(man);
        pt.exported = (uint32_t)m.contacts.size();
      }
    }
  }
};

}  // namespace sdfo
