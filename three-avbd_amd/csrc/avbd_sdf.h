// SDF collision backend on CDNA4: SdfCollisionBridge (src/avbd3d/gpu/sdf-bridge.ts) as HIP kernels.
//
// Installed on a solver it replaces the built-in broadphase and narrowphase: k_sdf_prepare poses
// every body's field (sdfPrepareWGSL, wgsl-sdf-bridge.ts), candidates come from the caller
// (k_sdf_filter, sdfFilterWGSL) or from a spatial grid (k_sdf_cells .. k_sdf_fill, the semantics of
// wgsl-sdf-broadphase.ts), k_sdf_narrow runs Algorithm 1 per pair (wgsl-sdf.ts with its three
// optimizers and the mesh vertex witnesses of wgsl-sdf-supports.ts) and k_sdf_export writes one
// manifold per active point into the solver's storage (sdfExportSource). The field kernels follow
// wgsl-sdf-fields.ts.
//
// Mapping: one lane per pair. Algorithm 1 is a chain of dependent optimiser steps (each point's
// objective repels from the points before it), so a pair has no parallelism worth a wavefront;
// pairs are independent. The points of the pair being built live in a lane-interleaved staging
// buffer (coalesced across the wavefront), and their positions, read by every objective
// evaluation's repulsion sum, in LDS. The grid is a counting sort of bucket keys with a count
// pass and a fill pass over the same predicate, so candidates need no same-address atomics.
//
// sin, cos and atan2 are written out (Cody-Waite reduction with fused multiply-adds, truncated
// Taylor series): every operation is then the same on the CPU restatement that checks them.
#pragma once
#include "avbd_kernels.h"

namespace avbd {
namespace sdfk {

static const uint32_t EMPTY = 0xffffffffu;
static const uint32_t WITNESS_ID = 0x40000000u;
static const uint32_t BODY_F4 = 7;     // SdfBody: 7 float4 (wgsl-sdf-bridge.ts SdfBody)
static const uint32_t POINT_F4 = 7;    // Point: 7 float4
static const uint32_t RESULT_HEAD = 2; // identity + stats uint4
enum { K_BOX = 0, K_SPHERE = 1, K_CAPSULE = 2, K_CYLINDER = 3, K_CONE = 4, K_TORUS = 5, K_WARP = 6, K_VOXEL = 7,
       K_SPIKE_PLATE = 8, K_SPIKY_WHEEL = 9 };

/** The pass parameters (packSdfParams, sdf-layout.ts:100-118) plus the bridge's distances. */
struct SdfParams {
  uint32_t pairCount, step, cacheSize, randomSamples;
  float dt, epsilon, alpha, tol;
  float margin, replacement;
  uint32_t signedSep, meshVertex;
  uint32_t maxIt, warmStart, curvature, skipConvex;
  uint32_t optimizer; float stepSize; uint32_t sgdStopOnly;
  uint32_t points;          // points per pair result
  float allowedPen, refresh;  // bridge distances: allowed penetration, refresh tolerance (0: off)
};
/** Device buffers of one bridge. */
struct SdfDev {
  const uint4 *atlas;
  float4 *bodies;            // n x 7 float4
  const uint4 *config;       // per body: field header (uint4 index), scale bits, ray slope bits, 0
  const float4 *configRot;   // per body: the field's rotation in the body frame
  uint4 *pairs;              // a, b, key, previous slot
  const uint4 *prev;         // last step's results
  uint4 *cur;                // this step's results
  float4 *stage;             // points being built: (point * 7 + word) * cap + slot
  uint32_t cap;              // pair capacity (result slots)
  uint32_t *live;            // automatic: [0] pairs found, [1] candidates tested, [2] large bodies
  SdfParams p;
  // automatic grid
  float cellSize;
  uint32_t tableMask, cacheMask;
  uint32_t *bucketCount, *bucketStart, *bucketCursor, *sorted, *largeList, *largeCount;
  int4 *cellKey;
  uint32_t *history;         // open-addressing table: slot of the previous result of a pair
  uint32_t *pairCount, *pairOffset;  // per body: accepted candidates, their exclusive scan
};

__device__ __forceinline__ uint32_t mix(uint32_t x) {
  x = (x ^ (x >> 16)) * 0x7feb352du; x = (x ^ (x >> 15)) * 0x846ca68bu; return x ^ (x >> 16);
}
__device__ __forceinline__ uint32_t pairHash(uint32_t a, uint32_t b) { return mix(a ^ mix(b)); }

// ---- trigonometry ----------------------------------------------------------------------------
__device__ __forceinline__ float reduceArg(float x, int &quadrant) {
  float k = rintf(x * 0.636619746685028076171875f);
  float r = fmaf(-k, 1.57079637050628662109375f, x);
  r = fmaf(-k, -4.371138828673793e-8f, r);
  r = fmaf(-k, -1.7151245100058819e-15f, r);
  quadrant = (int)((long long)k & 3);
  return r;
}
__device__ __forceinline__ float sinPoly(float r) {
  float z = r * r;
  float p = fmaf(z, 2.75573192239858907e-6f, -1.98412698412698413e-4f);
  p = fmaf(z, p, 8.33333333333333333e-3f);
  p = fmaf(z, p, -1.66666666666666667e-1f);
  return fmaf(r * z, p, r);
}
__device__ __forceinline__ float cosPoly(float r) {
  float z = r * r;
  float p = fmaf(z, -2.75573192239858907e-7f, 2.48015873015873016e-5f);
  p = fmaf(z, p, -1.38888888888888889e-3f);
  p = fmaf(z, p, 4.16666666666666667e-2f);
  p = fmaf(z, p, -0.5f);
  return fmaf(z, p, 1.0f);
}
__device__ __forceinline__ float ssin(float x) {
  int q; float r = reduceArg(x, q);
  float s = sinPoly(r), c = cosPoly(r);
  return q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
}
__device__ __forceinline__ float scos(float x) {
  int q; float r = reduceArg(x, q);
  float s = sinPoly(r), c = cosPoly(r);
  return q == 0 ? c : (q == 1 ? -s : (q == 2 ? -c : s));
}
__device__ __forceinline__ float atanSmall(float u) {
  float z = u * u;
  float p = fmaf(z, -0.0588235294117647059f, 0.0666666666666666667f);
  p = fmaf(z, p, -0.0769230769230769231f);
  p = fmaf(z, p, 0.0909090909090909091f);
  p = fmaf(z, p, -0.111111111111111111f);
  p = fmaf(z, p, 0.142857142857142857f);
  p = fmaf(z, p, -0.2f);
  p = fmaf(z, p, 0.333333333333333333f);
  return fmaf(-u * z, p, u);
}
__device__ __forceinline__ float satan2(float y, float x) {
  float ax = fabsf(x), ay = fabsf(y);
  float hi = mx(ax, ay), lo = mn(ax, ay);
  if (!(hi > 0.0f)) return 0.0f;
  float t = lo / hi;
  float a = t > 0.414213562373095049f ? 0.785398163397448310f + atanSmall((t - 1.0f) / (t + 1.0f)) : atanSmall(t);
  if (ay > ax) a = 1.57079632679489662f - a;
  if (x < 0.0f) a = 3.14159265358979324f - a;
  if (y < 0.0f) a = -a;
  return a;
}

// ---- fields ----------------------------------------------------------------------------------
struct Sample { float phi; V g; };
__device__ __forceinline__ float signnz(float x) { return x < 0.0f ? -1.0f : 1.0f; }
__device__ __forceinline__ V unit(V a) { float l = length(a); if (l > 1e-12f) return divs(a, l); return vv(0.0f, 1.0f, 0.0f); }
__device__ __forceinline__ V invrot(Q q, V a) { return qrotate(qq(-q.x, -q.y, -q.z, q.w), a); }
__device__ __forceinline__ V vmn(V a, V b) { return vv(mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z)); }
__device__ __forceinline__ V vmx(V a, V b) { return vv(mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z)); }
__device__ __forceinline__ V vclamp(V a, V lo, V hi) { return vmn(vmx(a, lo), hi); }
__device__ __forceinline__ V radial(V p, float r) { if (r > 1e-9f) return vv(p.x / r, 0.0f, p.z / r); return vv(1.0f, 0.0f, 0.0f); }
__device__ __forceinline__ Sample boxField(V p, V h) {
  V q = sub(vabs(p), h);
  V o = vmx(q, vv(0.0f, 0.0f, 0.0f));
  float d = length(o);
  V g = vv(0.0f, 0.0f, 0.0f);
  if (d > 0.0f) {
    g = vv((o.x / d) * signnz(p.x), (o.y / d) * signnz(p.y), (o.z / d) * signnz(p.z));
  } else {
    int axis = 2;
    if (q.x > q.y) { if (q.x > q.z) axis = 0; } else if (q.y > q.z) axis = 1;
    setComp(g, axis, signnz(comp(p, axis)));
  }
  Sample s; s.phi = d + mn(mx(q.x, mx(q.y, q.z)), 0.0f); s.g = g;
  return s;
}
__device__ __forceinline__ Sample sphereField(V p, float r) {
  float d = length(p);
  V g = vv(0.0f, 1.0f, 0.0f);
  if (d > 1e-9f) g = divs(p, d);
  Sample s; s.phi = d - r; s.g = g;
  return s;
}
__device__ __forceinline__ Sample cylinderField(V p, float r, float h) {
  float rl = sqrtf(p.x * p.x + p.z * p.z);
  float dx = rl - r, dy = fabsf(p.y) - h;
  float ox = mx(dx, 0.0f), oy = mx(dy, 0.0f);
  float l = sqrtf(ox * ox + oy * oy);
  V e = radial(p, rl);
  V g = vv(0.0f, signnz(p.y), 0.0f);
  if (dx > dy) g = e;
  if (l > 0.0f) g = add(mul(e, ox / l), vv(0.0f, signnz(p.y) * oy / l, 0.0f));
  Sample s; s.phi = l + mn(mx(dx, dy), 0.0f); s.g = g;
  return s;
}
__device__ __forceinline__ Sample coneField(V p, float radius, float h) {
  float qx = sqrtf(p.x * p.x + p.z * p.z), qy = p.y;
  float cax = qx - mn(qx, qy < 0.0f ? radius : 0.0f), cay = fabsf(qy) - h;
  float kx = -radius, ky = 2.0f * h;
  float t = clampf((qx * radius + (h - qy) * ky) / (kx * kx + ky * ky), 0.0f, 1.0f);
  float cbx = qx + kx * t, cby = (qy - h) + ky * t;
  float s = (cbx < 0.0f && cay < 0.0f) ? -1.0f : 1.0f;
  float da = cax * cax + cay * cay, db = cbx * cbx + cby * cby;
  float gx = cbx, gy = cby;
  if (da < db) { gx = cax; gy = cay * signnz(qy); }
  float l = sqrtf(gx * gx + gy * gy);
  if (l > 1e-9f) { float f = s / l; gx = gx * f; gy = gy * f; }
  else if (da < db) { gx = 0.0f; gy = signnz(qy); }
  else { float nl = sqrtf(ky * ky + radius * radius); gx = ky / nl; gy = radius / nl; }
  Sample o; o.phi = s * sqrtf(mn(da, db)); o.g = add(mul(radial(p, qx), gx), vv(0.0f, gy, 0.0f));
  return o;
}
__device__ __forceinline__ Sample torusField(V p, float major, float minor) {
  float r = sqrtf(p.x * p.x + p.z * p.z);
  float qx = r - major, qy = p.y;
  float l = sqrtf(qx * qx + qy * qy);
  V e = radial(p, r), g = e;
  if (l > 1e-9f) g = add(mul(e, qx / l), vv(0.0f, qy / l, 0.0f));
  Sample s; s.phi = l - minor; s.g = g;
  return s;
}
__device__ __forceinline__ Sample warpField(V p, V h, float k, uint32_t kind) {
  if (kind == 2u) {
    V s = vv(ssin(p.x * 7.0f), ssin(p.y * 7.0f), ssin(p.z * 7.0f));
    V c = vv(scos(p.x * 7.0f), scos(p.y * 7.0f), scos(p.z * 7.0f));
    Sample b = boxField(p, h);
    float k7 = 7.0f * k;
    V w = vv((c.x * s.y) * s.z, (s.x * c.y) * s.z, (s.x * s.y) * c.z);
    Sample o; o.phi = b.phi + ((k * s.x) * s.y) * s.z; o.g = add(b.g, mul(w, k7));
    return o;
  }
  float angle = k * (kind == 0u ? p.x : p.y);
  float c = scos(angle), s = ssin(angle);
  V q = kind == 0u ? vv(c * p.x - s * p.y, s * p.x + c * p.y, p.z) : vv(c * p.x - s * p.z, s * p.x + c * p.z, p.y);
  Sample b = boxField(q, h);
  V g = b.g;
  Sample o; o.phi = b.phi;
  if (kind == 0u) o.g = vv(dot(vv(c - k * q.y, s + k * q.x, 0.0f), g), dot(vv(-s, c, 0.0f), g), g.z);
  else o.g = vv(dot(vv(c, s, 0.0f), g), dot(vv(-(k * q.y), k * q.x, 1.0f), g), dot(vv(-s, c, 0.0f), g));
  return o;
}
__device__ __forceinline__ float shapeHash01(uint32_t x) {
  x = (x ^ (x >> 16)) * 0x7feb352du; x = (x ^ (x >> 15)) * 0x846ca68bu; x = x ^ (x >> 16);
  return (float)(x & 0xffffffu) / 16777216.0f;
}
__device__ __forceinline__ int floorIndex(float x, int lo, int hi) {
  float f = floorf(x);
  f = clampf(f, (float)(lo - 1), (float)(hi + 1));
  return (int)f;
}
__device__ Sample spikePlateField(V p, float4 a, float4 b) {
  V h = vv(a.x, a.y, a.z);
  int count = (int)a.w;
  float radius = b.x;
  V pitch = mul(h, 2.0f / a.w);
  int ci = floorIndex((p.x + h.x) / pitch.x, 0, count - 1), cj = floorIndex((p.z + h.z) / pitch.z, 0, count - 1);
  ci = ci < 0 ? 0 : (ci > count - 1 ? count - 1 : ci);
  cj = cj < 0 ? 0 : (cj > count - 1 ? count - 1 : cj);
  Sample result = boxField(p, h);
  int j1 = cj + 1 < count - 1 ? cj + 1 : count - 1, i1 = ci + 1 < count - 1 ? ci + 1 : count - 1;
  for (int j = cj - 1 > 0 ? cj - 1 : 0; j <= j1; j++)
    for (int i = ci - 1 > 0 ? ci - 1 : 0; i <= i1; i++) {
      float height = b.y + (b.z - b.y) * shapeHash01(17u * ((uint32_t)i + (uint32_t)j * (uint32_t)count) + 11u);
      V center = vv(-h.x + ((float)i + 0.5f) * pitch.x, (h.y + 0.5f * height) - 0.5f * radius, -h.z + ((float)j + 0.5f) * pitch.z);
      Sample c = coneField(sub(p, center), (radius * (height + radius)) / height, 0.5f * (height + radius));
      if (c.phi < result.phi) result = c;
    }
  return result;
}
__device__ Sample spikyWheelField(V p, float4 a, float spikes) {
  float radius = a.x, halfWidth = a.y, spikeRadius = a.z, spikeHeight = a.w;
  Sample result = cylinderField(p, radius, halfWidth);
  float step = (2.0f * 3.1415927410125732f) / spikes, ring = halfWidth - spikeRadius;
  int nearest = (int)floorf(clampf(satan2(p.z, p.x) / step + 0.5f, -1.0e6f, 1.0e6f));
  for (int k = nearest - 1; k <= nearest + 1; k++) {
    float ang = (float)k * step;
    float sa = ssin(ang), ca = scos(ang);
    V side = vv(-sa, 0.0f, ca), dir = vv(ca, 0.0f, sa);
    for (int j = 0; j < 2; j++) {
      float y = j == 1 ? ring : -ring;
      V r = sub(sub(p, mul(dir, radius + spikeHeight)), vv(0.0f, y, 0.0f));
      Sample c = coneField(vv(dot(r, side), dot(r, dir), r.y), spikeRadius, spikeHeight);
      if (c.phi < result.phi) { result.phi = c.phi; result.g = add(add(mul(side, c.g.x), mul(dir, c.g.y)), vv(0.0f, c.g.z, 0.0f)); }
    }
  }
  return result;
}
__device__ Sample voxelField(const uint4 *atlas, uint32_t h, V p) {
  uint4 grid = atlas[h + 1];
  uint32_t nx = grid.y, ny = grid.z, nz = grid.w;
  float4 org = __builtin_bit_cast(float4, atlas[h + 7]);
  float cell = __builtin_bit_cast(float4, atlas[h + 2]).w;
  V origin = vv(org.x, org.y, org.z);
  V hiV = add(origin, mul(vv((float)(nx - 1), (float)(ny - 1), (float)(nz - 1)), cell));
  V q = vclamp(p, origin, hiV);
  V coord = divs(sub(q, origin), cell);
  uint32_t bx = (uint32_t)coord.x, by = (uint32_t)coord.y, bz = (uint32_t)coord.z;
  bx = bx < nx - 2 ? bx : nx - 2; by = by < ny - 2 ? by : ny - 2; bz = bz < nz - 2 ? bz : nz - 2;
  V t = sub(coord, vv((float)bx, (float)by, (float)bz));
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
#pragma unroll
  for (uint32_t i = 0; i < 8; i++) {
    uint32_t ox = i & 1u, oy = (i >> 1) & 1u, oz = (i >> 2) & 1u;
    float wx = ox ? t.x : 1.0f - t.x, wy = oy ? t.y : 1.0f - t.y, wz = oz ? t.z : 1.0f - t.z;
    float w = (wx * wy) * wz;
    float4 node = __builtin_bit_cast(float4, atlas[grid.x + (bx + ox) + nx * ((by + oy) + ny * (bz + oz))]);
    a0 = a0 + node.x * w; a1 = a1 + node.y * w; a2 = a2 + node.z * w; a3 = a3 + node.w * w;
  }
  V outside = sub(p, q);
  float d = length(outside);
  V g = vv(a1, a2, a3);
  if (d > 0.0f) g = unit(add(g, divs(outside, d)));
  Sample s; s.phi = a0 + d; s.g = g;
  return s;
}
/** A body's field as the narrowphase evaluates it: pose, scale, the header's words it reads. */
struct FieldBody {
  V pos; float scale; Q rot;
  uint32_t h, kind;
  V com;
  float4 p0, p1;  // atlas words 20..27
};
__device__ __forceinline__ Sample localField(const uint4 *atlas, const FieldBody &f, V p) {
  float4 a = f.p0;
  switch (f.kind) {
    case K_BOX: return boxField(p, vv(a.x, a.y, a.z));
    case K_SPHERE: return sphereField(p, a.x);
    case K_CAPSULE: return sphereField(vv(p.x, p.y - clampf(p.y, -a.y, a.y), p.z), a.x);
    case K_CYLINDER: return cylinderField(p, a.x, a.y);
    case K_CONE: return coneField(p, a.x, a.y);
    case K_TORUS: return torusField(p, a.x, a.y);
    case K_WARP: return warpField(p, vv(a.x, a.y, a.z), a.w, __float_as_uint(f.p1.x));
    case K_SPIKE_PLATE: return spikePlateField(p, a, f.p1);
    case K_SPIKY_WHEEL: return spikyWheelField(p, a, f.p1.x);
    default: return voxelField(atlas, f.h, p);
  }
}
__device__ __forceinline__ V localPoint(const FieldBody &b, V p) { return add(divs(invrot(b.rot, sub(p, b.pos)), b.scale), b.com); }
__device__ __forceinline__ V worldPoint(const FieldBody &b, V p) { return add(b.pos, qrotate(b.rot, mul(sub(p, b.com), b.scale))); }
__device__ __forceinline__ Sample worldField(const uint4 *atlas, const FieldBody &b, V p) {
  Sample s = localField(atlas, b, localPoint(b, p));
  Sample o; o.phi = s.phi * b.scale; o.g = qrotate(b.rot, s.g);
  return o;
}
__device__ __forceinline__ FieldBody fieldBody(const SdfDev &g, uint32_t i) {
  FieldBody f;
  float4 p = g.bodies[BODY_F4 * i], r = g.bodies[BODY_F4 * i + 1], o = g.bodies[BODY_F4 * i + 3];
  f.pos = xyz(p); f.scale = p.w; f.rot = q4(r);
  f.h = __float_as_uint(o.w);
  uint4 hd = g.atlas[f.h];
  f.kind = hd.x;
  f.com = xyz(__builtin_bit_cast(float4, g.atlas[f.h + 4]));
  f.p0 = __builtin_bit_cast(float4, g.atlas[f.h + 5]);
  f.p1 = __builtin_bit_cast(float4, g.atlas[f.h + 6]);
  return f;
}

// ---- preparation and the explicit filter ----------------------------------------------------------
/** prepare (sdfPrepareWGSL): the field's world pose (body rotation x field rotation), velocity with
 *  the static flag, angular velocity with the header, and the bounds of its box corners. */
__global__ void k_sdf_prepare(Dev d, SdfDev g) {
  GRID_STRIDE(i, d.n) {
    float4 pos = d.b.pos[i];
    Q rotation = qmul(q4(d.b.rot[i]), q4(g.configRot[i]));
    uint4 c = g.config[i];
    uint32_t h = c.x;
    float scale = __uint_as_float(c.y);
    V llo = xyz(__builtin_bit_cast(float4, g.atlas[h + 2])), lhi = xyz(__builtin_bit_cast(float4, g.atlas[h + 3]));
    V com = xyz(__builtin_bit_cast(float4, g.atlas[h + 4]));
    V lo = vv(3.4e38f, 3.4e38f, 3.4e38f), hi = vv(-3.4e38f, -3.4e38f, -3.4e38f);
    V P = xyz(pos);
#pragma unroll
    for (uint32_t k = 0; k < 8; k++) {
      V q = vv((k & 1u) ? lhi.x : llo.x, (k & 2u) ? lhi.y : llo.y, (k & 4u) ? lhi.z : llo.z);
      V w = add(P, qrotate(rotation, mul(sub(q, com), scale)));
      lo = vmn(lo, w); hi = vmx(hi, w);
    }
    float4 *o = g.bodies + BODY_F4 * i;
    o[0] = f4(P, scale);
    o[1] = f4q(rotation);
    o[2] = f4(xyz(d.b.vel[i]), d.b.size[i].w <= 0.0f ? 1.0f : 0.0f);
    o[3] = f4(xyz(d.b.angvel[i]), __uint_as_float(h));
    o[4] = f4(lo, __uint_as_float(c.z));
    o[5] = f4(hi, 0.0f);
    o[6] = g.configRot[i];
  }
}
/** The pair filter (sdfFilterWGSL / the spatial grid's candidate test): not both static, the
 *  collision masks, bounds within the margin, not ignored (joints, IgnoreCollision). a > b. */
__device__ __forceinline__ bool sdfPasses(const Dev &d, const SdfDev &g, uint32_t a, uint32_t b) {
  float4 va = g.bodies[BODY_F4 * a + 2], vb = g.bodies[BODY_F4 * b + 2];
  if (va.w != 0.0f && vb.w != 0.0f) return false;
  uint2 fa = d.filters[a], fb = d.filters[b];
  if ((fa.x & fb.y) == 0u || (fb.x & fa.y) == 0u) return false;
  V alo = xyz(g.bodies[BODY_F4 * a + 4]), ahi = xyz(g.bodies[BODY_F4 * a + 5]);
  V blo = xyz(g.bodies[BODY_F4 * b + 4]), bhi = xyz(g.bodies[BODY_F4 * b + 5]);
  float m = g.p.margin;
  if (alo.x > bhi.x + m || alo.y > bhi.y + m || alo.z > bhi.z + m) return false;
  if (blo.x > ahi.x + m || blo.y > ahi.y + m || blo.z > ahi.z + m) return false;
  if (ignored(d, a, b)) return false;
  return true;
}
__global__ void k_sdf_filter(Dev d, SdfDev g) {
  GRID_STRIDE(slot, g.p.pairCount) {
    uint4 pr = g.pairs[slot];
    if (!sdfPasses(d, g, pr.x, pr.y)) g.pairs[slot].x = EMPTY;
  }
}

// ---- the spatial grid (automatic candidates) -----------------------------------------------------
__device__ __forceinline__ void cellOf(const SdfDev &g, uint32_t i, int c[3], bool &large) {
  V lo = xyz(g.bodies[BODY_F4 * i + 4]), hi = xyz(g.bodies[BODY_F4 * i + 5]);
  V p = divs(add(mul(lo, 0.5f), mul(hi, 0.5f)), g.cellSize);
  c[0] = (int)clampf(floorf(p.x), -1e9f, 1e9f);
  c[1] = (int)clampf(floorf(p.y), -1e9f, 1e9f);
  c[2] = (int)clampf(floorf(p.z), -1e9f, 1e9f);
  V ext = add(sub(hi, lo), vv(g.p.margin, g.p.margin, g.p.margin));
  large = ext.x > g.cellSize || ext.y > g.cellSize || ext.z > g.cellSize;
}
__device__ __forceinline__ uint32_t bucketOf(const SdfDev &g, int x, int y, int z) {
  return mix(((uint32_t)x * 0x8da6b343u) ^ ((uint32_t)y * 0xd8163841u) ^ ((uint32_t)z * 0xcb1ab31fu)) & g.tableMask;
}
/** Per body its centre cell (or the large list) and its bucket's count. */
__global__ void k_sdf_cells(Dev d, SdfDev g) {
  uint32_t n = d.n;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    uint32_t i = base + threadIdx.x;
    bool in = i < n, large = false;
    int c[3] = {0, 0, 0};
    if (in) {
      cellOf(g, i, c, large);
      g.cellKey[i] = make_int4(c[0], c[1], c[2], large ? 1 : 0);
      if (!large) atomicAdd(&g.bucketCount[bucketOf(g, c[0], c[1], c[2])], 1u);
    }
    uint32_t slot = waveAlloc(g.largeCount, in && large ? 1u : 0u);
    if (in && large) g.largeList[slot] = i;
  }
}
__global__ void k_sdf_scatter(Dev d, SdfDev g) {
  GRID_STRIDE(i, d.n) {
    int4 k = g.cellKey[i];
    if (k.w) continue;
    uint32_t b = bucketOf(g, k.x, k.y, k.z);
    g.sorted[g.bucketStart[b] + atomicAdd(&g.bucketCursor[b], 1u)] = i;
  }
}
/** cachePairs: the previous results by pair, open addressing on pairHash. */
__global__ void k_sdf_history(Dev d, SdfDev g) {
  GRID_STRIDE(i, g.cap) {
    uint4 id = g.prev[(RESULT_HEAD + POINT_F4 * g.p.points) * i];
    if (id.x >= d.n || id.x <= id.y) continue;
    uint32_t h = pairHash(id.x, id.y) & g.cacheMask;
    bool placed = false;
    for (uint32_t probe = 0; probe <= g.cacheMask; probe++) {
      uint32_t old = atomicCAS(&g.history[h], EMPTY, i);
      if (old == EMPTY) { placed = true; break; }
      h = (h + 1u) & g.cacheMask;
    }
    if (!placed) atomicOr(&d.counters[C_OVERFLOW], 1u);
  }
}
__device__ __forceinline__ uint32_t previousSlot(const SdfDev &g, uint32_t a, uint32_t b) {
  uint32_t h = pairHash(a, b) & g.cacheMask;
  for (uint32_t probe = 0; probe <= g.cacheMask; probe++) {
    uint32_t i = g.history[h];
    if (i == EMPTY) return EMPTY;
    uint4 id = g.prev[(RESULT_HEAD + POINT_F4 * g.p.points) * i];
    if (id.x == a && id.y == b) return i;
    h = (h + 1u) & g.cacheMask;
  }
  return EMPTY;
}
/** Candidates of body i (findPairs): a regular body probes the 27 cells around its centre cell for
 *  regular bodies of lower index whose centre cell it is, then the large list; a large body meets
 *  the large bodies of lower index. FILL false counts the accepted ones, true writes them. */
template <bool FILL>
__global__ void k_sdf_find(Dev d, SdfDev g) {
  GRID_STRIDE(i, d.n) {
    int4 k = g.cellKey[i];
    bool large = k.w != 0;
    uint32_t found = 0, tested = 0, out = FILL ? g.pairOffset[i] : 0u;
    auto candidate = [&](uint32_t j) {
      uint32_t a = i > j ? i : j, b = i > j ? j : i;
      tested++;
      if (!sdfPasses(d, g, a, b)) return;
      if (FILL) {
        uint32_t slot = out + found;
        if (slot < g.cap) g.pairs[slot] = make_uint4(a, b, pairHash(a, b), previousSlot(g, a, b));
      }
      found++;
    };
    if (!large) {
      for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
          for (int x = -1; x <= 1; x++) {
            int cx = k.x + x, cy = k.y + y, cz = k.z + z;
            uint32_t b = bucketOf(g, cx, cy, cz);
            uint32_t e = g.bucketStart[b] + g.bucketCount[b];
            for (uint32_t s = g.bucketStart[b]; s < e; s++) {
              uint32_t j = g.sorted[s];
              if (j >= i) continue;
              int4 kj = g.cellKey[j];
              if (kj.x == cx && kj.y == cy && kj.z == cz) candidate(j);
            }
          }
    }
    uint32_t nl = *g.largeCount;
    for (uint32_t s = 0; s < nl; s++) {
      uint32_t j = g.largeList[s];
      if (j == i) continue;
      if (!large || i > j) candidate(j);
    }
    if (!FILL) {
      g.pairCount[i] = found;
      atomicAdd(&g.live[1], tested);
    }
  }
}
/** After the fill: pairs found, the overflow (bit 1, counted into the solver's pairs as the
 *  reference's candidate test does), the empty slots past them. */
__global__ void k_sdf_close(Dev d, SdfDev g) {
  uint32_t total = g.pairOffset[d.n];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    g.live[0] = total;
    g.live[2] = *g.largeCount;
    if (total > g.cap) {
      atomicOr(&d.counters[C_OVERFLOW], 1u);
      atomicAdd(&d.counters[C_PAIRS], total - g.cap);
    }
  }
  GRID_STRIDE(s, g.cap) if (s >= total) g.pairs[s] = make_uint4(EMPTY, EMPTY, EMPTY, EMPTY);
}

// ---- Algorithm 1 per pair -------------------------------------------------------------------------
struct Lane {
  const SdfDev &g;
  FieldBody A, B;
  V lo, hi;
  uint32_t count, queries, iters, idc, rolling, slot, rng;
  uint32_t cachedBits;
  float *lx, *ly, *lz, *ls;  // this lane's LDS columns: x xyz and softmax per point (stride 64)
  float E[3][3];
  __device__ Lane(const SdfDev &gg) : g(gg) {}

  __device__ float uniform() { rng = mix(rng + 0x9e3779b9u); return (float)(rng >> 8) / 16777216.0f; }
  __device__ uint32_t nextU32() { rng = mix(rng + 0x9e3779b9u); return rng; }
  __device__ float smoothMax(float a, float b) const { return 0.5f * ((a + b) + sqrtf((a - b) * (a - b) + g.p.epsilon)); }
  __device__ V px(uint32_t i) const { return vv(lx[64 * i], ly[64 * i], lz[64 * i]); }
  __device__ float objective(V p) {
    queries++;
    float value = smoothMax(worldField(g.atlas, A, p).phi, worldField(g.atlas, B, p).phi);
    for (uint32_t i = 0; i < count; i++) value = value - g.p.alpha * length(sub(p, px(i)));
    return value;
  }
  __device__ void objectiveGradient(V p, V &grad, float &value) {
    queries++;
    Sample a = worldField(g.atlas, A, p), b = worldField(g.atlas, B, p);
    float root = sqrtf((a.phi - b.phi) * (a.phi - b.phi) + g.p.epsilon);
    value = 0.5f * ((a.phi + b.phi) + root);
    float t = 0.0f;
    if (root > 0.0f) t = (a.phi - b.phi) / root;
    grad = add(mul(a.g, 0.5f * (1.0f + t)), mul(b.g, 0.5f * (1.0f - t)));
    for (uint32_t i = 0; i < count; i++) {
      V dx = sub(p, px(i));
      float len = length(dx);
      value = value - g.p.alpha * len;
      if (len > 1e-9f) grad = sub(grad, mul(dx, g.p.alpha / len));
    }
  }
  __device__ V ellipsoidMethod(V seed, float M[3][3]) {
    V x = seed;
    V best = vclamp(x, lo, hi);
    float bestValue = 3.4e38f;
    uint32_t it = 0;
    for (; it < g.p.maxIt; it++) {
      V grad = vv(0.0f, 0.0f, 0.0f);
      float value = 0.0f, violation = 0.0f;
      for (int i = 0; i < 3; i++) {
        float below = comp(lo, i) - comp(x, i), above = comp(x, i) - comp(hi, i);
        if (mx(below, above) > violation) { violation = mx(below, above); grad = vv(0.0f, 0.0f, 0.0f); setComp(grad, i, above > below ? 1.0f : -1.0f); }
      }
      if (violation == 0.0f) {
        objectiveGradient(x, grad, value);
        if (value < bestValue) { bestValue = value; best = x; }
        if (length(grad) < g.p.tol) break;
      }
      V Eg = vv((M[0][0] * grad.x + M[0][1] * grad.y) + M[0][2] * grad.z, (M[1][0] * grad.x + M[1][1] * grad.y) + M[1][2] * grad.z,
                (M[2][0] * grad.x + M[2][1] * grad.y) + M[2][2] * grad.z);
      float gEg = dot(grad, Eg);
      if (!(gEg > 1e-12f)) break;
      float s = sqrtf(gEg);
      float a = mn((violation > 0.0f ? violation : value - bestValue) / s, 0.5f);
      V step = mul(Eg, (1.0f + 3.0f * a) / (4.0f * s));
      x = sub(x, step);
      float cut = 2.0f * (1.0f + 3.0f * a) / ((4.0f * (1.0f + a)) * gEg);
      float shrink = 9.0f * (1.0f - a * a) / 8.0f;
      float N[3][3];
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) N[r][c] = (M[r][c] - (comp(Eg, r) * comp(Eg, c)) * cut) * shrink;
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) M[r][c] = (N[r][c] + N[c][r]) * 0.5f;
      if (length(step) < g.p.tol) break;
    }
    iters += it;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) E[r][c] = M[r][c];
    return best;
  }
  __device__ V onSphere() {
    float z = -1.0f + 2.0f * uniform();
    float angle = 6.2831855f * uniform();
    float r = sqrtf(mx(0.0f, 1.0f - z * z));
    return vv(r * scos(angle), r * ssin(angle), z);
  }
  __device__ V gradientDescent(V seed) {
    V x = vclamp(seed, lo, hi), best = x;
    float bestValue = 3.4e38f;
    uint32_t it = 0;
    for (; it < g.p.maxIt; it++) {
      V gr; float value;
      objectiveGradient(x, gr, value);
      if (value < bestValue) { bestValue = value; best = x; }
      float gl = length(gr);
      if (gl < g.p.tol) break;
      float decay = 1.0f / sqrtf((float)(it + 1u));
      V noise = mul(onSphere(), (0.25f * decay) * mx(gl, 1.0f));
      V next = vclamp(sub(x, mul(add(gr, noise), g.p.stepSize * decay)), lo, hi);
      bool converged = !g.p.sgdStopOnly && length(sub(next, x)) < g.p.tol;
      x = next;
      if (converged) break;
    }
    iters += it;
    return best;
  }
  __device__ V nelderMead(V seed, V step) {
    V x0, x1, x2, x3;
    float f0, f1, f2, f3;
    x0 = vclamp(seed, lo, hi);
    x1 = x0; x1.x = x0.x + ((x0.x + step.x <= hi.x) ? step.x : -step.x); x1 = vclamp(x1, lo, hi);
    x2 = x0; x2.y = x0.y + ((x0.y + step.y <= hi.y) ? step.y : -step.y); x2 = vclamp(x2, lo, hi);
    x3 = x0; x3.z = x0.z + ((x0.z + step.z <= hi.z) ? step.z : -step.z); x3 = vclamp(x3, lo, hi);
    f0 = objective(x0); f1 = objective(x1); f2 = objective(x2); f3 = objective(x3);
    uint32_t iterations = 0;
    for (; iterations < g.p.maxIt; iterations++) {
      // insertion sort of the four vertices by value (stable)
      if (f1 < f0) { float t = f0; f0 = f1; f1 = t; V u = x0; x0 = x1; x1 = u; }
      if (f2 < f1) { float t = f1; f1 = f2; f2 = t; V u = x1; x1 = x2; x2 = u;
        if (f1 < f0) { t = f0; f0 = f1; f1 = t; u = x0; x0 = x1; x1 = u; } }
      if (f3 < f2) { float t = f2; f2 = f3; f3 = t; V u = x2; x2 = x3; x3 = u;
        if (f2 < f1) { t = f1; f1 = f2; f2 = t; u = x1; x1 = x2; x2 = u;
          if (f1 < f0) { t = f0; f0 = f1; f1 = t; u = x0; x0 = x1; x1 = u; } } }
      float extent = mx(length(sub(x1, x0)), mx(length(sub(x2, x0)), length(sub(x3, x0))));
      if (extent < g.p.tol || f3 - f0 < 0.05f * g.p.tol) break;
      V centroid = divs(add(add(x0, x1), x2), 3.0f);
      V reflected = vclamp(add(centroid, sub(centroid, x3)), lo, hi);
      float fr = objective(reflected);
      if (fr < f0) {
        V expanded = vclamp(add(centroid, mul(sub(centroid, x3), 2.0f)), lo, hi);
        float fe = objective(expanded);
        x3 = fe < fr ? expanded : reflected;
        f3 = mn(fe, fr);
      } else if (fr < f2) {
        x3 = reflected; f3 = fr;
      } else {
        V contracted = add(centroid, mul(sub(fr < f3 ? reflected : x3, centroid), 0.5f));
        float fc = objective(contracted);
        if (fc < mn(fr, f3)) { x3 = contracted; f3 = fc; }
        else {
          x1 = add(x0, mul(sub(x1, x0), 0.5f)); f1 = objective(x1);
          x2 = add(x0, mul(sub(x2, x0), 0.5f)); f2 = objective(x2);
          x3 = add(x0, mul(sub(x3, x0), 0.5f)); f3 = objective(x3);
        }
      }
    }
    iters += iterations;
    V best = x0; float fb = f0;
    if (f1 < fb) { best = x1; fb = f1; }
    if (f2 < fb) { best = x2; fb = f2; }
    if (f3 < fb) { best = x3; fb = f3; }
    return best;
  }
  __device__ bool inside(V p) const { return p.x >= lo.x && p.y >= lo.y && p.z >= lo.z && p.x <= hi.x && p.y <= hi.y && p.z <= hi.z; }
  __device__ V seedPoint() {
    float rx = uniform(), ry = uniform(), rz = uniform();
    V p = add(lo, mulv(sub(hi, lo), vv(rx, ry, rz)));
    uint32_t n = 0;
    if (!g.p.curvature) return p;
    for (int side = 0; side < 2; side++) {
      const FieldBody &body = side == 0 ? A : B;
      uint4 hd = g.atlas[body.h];
      for (uint32_t i = 0; i < hd.w; i++) {
        V s = xyz(__builtin_bit_cast(float4, g.atlas[hd.z + i]));
        V cand = worldPoint(body, s);
        if (inside(cand)) { n++; if (nextU32() % n == 0u) p = cand; }
      }
    }
    return p;
  }
  __device__ float4 &stageAt(uint32_t point, uint32_t word) const { return g.stage[(size_t)(point * POINT_F4 + word) * g.cap + slot]; }
  /** addPoint: optimise from `seed`, then classify (depth, normal, separation, anchors, id). */
  __device__ void addPoint(uint32_t ia, uint32_t ib, const float4 *old, V seed, bool cached) {
    float4 eA = make_float4(1.0f, 0.0f, 0.0f, 1.0f), eB = make_float4(0.0f, 1.0f, 0.0f, 0.0f);
    V oldX = cached ? xyz(old[0]) : vv(0.0f, 0.0f, 0.0f), oldN = cached ? xyz(old[1]) : vv(0.0f, 0.0f, 0.0f);
    V p;
    if (g.p.optimizer == 1u) {
      V extent = add(mul(sub(hi, lo), 0.5f), vabs(sub(seed, mul(add(lo, hi), 0.5f))));
      V dd = mul(mulv(extent, extent), 3.0f);
      float M[3][3] = {{dd.x, 0.0f, 0.0f}, {0.0f, dd.y, 0.0f}, {0.0f, 0.0f, dd.z}};
      if (cached) {
        float4 a5 = old[5], b6 = old[6];
        float C[3][3] = {{a5.x, a5.y, a5.z}, {a5.y, a5.w, b6.x}, {a5.z, b6.x, b6.y}};
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 3; c++) M[r][c] = C[r][c] * 0.75f + M[r][c] * 0.25f;
      }
      p = ellipsoidMethod(seed, M);
      eA = make_float4(E[0][0], E[0][1], E[0][2], E[1][1]);
      eB = make_float4(E[1][2], E[2][2], 0.0f, 0.0f);
    } else if (g.p.optimizer == 2u) {
      p = gradientDescent(seed);
    } else {
      float h = 0.5f * g.p.tol;
      V step = vmx(mul(sub(hi, lo), 0.1f), vv(h, h, h));
      if (cached) { float e = mx(length(sub(seed, oldX)), 5.0f * g.p.tol); step = vmn(step, vv(e, e, e)); }
      p = nelderMead(seed, step);
    }
    Sample a = worldField(g.atlas, A, p), b = worldField(g.atlas, B, p);
    queries++;
    V na = unit(a.g), nb = neg(unit(b.g));
    V normal = a.phi < b.phi ? na : nb;
    if (cached) normal = dot(na, oldN) > dot(nb, oldN) ? na : nb;
    V offA = mul(a.g, a.phi / mx(dot(a.g, a.g), 1e-8f)), offB = mul(b.g, b.phi / mx(dot(b.g, b.g), 1e-8f));
    float sep = a.phi + b.phi;
    if (a.phi > 0.0f || b.phi > 0.0f) {
      V gap = sub(offA, offB);
      sep = length(gap);
      if (g.p.signedSep && (a.phi <= 0.0f || b.phi <= 0.0f)) {
        bool fixedA = g.bodies[BODY_F4 * ia + 2].w != 0.0f && A.kind == K_BOX;
        bool fixedB = g.bodies[BODY_F4 * ib + 2].w != 0.0f && B.kind == K_BOX;
        if (fixedA || fixedB) {
          normal = fixedA ? na : nb;
          sep = dot(gap, normal);
          V moving = fixedA ? b.g : a.g;
          if (dot(moving, moving) < 0.25f) sep = a.phi + b.phi;
        }
      }
      if (a.phi > 0.0f && b.phi > 0.0f && sep > 1e-6f && !(cached && dot(gap, oldN) < 0.0f)) normal = divs(gap, sep);
    }
    uint32_t id = cached ? __float_as_uint(old[4].z) : 0u;
    if (!cached) {
      idc++;
      if (idc == SPHERE_FEATURE) idc++;
      if (idc == 0u || idc >= SDF_ROLLING_BIT) idc = 1u;
      id = idc;
    }
    Q arA = q4(g.bodies[BODY_F4 * ia + 6]), arB = q4(g.bodies[BODY_F4 * ib + 6]);
    V rA = qrotate(arA, invrot(A.rot, sub(sub(p, offA), A.pos)));
    V rB = qrotate(arB, invrot(B.rot, sub(sub(p, offB), B.pos)));
    float soft = smoothMax(a.phi, b.phi);
    uint32_t k = count;
    stageAt(k, 0) = f4(p, mn(a.phi, b.phi));
    stageAt(k, 1) = f4(normal, sep);
    stageAt(k, 2) = f4(rA, soft);
    stageAt(k, 3) = f4(rB, 0.0f);
    stageAt(k, 4) = make_float4(__uint_as_float(sep < g.p.margin ? 1u : 0u), __uint_as_float(cached ? 1u : 0u), __uint_as_float(id), 0.0f);
    stageAt(k, 5) = eA;
    stageAt(k, 6) = eB;
    lx[64 * k] = p.x; ly[64 * k] = p.y; lz[64 * k] = p.z; ls[64 * k] = soft;
    if (cached) cachedBits |= 1u << k;
    count = k + 1;
  }
};

/** One lane per result slot: the pair's Algorithm 1, then the result (identity, stats, points in
 *  rank order) written to its slot. Dynamic LDS: 4 floats x points x 64 lanes (+ 2 bytes x points
 *  x 64 of order indices). */
__global__ void __launch_bounds__(64) k_sdf_narrow(Dev d, SdfDev g) {
  extern __shared__ float lds[];
  const uint32_t P = g.p.points, lane = threadIdx.x;
  float *col = lds + lane;
  uint8_t *order = (uint8_t *)(lds + 4 * P * 64) + lane;
  const uint32_t stride = RESULT_HEAD + POINT_F4 * P;
  for (uint32_t base = blockIdx.x * 64u; base < g.cap; base += gridDim.x * 64u) {
    uint32_t slot = base + lane;
    if (slot >= g.cap) continue;
    uint4 *res = g.cur + (size_t)stride * slot;
    uint4 pr = g.pairs[slot];
    if (slot >= g.p.pairCount || pr.x == EMPTY) {
      res[0] = make_uint4(EMPTY, pr.y, pr.z, 0u);
      res[1] = make_uint4(0u, 0u, 0u, 0u);
      continue;
    }
    Lane L(g);
    L.slot = slot;
    L.lx = col; L.ly = col + 64 * P; L.lz = col + 128 * P; L.ls = col + 192 * P;
    L.A = fieldBody(g, pr.x);
    L.B = fieldBody(g, pr.y);
    L.count = 0; L.queries = 0; L.iters = 0; L.idc = 0; L.cachedBits = 0;
    bool rollA = L.A.kind >= 1u && L.A.kind <= 5u, rollB = L.B.kind >= 1u && L.B.kind <= 5u;
    L.rolling = (rollA || rollB) ? 1u : 0u;
    float m = 0.5f * g.p.margin;
    V mm = vv(m, m, m);
    V alo = xyz(g.bodies[BODY_F4 * pr.x + 4]), ahi = xyz(g.bodies[BODY_F4 * pr.x + 5]);
    V blo = xyz(g.bodies[BODY_F4 * pr.y + 4]), bhi = xyz(g.bodies[BODY_F4 * pr.y + 5]);
    L.lo = vmx(sub(alo, mm), sub(blo, mm));
    L.hi = vmn(add(ahi, mm), add(bhi, mm));
    bool fixedA = g.bodies[BODY_F4 * pr.x + 2].w != 0.0f, fixedB = g.bodies[BODY_F4 * pr.y + 2].w != 0.0f;
    if (L.lo.x > L.hi.x || L.lo.y > L.hi.y || L.lo.z > L.hi.z || (fixedA && fixedB)) {
      res[0] = make_uint4(pr.x, pr.y, pr.z, 0u);
      res[1] = make_uint4(0u, 0u, 0u, L.rolling);
      continue;
    }
    L.rng = mix(pr.z ^ mix(g.p.step) ^ mix(pr.x) ^ mix(pr.y + 0x9e3779b9u));
    uint32_t cached = 0, active = 0;
    const uint4 *prev = nullptr;
    if (g.p.warmStart && pr.w != EMPTY && pr.w < g.cap) {
      const uint4 *q = g.prev + (size_t)stride * pr.w;
      uint4 qid = q[0];
      if (qid.x == pr.x && qid.y == pr.y && qid.z == pr.z) prev = q;
    }
    if (prev) {
      L.idc = prev[1].z;
      uint32_t pc = prev[0].w < P - 1u ? prev[0].w : P - 1u;
      V velA = xyz(g.bodies[BODY_F4 * pr.x + 2]), omA = xyz(g.bodies[BODY_F4 * pr.x + 3]);
      V velB = xyz(g.bodies[BODY_F4 * pr.y + 2]), omB = xyz(g.bodies[BODY_F4 * pr.y + 3]);
      for (uint32_t i = 0; i < pc; i++) {
        const float4 *pt = (const float4 *)(prev + RESULT_HEAD + POINT_F4 * i);
        float4 old[7];
#pragma unroll
        for (int w = 0; w < 7; w++) old[w] = pt[w];
        if (__float_as_uint(old[4].z) >= WITNESS_ID) continue;
        if (cached >= g.p.cacheSize) break;
        cached++; active += __float_as_uint(old[4].x);
        V x = xyz(old[0]);
        V va = add(velA, cross(omA, sub(x, L.A.pos))), vb = add(velB, cross(omB, sub(x, L.B.pos)));
        V vel = mul(add(va, vb), 0.5f);
        if (fixedA) vel = vb; else if (fixedB) vel = va;
        L.addPoint(pr.x, pr.y, old, add(x, mul(vel, g.p.dt)), true);
      }
    }
    bool convexA = g.atlas[L.A.h].y != 0u, convexB = g.atlas[L.B.h].y != 0u;
    bool skip = g.p.skipConvex && g.p.optimizer != 2u && active > 0 && active == g.p.cacheSize && convexA && convexB;
    uint32_t samples = (skip ? 0u : g.p.randomSamples) + g.p.cacheSize - cached;
    for (uint32_t i = 0; i < samples; i++) { V s = L.seedPoint(); L.addPoint(pr.x, pr.y, nullptr, s, false); }
    // first-wins duplicate rejection (the kept points in creation order), then the rank of each and
    // a stable insertion sort by rank
    float tol2 = g.p.tol * g.p.tol;
    uint32_t unique = 0;
    for (uint32_t i = 0; i < L.count; i++) {
      V xi = L.px(i);
      bool dup = false;
      for (uint32_t j = 0; j < unique; j++) { V dd = sub(xi, L.px(order[64 * j])); if (dot(dd, dd) < tol2) dup = true; }
      if (!dup) { order[64 * unique] = (uint8_t)i; unique++; }
    }
    for (uint32_t i = 0; i < unique; i++) {
      uint32_t oi = order[64 * i];
      V xi = L.px(oi);
      float rank = L.ls[64 * oi] + (((L.cachedBits >> oi) & 1u) ? 0.0f : g.p.replacement);
      for (uint32_t j = 0; j < unique; j++) rank = rank - g.p.alpha * length(sub(xi, L.px(order[64 * j])));
      L.ls[64 * oi] = rank;  // the softmax is no longer read
    }
    for (uint32_t i = 1; i < unique; i++)
      for (uint32_t j = i; j > 0; j--) {
        uint8_t oj = order[64 * j], op = order[64 * (j - 1)];
        if (!(L.ls[64 * oj] < L.ls[64 * op])) break;
        order[64 * j] = op; order[64 * (j - 1)] = oj;
      }
    // mesh vertex witnesses (meshPlaneContacts, else meshMeshContacts)
    bool witness = false;
    float4 wp[7];
    if (g.p.meshVertex) {
      float lowest = 1e30f;
      uint32_t wid = 0;
      V wpos = vv(0.0f, 0.0f, 0.0f), wn = vv(0.0f, 0.0f, 0.0f), wa = wpos, wb = wpos;
      bool boxA = fixedA && L.A.kind == K_BOX, boxB = fixedB && L.B.kind == K_BOX;
      if (boxA || boxB) {
        const FieldBody &mesh = boxA ? L.B : L.A, &plane = boxA ? L.A : L.B;
        uint32_t meshIdx = boxA ? pr.y : pr.x;
        bool meshFixed = boxA ? fixedB : fixedA;
        uint4 mh = g.atlas[mesh.h];
        uint32_t cnt = g.atlas[mesh.h + 3].w;
        bool ok = !meshFixed && cnt > 0;
        V half = vv(plane.p0.x, plane.p0.y, plane.p0.z);
        if (ok) {
          V mlo = xyz(g.bodies[BODY_F4 * meshIdx + 4]), mhi = xyz(g.bodies[BODY_F4 * meshIdx + 5]);
          for (uint32_t i = 0; i < 8 && ok; i++) {
            V corner = vv((i & 1u) ? mhi.x : mlo.x, (i & 2u) ? mhi.y : mlo.y, (i & 4u) ? mhi.z : mlo.z);
            V q = localPoint(plane, corner);
            if (fabsf(q.x) >= half.x || fabsf(q.z) >= half.z || q.y <= -half.y) ok = false;
          }
        }
        if (ok) {
          V up = qrotate(plane.rot, vv(0.0f, 1.0f, 0.0f));
          V normal = boxA ? up : neg(up);
          uint32_t best = 0;
          for (uint32_t i = 0; i < cnt; i++) {
            V q = worldPoint(mesh, xyz(__builtin_bit_cast(float4, g.atlas[mh.z + mh.w + i])));
            float s = (localPoint(plane, q).y - half.y) * plane.scale;
            if (s < lowest) { lowest = s; best = i; }
          }
          L.queries += cnt;
          if (lowest < mn(g.p.margin, 0.5f * g.p.tol)) {
            V q = worldPoint(mesh, xyz(__builtin_bit_cast(float4, g.atlas[mh.z + mh.w + best])));
            V projected = sub(q, mul(up, lowest));
            wpos = q; wn = normal; wa = boxA ? projected : q; wb = boxA ? q : projected;
            wid = WITNESS_ID + best;
            witness = true;
          }
        }
      } else if (!fixedA && !fixedB) {
        float bw = 1e30f;
        V bp = vv(0.0f, 0.0f, 0.0f);
        uint32_t id = 0;
        for (int side = 0; side < 2; side++) {
          const FieldBody &mesh = side == 0 ? L.A : L.B, &other = side == 0 ? L.B : L.A;
          uint4 mh = g.atlas[mesh.h];
          uint32_t cnt = g.atlas[mesh.h + 3].w;
          cnt = cnt < 0x08000000u ? cnt : 0x08000000u;
          for (uint32_t i = 0; i < cnt; i++) {
            V q = worldPoint(mesh, xyz(__builtin_bit_cast(float4, g.atlas[mh.z + mh.w + i])));
            if (!L.inside(q)) continue;
            Sample s = worldField(g.atlas, other, q);
            L.queries++;
            if (s.phi < bw && dot(s.g, s.g) >= 0.25f) { bw = s.phi; bp = q; id = WITNESS_ID | ((uint32_t)side << 27) | i; }
          }
        }
        if (id != 0u && bw < mn(g.p.margin, 0.5f * g.p.tol)) {
          lowest = bw; wpos = bp; wa = bp; wb = bp; wid = id;
          if (((id >> 27) & 1u) != 0u) { V n = unit(worldField(g.atlas, L.A, bp).g); wn = n; wa = sub(bp, mul(n, bw)); }
          else { V n = unit(worldField(g.atlas, L.B, bp).g); wn = neg(n); wb = sub(bp, mul(n, bw)); }
          witness = true;
        }
      }
      if (witness) {
        uint32_t wc = 0;
        if (prev) for (uint32_t j = 0; j < prev[0].w; j++) if (__float_as_uint(((const float4 *)(prev + RESULT_HEAD + POINT_F4 * j))[4].z) == wid) wc = 1u;
        Q arA = q4(g.bodies[BODY_F4 * pr.x + 6]), arB = q4(g.bodies[BODY_F4 * pr.y + 6]);
        wp[0] = f4(wpos, mn(0.0f, lowest));
        wp[1] = f4(wn, lowest);
        wp[2] = f4(qrotate(arA, invrot(L.A.rot, sub(wa, L.A.pos))), 0.0f);
        wp[3] = f4(qrotate(arB, invrot(L.B.rot, sub(wb, L.B.pos))), 0.0f);
        wp[4] = make_float4(__uint_as_float(1u), __uint_as_float(wc), __uint_as_float(wid), 0.0f);
        wp[5] = make_float4(1.0f, 0.0f, 0.0f, 1.0f);
        wp[6] = make_float4(0.0f, 1.0f, 0.0f, 0.0f);
      }
    }
    uint32_t outCount = unique, wslot = P;
    if (witness) { wslot = unique < P - 1u ? unique : P - 1u; outCount = wslot + 1u; }
    res[0] = make_uint4(pr.x, pr.y, pr.z, outCount);
    res[1] = make_uint4(L.queries, L.iters, L.idc, L.rolling);
    float4 *outp = (float4 *)(res + RESULT_HEAD);
    for (uint32_t k = 0; k < outCount; k++) {
      float4 *o = outp + POINT_F4 * k;
      if (k == wslot) {
#pragma unroll
        for (int w = 0; w < 7; w++) o[w] = wp[w];
        continue;
      }
      uint32_t src = order[64 * k];
#pragma unroll
      for (int w = 0; w < 7; w++) o[w] = L.stageAt(src, w);
      o[3].w = L.ls[64 * src];
    }
  }
}

// ---- export (sdfExportSource) ---------------------------------------------------------------------
__global__ void k_sdf_export(Dev d, SdfDev g) {
  const uint32_t P = g.p.points, stride = RESULT_HEAD + POINT_F4 * P;
  const float ag = d.p.alpha * d.p.gamma;
  for (uint32_t base = blockIdx.x * blockDim.x; base < g.cap; base += gridDim.x * blockDim.x) {
    uint32_t slot = base + threadIdx.x;
    uint4 *res = g.cur + (size_t)stride * (slot < g.cap ? slot : 0u);
    uint4 id = slot < g.cap ? res[0] : make_uint4(EMPTY, 0u, 0u, 0u);
    bool valid = slot < g.cap && id.x != EMPTY;
    uint32_t n = valid ? id.w : 0u;
    uint32_t want = 0;
    for (uint32_t i = 0; i < n; i++) {
      float4 *pt = (float4 *)(res + RESULT_HEAD + POINT_F4 * i);
      pt[4].w = __uint_as_float(0u);
      want += __float_as_uint(pt[4].x) != 0u ? 1u : 0u;
    }
    (void)waveAlloc(&d.counters[C_PAIRS], valid ? 1u : 0u);
    uint32_t mBase = waveAlloc(&d.counters[C_MANIFOLDS], want);
    uint32_t fit = mBase >= d.manCap ? 0u : (d.manCap - mBase < want ? d.manCap - mBase : want);
    uint32_t cBase = waveAlloc(&d.counters[C_CONTACTS], fit);
    if (!valid || want == 0) continue;
    uint32_t a = id.x, b = id.y;
    uint32_t rolling = res[1].w;
    V pA = xyz(d.b.pos[a]), pB = xyz(d.b.pos[b]);
    Q qA = q4(d.b.rot[a]), qB = q4(d.b.rot[b]);
    float friction = sqrtf(d.b.pos[a].w * d.b.pos[b].w);
    float pmin = contactPenaltyMin(d, a, b);
    const uint4 *prev = nullptr;
    uint4 pid = make_uint4(EMPTY, 0u, 0u, 0u);
    uint32_t pslot = g.pairs[slot].w;
    if (pslot != EMPTY && pslot < g.cap) {
      prev = g.prev + (size_t)stride * pslot;
      pid = prev[0];
      if (!(pid.x == id.x && pid.y == id.y && pid.z == id.z)) prev = nullptr;
    }
    uint32_t prevContacts = d.counters[C_PREV_CONTACTS];
    uint32_t k = 0;
    for (uint32_t i = 0; i < n; i++) {
      float4 *pt = (float4 *)(res + RESULT_HEAD + POINT_F4 * i);
      float4 n4 = pt[1], a4 = pt[2], b4 = pt[3], f4w = pt[4];
      if (__float_as_uint(f4w.x) == 0u) continue;
      uint32_t m = mBase + k;
      k++;
      if (m >= d.manCap) { atomicOr(&d.counters[C_OVERFLOW], 4u); continue; }
      uint32_t c = cBase + (m - mBase);
      V nrm = neg(xyz(n4));
      V basis[3];
      orthonormal(nrm, basis);
      V rA = xyz(a4), rB = xyz(b4);
      V lam = vv(0.0f, 0.0f, 0.0f), pen = vv(0.0f, 0.0f, 0.0f);
      uint32_t stick = 0;
      uint32_t pointId = __float_as_uint(f4w.z);
      if (__float_as_uint(f4w.y) != 0u && prev) {
        for (uint32_t j = 0; j < pid.w; j++) {
          const float4 *prior = (const float4 *)(prev + RESULT_HEAD + POINT_F4 * j);
          float4 pf = prior[4];
          uint32_t ex = __float_as_uint(pf.w);
          if (__float_as_uint(pf.z) == pointId && ex > 0u && ex <= prevContacts) {
            Contact kk = d.pcon[ex - 1u];
            V ob[3];
            orthonormal(neg(xyz(prior[1])), ob);
            V kl = xyz(kk.lam);
            V w = add(add(mul(ob[0], kl.x), mul(ob[1], kl.y)), mul(ob[2], kl.z));
            lam = vv(dot(basis[0], w), dot(basis[1], w), dot(basis[2], w));
            pen = xyz(kk.pen);
            stick = rolling ? 0u : (__float_as_uint(kk.rA.w) & STICK_BIT);
            V krA = xyz(kk.rA), krB = xyz(kk.rB);
            if (stick != 0u && g.p.refresh > 0.0f) {
              V retained = sub(sub(add(pA, qrotate(qA, krA)), pB), qrotate(qB, krB));
              float moved = mx(length(sub(rA, krA)), length(sub(rB, krB)));
              if (moved > 3.0f * g.p.refresh && n4.w < dot(nrm, retained) - g.p.refresh) stick = 0u;
            }
            if (stick != 0u) { rA = krA; rB = krB; }
            break;
          }
        }
      }
      V dd = sub(sub(add(pA, qrotate(qA, rA)), pB), qrotate(qB, rB));
      float sep = stick != 0u ? dot(nrm, dd) : n4.w;
      float c0x = (sep - g.p.margin) + g.p.allowedPen;
      float c0y = stick != 0u ? dot(basis[1], dd) : 0.0f, c0z = stick != 0u ? dot(basis[2], dd) : 0.0f;
      lam = mul(lam, ag);
      pen = vv(clampf(pen.x * d.p.gamma, PENALTY_MIN, PENALTY_MAX), clampf(pen.y * d.p.gamma, PENALTY_MIN, PENALTY_MAX),
               clampf(pen.z * d.p.gamma, PENALTY_MIN, PENALTY_MAX));
      if (d.p.flags & FLAG_MASS_PENALTY) pen.x = mx(pen.x, pmin);
      if (c >= d.conCap) {
        atomicOr(&d.counters[C_OVERFLOW], 2u);
        d.mIds[m] = make_uint4(a, b, 0u, 0u);
        d.mGeo[m] = f4(nrm, 0.0f);
        continue;
      }
      Contact out;
      out.rA = f4(rA, __uint_as_float((pointId & ~STICK_BIT) | stick | (rolling ? SDF_ROLLING_BIT : 0u)));
      out.rB = f4(rB, c0x);
      out.pen = f4(pen, c0y);
      out.lam = f4(lam, c0z);
      d.con[c] = out;
      d.cman[c] = m;
      d.mIds[m] = make_uint4(a, b, c, 1u | (d.counters[C_STEP] << 4));
      d.mGeo[m] = f4(nrm, friction);
      pt[4].w = __uint_as_float(c + 1u);
    }
  }
}

/** copyHistory (wgsl-sdf-history.ts): results carried into grown storage (every slot retained). */
__global__ void k_sdf_copy_history(const uint4 *src, uint32_t srcCap, uint4 *dst, uint32_t dstCap, uint32_t stride) {
  GRID_STRIDE(i, dstCap) {
    uint4 *o = dst + (size_t)stride * i;
    if (i >= srcCap) { o[0] = make_uint4(EMPTY, EMPTY, EMPTY, 0u); o[1] = make_uint4(0u, 0u, 0u, 0u); continue; }
    const uint4 *s = src + (size_t)stride * i;
    for (uint32_t w = 0; w < stride; w++) o[w] = s[w];
  }
}

}  // namespace sdfk
}  // namespace avbd
