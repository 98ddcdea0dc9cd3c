// Kernels of one AVBD step on CDNA4 (gfx950). Each is a grid-stride loop over a count that
// lives in device memory, so no step ever waits on the host and the whole step can be captured
// in a hipGraph. Mirrors GpuSolver3D.step's phases (src/avbd3d/gpu/solver.ts:2174-2261):
// collision (refs, hash of last step's pairs, uniform-grid broadphase, narrowphase with warm
// start), adjacency (CSR + per-body sort), Jones-Plassmann colouring with stable buckets, and
// the solve (warm starts, iterations x (per-colour primal, dual), velocities).
#pragma once
#include "avbd_device.h"
#include "avbd_hull.h"

namespace avbd {

struct Dev {
  uint32_t n, nj;
  Bodies b;
  float4 *refPos, *refRot;
  uint2 *filters;
  float4 *joints;  // 10 float4 per joint (layout.ts:232-243)
  // per body (null until avbd_set_body_aux): gravity scale, the per-step velocity decay factors
  // exp(-rate dt) (linear, angular), 1 when either rate is nonzero (the reference's aux record)
  const float4 *aux;
  int4 *jinfo;     // type, a (-1 world), b, 0
  uint32_t *statics;  // large bodies [largeCount] | no-collide (hi, lo, slot) triples sorted by (hi, lo)
  uint32_t largeCount, noCollideCount;
  const uint32_t *ncFirst;  // n + 1: body b's no-collide entries (as the higher body) are [ncFirst[b], ncFirst[b + 1])
  // worlds: bodies [worldStart[w], worldStart[w+1]) never touch another world's; large bodies are
  // listed by world (statics[largeStart[w] .. largeStart[w+1]))
  uint32_t *worldOf, *worldStart, *largeStart;
  float *largeRadius;  // per world: bodies with a larger bounding radius are tested against all of theirs
  uint32_t worlds;
  const uint4 *hulls;
  // uniform grid, two levels (uploadStatics, solver.ts:1263-1290): small bodies with a bounding
  // radius above fineRadius use cellSize cells (level 1), the rest fineCell cells (level 0);
  // hasFine 0: one level (level 0, cellSize cells). coarseMovers: level 1 holds dynamic bodies.
  float cellSize, fineCell, fineRadius;
  uint32_t hasFine, coarseMovers;
  uint32_t tableMask;
  uint32_t *cellStart, *cellSorted;
  // the cell sort: keys (cell hashes; tableMask + 1 for bodies not in the grid) and body indexes,
  // ping-pong, and the per-block digit histograms
  uint32_t *sortKey[2], *sortIdx[2], *sortHist;
  int4 *cellOf;
  float4 *sortedPos;  // pos.xyz, bounding radius of cellSorted[k]: the candidate scan's stream
  int4 *sortedKey;    // cell xyz, world | 0x80000000 when dynamic, of cellSorted[k]
  // pairs, manifolds (current / previous ping-pong), contacts
  uint2 *pairs;
  uint32_t pairCap;
  uint4 *mIds, *pmIds;
  float4 *mGeo, *pmGeo;
  Contact *con, *pcon;
  uint32_t *cman;  // per contact of this step, its manifold (the dual's thread-per-contact map)
  int *src;  // when recording: per contact, the previous step's contact it warm-started from (-1: none)
  uint2 *hullPairs;  // pairs with a hull on either side, for the wavefront-per-pair narrowphase (null: inline)
  uint2 *boxPairs;   // box-box pairs, collided in wavefronts of their own (null: inline)
  float4 *hullOut;   // per hull pair, its contact for the emit pass (HULL_OUT float4)
  uchar4 *bits;  // when recording: this iteration's branch bits per contact (x: body a's update, y: b's, z: dual)
  uint32_t manCap, conCap;
  uint32_t *table;
  uint32_t hashMask;
  // adjacency
  uint32_t *adjCount, *adjStart, *adjFill, *adjList;
  uint32_t *adjTmp, *adjOwner;  // the fill's unsorted entries and their bodies
  uint2 *adjKeys;               // and their sort keys
  int *adjOther;                // per sorted entry, its other body when dynamic, else -1 (the colouring's neighbour)
  uint32_t *prio;               // per body, its colouring priority (fixed at creation)
  uint32_t adjCap;
  // the solve's flattened views: per adjacency entry its manifold record (ids.x = 0xffffffff marks
  // a joint, id in ids.y); per body its walk items (one per joint entry and per contact point of
  // its manifold entries, adjacency order: contact or joint slot, entry, partner, bit 0 the body is
  // A, bit 1 a joint) from itemStart; per colour-bucket slot its body, item range and whether it is
  // enabled
  uint4 *adjIds;
  float4 *adjGeo;
  uint32_t *itemStart;
  uint4 *items;
  uint32_t itemCap;
  uint4 *colorItems;
  // colouring
  uint32_t *colA, *colB, *colorStart, *colorBodies, *colorRank, *hist;
  uint32_t colorGroups;
  // the dataflow sweep: per body its order key (colour << 26 | its chunk within a serial colour)
  // and the iterations of this substep's sweep it has completed
  uint32_t *bodyKey, *flowDone;
  uint32_t *flowOrphan;  // per joint: 1 when no body walks it (its duals run apart from the sweep)
  uint32_t *counters;
  Params p;
};

#define GRID_STRIDE(i, count) for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (count); i += gridDim.x * blockDim.x)

__device__ __forceinline__ bool isDynamic(const Dev &d, int i) { return i >= 0 && d.b.size[i].w > 0.0f; }
__device__ __forceinline__ uint32_t manifoldCount(const Dev &d) { uint32_t m = d.counters[C_MANIFOLDS]; return m < d.manCap ? m : d.manCap; }
__device__ __forceinline__ uint32_t prevManifoldCount(const Dev &d) { return d.counters[C_PREV_MANIFOLDS]; }

// ---------------------------------------------------------------------------------------------
// Utilities
/** Fills p[0, n) with v; nothing when skip is given and its bit 0 is set (a kept adjacency). */
__global__ void k_fill(uint32_t *p, uint32_t n, uint32_t v, const uint32_t *skip) {
  if (skip && (*skip & 1u)) return;
  GRID_STRIDE(i, n) p[i] = v;
}

/** beginFrame (wgsl-collision.ts:97-107): last step's counts become "previous". */
__global__ void k_begin(Dev d, int reuseManifolds, int mayKeepAdjacency) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t *c = d.counters;
  uint32_t m = c[C_MANIFOLDS], k = c[C_CONTACTS];
  c[C_STICKY] |= c[C_OVERFLOW] | (c[C_CLASHES] ? 0x10000u : 0u);
  if (m > c[C_MAX_MANIFOLDS]) c[C_MAX_MANIFOLDS] = m;
  if (k > c[C_MAX_CONTACTS]) c[C_MAX_CONTACTS] = k;
  c[C_PREV_MANIFOLDS] = m < d.manCap ? m : d.manCap;
  c[C_PREV_CONTACTS] = k < d.conCap ? k : d.conCap;
  c[C_MANIFOLDS] = 0;
  c[C_CONTACTS] = 0;
  if (!reuseManifolds) c[C_PAIRS] = 0;
  c[C_PREV_OVERFLOW] = c[C_OVERFLOW];
  c[C_OVERFLOW] = 0;
  c[C_STEP] += 1u;  // the step clock lives on the device, so a captured substep replays as it ran
  c[C_TOPO] &= ~1u;  // the adjacency is rebuilt unless this substep's reuse keeps it
  if (!mayKeepAdjacency) c[C_TOPO] |= 2u;  // the host's word: rebuild (buffers regrown, or a lone step)
  c[C_HULL_PAIRS] = 0;
  c[C_BOX_PAIRS] = 0;
  c[C_CLASHES] = 0;
  c[C_NUM_COLORS] = 0;
  // this step's colouring marks its clashing colours afresh (colorStarts clears them,
  // wgsl-topology.ts:313-315): a colour that clashed once must not stay serial
  c[C_CLASHING] = 0;
  c[C_CLASHING + 1] = 0;
}

/** lifecycleBits (layout.ts:26-31): the flags in x-.w when it holds a whole number 0..7 (1 rewritten
 *  slot, 2 disabled, 4 cold); bodyEnabled / bodyCold (layout.ts:29-32). A disabled body is left
 *  out of the grid, the pairs, the warm starts, the solve and the velocities; its flag stays
 *  until the host clears it. A step's warm start consumes the others. */
__device__ __forceinline__ uint32_t lifecycleBits(float w) { return (w >= 0.0f && w <= 7.0f) ? (uint32_t)w : 0u; }
__device__ __forceinline__ bool bodyEnabled(const Dev &d, uint32_t i) { return (lifecycleBits(d.b.ipos[i].w) & 2u) == 0u; }
__device__ __forceinline__ bool bodyCold(const Dev &d, uint32_t i) { return (lifecycleBits(d.b.ipos[i].w) & 4u) != 0u; }
/** updateRefs (wgsl-collision.ts:34-48): reference poses for contact reuse. */
__global__ void k_update_refs(Dev d) {
  GRID_STRIDE(i, d.n) {
    if (!bodyEnabled(d, i)) continue;
    V pos = xyz(d.b.pos[i]);
    Q rot = q4(d.b.rot[i]);
    float4 lp = d.refPos[i], lr = d.refRot[i];
    Q last = q4(lr);
    bool moved = dot4(last, last) == 0.0f || length(sub(pos, xyz(lp))) > 0.002f || length(qsub(rot, last)) > 0.002f;
    if (moved) {
      d.refPos[i] = f4(pos, 0.0f);
      d.refRot[i] = f4q(rot);
      float4 ip = d.b.inpos[i];
      ip.w = __uint_as_float(d.counters[C_STEP]);
      d.b.inpos[i] = ip;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Hash of last step's pairs (wgsl-collision.ts:344-381), compare-and-swap linear probing.
__device__ __forceinline__ uint32_t pairHash(uint32_t a, uint32_t b) { return hash32((a * 0x9e3779b1u) ^ hash32(b)); }

__global__ void k_hash_insert(Dev d) {
  uint32_t count = prevManifoldCount(d);
  GRID_STRIDE(m, count) {
    uint4 ids = d.pmIds[m];
    // Rewritten slots have no contact history
    if (d.b.ipos[ids.x].w != 0.0f || d.b.ipos[ids.y].w != 0.0f) continue;
    uint32_t h = pairHash(ids.x, ids.y) & d.hashMask;
    for (uint32_t probe = 0; probe <= d.hashMask; probe++) {
      if (atomicCAS(&d.table[h], 0u, m + 1) == 0u) break;
      h = (h + 1) & d.hashMask;
    }
  }
}
__device__ __forceinline__ int hashFind(const Dev &d, uint32_t a, uint32_t b) {
  uint32_t h = pairHash(a, b) & d.hashMask;
  for (uint32_t probe = 0; probe <= d.hashMask; probe++) {
    uint32_t v = d.table[h];
    if (v == 0) return -1;
    uint4 ids = d.pmIds[v - 1];
    if (ids.x == a && ids.y == b) return (int)(v - 1);
    h = (h + 1) & d.hashMask;
  }
  return -1;
}

// ---------------------------------------------------------------------------------------------
// Broadphase: a hashed uniform grid over bounding spheres, in two levels when the scene's small
// bodies separate by size (cells 2x the largest radius of each level), a list of large bodies
// tested against every body, and the reference's exact pair predicate (wgsl-collision.ts:150-234).
// The predicate decides; the grid only has to be conservative. A cell key is (x, y, z, slot) with
// slot = world * 4 + level * 2 + fixed: moving and fixed bodies sit in separate cells of the table.
__device__ __forceinline__ bool isLarge(const Dev &d, uint32_t i) { return d.b.moment[i].w > d.largeRadius[d.worldOf[i]]; }
__device__ __forceinline__ uint32_t levelOf(const Dev &d, uint32_t i) { return d.hasFine && d.b.moment[i].w > d.fineRadius ? 1u : 0u; }
__device__ __forceinline__ int4 cellAt(const Dev &d, V pos, uint32_t level) {
  float size = level || !d.hasFine ? d.cellSize : d.fineCell;
  return make_int4((int)floorf(pos.x / size), (int)floorf(pos.y / size), (int)floorf(pos.z / size), 0);
}
__device__ __forceinline__ uint32_t cellHash(const Dev &d, int4 c) {
  return (((uint32_t)c.x * 73856093u) ^ ((uint32_t)c.y * 19349663u) ^ ((uint32_t)c.z * 83492791u) ^ ((uint32_t)c.w * 2654435769u)) & d.tableMask;
}
/** A small body's cell key at its own level. */
__device__ __forceinline__ int4 ownCell(const Dev &d, uint32_t i) {
  uint32_t level = levelOf(d, i);
  int4 c = cellAt(d, xyz(d.b.pos[i]), level);
  c.w = (int)(d.worldOf[i] * 4u + level * 2u + (d.b.size[i].w > 0.0f ? 0u : 1u));
  return c;
}
/** The grid's sort keys: each small enabled body's cell (at its level) and the cell's hash, the
 *  rest a key past every hash; body indexes as the payload. */
__global__ void k_grid_keys(Dev d) {
  GRID_STRIDE(i, d.n) {
    uint32_t key = d.tableMask + 1u;
    if (bodyEnabled(d, i) && !isLarge(d, i)) {
      int4 c = ownCell(d, i);
      d.cellOf[i] = c;
      key = cellHash(d, c);
    }
    d.sortKey[0][i] = key;
    d.sortIdx[0][i] = i;
  }
}
/** One pass of an LSD radix sort of (key, index) over `bits`-bit digits at `shift` (at most 10
 *  bits), 256 elements per workgroup staged through LDS. Each element's rank among the wavefront's
 *  elements of its digit comes from one ballot per digit bit (the lanes agreeing on every bit);
 *  the wavefronts' digit counts sit in LDS. Without SCATTER the workgroup writes its digit
 *  histogram (digit-major, for an exclusive scan over digits and workgroups); with SCATTER each
 *  element goes to the scanned offset of its digit and workgroup plus its rank in the workgroup.
 *  Equal digits keep their order: the sort is stable, so after the last pass bodies are ordered
 *  by key and, within a key, by index. */
template <bool SCATTER>
__global__ void __launch_bounds__(256) k_radix(const uint32_t *kin, const uint32_t *iin, uint32_t *kout, uint32_t *iout, uint32_t *hist,
                                               uint32_t n, uint32_t shift, uint32_t bits) {
  __shared__ uint32_t waveCount[4][1024];
  const uint32_t bins = 1u << bits, t = threadIdx.x, w = t >> 6, lane = t & 63u, e = blockIdx.x * 256u + t;
  for (uint32_t k = t; k < 4u * bins; k += 256u) waveCount[k >> bits][k & (bins - 1u)] = 0u;
  __syncthreads();
  bool valid = e < n;
  uint32_t key = valid ? kin[e] : 0u, digit = (key >> shift) & (bins - 1u);
  uint64_t same = __ballot(valid);
  for (uint32_t b = 0; b < bits; b++) {
    uint64_t ones = __ballot((digit >> b) & 1u);
    same &= ((digit >> b) & 1u) ? ones : ~ones;
  }
  uint32_t rank = (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
  if (valid && rank == 0u) waveCount[w][digit] = (uint32_t)__popcll(same);
  __syncthreads();
  if (!SCATTER) {
    for (uint32_t k = t; k < bins; k += 256u) hist[k * gridDim.x + blockIdx.x] = waveCount[0][k] + waveCount[1][k] + waveCount[2][k] + waveCount[3][k];
    if (blockIdx.x == 0 && t == 0) hist[bins * gridDim.x] = 0u;
    return;
  }
  if (valid) {
    uint32_t at = hist[digit * gridDim.x + blockIdx.x] + rank;
    for (uint32_t v = 0; v < w; v++) at += waveCount[v][digit];
    kout[at] = key;
    iout[at] = iin[e];
  }
}
/** Every hash bucket's start in the sorted keys: the first position holding a key at least the
 *  bucket (a binary search each), so bucket h holds [start[h], start[h + 1]) and start[tableMask + 1]
 *  is the count of bodies in the grid. */
__global__ void k_cell_starts(Dev d, const uint32_t *ks) {
  GRID_STRIDE(h, d.tableMask + 2u) {
    uint32_t lo = 0, hi = d.n;
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      if (ks[mid] < h) lo = mid + 1u; else hi = mid;
    }
    d.cellStart[h] = lo;
  }
}
/** The grid's streams in sorted order: body, position and radius, cell key. */
__global__ void k_grid_place(Dev d, const uint32_t *ks, const uint32_t *is) {
  GRID_STRIDE(k, d.n) {
    if (ks[k] > d.tableMask) continue;
    uint32_t i = is[k];
    d.cellSorted[k] = i;
    float4 p = d.b.pos[i];
    d.sortedPos[k] = make_float4(p.x, p.y, p.z, d.b.moment[i].w);
    d.sortedKey[k] = d.cellOf[i];
  }
}

/** The broadphase's view of a body: world axes and half extents (a sphere: its radius). */
struct Shape { V ax[3]; V h; bool sphere; };
__device__ __forceinline__ Shape shapeOf(const Dev &d, uint32_t i) {
  Q q = q4(d.b.rot[i]);
  Shape s;
  s.h = mul(xyz(d.b.size[i]), 0.5f);
  s.sphere = d.b.angvel[i].w == SHAPE_SPHERE;
  s.ax[0] = qrotate(q, vv(1.0f, 0.0f, 0.0f));
  s.ax[1] = qrotate(q, vv(0.0f, 1.0f, 0.0f));
  s.ax[2] = qrotate(q, vv(0.0f, 0.0f, 1.0f));
  return s;
}
__device__ __forceinline__ float extent(const Shape &s, V n) {
  if (s.sphere) return s.h.x;
  return s.h.x * fabsf(dot(n, s.ax[0])) + s.h.y * fabsf(dot(n, s.ax[1])) + s.h.z * fabsf(dot(n, s.ax[2]));
}
__device__ __forceinline__ V aabbHalf(const Shape &s) {
  if (s.sphere) return vv(s.h.x, s.h.x, s.h.x);
  return add(add(mul(vabs(s.ax[0]), s.h.x), mul(vabs(s.ax[1]), s.h.y)), mul(vabs(s.ax[2]), s.h.z));
}
__device__ __forceinline__ bool faceSeparated(const Shape &A, const Shape &B, V dd) {
  for (int k = 0; k < 3; k++) {
    if (fabsf(dot(dd, A.ax[k])) - extent(A, A.ax[k]) - extent(B, A.ax[k]) > 1e-3f) return true;
    if (fabsf(dot(dd, B.ax[k])) - extent(A, B.ax[k]) - extent(B, B.ax[k]) > 1e-3f) return true;
  }
  return false;
}
/** The (hi, lo) entries of the no-collide list (sorted by (hi, lo); body hi's range indexed by
 *  ncFirst); an entry holds while its joint does. */
__device__ __forceinline__ bool ignored(const Dev &d, uint32_t hi, uint32_t lo) {
  const uint32_t *nc = d.statics + d.largeCount;
  uint32_t end = d.ncFirst[hi + 1];
  for (uint32_t k = d.ncFirst[hi]; k < end; k++) {
    uint32_t ml = nc[3 * k + 1];
    if (ml < lo) continue;
    if (ml > lo) break;
    uint32_t j = nc[3 * k + 2];
    if (j == IGNORE_SLOT) return true;
    if (d.joints[10 * j].w != 0.0f || d.joints[10 * j + 1].w != 0.0f) return true;
  }
  return false;
}
struct Probe { uint32_t index; V pos; float radius; V aabb; bool dynamic; Shape shape; uint2 filter; };
__device__ __forceinline__ Probe probeOf(const Dev &d, uint32_t i) {
  Probe P;
  P.index = i;
  P.pos = xyz(d.b.pos[i]);
  P.radius = d.b.moment[i].w;
  P.shape = shapeOf(d, i);
  P.aabb = aabbHalf(P.shape);
  P.dynamic = d.b.size[i].w > 0.0f;
  P.filter = d.filters[i];
  return P;
}
__device__ __forceinline__ void emitPair(const Dev &d, uint32_t a, uint32_t b) {
  uint32_t slot = atomicAdd(&d.counters[C_PAIRS], 1u);
  if (slot >= d.pairCap) { atomicOr(&d.counters[C_OVERFLOW], 1u); return; }
  d.pairs[slot] = make_uint2(a, b);
}
/** testPair (wgsl-collision.ts:208-234), given body j's position and radius already loaded:
 *  whether the pair is a candidate, and the pair (higher index first). */
__device__ __forceinline__ bool testPair(const Dev &d, const Probe &P, uint32_t j, float4 pj, uint2 &pair) {
  // a conjunction of pure tests: the bounding-sphere one first, from what is loaded already
  V dd = sub(P.pos, xyz(pj));
  float r = P.radius + pj.w;
  if (dot(dd, dd) > r * r) return false;
  if (!P.dynamic && d.b.size[j].w <= 0.0f) return false;
  uint2 fj = d.filters[j];
  if ((P.filter.x & fj.y) == 0u || (fj.x & P.filter.y) == 0u) return false;
  if (!bodyEnabled(d, j)) return false;  // (the predicate is pure: checked where few candidates remain)
  Shape B = shapeOf(d, j);
  V hb = aabbHalf(B);
  V lim = add(add(P.aabb, hb), vv(1e-4f, 1e-4f, 1e-4f));
  V ad = vabs(dd);
  if (ad.x > lim.x || ad.y > lim.y || ad.z > lim.z) return false;
  if (faceSeparated(P.shape, B, dd)) return false;
  uint32_t a = P.index > j ? P.index : j, b = P.index > j ? j : P.index;
  if (ignored(d, a, b)) return false;
  pair = make_uint2(a, b);
  return true;
}
/** A workgroup's pairs gather in LDS (an LDS atomic per pair) and go out together, one global
 *  atomic per flush: a returning global atomic per pair would stall the whole wavefront each time
 *  one of its lanes finds a pair. A full queue falls back to the global append. */
struct PairQueue {
  static const uint32_t CAP = 1024;
  uint2 item[CAP];
  uint32_t count, base;
};
__device__ __forceinline__ void queuePair(const Dev &d, PairQueue &q, uint2 pair) {
  uint32_t s = atomicAdd(&q.count, 1u);
  if (s < PairQueue::CAP) q.item[s] = pair;
  else emitPair(d, pair.x, pair.y);
}
/** Every thread of the workgroup calls this (it synchronises). */
__device__ __forceinline__ void flushPairs(const Dev &d, PairQueue &q) {
  __syncthreads();
  uint32_t cnt = q.count < PairQueue::CAP ? q.count : PairQueue::CAP;
  if (threadIdx.x == 0 && cnt > 0) q.base = atomicAdd(&d.counters[C_PAIRS], cnt);
  __syncthreads();
  if (cnt > 0) {
    uint32_t base = q.base;
    for (uint32_t t = threadIdx.x; t < cnt; t += blockDim.x) {
      if (base + t < d.pairCap) d.pairs[base + t] = q.item[t];
      else atomicOr(&d.counters[C_OVERFLOW], 1u);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) q.count = 0;
  __syncthreads();
}
/** Threads per body of the broadphase probe: they split its cell probes and the large list. */
#ifndef PROBE_LANES
#define PROBE_LANES 4
#endif
/** The cell offset of probe q of a half stencil: q = 0 the cell itself, 1..13 the neighbours after
 *  it in (z, y, x) order. Of two neighbouring cells, exactly one has the other in its half. */
__device__ __forceinline__ int4 halfOffset(uint32_t q) {
  uint32_t l = q == 0 ? 13u : 13u + q;
  return make_int4((int)(l % 3) - 1, (int)(l / 3 % 3) - 1, (int)(l / 9) - 1, 0);
}
__device__ __forceinline__ int4 fullOffset(uint32_t q) { return make_int4((int)(q % 3) - 1, (int)(q / 3 % 3) - 1, (int)(q / 9) - 1, 0); }
/** Tests body P against the entries of cell `cell` (its hashed bucket holds other cells too, told
 *  apart by key), four entries' key, position and body loaded together; `below`: only bodies of a
 *  lower index (a pair within one cell is found once). */
template <class Emit>
__device__ __forceinline__ void probeCell(const Dev &d, const Probe &P, int4 cell, uint32_t from, uint32_t to, bool below, Emit emit) {
  uint2 pair;
  for (uint32_t e = from; e < to; e += 4) {
    int4 key[4];
    float4 pe[4];
    uint32_t je[4];
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
      uint32_t x = e + k < to ? e + k : to - 1;
      key[k] = d.sortedKey[x];
      pe[k] = d.sortedPos[x];
      je[k] = d.cellSorted[x];
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
      if (e + k >= to) break;
      if (key[k].x != cell.x || key[k].y != cell.y || key[k].z != cell.z || key[k].w != cell.w) continue;
      if (below && je[k] >= P.index) continue;
      if (testPair(d, P, je[k], pe[k], pair)) emit(pair);
    }
  }
}
/** Body at position k of the grid's cell order (the small bodies as sorted, then the large list),
 *  when dynamic, probes with PROBE_LANES threads (lane l takes probes l, l + PROBE_LANES, ...):
 *  - a small one, at its own level, the moving bodies of its cell (lower indexes) and of the 13
 *    cells of its half stencil; fixed small bodies find their movers themselves (k_find_fixed);
 *  - a fine one, at the coarse level, every body (moving or fixed) of the 27 cells around it:
 *    pairs across the levels belong to the finer body (wgsl-collision.ts:244-246);
 *  - every one the large list (a dynamic large-large pair once).
 *  Neighbouring bodies take neighbouring threads, so a wavefront walks the same and adjacent cells. */
__device__ __forceinline__ void findPairsOf(const Dev &d, PairQueue &q, uint32_t k, uint32_t lane, uint32_t smallCount) {
  uint32_t i;
  if (k < smallCount) i = d.cellSorted[k];
  else if (k < d.n && k - smallCount < d.largeCount) i = d.statics[k - smallCount];
  else return;
  if (d.b.size[i].w <= 0.0f || !bodyEnabled(d, i)) return;
  bool small = !isLarge(d, i);
  Probe P = probeOf(d, i);
  uint32_t w = d.worldOf[i];
  auto emit = [&](uint2 pair) { queuePair(d, q, pair); };
  if (small) {
    int4 c = d.cellOf[i];
    // own level: the half stencil's 14 cells, this lane's ranges loaded together first
    const uint32_t NH = (14 + PROBE_LANES - 1) / PROBE_LANES;
    uint32_t from[NH], to[NH];
#pragma unroll
    for (uint32_t u = 0; u < NH; u++) {
      uint32_t p = lane + u * PROBE_LANES;
      int4 o = halfOffset(p < 14 ? p : 0u);
      uint32_t h = cellHash(d, make_int4(c.x + o.x, c.y + o.y, c.z + o.z, c.w));
      from[u] = p < 14 ? d.cellStart[h] : 0u;
      to[u] = p < 14 ? d.cellStart[h + 1] : 0u;
    }
#pragma unroll
    for (uint32_t u = 0; u < NH; u++) {
      uint32_t p = lane + u * PROBE_LANES;
      int4 o = halfOffset(p < 14 ? p : 0u);
      if (from[u] < to[u]) probeCell(d, P, make_int4(c.x + o.x, c.y + o.y, c.z + o.z, c.w), from[u], to[u], p == 0, emit);
    }
    if (((c.w >> 1) & 1) == 0 && d.hasFine) {
      // a fine body, at the coarse level: moving (r < 27) and fixed (r >= 27) bodies of 27 cells
      int4 cc = cellAt(d, P.pos, 1);
      for (uint32_t r = lane; r < 54; r += PROBE_LANES) {
        int4 o = fullOffset(r % 27);
        int4 cell = make_int4(cc.x + o.x, cc.y + o.y, cc.z + o.z, (int)(w * 4u + 2u + (r < 27 ? 0u : 1u)));
        uint32_t h = cellHash(d, cell);
        uint32_t f = d.cellStart[h], t = d.cellStart[h + 1];
        if (f < t) probeCell(d, P, cell, f, t, false, emit);
      }
    }
  }
  uint2 pair;
  for (uint32_t e = d.largeStart[w] + lane; e < d.largeStart[w + 1]; e += PROBE_LANES) {
    uint32_t l = d.statics[e];
    if (l == i || (!small && l < i && d.b.size[l].w > 0.0f)) continue;
    float4 pl = d.b.pos[l];
    if (testPair(d, P, l, make_float4(pl.x, pl.y, pl.z, d.b.moment[l].w), pair)) queuePair(d, q, pair);
  }
}
/** Launched with 64-thread workgroups over PROBE_LANES threads per body (1024-thread workgroups,
 *  with a sixteenth of the queue atomics, took longer: every pass waits for its slowest wavefront). */
__global__ void __launch_bounds__(64) k_find_pairs(Dev d) {
  __shared__ PairQueue q;
  if (threadIdx.x == 0) q.count = 0;
  __syncthreads();
  uint32_t smallCount = d.cellStart[d.tableMask + 1];
  const uint32_t per = 64 / PROBE_LANES;
  uint32_t slot = threadIdx.x / PROBE_LANES, lane = threadIdx.x % PROBE_LANES;
  // the workgroup walks its bodies together so that it can flush between passes
  for (uint32_t base = blockIdx.x * per; base < d.n; base += gridDim.x * per) {
    findPairsOf(d, q, base + slot, lane, smallCount);
    flushPairs(d, q);
  }
}
/** Fixed bodies find their moving partners: a small one the moving bodies of the 27 cells around
 *  it at its own level and, when fine, at the coarse level (a coarse mover does not probe fine
 *  cells: wgsl-collision.ts:288-307); every small one the large dynamic bodies (findFixedLarge,
 *  wgsl-collision.ts:285-322). Fixed large bodies are every mover's large-list partners. */
__global__ void k_find_fixed(Dev d) {
  GRID_STRIDE(i, d.n) {
    if (d.b.size[i].w > 0.0f || !bodyEnabled(d, i) || isLarge(d, i)) continue;
    Probe P = probeOf(d, i);
    uint32_t w = d.worldOf[i];
    auto emit = [&](uint2 pair) { emitPair(d, pair.x, pair.y); };
    int4 c = d.cellOf[i];
    uint32_t level = ((uint32_t)c.w >> 1) & 1u;
    for (uint32_t pass = 0; pass < 2; pass++) {
      if (pass == 1 && !(level == 0 && d.hasFine && d.coarseMovers)) break;
      int4 base = pass == 0 ? c : cellAt(d, P.pos, 1);
      int slotW = (int)(w * 4u + (pass == 0 ? level : 1u) * 2u);
      for (uint32_t q27 = 0; q27 < 27; q27++) {
        int4 o = fullOffset(q27);
        int4 cell = make_int4(base.x + o.x, base.y + o.y, base.z + o.z, slotW);
        uint32_t h = cellHash(d, cell);
        uint32_t from = d.cellStart[h], to = d.cellStart[h + 1];
        if (from < to) probeCell(d, P, cell, from, to, false, emit);
      }
    }
    uint2 pair;
    for (uint32_t k = d.largeStart[w]; k < d.largeStart[w + 1]; k++) {
      uint32_t l = d.statics[k];
      if (!bodyEnabled(d, l) || d.b.size[l].w <= 0.0f) continue;
      float4 pl = d.b.pos[l];
      if (testPair(d, P, l, make_float4(pl.x, pl.y, pl.z, d.b.moment[l].w), pair)) emitPair(d, pair.x, pair.y);
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Narrowphase + warm start (wgsl-collision.ts:685-892)
__device__ __forceinline__ float contactPenaltyMin(const Dev &d, uint32_t a, uint32_t b) {
  if (!(d.p.flags & FLAG_MASS_PENALTY)) return PENALTY_MIN;
  float ma = d.b.size[a].w, mb = d.b.size[b].w;
  float mA = ma <= 0.0f ? 3.4e38f : ma, mB = mb <= 0.0f ? 3.4e38f : mb;
  return clampf(mn(mA, mB) / (d.p.dt * d.p.dt), PENALTY_MIN, PENALTY_MAX);
}
/** Exclusive prefix sum over the wavefront of v, and the wavefront's total. Every lane must take part. */
__device__ __forceinline__ uint32_t waveExclusiveSum(uint32_t v, uint32_t &total) {
  uint32_t lane = threadIdx.x & 63u, x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += t;
  }
  total = __shfl(x, 63, 64);
  return x - v;
}
/** Bump allocation of `v` slots of *counter per lane with one atomic per wavefront (lanes take
 *  consecutive ranges in lane order); returns the lane's first slot. Every lane must take part. */
__device__ __forceinline__ uint32_t waveAlloc(uint32_t *counter, uint32_t v) {
  uint32_t total, pre = waveExclusiveSum(v, total), base = 0;
  if ((threadIdx.x & 63u) == 0 && total) base = atomicAdd(counter, total);
  return __shfl(base, 0, 64) + pre;
}
/** Bump allocation for a whole workgroup (at most 16 wavefronts): v[k] slots of counter k per
 *  thread, k < K, with one atomic per counter per workgroup (the wavefronts' totals prefixed in LDS);
 *  each thread gets its first slots in out[k], threads in order. Every thread must take part. A
 *  returning atomic on one word from every wavefront of a pass serialises across the chip: at 100k
 *  the narrowphase's four counters cost a few hundred microseconds that way. */
template <uint32_t K>
__device__ __forceinline__ void blockAlloc(uint32_t *const *counter, const uint32_t *v, uint32_t *out) {
  __shared__ uint32_t wsum[K][17];
  const uint32_t nw = blockDim.x >> 6, w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  uint32_t pre[K];
#pragma unroll
  for (uint32_t k = 0; k < K; k++) {
    uint32_t total;
    pre[k] = waveExclusiveSum(v[k], total);
    if (lane == 0) wsum[k][w] = total;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    uint32_t k = threadIdx.x, sum = 0;
    for (uint32_t x = 0; x < nw; x++) {
      uint32_t t = wsum[k][x];
      wsum[k][x] = sum;
      sum += t;
    }
    wsum[k][16] = sum ? atomicAdd(counter[k], sum) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < K; k++) out[k] = wsum[k][16] + wsum[k][w] + pre[k];
  __syncthreads();  // the words are reused by the next call
}
/** copyContacts (wgsl-collision.ts:721-755) with the manifold slot m and contact range base already
 *  allocated. */
__device__ __forceinline__ void copyContactsAt(const Dev &d, uint4 pids, float4 pgeo, uint32_t m, uint32_t base) {
  uint32_t a = pids.x, b = pids.y, count = pids.w & 15u;
  if (m >= d.manCap) { atomicOr(&d.counters[C_OVERFLOW], 4u); return; }
  if (base + count > d.conCap) {
    atomicOr(&d.counters[C_OVERFLOW], 2u);
    d.mIds[m] = make_uint4(a, b, base, pids.w & ~15u);
    d.mGeo[m] = pgeo;
    return;
  }
  d.mIds[m] = make_uint4(a, b, base, pids.w);
  d.mGeo[m] = pgeo;
  V pA = xyz(d.b.pos[a]), pB = xyz(d.b.pos[b]);
  Q qA = q4(d.b.rot[a]), qB = q4(d.b.rot[b]);
  V basis[3];
  orthonormal(xyz(pgeo), basis);
  float pmin = contactPenaltyMin(d, a, b);
  for (uint32_t i = 0; i < count; i++) {
    Contact k = d.pcon[pids.z + i];
    V dd = sub(add(qrotate(qA, xyz(k.rA)), pA), add(qrotate(qB, xyz(k.rB)), pB));
    k.rB.w = dot(basis[0], dd) + d.p.margin;
    k.pen.w = dot(basis[1], dd);
    k.lam.w = dot(basis[2], dd);
    V lam = mul(mul(xyz(k.lam), d.p.alpha), d.p.gamma);
    V pen = xyz(k.pen);
    pen = vv(clampf(pen.x * d.p.gamma, PENALTY_MIN, PENALTY_MAX), clampf(pen.y * d.p.gamma, PENALTY_MIN, PENALTY_MAX), clampf(pen.z * d.p.gamma, PENALTY_MIN, PENALTY_MAX));
    pen.x = mx(pen.x, pmin);
    k.lam = f4(lam, k.lam.w);
    k.pen = f4(pen, k.pen.w);
    d.con[base + i] = k;
    d.cman[base + i] = m;
    if (d.src) d.src[base + i] = (int)(pids.z + i);
  }
}
/** reuseManifolds (wgsl-collision.ts:757-764). When the last step left no manifold empty (no contact
 *  overflow: a manifold is only emitted or copied with points), every previous manifold is kept, so
 *  manifold m keeps slot m and its contact range, and the counts are the last step's: no atomics
 *  (same-address atomics from every wavefront cost ~90 us a launch at 100k). Otherwise each kept
 *  manifold's slot and contact range come from one atomic per wavefront (waveAlloc):
 *  which slot a manifold lands in is arbitrary either way (one atomic each lands them in
 *  arrival order), and no later pass depends on it. */
__global__ void __launch_bounds__(256) k_reuse_manifolds(Dev d) {
  uint32_t count = prevManifoldCount(d);
  bool keepAll = (d.counters[C_PREV_OVERFLOW] & 2u) == 0u;
  uint32_t first = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u, stride = gridDim.x * blockDim.x;
  if (keepAll && first == 0 && threadIdx.x == 0) {
    d.counters[C_MANIFOLDS] = count;
    d.counters[C_CONTACTS] = d.counters[C_PREV_CONTACTS];
    // same manifolds in the same slots, same records: the adjacency built from the last list still
    // holds unless a joint broke since (its entry leaves the lists)
    if (!(d.counters[C_TOPO] & 2u)) d.counters[C_TOPO] |= 1u;
  }
  for (uint32_t w = first; w < count; w += stride) {  // wavefront-uniform: every lane takes part
    uint32_t m = w + (threadIdx.x & 63u);
    uint4 ids = m < count ? d.pmIds[m] : make_uint4(0u, 0u, 0u, 0u);
    uint32_t cnt = ids.w & 15u, slot = m, base = ids.z;
    if (!keepAll) {
      slot = waveAlloc(&d.counters[C_MANIFOLDS], cnt > 0 ? 1u : 0u);
      base = waveAlloc(&d.counters[C_CONTACTS], cnt > 0 && slot < d.manCap ? cnt : 0u);
    }
    if (cnt > 0) copyContactsAt(d, ids, d.pmGeo[m], slot, base);
  }
}
/** reuseContacts' test (wgsl-collision.ts:706-719): the previous step's manifold of pair (a, b)
 *  when its points were computed no earlier than both bodies last moved beyond the reuse
 *  tolerance, else -1. */
__device__ __forceinline__ int reusableManifold(const Dev &d, uint32_t a, uint32_t b) {
  int pm = hashFind(d, a, b);
  if (pm < 0) return -1;
  uint4 ids = d.pmIds[pm];
  uint32_t count = ids.w & 15u, generated = ids.w >> 4;
  uint32_t sa = __float_as_uint(d.b.inpos[a].w), sb = __float_as_uint(d.b.inpos[b].w);
  uint32_t moved = sa > sb ? sa : sb;
  if (count == 0 || generated < moved) return -1;
  return pm;
}

/** emitManifold (wgsl-collision.ts:685-757), the manifold of pair (a, b) from its contact points,
 *  given its slot m and contact range base (allocated by the caller, one atomic per wavefront):
 *  normal basis, per point the body-local anchors, warm start from last step's contact with the
 *  same feature (or the nearest anchor), C0 and the decayed λ/penalty. */
__device__ __forceinline__ void emitManifoldAt(const Dev &d, uint32_t m, uint32_t base, uint32_t a, uint32_t b, const Box &A, const Box &B, Q qA,
                                               Q qB, float4 posA, float4 posB, bool anySphere, const Sat &sat, const Found &found) {
  if (m >= d.manCap) { atomicOr(&d.counters[C_OVERFLOW], 4u); return; }
  uint32_t cnt = found.count;
  if (base + cnt > d.conCap) { atomicOr(&d.counters[C_OVERFLOW], 2u); cnt = 0; }
  V n = neg(sat.n);
  d.mIds[m] = make_uint4(a, b, base, cnt | (d.counters[C_STEP] << 4));
  d.mGeo[m] = f4(n, sqrtf(posA.w * posB.w));
  if (cnt == 0) return;
  V basis[3];
  orthonormal(n, basis);
  uint32_t prevFirst = 0, prevCount = 0;
  int pm = hashFind(d, a, b);
  if (pm >= 0) { uint4 pids = d.pmIds[pm]; prevFirst = pids.z; prevCount = pids.w & 15u; }
  bool matchNearest = (d.p.flags & FLAG_MATCH_NEAREST) != 0;
  float minSide = mn(mn(mn(A.h.x, A.h.y), A.h.z), mn(mn(B.h.x, B.h.y), B.h.z)) * 2.0f;
  float penaltyMin = contactPenaltyMin(d, a, b);
  for (uint32_t i = 0; i < cnt; i++) {
    V rA = qrotate(qconj(qA), sub(found.xA[i], A.c));
    V rB = qrotate(qconj(qB), sub(found.xB[i], B.c));
    V pen = vv(0.0f, 0.0f, 0.0f), lam = pen;
    uint32_t stick = 0;
    int j = -1;
    for (uint32_t c = prevFirst; c < prevFirst + prevCount; c++)
      if ((__float_as_uint(d.pcon[c].rA.w) & ~STICK_BIT) == found.feature[i]) { j = (int)c; break; }
    if (j < 0 && matchNearest) {
      float bestDist = 0.05f * minSide;
      for (uint32_t c = prevFirst; c < prevFirst + prevCount; c++) {
        float dist = length(sub(xyz(d.pcon[c].rA), rA));
        if (dist <= bestDist) { bestDist = dist; j = (int)c; }
      }
    }
    if (j >= 0) {
      Contact pc = d.pcon[j];
      pen = xyz(pc.pen); lam = xyz(pc.lam); stick = __float_as_uint(pc.rA.w) & STICK_BIT;
      if (stick != 0 && !anySphere) { rA = xyz(pc.rA); rB = xyz(pc.rB); }
    }
    V dd = sub(add(qrotate(qA, rA), A.c), add(qrotate(qB, rB), B.c));
    Contact k;
    k.rA = f4(rA, __uint_as_float(found.feature[i] | stick));
    k.rB = f4(rB, dot(basis[0], dd) + d.p.margin);
    lam = mul(mul(lam, d.p.alpha), d.p.gamma);
    pen = vv(clampf(pen.x * d.p.gamma, PENALTY_MIN, PENALTY_MAX), clampf(pen.y * d.p.gamma, PENALTY_MIN, PENALTY_MAX), clampf(pen.z * d.p.gamma, PENALTY_MIN, PENALTY_MAX));
    pen.x = mx(pen.x, penaltyMin);
    k.pen = f4(pen, dot(basis[1], dd));
    k.lam = f4(lam, dot(basis[2], dd));
    d.con[base + i] = k;
    d.cman[base + i] = m;
    if (d.src) d.src[base + i] = j;
  }
}

/** narrowphase (wgsl-collision.ts:767) of every broadphase pair, one thread per pair: kept points
 *  (contact reuse), a deferral, or a fresh manifold. The slots of all of them come from one atomic
 *  per wavefront each (waveAlloc) at a point every lane reaches: same-address atomics per lane
 *  serialise across the chip. Which slot a pair lands in is arbitrary either way.
 *  INLINE_HULLS: every pair is collided here by one thread (the solver built without the hull
 *  list). Otherwise hull pairs go to the list for k_narrow_hull (a wavefront per pair) and box-box
 *  pairs to the list for k_narrow_box, and this kernel carries the sphere pairs and the reuse
 *  alone: no private memory, and no wavefront runs the box clip for a few of its lanes. */
template <bool INLINE_HULLS>
__global__ void __launch_bounds__(1024) k_narrowphase(Dev d) {
  uint32_t count = d.counters[C_PAIRS];
  if (count > d.pairCap) count = d.pairCap;
  bool reuseOn = (d.p.flags & FLAG_REUSE_CONTACTS) != 0;
  for (uint32_t w = blockIdx.x * blockDim.x; w < count; w += gridDim.x * blockDim.x) {  // workgroup-uniform
    uint32_t p = w + threadIdx.x;
    bool live = p < count;
    uint2 pr = live ? d.pairs[p] : make_uint2(0u, 0u);
    uint32_t a = pr.x, b = pr.y;
    int pm = live && reuseOn ? reusableManifold(d, a, b) : -1;
    bool defer = false, deferBox = false, sphereA = false, sphereB = false;
    float4 posA = make_float4(0.0f, 0.0f, 0.0f, 0.0f), posB = posA;
    Q qA = qq(0.0f, 0.0f, 0.0f, 1.0f), qB = qA;
    Box A, B;
    Sat sat;
    sat.n = vv(0.0f, 0.0f, 0.0f);
    Found found;
    found.count = 0;
    if (live && pm < 0) {
      posA = d.b.pos[a]; posB = d.b.pos[b];
      float shA = d.b.angvel[a].w, shB = d.b.angvel[b].w;
      sphereA = shA == SHAPE_SPHERE; sphereB = shB == SHAPE_SPHERE;
      bool hullA = shA >= SHAPE_HULL, hullB = shB >= SHAPE_HULL;
      if ((hullA || hullB) && !INLINE_HULLS) {
        defer = true;
      } else if (!INLINE_HULLS && !sphereA && !sphereB) {
        deferBox = true;  // the box SAT and clip, in wavefronts of box pairs alone (k_narrow_box)
      } else {
        qA = q4(d.b.rot[a]); qB = q4(d.b.rot[b]);
        A = makeBox(xyz(posA), qA, xyz(d.b.size[a])); B = makeBox(xyz(posB), qB, xyz(d.b.size[b]));
        if (sphereA && sphereB) sphereSphere(A, B, sat, found);
        else if (INLINE_HULLS && (hullA || hullB))
          collideHullPair(d.hulls, A, qA, shA, B, qB, shB, sphereA, sphereB, (d.p.flags & FLAG_FACE_BIAS) != 0, sat, found);
        else if (sphereA || sphereB) {
          if (sphereA) {
            SphereHit h = sphereBox(A, B, qB);
            sat.n = h.n;
            if (h.hit) addFound(found, h.onSphere, h.onBox, SPHERE_FEATURE);
          } else {
            SphereHit h = sphereBox(B, A, qA);
            sat.n = neg(h.n);
            if (h.hit) addFound(found, h.onBox, h.onSphere, SPHERE_FEATURE);
          }
        } else if (INLINE_HULLS) collideBoxes(A, B, (d.p.flags & FLAG_FACE_BIAS) != 0, sat, found);
      }
    }
    uint4 pids = pm >= 0 ? d.pmIds[pm] : make_uint4(0u, 0u, 0u, 0u);
    bool wants = pm >= 0 || found.count > 0;
    uint32_t points = pm >= 0 ? (pids.w & 15u) : found.count;
    uint32_t *const ctr[3] = {&d.counters[C_HULL_PAIRS], &d.counters[C_BOX_PAIRS], &d.counters[C_MANIFOLDS]};
    const uint32_t need[3] = {defer ? 1u : 0u, deferBox ? 1u : 0u, wants ? 1u : 0u};
    uint32_t got[3];
    blockAlloc<3>(ctr, need, got);
    uint32_t hslot = got[0], bslot = got[1], m = got[2];
    uint32_t *const cctr[1] = {&d.counters[C_CONTACTS]};
    const uint32_t cneed[1] = {wants && m < d.manCap ? points : 0u};
    uint32_t base;
    blockAlloc<1>(cctr, cneed, &base);
    if (defer) d.hullPairs[hslot] = pr;
    if (deferBox) d.boxPairs[bslot] = pr;
    if (pm >= 0) copyContactsAt(d, pids, d.pmGeo[pm], m, base);
    else if (found.count > 0) emitManifoldAt(d, m, base, a, b, A, B, qA, qB, posA, posB, sphereA || sphereB, sat, found);
  }
}

/** The box-box pairs the narrowphase deferred: collideBoxes a thread per pair as it runs inline,
 *  the wavefronts now all taking the SAT and clip path (mixed with sphere pairs, every wavefront
 *  ran every shape's path). */
__global__ void __launch_bounds__(256) k_narrow_box(Dev d) {
  uint32_t count = d.counters[C_BOX_PAIRS];
  if (count > d.pairCap) count = d.pairCap;
  for (uint32_t w = blockIdx.x * blockDim.x; w < count; w += gridDim.x * blockDim.x) {  // workgroup-uniform
    uint32_t p = w + threadIdx.x;
    bool live = p < count;
    uint2 pr = live ? d.boxPairs[p] : make_uint2(0u, 0u);
    uint32_t a = pr.x, b = pr.y;
    float4 posA = make_float4(0.0f, 0.0f, 0.0f, 0.0f), posB = posA;
    Q qA = qq(0.0f, 0.0f, 0.0f, 1.0f), qB = qA;
    Box A, B;
    Sat sat;
    sat.n = vv(0.0f, 0.0f, 0.0f);
    Found found;
    found.count = 0;
    if (live) {
      posA = d.b.pos[a]; posB = d.b.pos[b];
      qA = q4(d.b.rot[a]); qB = q4(d.b.rot[b]);
      A = makeBox(xyz(posA), qA, xyz(d.b.size[a])); B = makeBox(xyz(posB), qB, xyz(d.b.size[b]));
      collideBoxes(A, B, (d.p.flags & FLAG_FACE_BIAS) != 0, sat, found);
    }
    bool wants = found.count > 0;
    uint32_t *const mctr[1] = {&d.counters[C_MANIFOLDS]}, *const cctr[1] = {&d.counters[C_CONTACTS]};
    const uint32_t mneed[1] = {wants ? 1u : 0u};
    uint32_t m, base;
    blockAlloc<1>(mctr, mneed, &m);
    const uint32_t cneed[1] = {wants && m < d.manCap ? found.count : 0u};
    blockAlloc<1>(cctr, cneed, &base);
    if (wants) emitManifoldAt(d, m, base, a, b, A, B, qA, qB, posA, posB, false, sat, found);
  }
}

/** The pairs the narrowphase deferred (a hull on either side), one 64-wide wavefront per pair:
 *  the SAT's face and edge-pair queries are spread over the lanes, lane 0 clips and emits. */
/** float4 per hull pair in hullOut: sat.n and the point count, then per point xA with the feature
 *  bits, then per point xB. */
static const uint32_t HULL_OUT = 17;
__global__ void __launch_bounds__(256) k_narrow_hull(Dev d) {
  uint32_t count = d.counters[C_HULL_PAIRS];
  if (count > d.pairCap) count = d.pairCap;
  uint32_t lane = threadIdx.x & 63u;
  uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, waves = (gridDim.x * blockDim.x) >> 6;
  __shared__ HullStage stage[4];  // one per wavefront of the 256-thread workgroup
  HullStage *st = &stage[threadIdx.x >> 6];
  for (uint32_t p = wave; p < count; p += waves) {
    uint2 pr = d.hullPairs[p];
    uint32_t a = pr.x, b = pr.y;
    float4 posA = d.b.pos[a], posB = d.b.pos[b];
    Q qA = q4(d.b.rot[a]), qB = q4(d.b.rot[b]);
    float shA = d.b.angvel[a].w, shB = d.b.angvel[b].w;
    Box A = makeBox(xyz(posA), qA, xyz(d.b.size[a])), B = makeBox(xyz(posB), qB, xyz(d.b.size[b]));
    bool sphereA = shA == SHAPE_SPHERE, sphereB = shB == SHAPE_SPHERE;
    Sat sat;
    sat.n = vv(0.0f, 0.0f, 0.0f);
    Found found;
    found.count = 0;
    collideHullPairWave(d.hulls, A, qA, shA, B, qB, shB, sphereA, sphereB, (d.p.flags & FLAG_FACE_BIAS) != 0, lane, st, sat, found);
    if (lane == 0) {  // the contact, for k_hull_emit (no allocation here: see there)
      float4 *o = d.hullOut + (size_t)p * HULL_OUT;
      o[0] = f4(sat.n, __uint_as_float(found.count));
      for (uint32_t i = 0; i < found.count; i++) {
        o[1 + i] = f4(found.xA[i], __uint_as_float(found.feature[i]));
        o[9 + i] = f4(found.xB[i], 0.0f);
      }
    }
  }
}
/** The hull pairs' manifolds (emitManifold, wgsl-collision.ts:685-757), a thread per pair: the boxes
 *  rebuilt from the same body data as in k_narrow_hull, the contact as it left it. Slots come
 *  from one atomic per wavefront each (waveAlloc): an atomic per pair from the pair's wavefront
 *  cost about a quarter of the hull kernel. */
__global__ void __launch_bounds__(1024) k_hull_emit(Dev d) {
  uint32_t count = d.counters[C_HULL_PAIRS];
  if (count > d.pairCap) count = d.pairCap;
  for (uint32_t w = blockIdx.x * blockDim.x; w < count; w += gridDim.x * blockDim.x) {  // workgroup-uniform
    uint32_t p = w + threadIdx.x;
    const float4 *o = d.hullOut + (size_t)p * HULL_OUT;
    float4 h0 = p < count ? o[0] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t points = p < count ? __float_as_uint(h0.w) : 0u;
    uint32_t *const mctr[1] = {&d.counters[C_MANIFOLDS]}, *const cctr[1] = {&d.counters[C_CONTACTS]};
    const uint32_t mneed[1] = {points > 0 ? 1u : 0u};
    uint32_t m, base;
    blockAlloc<1>(mctr, mneed, &m);
    const uint32_t cneed[1] = {points > 0 && m < d.manCap ? points : 0u};
    blockAlloc<1>(cctr, cneed, &base);
    if (points == 0) continue;
    uint2 pr = d.hullPairs[p];
    uint32_t a = pr.x, b = pr.y;
    float4 posA = d.b.pos[a], posB = d.b.pos[b];
    Q qA = q4(d.b.rot[a]), qB = q4(d.b.rot[b]);
    bool anySphere = d.b.angvel[a].w == SHAPE_SPHERE || d.b.angvel[b].w == SHAPE_SPHERE;
    Box A = makeBox(xyz(posA), qA, xyz(d.b.size[a])), B = makeBox(xyz(posB), qB, xyz(d.b.size[b]));
    Sat sat;
    sat.n = xyz(h0);
    Found found;
    found.count = points;
    for (uint32_t i = 0; i < points; i++) {
      float4 x = o[1 + i];
      found.xA[i] = xyz(x);
      found.feature[i] = __float_as_uint(x.w);
      found.xB[i] = xyz(o[9 + i]);
    }
    emitManifoldAt(d, m, base, a, b, A, B, qA, qB, posA, posB, anySphere, sat, found);
  }
}

// ---------------------------------------------------------------------------------------------
// Adjacency (wgsl-topology.ts:63-153): constraint ids are joints 0..J-1, then manifolds.
__device__ __forceinline__ bool jointActive(const Dev &d, uint32_t j) {
  return d.jinfo[j].x != (int)T_NONE && (d.joints[10 * j].w != 0.0f || d.joints[10 * j + 1].w != 0.0f);
}
__device__ __forceinline__ int2 endpoints(const Dev &d, uint32_t id) {
  if (id < d.nj) { int4 in = d.jinfo[id]; return make_int2(in.y, in.z); }
  uint4 ids = d.mIds[id - d.nj];
  return make_int2((int)ids.x, (int)ids.y);
}
__global__ void k_degree(Dev d) {
  if (d.counters[C_TOPO] & 1u) return;  // a kept adjacency
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAnd(&d.counters[C_TOPO], ~2u);  // rebuilt from here on
  uint32_t total = d.nj + manifoldCount(d);
  GRID_STRIDE(id, total) {
    if (id < d.nj && !jointActive(d, id)) continue;
    int2 e = endpoints(d, id);
    if (isDynamic(d, e.x)) atomicAdd(&d.adjCount[e.x], 1u);
    if (isDynamic(d, e.y)) atomicAdd(&d.adjCount[e.y], 1u);
  }
}
/** An adjacency entry's sort key: joints first (by id), then manifolds by partner and first
 *  contact's feature. */
__device__ __forceinline__ uint2 adjKey(const Dev &d, uint32_t b, uint32_t id) {
  if (id < d.nj) return make_uint2(0u, id);
  uint4 ids = d.mIds[id - d.nj];
  uint32_t other = ids.x == b ? ids.y : ids.x;
  uint32_t sec = (ids.w & 15u) == 0 ? 0u : (__float_as_uint(d.con[ids.z].rA.w) & ~STICK_BIT);
  return make_uint2(other + 2, sec);
}
__device__ __forceinline__ bool keyLess(uint2 a, uint2 b) { return a.x < b.x || (a.x == b.x && a.y < b.y); }
__global__ void k_fill_adj(Dev d) {
  if (d.counters[C_TOPO] & 1u) return;  // a kept adjacency
  uint32_t total = d.nj + manifoldCount(d);
  GRID_STRIDE(id, total) {
    if (id < d.nj && !jointActive(d, id)) continue;
    int2 e = endpoints(d, id);
    int ends[2] = {e.x, e.y};
    for (int s = 0; s < 2; s++) {
      int x = ends[s];
      if (!isDynamic(d, x)) continue;
      uint32_t slot = d.adjStart[x] + atomicAdd(&d.adjFill[x], 1u);
      if (slot < d.adjCap) {
        d.adjTmp[slot] = id;
        d.adjOwner[slot] = (uint32_t)x;
        d.adjKeys[slot] = adjKey(d, (uint32_t)x, id);
      }
    }
  }
}
/** Where id goes in body b's list: joints by id, then manifolds by (other + 2, first key). */
/** Each body's entries in key order (joints first by id, then manifolds by partner and first
 *  contact's feature; keys are distinct): entry e of the fill goes to its body's range at the
 *  number of that body's keys below its own, with its manifold record beside it for the solve's
 *  flattened walk (a joint's record: ids.x = 0xffffffff, id in ids.y). */
__global__ void k_sort_adj(Dev d) {
  if (d.counters[C_TOPO] & 1u) return;  // a kept adjacency
  uint32_t total = d.adjStart[d.n];
  if (total > d.adjCap) total = d.adjCap;
  GRID_STRIDE(e, total) {
    uint32_t b = d.adjOwner[e];
    uint32_t lo = d.adjStart[b], hi = d.adjStart[b + 1];
    if (hi > d.adjCap) hi = d.adjCap;
    uint2 key = d.adjKeys[e];
    uint32_t rank = 0;
    for (uint32_t k = lo; k < hi; k++) rank += keyLess(d.adjKeys[k], key) ? 1u : 0u;
    uint32_t id = d.adjTmp[e], at = lo + rank;
    d.adjList[at] = id;
    int other;
    if (id < d.nj) {
      d.adjIds[at] = make_uint4(0xffffffffu, id, 0u, 0u);
      int4 in = d.jinfo[id];
      other = in.y == (int)b ? in.z : in.y;
    } else {
      uint4 ids = d.mIds[id - d.nj];
      d.adjIds[at] = ids;
      d.adjGeo[at] = d.mGeo[id - d.nj];
      other = ids.x == b ? (int)ids.y : (int)ids.x;
    }
    d.adjOther[at] = isDynamic(d, other) ? other : -1;
  }
}

// ---------------------------------------------------------------------------------------------
// Colouring: incremental Jones-Plassmann with hashed priorities (wgsl-topology.ts:155-262).
/** lowbias32 priority of a body's index within its world (its global index in a single world). */
__device__ __forceinline__ uint32_t priority(const Dev &d, uint32_t i) { return hash32(i - d.worldStart[d.worldOf[i]] + 0x9e3779b9u); }
/** Every body's priority, once at creation (bodies and worlds are fixed). */
__global__ void k_priorities(Dev d) { GRID_STRIDE(i, d.n) d.prio[i] = priority(d, i); }
/** i beats j (always the same world: ties by index there, which the global index orders alike). */
__device__ __forceinline__ bool beats(const Dev &d, uint32_t i, uint32_t j) {
  uint32_t pi = d.prio[i], pj = d.prio[j];
  return pi > pj || (pi == pj && i > j);
}
/** Entry e's other body when dynamic, else -1 (written by k_sort_adj: one load, not the record chain). */
__device__ __forceinline__ int neighbour(const Dev &d, uint32_t b, uint32_t e) { return d.adjOther[e]; }
__device__ __forceinline__ uint32_t smallestFree(const Dev &d, uint32_t b, const uint32_t *src, bool skipPending) {
  uint32_t lo = 0, hi = 0;
  for (uint32_t e = d.adjStart[b]; e < d.adjStart[b + 1]; e++) {
    int j = neighbour(d, b, e);
    if (j < 0) continue;
    uint32_t s = src[j];
    if (skipPending && (s & PENDING)) continue;
    uint32_t col = s & 0xffu;
    if (col == NO_COLOR) continue;
    if (col < 32) lo |= 1u << col; else hi |= 1u << (col - 32);
  }
  if (~lo != 0u) return (uint32_t)__builtin_ctz(~lo);
  if (~hi != 0u) return 32u + (uint32_t)__builtin_ctz(~hi);
  return MAX_COLORS;
}
/** Neighbour entries a colouring pass fetches at a time: the entries' loads, then the loads they
 *  index, go out together instead of one dependent pair per entry. */
#define COLOR_FETCH 4u
/** colorCompact (avbd2d/gpu/wgsl-topology.ts:201-216): a coloured body whose priority beats every
 *  neighbour's takes the smallest colour no neighbour holds, if lower. The beats test and the
 *  neighbours' colours come from one pass over the entries. */
__global__ void k_color_compact(Dev d) {
  GRID_STRIDE(b, d.n) {
    if (!isDynamic(d, b)) { d.colB[b] = NO_COLOR; continue; }
    uint32_t col = d.colA[b] & 0xffu;
    if (col != NO_COLOR) {
      bool isMax = true;
      uint32_t lo = 0, hi = 0, pb = d.prio[b], e1 = d.adjStart[b + 1];
      for (uint32_t e = d.adjStart[b]; e < e1; e += COLOR_FETCH) {
        int j[COLOR_FETCH];
        uint32_t sj[COLOR_FETCH], pj[COLOR_FETCH];
#pragma unroll
        for (uint32_t k = 0; k < COLOR_FETCH; k++) j[k] = e + k < e1 ? d.adjOther[e + k] : -1;
#pragma unroll
        for (uint32_t k = 0; k < COLOR_FETCH; k++) {
          sj[k] = j[k] >= 0 ? d.colA[j[k]] : NO_COLOR;
          pj[k] = j[k] >= 0 ? d.prio[j[k]] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < COLOR_FETCH; k++) {
          if (j[k] < 0) continue;
          if (!(pb > pj[k] || (pb == pj[k] && b > (uint32_t)j[k]))) isMax = false;  // !beats(b, j)
          uint32_t c = sj[k] & 0xffu;
          if (c == NO_COLOR) continue;
          if (c < 32) lo |= 1u << c; else hi |= 1u << (c - 32);
        }
      }
      if (isMax) {
        uint32_t f = ~lo != 0u ? (uint32_t)__builtin_ctz(~lo) : ~hi != 0u ? 32u + (uint32_t)__builtin_ctz(~hi) : MAX_COLORS;
        col = col < f ? col : f;
      }
    }
    d.colB[b] = col;
  }
}
/** colorMark (avbd2d/gpu/wgsl-topology.ts:220-232): a body is pending when uncoloured or when a
 *  neighbour of the same colour beats it. */
__global__ void k_color_mark(Dev d) {
  GRID_STRIDE(b, d.n) {
    uint32_t col = d.colB[b] & 0xffu;
    if (!isDynamic(d, b)) { d.colA[b] = NO_COLOR; continue; }
    bool pending = col == NO_COLOR;
    uint32_t pb = d.prio[b], e1 = d.adjStart[b + 1];
    for (uint32_t e = d.adjStart[b]; e < e1 && !pending; e += COLOR_FETCH) {
      int j[COLOR_FETCH];
      uint32_t cj[COLOR_FETCH], pj[COLOR_FETCH];
#pragma unroll
      for (uint32_t k = 0; k < COLOR_FETCH; k++) j[k] = e + k < e1 ? d.adjOther[e + k] : -1;
#pragma unroll
      for (uint32_t k = 0; k < COLOR_FETCH; k++) {
        cj[k] = j[k] >= 0 ? d.colB[j[k]] & 0xffu : NO_COLOR;
        pj[k] = j[k] >= 0 ? d.prio[j[k]] : 0u;
      }
#pragma unroll
      for (uint32_t k = 0; k < COLOR_FETCH; k++)
        if (j[k] >= 0 && cj[k] == col && (pj[k] > pb || (pj[k] == pb && (uint32_t)j[k] > b))) pending = true;  // beats(j, b)
    }
    d.colA[b] = col | (pending ? PENDING : 0u);
  }
}
__global__ void k_color_round(Dev d, const uint32_t *src, uint32_t *dst) {
  GRID_STRIDE(b, d.n) {
    uint32_t s = src[b];
    if (!(s & PENDING)) { dst[b] = s; continue; }
    bool isMax = true;
    for (uint32_t e = d.adjStart[b]; e < d.adjStart[b + 1] && isMax; e++) {
      int j = neighbour(d, b, e);
      if (j >= 0 && (src[j] & PENDING) && !beats(d, b, j)) isMax = false;
    }
    if (isMax) {
      uint32_t f = smallestFree(d, b, src, true);
      dst[b] = f >= d.p.colorCap ? ((d.p.colorCap - 1) | PENDING) : f;
    } else dst[b] = s;
  }
}
/** colorCount (wgsl-topology.ts:276-304): per 256-body chunk, colour histogram and stable ranks
 *  (a body's rank: the chunk's bodies of its colour at lower indexes). Within a wavefront the rank
 *  comes from one ballot per distinct colour; the earlier wavefronts' counts of it are added. */
__global__ void __launch_bounds__(256) k_color_count(Dev d) {
  __shared__ uint32_t waveHist[4][64];
  uint32_t tid = threadIdx.x, b = blockIdx.x * 256 + tid, w = tid >> 6, lane = tid & 63u;
  waveHist[w][lane] = 0;
  uint32_t col = NO_COLOR;
  if (b < d.n && isDynamic(d, b)) {
    uint32_t s = d.colA[b];
    col = s & 0xffu;
    if (s & PENDING) atomicAdd(&d.counters[C_CLASHES], 1u);
    if (col == NO_COLOR) col = d.p.colorCap - 1;
    d.colA[b] = col | (s & PENDING);
    if (s & PENDING) atomicOr(&d.counters[C_CLASHING + (col >> 5)], 1u << (col & 31u));
  }
  // wavefront-uniform loop: each round takes the colour of the lowest lane not yet ranked
  uint64_t below = (1ull << lane) - 1ull;
  uint32_t rank = 0;
  bool done = col == NO_COLOR;
  for (;;) {
    uint64_t left = __ballot(!done);
    if (left == 0) break;
    uint32_t leader = (uint32_t)__ffsll((unsigned long long)left) - 1u;
    uint32_t c = (uint32_t)__shfl((int)col, (int)leader);
    bool mine = !done && col == c;
    uint64_t same = __ballot(mine);
    if (mine) { rank = (uint32_t)__popcll(same & below); done = true; }
    if (lane == leader) waveHist[w][c] = (uint32_t)__popcll(same);
  }
  __syncthreads();
  if (col != NO_COLOR) {
    for (uint32_t v = 0; v < w; v++) rank += waveHist[v][col];
    d.colorRank[b] = rank;
  }
  if (tid < 64) d.hist[tid * d.colorGroups + blockIdx.x] = waveHist[0][tid] + waveHist[1][tid] + waveHist[2][tid] + waveHist[3][tid];
  if (b == 0) d.hist[64 * d.colorGroups] = 0;
}
/** Colour starts from the scanned histogram, and the colours in use (one past the highest colour
 *  holding a body) from the same counts: no atomic from the counting pass. One wavefront. */
__global__ void k_color_starts(Dev d) {
  uint32_t c = threadIdx.x;
  uint32_t at = c < 64 ? d.hist[c * d.colorGroups] : 0u, next = c < 64 ? d.hist[(c + 1) * d.colorGroups] : 0u;
  if (c < 64) d.colorStart[c] = at;
  if (c == 0) d.colorStart[64] = d.hist[64 * d.colorGroups];
  uint64_t used = __ballot(c < 64 && next > at);
  if (c == 0) d.counters[C_NUM_COLORS] = used ? 64u - (uint32_t)__clzll((long long)used) : 0u;
}
/** Threads per body of a colour's primal, from the colour's size (the reference's primal_lanes rule). */
__device__ __forceinline__ uint32_t lanesFor(const Params &p, uint32_t count) {
  if (count >= p.lanes1) return 1;
  if (count >= p.lanes2) return 2;
  if (count >= p.lanes4) return 4;
  if (count >= p.lanes8) return 8;
  return 16;  // small colours: their time is the longest adjacency walk, so split it further
}
__global__ void __launch_bounds__(256) k_color_scatter(Dev d) {
  uint32_t b = blockIdx.x * 256 + threadIdx.x;
  if (b >= d.n || !isDynamic(d, b)) return;
  uint32_t col = d.colA[b] & 0xffu;
  uint32_t slot = d.hist[col * d.colorGroups + blockIdx.x] + d.colorRank[b];
  d.colorBodies[slot] = b;
  d.colorItems[slot] = make_uint4(b, d.itemStart[b], d.itemStart[b + 1], bodyEnabled(d, b) ? 1u : 0u);
  if (d.bodyKey) {
    // a serial (clashing) colour runs its bucket in chunks of one wavefront's bodies, in order
    uint32_t start = d.colorStart[col], chunk = 0;
    if ((d.counters[C_CLASHING + (col >> 5)] >> (col & 31u)) & 1u) chunk = (slot - start) / (64u / lanesFor(d.p, d.colorStart[col + 1] - start));
    d.bodyKey[b] = (col << 26) | chunk;
  }
}
/** The primal's walk items of every body: counted (then scanned into itemStart) and filled in
 *  adjacency order. Nothing when the substep keeps the last adjacency (C_TOPO bit 0). */
__global__ void __launch_bounds__(256) k_item_count(Dev d) {
  if (d.counters[C_TOPO] & 1u) return;
  GRID_STRIDE(i, d.n + 1) {
    uint32_t c = 0;
    if (i < d.n)
      for (uint32_t e = d.adjStart[i], e1 = d.adjStart[i + 1]; e < e1; e++) {
        uint4 ids = d.adjIds[e];
        c += ids.x == 0xffffffffu ? 1u : (ids.w & 15u);
      }
    d.itemStart[i] = c;
  }
}
__global__ void __launch_bounds__(256) k_item_fill(Dev d) {
  if (d.counters[C_TOPO] & 1u) return;
  GRID_STRIDE(i, d.n) {
    uint32_t o = d.itemStart[i];
    for (uint32_t e = d.adjStart[i], e1 = d.adjStart[i + 1]; e < e1; e++) {
      uint4 ids = d.adjIds[e];
      if (ids.x == 0xffffffffu) {
        if (o < d.itemCap) d.items[o] = make_uint4(ids.y, e, 0u, 2u);
        o++;
        continue;
      }
      bool isA = ids.x == i;
      uint32_t partner = isA ? ids.y : ids.x, cnt = ids.w & 15u;
      for (uint32_t k = 0; k < cnt; k++, o++)
        if (o < d.itemCap) d.items[o] = make_uint4(ids.z + k, e, partner, isA ? 1u : 0u);
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Solve (wgsl-solve.ts)
/** How the solve's helpers touch state that a sweep updates (poses, contact multipliers and
 *  penalties, joint records). Plain: ordinary loads and stores; a kernel boundary orders them.
 *  Coherent: device-scope relaxed atomics (global_load/store ... sc1), which read and write past
 *  the per-XCD caches, so that within one launch a workgroup on one XCD reads what a workgroup on
 *  another wrote once that store completed (the dataflow sweep orders them by per-body counters). */
struct Plain {
  static __device__ __forceinline__ float4 ld(const float4 &r) { return r; }
  static __device__ __forceinline__ void st(float4 &r, float4 v) { r = v; }
  static __device__ __forceinline__ void stw(float &r, float v) { r = v; }
};
struct Coherent {
  static __device__ __forceinline__ float4 ld(const float4 &r) {
    uint64_t *q = reinterpret_cast<uint64_t *>(const_cast<float4 *>(&r));
    uint64_t a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return make_float4(__uint_as_float((uint32_t)a), __uint_as_float((uint32_t)(a >> 32)), __uint_as_float((uint32_t)b),
                       __uint_as_float((uint32_t)(b >> 32)));
  }
  static __device__ __forceinline__ void st(float4 &r, float4 v) {
    uint64_t *q = reinterpret_cast<uint64_t *>(&r);
    __hip_atomic_store(q, (uint64_t)__float_as_uint(v.x) | ((uint64_t)__float_as_uint(v.y) << 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, (uint64_t)__float_as_uint(v.z) | ((uint64_t)__float_as_uint(v.w) << 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  static __device__ __forceinline__ void stw(float &r, float v) {
    __hip_atomic_store(reinterpret_cast<uint32_t *>(&r), __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
};
template <class M>
__device__ __forceinline__ Contact loadContact(const Contact &c) {
  Contact k;
  k.rA = M::ld(c.rA); k.rB = M::ld(c.rB); k.pen = M::ld(c.pen); k.lam = M::ld(c.lam);
  return k;
}
template <class M>
__device__ __forceinline__ bool enabledM(const Dev &d, uint32_t i) { return (lifecycleBits(M::ld(d.b.ipos[i]).w) & 2u) == 0u; }
/** One body's Newton system: lin/ang lower triangles (00 10 11 20 21 22), cross (row = angular). */
struct Acc { float lin[6], ang[6], cr[9], rl[3], ra[3]; };
__device__ __forceinline__ void accZero(Acc &a) {
#pragma unroll
  for (int i = 0; i < 6; i++) { a.lin[i] = 0.0f; a.ang[i] = 0.0f; }
#pragma unroll
  for (int i = 0; i < 9; i++) a.cr[i] = 0.0f;
#pragma unroll
  for (int i = 0; i < 3; i++) { a.rl[i] = 0.0f; a.ra[i] = 0.0f; }
}
/** Fold one constraint row in: Jacobian (l, a), stiffness k, force f. The row's scaled Jacobian
 *  (k l, k a) is formed once and each entry of the system takes one fused multiply-add
 *  (wgsl-solve.ts:69-113 folds k J J^T the same way, as sums of products). */
__device__ __forceinline__ void addRow(Acc &acc, V l, V a, float k, float f) {
  float L[3] = {l.x, l.y, l.z}, A[3] = {a.x, a.y, a.z};
  float kl[3] = {L[0] * k, L[1] * k, L[2] * k}, ka[3] = {A[0] * k, A[1] * k, A[2] * k};
  const int I[6] = {0, 1, 1, 2, 2, 2}, J[6] = {0, 0, 1, 0, 1, 2};
#pragma unroll
  for (int e = 0; e < 6; e++) { acc.lin[e] = fmaf(kl[I[e]], L[J[e]], acc.lin[e]); acc.ang[e] = fmaf(ka[I[e]], A[J[e]], acc.ang[e]); }
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) acc.cr[i * 3 + j] = fmaf(ka[i], L[j], acc.cr[i * 3 + j]);
#pragma unroll
  for (int i = 0; i < 3; i++) { acc.rl[i] = fmaf(L[i], f, acc.rl[i]); acc.ra[i] = fmaf(A[i], f, acc.ra[i]); }
}
/** a . b as a chain of fused multiply-adds (x first). */
__device__ __forceinline__ float fdot(V a, V b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
template <class M = Plain>
__device__ __forceinline__ Q rotAOf(const Dev &d, int a) { return a < 0 ? qq(0.0f, 0.0f, 0.0f, 1.0f) : q4(M::ld(d.b.rot[a])); }
template <class M = Plain>
__device__ __forceinline__ V jointLinC(const Dev &d, const float4 *k, int a, int b) {
  V anchor = a < 0 ? xyz(M::ld(k[6])) : add(qrotate(q4(M::ld(d.b.rot[a])), xyz(M::ld(k[6]))), xyz(M::ld(d.b.pos[a])));
  return sub(anchor, add(qrotate(q4(M::ld(d.b.rot[b])), xyz(M::ld(k[7]))), xyz(M::ld(d.b.pos[b]))));
}
template <class M = Plain>
__device__ __forceinline__ V jointAngC(const Dev &d, const float4 *k, int a, int b) {
  Q qb = q4(M::ld(d.b.rot[b]));
  float dd = dot4(qb, qb);
  Q dq = qmul(qmul(rotAOf<M>(d, a), q4(M::ld(k[8]))), qq(-qb.x / dd, -qb.y / dd, -qb.z / dd, qb.w / dd));
  if (dq.w < 0.0f) dq = qq(-dq.x, -dq.y, -dq.z, -dq.w);
  return mul(mul(vv(dq.x, dq.y, dq.z), 2.0f), M::ld(k[3]).w);
}
template <class M = Plain>
__device__ __forceinline__ V jointLinRows(const Dev &d, const float4 *k, int a, int b, float alpha) {
  return sub(jointLinC<M>(d, k, a, b), M::ld(k[0]).w >= HARD ? mul(xyz(M::ld(k[4])), alpha) : vv(0.0f, 0.0f, 0.0f));
}
template <class M = Plain>
__device__ __forceinline__ V jointAngRows(const Dev &d, const float4 *k, int a, int b, float alpha) {
  return sub(jointAngC<M>(d, k, a, b), M::ld(k[1]).w >= HARD ? mul(xyz(M::ld(k[5])), alpha) : vv(0.0f, 0.0f, 0.0f));
}
/** addJoint (wgsl-solve.ts:115-165): spring row, ball-socket rows with lumped geometric
 *  stiffness, angle-lock rows. */
template <class M = Plain>
__device__ __forceinline__ void addJoint(const Dev &d, Acc &acc, uint32_t j, float alpha, uint32_t i) {
  int4 in = d.jinfo[j];
  int a = in.y, b = in.z;
  const float4 *k = d.joints + 10 * j;
  bool isA = (int)i == a;
  float sg = isA ? 1.0f : -1.0f;
  if (in.x == (int)T_SPRING) {
    V rAW = qrotate(q4(M::ld(d.b.rot[a])), xyz(M::ld(k[6]))), rBW = qrotate(q4(M::ld(d.b.rot[b])), xyz(M::ld(k[7])));
    V dd = sub(add(rAW, xyz(M::ld(d.b.pos[a]))), add(rBW, xyz(M::ld(d.b.pos[b]))));
    float len = length(dd);
    if (len <= 1.0e-6f) return;
    V n = divs(dd, len);
    float st = M::ld(k[0]).w;
    V r = isA ? rAW : rBW;
    addRow(acc, mul(n, sg), mul(cross(r, n), sg), st, st * (len - M::ld(k[6]).w));
    return;
  }
  V penLin = xyz(M::ld(k[0]));
  if (dot(penLin, penLin) > 0.0f) {
    V F = add(mulv(penLin, jointLinRows<M>(d, k, a, b, alpha)), xyz(M::ld(k[2])));
    V r = isA ? qrotate(q4(M::ld(d.b.rot[a])), xyz(M::ld(k[6]))) : neg(qrotate(q4(M::ld(d.b.rot[b])), xyz(M::ld(k[7]))));
    V v = neg(r);
    addRow(acc, vv(sg, 0.0f, 0.0f), vv(0.0f, -v.z, v.y), penLin.x, F.x);
    addRow(acc, vv(0.0f, sg, 0.0f), vv(v.z, 0.0f, -v.x), penLin.y, F.y);
    addRow(acc, vv(0.0f, 0.0f, sg), vv(-v.y, v.x, 0.0f), penLin.z, F.z);
    float rf = dot(r, F);
    float g0 = length(sub(mul(r, F.x), vv(rf, 0.0f, 0.0f)));
    float g1 = length(sub(mul(r, F.y), vv(0.0f, rf, 0.0f)));
    float g2 = length(sub(mul(r, F.z), vv(0.0f, 0.0f, rf)));
    acc.ang[0] += g0; acc.ang[1] += 0.0f; acc.ang[2] += g1; acc.ang[3] += 0.0f; acc.ang[4] += 0.0f; acc.ang[5] += g2;
  }
  V penAng = xyz(M::ld(k[1]));
  if (dot(penAng, penAng) > 0.0f) {
    float4 lamAng = M::ld(k[3]);
    V F = add(mulv(penAng, jointAngRows<M>(d, k, a, b, alpha)), xyz(lamAng));
    float s = sg * lamAng.w;
    V z = vv(0.0f, 0.0f, 0.0f);
    addRow(acc, z, vv(s, 0.0f, 0.0f), penAng.x, F.x);
    addRow(acc, z, vv(0.0f, s, 0.0f), penAng.y, F.y);
    addRow(acc, z, vv(0.0f, 0.0f, s), penAng.z, F.z);
  }
}
/** A contact partner's pose, its rotation at the substep's start (x-, a sphere contact's frame) and
 *  its motion since then. */
struct PairBody { Q rot, irot; V dLin, dAng; };
template <class M = Plain>
__device__ __forceinline__ PairBody pairBody(const Dev &d, uint32_t i) {
  PairBody P;
  P.rot = q4(M::ld(d.b.rot[i]));
  P.irot = q4(M::ld(d.b.irot[i]));
  P.dLin = sub(xyz(M::ld(d.b.pos[i])), xyz(M::ld(d.b.ipos[i])));
  P.dAng = qsub(P.rot, P.irot);
  return P;
}
struct Eval { V rAW, rBW, C, F; float frictionScale, bounds; V J[3]; };
/** Branch bits of an evaluation: bit0 normal force clamped to 0, bit1 friction on the cone. */
__device__ __forceinline__ unsigned char evalBits(const Eval &e) {
  return (unsigned char)((e.F.x < 0.0f ? 0u : 1u) | ((e.frictionScale > e.bounds && e.frictionScale > 0.0f) ? 2u : 0u));
}
/** evalContact (wgsl-solve.ts:196-231): Taylor-expanded C about x-, cone-clamped force. Each row's
 *  C = c0 (1 - alpha) + l . dxA - l . dxB + (rA x l) . dthA + (rB x -l) . dthB is taken as
 *  c0 (1 - alpha) + l . g, with g = (dxA - dxB) + (dthA x rA - dthB x rB) by the scalar triple
 *  product: one vector per point instead of two cross products and four dot products per row.
 *  With WANT_J, also each row's angular Jacobian (e.J) for the body on side A (isA) or B: r_A x l,
 *  or r_B x (-l) (a basis row times sg = +-1 is the row or its exact negation). */
template <bool WANT_J = false>
__device__ __forceinline__ Eval evalContact(const Contact &k, const V *basis, float friction, const PairBody &A, const PairBody &B,
                                            float alpha, bool isA = true) {
  Q rotA = A.rot, rotB = B.rot;
  if ((__float_as_uint(k.rA.w) & ~STICK_BIT) == SPHERE_FEATURE || (__float_as_uint(k.rA.w) & SDF_ROLLING_BIT) != 0u) { rotA = A.irot; rotB = B.irot; }
  Eval e;
  e.rAW = qrotate(rotA, xyz(k.rA));
  e.rBW = qrotate(rotB, xyz(k.rB));
  float keep = 1.0f - alpha;
  float c0[3] = {k.rB.w, k.pen.w, k.lam.w};
  float C[3];
  V g = add(sub(A.dLin, B.dLin), sub(cross(A.dAng, e.rAW), cross(B.dAng, e.rBW)));
#pragma unroll
  for (int r = 0; r < 3; r++) {
    C[r] = fmaf(c0[r], keep, fdot(basis[r], g));
    if (WANT_J) e.J[r] = isA ? cross(e.rAW, basis[r]) : cross(e.rBW, neg(basis[r]));
  }
  e.C = vv(C[0], C[1], C[2]);
  V F = add(mulv(xyz(k.pen), e.C), xyz(k.lam));
  F.x = mn(F.x, 0.0f);
  e.bounds = fabsf(F.x) * friction;
  e.frictionScale = sqrtf(F.y * F.y + F.z * F.z);
  if (e.frictionScale > e.bounds && e.frictionScale > 0.0f) {
    float s = e.bounds / e.frictionScale;
    F.y = F.y * s; F.z = F.z * s;
  }
  e.F = F;
  return e;
}
/** accumulate (wgsl-solve.ts:436-501): body i's system from its walk items lane, lane+lanes, ...
 *  (a joint entry, or one contact point of a manifold entry, in adjacency order). With one lane
 *  this is the reference's order; with more, lane l sums its items in order and the lanes are
 *  added by reduceLanes's tree, as the oracle does. A body's points spread over its lanes (not
 *  its manifolds), so a lane evaluates about one point however the points group into manifolds.
 *  Lane 0 also returns the body's pose (pos, rot) for the update: only the body's own update
 *  writes it, so the value read here is the one the update starts from. */
__device__ __forceinline__ void accumulate(const Dev &d, uint32_t i, uint32_t q0, uint32_t end, uint32_t lane, uint32_t lanes, float alpha, Acc &acc,
                                           float4 &pos, float4 &rot) {
  accZero(acc);
  // this lane's first item loads beside lane 0's body reads
  uint32_t q = q0 + lane;
  uint4 it = make_uint4(0u, 0u, 0u, 0u), next;
  if (q < end) it = d.items[q];
  float dt = d.p.dt, dt2 = dt * dt;
  if (lane == 0) {
    float4 sz = d.b.size[i], mo = d.b.moment[i];
    pos = d.b.pos[i];
    rot = d.b.rot[i];
    float m = sz.w / dt2;
    V I = divs(xyz(mo), dt2);
    acc.lin[0] = m; acc.lin[2] = m; acc.lin[5] = m;
    acc.ang[0] = I.x; acc.ang[2] = I.y; acc.ang[5] = I.z;
    V dx = sub(xyz(pos), xyz(d.b.inpos[i]));
    acc.rl[0] = m * dx.x; acc.rl[1] = m * dx.y; acc.rl[2] = m * dx.z;
    V dq = qsub(q4(rot), q4(d.b.inrot[i]));
    acc.ra[0] = I.x * dq.x; acc.ra[1] = I.y * dq.y; acc.ra[2] = I.z * dq.z;
  }
  // The walk is a chain of dependent loads (item -> contact, geometry, partner), so an item's loads
  // go out together and the next item's record loads beside them (its index clamped into the
  // range). The body's own pose terms are the same for every item: read once.
  PairBody self = pairBody(d, i);
  for (; q < end; q += lanes, it = next) {
    next = d.items[q + lanes < end ? q + lanes : end - 1];
    asm volatile("" : "+v"(it.x), "+v"(it.y), "+v"(it.z), "+v"(it.w));  // the record whole at the branch
    if (it.w & 2u) { addJoint(d, acc, it.x, alpha, i); continue; }
    Contact k = d.con[it.x];
    float4 geo = d.adjGeo[it.y];
    bool isA = (it.w & 1u) != 0u;
    PairBody other = pairBody(d, it.z);
    PairBody A = self, B = other;  // by value: a reference picked between two locals would put them in memory
    if (!isA) { A = other; B = self; }
    V basis[3];
    orthonormal(xyz(geo), basis);
    float sg = isA ? 1.0f : -1.0f;
    Eval ev = evalContact<true>(k, basis, geo.w, A, B, alpha, isA);  // ev.J[r] = r x l_r·sg
    if (d.bits) { if (isA) d.bits[it.x].x = evalBits(ev); else d.bits[it.x].y = evalBits(ev); }
    addRow(acc, mul(basis[0], sg), ev.J[0], k.pen.x, ev.F.x);
    addRow(acc, mul(basis[1], sg), ev.J[1], k.pen.y, ev.F.y);
    addRow(acc, mul(basis[2], sg), ev.J[2], k.pen.z, ev.F.z);
  }
}
/** finishBody (wgsl-solve.ts:504-558): LDLᵀ in registers, pivots floored at 1e-5 of the diagonal;
 *  (pos, rot) is the body's pose as accumulate read it, and returns the updated one. */
template <class M = Plain>
__device__ __forceinline__ void finishBody(const Dev &d, uint32_t i, const Acc &acc, float4 &pos, float4 &rot) {
  const float PF = 1e-5f;
  float A11 = acc.lin[0], A21 = acc.lin[1], A22 = acc.lin[2], A31 = acc.lin[3], A32 = acc.lin[4], A33 = acc.lin[5];
  float A41 = acc.cr[0], A42 = acc.cr[1], A43 = acc.cr[2], A44 = acc.ang[0];
  float A51 = acc.cr[3], A52 = acc.cr[4], A53 = acc.cr[5], A54 = acc.ang[1], A55 = acc.ang[2];
  float A61 = acc.cr[6], A62 = acc.cr[7], A63 = acc.cr[8], A64 = acc.ang[3], A65 = acc.ang[4], A66 = acc.ang[5];
  float D1 = A11;
  float L21 = A21 / D1, L31 = A31 / D1, L41 = A41 / D1, L51 = A51 / D1, L61 = A61 / D1;
  float D2 = mx(A22 - L21 * L21 * D1, A22 * PF);
  float L32 = (A32 - L21 * L31 * D1) / D2, L42 = (A42 - L21 * L41 * D1) / D2, L52 = (A52 - L21 * L51 * D1) / D2, L62 = (A62 - L21 * L61 * D1) / D2;
  float D3 = mx(A33 - (L31 * L31 * D1 + L32 * L32 * D2), A33 * PF);
  float L43 = (A43 - L31 * L41 * D1 - L32 * L42 * D2) / D3, L53 = (A53 - L31 * L51 * D1 - L32 * L52 * D2) / D3, L63 = (A63 - L31 * L61 * D1 - L32 * L62 * D2) / D3;
  float D4 = mx(A44 - (L41 * L41 * D1 + L42 * L42 * D2 + L43 * L43 * D3), A44 * PF);
  float L54 = (A54 - L41 * L51 * D1 - L42 * L52 * D2 - L43 * L53 * D3) / D4, L64 = (A64 - L41 * L61 * D1 - L42 * L62 * D2 - L43 * L63 * D3) / D4;
  float D5 = mx(A55 - (L51 * L51 * D1 + L52 * L52 * D2 + L53 * L53 * D3 + L54 * L54 * D4), A55 * PF);
  float L65 = (A65 - L51 * L61 * D1 - L52 * L62 * D2 - L53 * L63 * D3 - L54 * L64 * D4) / D5;
  float D6 = mx(A66 - (L61 * L61 * D1 + L62 * L62 * D2 + L63 * L63 * D3 + L64 * L64 * D4 + L65 * L65 * D5), A66 * PF);
  float y1 = acc.rl[0], y2 = acc.rl[1] - L21 * y1, y3 = acc.rl[2] - L31 * y1 - L32 * y2;
  float y4 = acc.ra[0] - L41 * y1 - L42 * y2 - L43 * y3, y5 = acc.ra[1] - L51 * y1 - L52 * y2 - L53 * y3 - L54 * y4;
  float y6 = acc.ra[2] - L61 * y1 - L62 * y2 - L63 * y3 - L64 * y4 - L65 * y5;
  float w3 = y6 / D6, w2 = y5 / D5 - L65 * w3, w1 = y4 / D4 - L54 * w2 - L64 * w3;
  float v3 = y3 / D3 - L43 * w1 - L53 * w2 - L63 * w3;
  float v2 = y2 / D2 - L32 * v3 - L42 * w1 - L52 * w2 - L62 * w3;
  float v1 = y1 / D1 - L21 * v2 - L31 * v3 - L41 * w1 - L51 * w2 - L61 * w3;
  pos = make_float4(pos.x - v1, pos.y - v2, pos.z - v3, pos.w);
  rot = f4q(qadd(q4(rot), neg(vv(w1, w2, w3))));
  M::st(d.b.pos[i], pos);
  M::st(d.b.rot[i], rot);
}
/** v of the lane S places up in the same 16-lane row (DPP row_shl:S); lanes whose source leaves
 *  the row read 0. A body's lanes are an aligned group of at most 16 (one row), so lane j < S of a
 *  group reads lane j + S of the same group. */
template <int S>
__device__ __forceinline__ float fromUp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x100 + S, 0xf, 0xf, true));
}
template <int S>
__device__ __forceinline__ void reduceStep(Acc &acc) {
#pragma unroll
  for (int e = 0; e < 6; e++) { acc.lin[e] += fromUp<S>(acc.lin[e]); acc.ang[e] += fromUp<S>(acc.ang[e]); }
#pragma unroll
  for (int e = 0; e < 9; e++) acc.cr[e] += fromUp<S>(acc.cr[e]);
#pragma unroll
  for (int e = 0; e < 3; e++) { acc.rl[e] += fromUp<S>(acc.rl[e]); acc.ra[e] += fromUp<S>(acc.ra[e]); }
}
/** Reduce the partial systems of a body's lanes onto lane 0: lanes s..2s-1 onto 0..s-1 for
 *  s = 8, 4, 2, 1 (those below the body's lane count), the sums lane 0 ends with being those of a
 *  butterfly over the group. Every lane of the wavefront must be active. */
__device__ __forceinline__ void reduceLanes(Acc &acc, uint32_t lanes) {
  if (lanes > 8) reduceStep<8>(acc);
  if (lanes > 4) reduceStep<4>(acc);
  if (lanes > 2) reduceStep<2>(acc);
  if (lanes > 1) reduceStep<1>(acc);
}
/** One colour's primal over blocks [block, block + nblocks·k): lanesFor(size) threads per body in
 *  64-wide wavefronts, lane partials reduced by cross-lane shuffles, lane 0 solves. A clashing
 *  colour (bodies that share a constraint) runs in block 0 alone, chunk after chunk, every read
 *  of a chunk before any of its writes (the reference's serial workgroup, wgsl-solve.ts:363-374). */
__device__ __forceinline__ void primalColor(const Dev &d, uint32_t c, uint32_t block, uint32_t nblocks) {
  // No early exit for an empty colour (the loops below run no chunk): a branch here would sink
  // every argument load behind the colour's bounds, one more dependent load at each launch's start.
  uint32_t start = d.colorStart[c], count = d.colorStart[c + 1] - start;
  bool serial = (d.counters[C_CLASHING + (c >> 5)] >> (c & 31u)) & 1u;
  uint32_t lanes = lanesFor(d.p, count);
  uint32_t slot = threadIdx.x / lanes, lane = threadIdx.x % lanes;
  if (serial) {
    // chunks of one 64-wide workgroup's bodies, in order, on block 0's first wavefront
    if (block != 0) return;
    uint32_t per = 64 / lanes;
    for (uint32_t base = 0; base < count; base += per) {
      uint32_t k = base + slot;
      bool live = threadIdx.x < 64 && k < count;
      uint4 item = live ? d.colorItems[start + k] : make_uint4(0u, 0u, 0u, 0u);
      uint32_t i = item.x;
      bool enabled = live && item.w != 0u;  // a disabled body has no contacts: only its update is skipped
      if (d.bits) live = enabled;           // (while recording, nothing of it is evaluated)
      Acc acc;
      float4 pos = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rot = pos;
      if (live) accumulate(d, i, item.y, item.z, lane, lanes, d.p.alpha, acc, pos, rot);
      else accZero(acc);
      reduceLanes(acc, lanes);
      __syncthreads();
      if (enabled && lane == 0) finishBody(d, i, acc, pos, rot);
      __threadfence_block();
      __syncthreads();
    }
    return;
  }
  uint32_t per = blockDim.x / lanes;
  // XCD-aware: workgroup b runs on XCD b mod 8, so give each XCD a contiguous range of the bucket
  // (bodies in index order, spatially coherent): its L2 then holds the range's neighbours.
  uint32_t active = (count + per - 1) / per;
  if (nblocks >= active && block < active) {
    uint32_t x = block & 7u, q = active >> 3, r = active & 7u;
    block = x * q + (x < r ? x : r) + (block >> 3);
  }
  for (uint32_t base = block * per; base < count; base += nblocks * per) {
    uint32_t k = base + slot;
    bool live = k < count;
    uint4 item = live ? d.colorItems[start + k] : make_uint4(0u, 0u, 0u, 0u);
    uint32_t i = item.x;
    bool enabled = live && item.w != 0u;  // bodyEnabled at the colouring
    if (d.bits) live = enabled;           // (while recording, nothing of it is evaluated)
    Acc acc;
    float4 pos = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rot = pos;
    if (live) accumulate(d, i, item.y, item.z, lane, lanes, d.p.alpha, acc, pos, rot);
    else accZero(acc);
    reduceLanes(acc, lanes);
    if (enabled && lane == 0) finishBody(d, i, acc, pos, rot);
  }
}
#ifndef PRIMAL_THREADS
#define PRIMAL_THREADS 64  // k_primal's workgroup size
#endif
__global__ void __launch_bounds__(PRIMAL_THREADS) k_primal(Dev d, uint32_t c) { primalColor(d, c, blockIdx.x, gridDim.x); }
/** The sweep's tail, colours [c0, c1), in one workgroup (primalSparse, wgsl-solve.ts:378-394, with
 *  a chosen first colour): colours stay ordered, each one's bodies are solved chunk by chunk by the
 *  workgroup, and every pose write is visible before the next colour reads it. For small trailing
 *  colours and the cap's empty spares one launch replaces a launch each. */
__global__ void __launch_bounds__(512) k_primal_tail(Dev d, uint32_t c0, uint32_t c1) {
  for (uint32_t c = c0; c < c1; c++) {
    primalColor(d, c, 0, 1);
    __threadfence_block();
    __syncthreads();
  }
}
/** dualJoint (wgsl-solve.ts:562-635): multipliers, penalty ramp, plastic yield and bend limit, fracture. */
template <class M = Plain>
__device__ __forceinline__ void dualJoint(const Dev &d, uint32_t j, float alpha) {
  int4 in = d.jinfo[j];
  if (in.x != (int)T_JOINT) return;
  int a = in.y, b = in.z;
  if ((a >= 0 && !enabledM<M>(d, (uint32_t)a)) || !enabledM<M>(d, (uint32_t)b)) return;
  float4 *k = d.joints + 10 * j;
  float4 penLin = M::ld(k[0]), penAng = M::ld(k[1]), lamLin = M::ld(k[2]), lamAng = M::ld(k[3]);
  if (dot(xyz(penLin), xyz(penLin)) > 0.0f) {
    V C = jointLinRows<M>(d, k, a, b, alpha);
    if (penLin.w >= HARD) lamLin = f4(add(mulv(xyz(penLin), C), xyz(lamLin)), lamLin.w);
    float cap = mn(penLin.w, PENALTY_MAX);
    V np = add(xyz(penLin), mul(vabs(C), d.p.betaLin));
    penLin = make_float4(mn(np.x, cap), mn(np.y, cap), mn(np.z, cap), penLin.w);
  }
  float frac = lamLin.w, limit = fabsf(frac);
  bool yielded = false;
  if (dot(xyz(penAng), xyz(penAng)) > 0.0f) {
    V C = jointAngRows<M>(d, k, a, b, alpha);
    if (penAng.w >= HARD) {
      lamAng = f4(add(mulv(xyz(penAng), C), xyz(lamAng)), lamAng.w);
      // a plastic joint (yield in c0Ang.w, wgsl-solve.ts:586-599): past yield, short of fracture,
      // the angular force is cut back to yield and the bend becomes the new rest
      float yieldForce = M::ld(k[5]).w;
      if (yieldForce < BIG) {
        float force = length(xyz(lamAng));
        if (force > yieldForce && !(limit < BIG && force > limit)) {
          lamAng = f4(mul(xyz(lamAng), yieldForce / force), lamAng.w);
          yielded = true;
        }
      }
    }
    if (!yielded) {  // a yielding row is at its bound: its penalty is not ramped
      float cap = mn(penAng.w, PENALTY_MAX);
      V np = add(xyz(penAng), mul(vabs(C), d.p.betaAng));
      penAng = make_float4(mn(np.x, cap), mn(np.y, cap), mn(np.z, cap), penAng.w);
    }
  }
  // the rest moves to where the bodies are; past the bend limit (rB.w) from the rest it started
  // with, the joint breaks (wgsl-solve.ts:606-614)
  Q rest = q4(M::ld(k[8]));
  bool bentOff = false;
  if (yielded) {
    rest = qnormalize(qmul(qconj(rotAOf<M>(d, a)), q4(M::ld(d.b.rot[b]))));
    if (M::ld(k[7]).w < BIG) {
      Q dq = qmul(qconj(q4(M::ld(k[9]))), rest);
      bentOff = 2.0f * atan2f(length(vv(dq.x, dq.y, dq.z)), fabsf(dq.w)) > M::ld(k[7]).w;
    }
  }
  bool linear = frac < 0.0f && dot(xyz(lamLin), xyz(lamLin)) > limit * limit;
  if ((limit < BIG && (dot(xyz(lamAng), xyz(lamAng)) > limit * limit || linear)) || bentOff) {
    atomicOr(&d.counters[C_TOPO], 2u);  // broken: the next adjacency drops it
    penLin = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    penAng = penLin;
    lamLin = make_float4(0.0f, 0.0f, 0.0f, frac);
    lamAng = make_float4(0.0f, 0.0f, 0.0f, lamAng.w);
  }
  M::st(k[0], penLin); M::st(k[1], penAng); M::st(k[2], lamLin); M::st(k[3], lamAng);
  if (yielded) {
    M::st(k[8], f4q(rest));
    M::st(k[5], make_float4(0.0f, 0.0f, 0.0f, M::ld(k[5]).w));
  }
}
/** dualManifold's update of one contact point c (wgsl-solve.ts:637-662), given its manifold's basis,
 *  friction and both bodies' terms: penalty ramp, stick flag, λ. Written back: the record's upper
 *  32 bytes (penalties and multipliers) and, when the stick bit flipped, the key word. */
template <class M = Plain>
__device__ __forceinline__ void dualContactOf(const Dev &d, uint32_t c, const Contact &k, const V *basis, float friction, const PairBody &A,
                                              const PairBody &B, float alpha) {
  Eval e = evalContact(k, basis, friction, A, B, alpha);
  if (d.bits) d.bits[c].z = evalBits(e);
  V pen = xyz(k.pen);
  float BL = d.p.betaLin;
  if (e.F.x < 0.0f) pen.x = mn(pen.x + BL * fabsf(e.C.x), PENALTY_MAX);
  uint32_t key = __float_as_uint(k.rA.w);
  if (e.frictionScale <= e.bounds) {
    pen.y = mn(pen.y + BL * fabsf(e.C.y), PENALTY_MAX);
    pen.z = mn(pen.z + BL * fabsf(e.C.z), PENALTY_MAX);
    bool stick = sqrtf(e.C.y * e.C.y + e.C.z * e.C.z) < 0.00001f;
    key = (key & ~STICK_BIT) | (stick ? STICK_BIT : 0u);
  }
  // only what changes goes out: penalties and multipliers (the record's upper 32 bytes), and the
  // key word when its stick bit flipped (anchors and C0 stay)
  float4 *rec = reinterpret_cast<float4 *>(d.con + c);
  if (key != __float_as_uint(k.rA.w)) M::stw(reinterpret_cast<float *>(rec)[3], __uint_as_float(key));
  M::st(rec[2], f4(pen, k.pen.w));
  M::st(rec[3], f4(e.F, k.lam.w));
}
template <class M = Plain>
__device__ __forceinline__ void dualContact(const Dev &d, uint32_t c, const V *basis, float friction, const PairBody &A, const PairBody &B,
                                            float alpha) {
  dualContactOf<M>(d, c, loadContact<M>(d.con[c]), basis, friction, A, B, alpha);
}
/** dualManifold (wgsl-solve.ts:637-662) for contact point c alone: every point's update is its own,
 *  so a thread per point gives the manifold loop's results while no wavefront waits on its widest
 *  manifold. The point's manifold comes from the map the contact's writer left (cman); a slot no
 *  manifold of this step covers (dropped on overflow) is skipped. */
__device__ __forceinline__ void dualPoint(const Dev &d, uint32_t c, float alpha) {
  uint32_t m = d.cman[c];
  if (m >= manifoldCount(d)) return;
  uint4 ids = d.mIds[m];
  if (c < ids.z || c >= ids.z + (ids.w & 15u)) return;
  float4 geo = d.mGeo[m];
  PairBody A = pairBody(d, ids.x), B = pairBody(d, ids.y);
  V basis[3];
  orthonormal(xyz(geo), basis);
  dualContact(d, c, basis, geo.w, A, B, alpha);
}
__device__ __forceinline__ void dualAll(const Dev &d, uint32_t t0, uint32_t nt) {
  uint32_t points = d.counters[C_CONTACTS];
  if (points > d.conCap) points = d.conCap;
  uint32_t total = d.nj + points;
  for (uint32_t id = t0; id < total; id += nt) {
    if (id < d.nj) dualJoint(d, id, d.p.alpha);
    else dualPoint(d, id - d.nj, d.p.alpha);
  }
}
__global__ void __launch_bounds__(256) k_dual(Dev d) { dualAll(d, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x); }
/** warmStartJoints (wgsl-solve.ts:236-260). */
__device__ __forceinline__ void warmJoints(const Dev &d, uint32_t t0, uint32_t nt) {
  for (uint32_t j = t0; j < d.nj; j += nt) {
    int4 in = d.jinfo[j];
    if (in.x != (int)T_JOINT) continue;
    if ((in.y >= 0 && !bodyEnabled(d, (uint32_t)in.y)) || !bodyEnabled(d, (uint32_t)in.z)) continue;
    float4 *k = d.joints + 10 * j;
    k[4] = f4(jointLinC(d, k, in.y, in.z), 0.0f);
    k[5] = f4(jointAngC(d, k, in.y, in.z), k[5].w);
    if ((in.y >= 0 && bodyCold(d, (uint32_t)in.y)) || bodyCold(d, (uint32_t)in.z)) {
      // a cold body's joints restart: no multipliers, penalties at their floor
      float4 pl = k[0], pa = k[1], ll = k[2], la = k[3];
      k[2] = make_float4(0.0f, 0.0f, 0.0f, ll.w);
      k[3] = make_float4(0.0f, 0.0f, 0.0f, la.w);
      float fl = mn(PENALTY_MIN, pl.w), fa = mn(PENALTY_MIN, pa.w);
      k[0] = make_float4(fl, fl, fl, pl.w);
      k[1] = make_float4(fa, fa, fa, pa.w);
      continue;
    }
    float decay = d.p.alpha * d.p.gamma, g = d.p.gamma;
    float4 pl = k[0], pa = k[1], ll = k[2], la = k[3];
    k[2] = f4(mul(xyz(ll), decay), ll.w);
    k[3] = f4(mul(xyz(la), decay), la.w);
    k[0] = make_float4(mn(clampf(pl.x * g, PENALTY_MIN, PENALTY_MAX), pl.w), mn(clampf(pl.y * g, PENALTY_MIN, PENALTY_MAX), pl.w), mn(clampf(pl.z * g, PENALTY_MIN, PENALTY_MAX), pl.w), pl.w);
    k[1] = make_float4(mn(clampf(pa.x * g, PENALTY_MIN, PENALTY_MAX), pa.w), mn(clampf(pa.y * g, PENALTY_MIN, PENALTY_MAX), pa.w), mn(clampf(pa.z * g, PENALTY_MIN, PENALTY_MAX), pa.w), pa.w);
  }
}
__global__ void k_warm_joints(Dev d) { warmJoints(d, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x); }
/** warmStartBodies (wgsl-solve.ts:293-335): inertial target, adaptive warm start, x-. */
__device__ __forceinline__ void warmBodies(const Dev &d, uint32_t t0, uint32_t nt) {
  for (uint32_t i = t0; i < d.n; i += nt) {
    if (!bodyEnabled(d, i)) continue;
    float4 pos = d.b.pos[i], vel4 = d.b.vel[i], av4 = d.b.angvel[i];
    Q rot = q4(d.b.rot[i]);
    V vel = xyz(vel4), angVel = xyz(av4);
    float dt = d.p.dt, g = d.p.gravity;
    bool dyn = d.b.size[i].w > 0.0f;
    if (d.aux) {
      // the aux variant (wgsl-solve.ts:293-305, DAMPING_WGSL :14-20): scaled gravity, and a damped
      // body's velocities decayed before the inertial target and the warm start read them
      float4 ax = d.aux[i];
      g = g * ax.x;
      if (dyn && ax.w != 0.0f) {
        vel = mul(vel, ax.y);
        angVel = mul(angVel, ax.z);
        d.b.vel[i] = f4(vel, vel4.w);
        d.b.angvel[i] = f4(angVel, av4.w);
      }
    }
    V up = vv(d.p.upx, d.p.upy, d.p.upz);
    V inpos = add(xyz(pos), mul(vel, dt));
    if (dyn) inpos = add(inpos, mul(up, g * (dt * dt)));
    d.b.inpos[i] = f4(inpos, d.b.inpos[i].w);
    d.b.inrot[i] = f4q(qadd(rot, mul(angVel, dt)));
    float w = 0.0f;
    if (fabsf(g) > 0.0f) {
      float sg = g > 0.0f ? 1.0f : (g < 0.0f ? -1.0f : 0.0f);
      w = clampf((dot(vel, up) - vel4.w) / dt * sg / fabsf(g), 0.0f, 1.0f);
    }
    bool slow = length(vel) + length(angVel) * d.b.moment[i].w < 0.1f;
    bool atRest = (d.p.flags & FLAG_START_AT_REST) && slow;
    d.b.ipos[i] = f4(xyz(pos), 0.0f);
    d.b.irot[i] = f4q(rot);
    if (dyn && !atRest) {
      d.b.pos[i] = f4(add(add(xyz(pos), mul(vel, dt)), mul(up, g * (w * dt * dt))), pos.w);
      d.b.rot[i] = f4q(qadd(rot, mul(angVel, dt)));
    }
    bool moving = vel.x != 0.0f || vel.y != 0.0f || vel.z != 0.0f || angVel.x != 0.0f || angVel.y != 0.0f || angVel.z != 0.0f;
    if (!dyn && moving) {
      d.b.pos[i] = f4(add(xyz(pos), mul(vel, dt)), pos.w);
      d.b.rot[i] = f4q(qadd(rot, mul(angVel, dt)));
    }
  }
}
__global__ void k_warm_bodies(Dev d) { warmBodies(d, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x); }
/** updateVelocities (wgsl-solve.ts:672-684): BDF1. */
__device__ __forceinline__ void velocities(const Dev &d, uint32_t t0, uint32_t nt) {
  for (uint32_t i = t0; i < d.n; i += nt) {
    if (!bodyEnabled(d, i)) continue;
    V up = vv(d.p.upx, d.p.upy, d.p.upz);
    float4 vel4 = d.b.vel[i];
    float prevUp = dot(xyz(vel4), up);
    if (d.b.size[i].w > 0.0f) {
      d.b.vel[i] = f4(divs(sub(xyz(d.b.pos[i]), xyz(d.b.ipos[i])), d.p.dt), prevUp);
      float4 av = d.b.angvel[i];
      d.b.angvel[i] = f4(divs(qsub(q4(d.b.rot[i]), q4(d.b.irot[i])), d.p.dt), av.w);
    } else {
      vel4.w = prevUp;
      d.b.vel[i] = vel4;
    }
  }
}
__global__ void k_velocities(Dev d) { velocities(d, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x); }

/** Grid-wide barrier for a cooperative (co-resident) grid. bar[0] counts arrivals over the whole
 *  launch (the host zeroes it before each launch): a workgroup arriving at its k-th barrier adds 1
 *  and waits until the count reaches k x gridDim.x, so a count never needs a reset between
 *  barriers, and a workgroup that stopped waiting still arrives exactly once at every later barrier.
 *  The wait is bounded: past `polls` polls it adds 1 to counters[C_BARRIER] (read by the host,
 *  which then drops to per-colour launches) and lets the workgroup go on, so a barrier can give a
 *  wrong result but never hang the device, nor spoil the barriers after it. */
__device__ __forceinline__ void gridBarrier(uint32_t *bar, uint32_t *counters, uint32_t &target, uint32_t polls) {
  __syncthreads();
  target += gridDim.x;
  if (threadIdx.x == 0) {
    __threadfence();
    atomicAdd(&bar[0], 1u);
    // relaxed polls (no cache invalidation per poll), one acquire fence once released
    uint32_t spins = 0;
    while (__hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(2);
      if (++spins > polls) { atomicAdd(&counters[C_BARRIER], 1u); break; }
    }
    __threadfence();
  }
  __syncthreads();
}

/** The whole solve of one substep in one cooperative launch (GpuSolver3D.step's solve pass,
 *  solver.ts:2226-2261): warm starts, iterations x (primal of each colour in use, dual), BDF1
 *  velocities, with a grid-wide barrier wherever the reference ends a dispatch. Sweeps only the
 *  colours the colouring used (the device's colour count), not the whole cap. One instance per
 *  workgroup size: the body walk holds its contact batch in registers, which a 1024-thread bound
 *  (128 per lane) would spill. */
template <uint32_t THREADS>
__global__ void __launch_bounds__(THREADS) k_solve(Dev d, uint32_t *bar, uint32_t polls) {
  uint32_t target = 0;
  uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
  warmJoints(d, t0, nt);
  gridBarrier(bar, d.counters, target, polls);
  warmBodies(d, t0, nt);
  gridBarrier(bar, d.counters, target, polls);
  uint32_t colors = d.counters[C_NUM_COLORS];
  if (colors > d.p.colorCap) colors = d.p.colorCap;
  for (uint32_t it = 0; it < d.p.iterations; it++) {
    for (uint32_t c = 0; c < colors; c++) {
      primalColor(d, c, blockIdx.x, gridDim.x);
      gridBarrier(bar, d.counters, target, polls);
    }
    dualAll(d, t0, nt);
    gridBarrier(bar, d.counters, target, polls);
  }
  velocities(d, t0, nt);
}

// ---------------------------------------------------------------------------------------------
// Exclusive scan of u32 (1024 elements per 256-thread block, then block sums, then add).
__device__ __forceinline__ uint32_t waveInclusive(uint32_t v) {
  uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(v, o, 64);
    if (lane >= (uint32_t)o) v += t;
  }
  return v;
}
__device__ __forceinline__ uint32_t blockExclusive(uint32_t v, uint32_t *total) {
  __shared__ uint32_t waves[17];  // per-wave offsets, then the block total in [16]
  uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t inc = waveInclusive(v);
  if (lane == 63) waves[w] = inc;
  __syncthreads();
  if (w == 0) {
    uint32_t x = lane < nw ? waves[lane] : 0u;
    uint32_t xi = waveInclusive(x);
    if (lane < nw) waves[lane] = xi - x;
    if (lane == nw - 1) waves[16] = xi;
  }
  __syncthreads();
  uint32_t r = inc - v + waves[w];
  *total = waves[16];
  __syncthreads();
  return r;
}
__global__ void __launch_bounds__(256) k_scan_local(const uint32_t *in, uint32_t *out, uint32_t *sums, uint32_t n, const uint32_t *skip) {
  if (skip && (*skip & 1u)) return;
  uint32_t base = blockIdx.x * 1024 + threadIdx.x * 4;
  uint32_t v[4];
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = base + k < n ? in[base + k] : 0u;
  uint32_t t = v[0] + v[1] + v[2] + v[3], total;
  uint32_t ex = blockExclusive(t, &total);
#pragma unroll
  for (int k = 0; k < 4; k++) { if (base + k < n) out[base + k] = ex; ex += v[k]; }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}
/** Adds to block b's outputs the totals of blocks 0..b-1, summed here (integer sums: any order gives
 *  the same result), so the block totals need no scan launch of their own. */
__global__ void __launch_bounds__(256) k_scan_add(uint32_t *out, const uint32_t *sums, uint32_t n, const uint32_t *skip) {
  if (skip && (*skip & 1u)) return;
  if (blockIdx.x == 0) return;  // nothing before it
  uint32_t part = 0, s;
  for (uint32_t k = threadIdx.x; k < blockIdx.x; k += blockDim.x) part += sums[k];
  (void)blockExclusive(part, &s);
  uint32_t base = blockIdx.x * 1024 + threadIdx.x * 4;
#pragma unroll
  for (int k = 0; k < 4; k++) if (base + k < n) out[base + k] += s;
}

// ---------------------------------------------------------------------------------------------
// Record interchange: SoA streams <-> 40-float body records (layout.ts:8-58).
__global__ void k_bodies_to_records(Dev d, float4 *rec, uint32_t first, uint32_t count) {
  GRID_STRIDE(k, count) {
    uint32_t i = first + k;
    float4 *o = rec + 10 * k;
    o[0] = d.b.pos[i]; o[1] = d.b.rot[i]; o[2] = d.b.ipos[i]; o[3] = d.b.irot[i]; o[4] = d.b.size[i];
    o[5] = d.b.moment[i]; o[6] = d.b.inpos[i]; o[7] = d.b.inrot[i]; o[8] = d.b.vel[i]; o[9] = d.b.angvel[i];
  }
}
__global__ void k_records_to_bodies(Dev d, const float4 *rec, uint32_t first, uint32_t count) {
  GRID_STRIDE(k, count) {
    uint32_t i = first + k;
    const float4 *o = rec + 10 * k;
    d.b.pos[i] = o[0]; d.b.rot[i] = o[1]; d.b.ipos[i] = o[2]; d.b.irot[i] = o[3]; d.b.size[i] = o[4];
    d.b.moment[i] = o[5]; d.b.inpos[i] = o[6]; d.b.inrot[i] = o[7]; d.b.vel[i] = o[8]; d.b.angvel[i] = o[9];
    d.refRot[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}
/** patchBody's masked writes (solver.ts:1103-1131) on one body: `fields` bit 1 position (pos, x-,
 *  inertial pos xyz), 2 rotation (rot, q-, inertial rot), 4 velocity xyz, 8 angular velocity xyz,
 *  16 mass and moments (a fixed/dynamic switch), 32 the lifecycle word of x- (rec[2].w). */
struct PatchRecord { float4 r[10]; };
__global__ void k_patch_body(Dev d, uint32_t i, uint32_t fields, PatchRecord rec) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const float4 *r = rec.r;
  auto xyzOf = [](float4 dst, float4 src) { return make_float4(src.x, src.y, src.z, dst.w); };
  if (fields & 1u) { d.b.pos[i] = xyzOf(d.b.pos[i], r[0]); d.b.ipos[i] = xyzOf(d.b.ipos[i], r[0]); d.b.inpos[i] = xyzOf(d.b.inpos[i], r[0]); }
  if (fields & 2u) { d.b.rot[i] = r[1]; d.b.irot[i] = r[1]; d.b.inrot[i] = r[1]; }
  if (fields & 4u) d.b.vel[i] = xyzOf(d.b.vel[i], r[8]);
  if (fields & 8u) d.b.angvel[i] = xyzOf(d.b.angvel[i], r[9]);
  if (fields & 16u) {
    float4 sz = d.b.size[i];
    d.b.size[i] = make_float4(sz.x, sz.y, sz.z, r[4].w);
    d.b.moment[i] = xyzOf(d.b.moment[i], r[5]);
  }
  if (fields & 32u) { float4 ip = d.b.ipos[i]; ip.w = r[2].w; d.b.ipos[i] = ip; }
}

}  // namespace avbd
