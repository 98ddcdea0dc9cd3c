// The dataflow sweep: a substep's iterations (every colour's primal, then the dual) in one
// cooperative launch, ordered by per-body completion counts instead of a kernel boundary per colour.
//
// The colour-by-colour sweep (GpuSolver3D.step's solve pass, solver.ts:2226-2261; wgsl-solve.ts
// primal/dual) fixes what every update reads: a body of colour c at iteration t reads the poses
// its neighbours of lower colours wrote at iteration t and those of higher colours wrote at t - 1,
// and the multipliers and penalties the dual of t - 1 left. Here each body waits for exactly those
// versions and no more: with flowDone[j] the iterations body j has completed (its update and the
// duals it owns), body i at iteration t waits until flowDone[j] >= t + 1 for every neighbour j
// ordered before it and flowDone[j] >= t for every neighbour ordered after it (and for itself).
// A neighbour ordered after it cannot update before i's update of t is complete (it waits for
// i), so nothing i reads is overwritten early. Bodies are ordered by colour, and within a
// clashing (serial) colour by the chunk of one wavefront's bodies the serial sweep takes them in;
// bodies of one chunk read before any of them writes, as the serial sweep does (one wavefront).
//
// The dual of a contact or joint runs right after the update of whichever of its bodies is
// ordered last (the higher index within one chunk; the dynamic one against a fixed body): both
// poses are then the iteration's final ones, and the next iteration's readers of its multipliers
// wait for that body. Joints that no dynamic body walks have no reader in the sweep: their duals,
// every iteration's in turn, run at the launch's start (marked beforehand by k_flow_prep, since a
// walked joint that breaks during the sweep stops counting as active).
//
// Chunks (one wavefront's bodies of one colour) are numbered in (iteration, colour, chunk) order
// and dealt round-robin to the launch's workgroups, each taking its own in increasing order.
// Every wait is on a lower-numbered chunk, so with all workgroups co-resident (a cooperative
// launch) the smallest unfinished chunk can always proceed: the sweep cannot deadlock. Waits are
// still bounded (counters[C_BARRIER], as the cooperative barrier's), so a fault in this reasoning
// shows as a reported give-up, never a hung device.
//
// Memory: poses, contact records and joint records, which the sweep writes, are read and written
// with device-scope atomics (Coherent): they bypass the per-XCD L2s, so a value stored by one XCD
// is seen by another once the store completed. A workgroup publishes a count only after all of its
// stores completed (s_waitcnt vmcnt(0)). Everything else the sweep reads is fixed during it.
#pragma once
#include "avbd_kernels.h"

namespace avbd {

// walk item flags written by k_flow_prep (bits 0 and 1 are k_item_fill's: body is A, joint)
static const uint32_t FLOW_DEP = 4u;    // the item's other body is dynamic: wait for it
static const uint32_t FLOW_LOWER = 8u;  // ... and it is ordered before this body
static const uint32_t FLOW_OWN = 16u;   // this body runs the item's dual (a point or a joint)

/** No compiler motion of memory operations across this point. */
__device__ __forceinline__ void flowOrder() { __atomic_signal_fence(__ATOMIC_SEQ_CST); }
/** Every memory operation this wavefront issued has completed. */
__device__ __forceinline__ void flowStoresDone() {
  flowOrder();
  __builtin_amdgcn_s_waitcnt(0);
  flowOrder();
}
__device__ __forceinline__ uint32_t flowLoad(const uint32_t *p) {
  return __hip_atomic_load(const_cast<uint32_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/** Joint j's dual has no reader in the sweep: not a joint the adjacency walks (neither body
 *  dynamic, or switched off). A switched-off joint's penalties are zero after its warm start, so
 *  its dual reads no pose; a joint between fixed bodies reads poses the sweep does not move. */
__device__ __forceinline__ bool flowOrphan(const Dev &d, uint32_t j) {
  int4 in = d.jinfo[j];
  if (in.x != (int)T_JOINT) return false;
  return !(jointActive(d, j) && (isDynamic(d, in.y) || isDynamic(d, in.z)));
}

/** The order key a sweep starting at colour `first` starts from (past every key: no colour). */
__host__ __device__ __forceinline__ uint32_t flowFirstKey(uint32_t first) { return first >= 64u ? 0xffffffffu : first << 26; }
/** Before the sweep: each walk item's flags from the bodies' order keys (and a joint item's other
 *  body in .z), and which joints no body walks. The sweep starts at colour `first` (0, or where the
 *  dataflow tail starts): bodies ordered before it are final when it starts, so their items need no
 *  flags and their partners no waits. */
__global__ void __launch_bounds__(256) k_flow_prep(Dev d, uint32_t first, uint32_t lead) {
  // the swept bodies are the buckets of colours [first, colours) (from colour 0 with `lead`: the
  // colours before the tail run as one launch too), contiguous in colorItems; sixteen threads a
  // body split its items (each item's flags are its own)
  const uint32_t firstKey = flowFirstKey(first), L = 16;
  uint32_t colors = d.counters[C_NUM_COLORS];
  if (colors > d.p.colorCap) colors = d.p.colorCap;
  uint32_t from = lead ? 0u : first;
  uint32_t b0 = from < colors ? d.colorStart[from] : 0u, b1 = from < colors ? d.colorStart[colors] : 0u;
  GRID_STRIDE(t, (b1 - b0) * L) {
    uint32_t i = d.colorItems[b0 + t / L].x;
    uint32_t key = d.bodyKey[i];
    for (uint32_t q = d.itemStart[i] + t % L, q1 = d.itemStart[i + 1]; q < q1; q += L) {
      uint4 it = d.items[q];
      int other = d.adjOther[it.y];
      uint32_t f = it.w & 3u;
      if (key < firstKey) {
        // a body of the leading launch: one iteration per launch, so it waits only for partners
        // ordered before it (their update of this iteration); its duals run in the tail launch
        if (other >= 0 && d.bodyKey[other] < key) {
          f |= FLOW_DEP | FLOW_LOWER;
          if (it.w & 2u) it.z = (uint32_t)other;
        }
        it.w = f;
        d.items[q] = it;
        continue;
      }
      if (other >= 0 && d.bodyKey[other] >= firstKey) {
        uint32_t ko = d.bodyKey[other];
        f |= FLOW_DEP;
        if (ko < key) f |= FLOW_LOWER | FLOW_OWN;
        else if (ko == key && i > (uint32_t)other) f |= FLOW_OWN;
        if (it.w & 2u) it.z = (uint32_t)other;
      } else {
        f |= FLOW_OWN;
      }
      it.w = f;
      d.items[q] = it;
    }
  }
  GRID_STRIDE(j, d.nj) d.flowOrphan[j] = flowOrphan(d, j) ? 1u : 0u;
}

/** Waits until body j has completed `need` iterations. Bounded: past `polls` polls it adds 1 to
 *  counters[C_BARRIER] and goes on, and once any wait has given up the others return at once, so
 *  the launch always ends (its result then void: the host reports it and drops the schedule). */
__device__ __forceinline__ void flowWait(const Dev &d, uint32_t j, uint32_t need, uint32_t polls) {
  if (flowLoad(&d.flowDone[j]) >= need) return;
  for (uint32_t spins = 1;; spins++) {
    __builtin_amdgcn_s_sleep(1);
    if (flowLoad(&d.flowDone[j]) >= need) return;
    if ((spins & 255u) == 0u && flowLoad(&d.counters[C_BARRIER]) != 0u) return;
    if (spins > polls) {
      atomicAdd(&d.counters[C_BARRIER], 1u);
      return;
    }
  }
}

/** A contact partner's terms: its pose (which the sweep writes) read coherently, its pose at the
 *  substep's start (fixed during the sweep) through the caches; pairBody's arithmetic. */
__device__ __forceinline__ PairBody flowPartner(const Dev &d, uint32_t j) {
  PairBody P;
  P.rot = q4(Coherent::ld(d.b.rot[j]));
  P.irot = q4(d.b.irot[j]);
  P.dLin = sub(xyz(Coherent::ld(d.b.pos[j])), xyz(d.b.ipos[j]));
  P.dAng = qsub(P.rot, P.irot);
  return P;
}
/** Body i's own terms from its pose (pos, rot), as pairBody forms them. */
__device__ __forceinline__ PairBody flowSelf(const Dev &d, uint32_t i, float4 pos, float4 rot) {
  PairBody P;
  P.rot = q4(rot);
  P.irot = q4(d.b.irot[i]);
  P.dLin = sub(xyz(pos), xyz(d.b.ipos[i]));
  P.dAng = qsub(P.rot, P.irot);
  return P;
}
/** lane `src`'s value of v (a body's lanes read their first lane's). */
__device__ __forceinline__ float4 flowShfl(float4 v, uint32_t src) {
  return make_float4(__shfl(v.x, (int)src), __shfl(v.y, (int)src), __shfl(v.z, (int)src), __shfl(v.w, (int)src));
}

/** accumulate's walk (same items, same order, same arithmetic) with each item's wait before its
 *  reads. The body's pose is read once (its first lane) and handed to its lanes. A contact's
 *  anchors and C0 are fixed during the sweep and come through the caches (the key's stick bit,
 *  which the dual flips, is masked by the evaluation), its penalties and multipliers coherently.
 *  The body's own last update needs no wait: the same workgroup ran it, stores completed. */
template <bool CACHED = false>
__device__ __forceinline__ void flowAccumulate(const Dev &d, uint32_t i, uint32_t q0, uint32_t end, uint32_t lane, uint32_t lanes, float alpha,
                                               uint32_t iter, uint32_t polls, Acc &acc, float4 &pos, float4 &rot) {
  accZero(acc);
  uint32_t q = q0 + lane;
  uint4 it = make_uint4(0u, 0u, 0u, 0u), next;
  if (q < end) it = d.items[q];
  float dt = d.p.dt, dt2 = dt * dt;
  if (lane == 0) {
    pos = CACHED ? d.b.pos[i] : Coherent::ld(d.b.pos[i]);
    rot = CACHED ? d.b.rot[i] : Coherent::ld(d.b.rot[i]);
    float4 sz = d.b.size[i], mo = d.b.moment[i], inp = d.b.inpos[i], inr = d.b.inrot[i];
    float m = sz.w / dt2;
    V I = divs(xyz(mo), dt2);
    acc.lin[0] = m; acc.lin[2] = m; acc.lin[5] = m;
    acc.ang[0] = I.x; acc.ang[2] = I.y; acc.ang[5] = I.z;
    V dx = sub(xyz(pos), xyz(inp));
    acc.rl[0] = m * dx.x; acc.rl[1] = m * dx.y; acc.rl[2] = m * dx.z;
    V dq = qsub(q4(rot), q4(inr));
    acc.ra[0] = I.x * dq.x; acc.ra[1] = I.y * dq.y; acc.ra[2] = I.z * dq.z;
  }
  uint32_t first = threadIdx.x - lane;
  pos = flowShfl(pos, first);
  rot = flowShfl(rot, first);
  PairBody self = flowSelf(d, i, pos, rot);
  for (; q < end; q += lanes, it = next) {
    next = d.items[q + lanes < end ? q + lanes : end - 1];
    asm volatile("" : "+v"(it.x), "+v"(it.y), "+v"(it.z), "+v"(it.w));  // the record whole at the branch
    float4 geo = d.adjGeo[it.y];
    Contact k;
    if (!(it.w & 2u)) { k.rA = d.con[it.x].rA; k.rB = d.con[it.x].rB; }
    if (it.w & FLOW_DEP) flowWait(d, it.z, iter + ((it.w & FLOW_LOWER) ? 1u : 0u), polls);
    flowOrder();
    if (it.w & 2u) { addJoint<Coherent>(d, acc, it.x, alpha, i); continue; }
    if (CACHED) {
      k.pen = d.con[it.x].pen;
      k.lam = d.con[it.x].lam;
    } else {
      k.pen = Coherent::ld(d.con[it.x].pen);
      k.lam = Coherent::ld(d.con[it.x].lam);
    }
    bool isA = (it.w & 1u) != 0u;
    bool fresh = !CACHED || (it.w & (FLOW_DEP | FLOW_LOWER)) == (FLOW_DEP | FLOW_LOWER);
    PairBody other = fresh ? flowPartner(d, it.z) : pairBody(d, it.z);
    PairBody A = self, B = other;
    if (!isA) { A = other; B = self; }
    V basis[3];
    orthonormal(xyz(geo), basis);
    float sg = isA ? 1.0f : -1.0f;
    Eval ev = evalContact<true>(k, basis, geo.w, A, B, alpha, isA);
    addRow(acc, mul(basis[0], sg), ev.J[0], k.pen.x, ev.F.x);
    addRow(acc, mul(basis[1], sg), ev.J[1], k.pen.y, ev.F.y);
    addRow(acc, mul(basis[2], sg), ev.J[2], k.pen.z, ev.F.z);
  }
}

/** The duals body i owns among its items (lane, lane + lanes, ...), as k_dual computes them, with
 *  the body's updated pose (pos, rot) from registers. A contact's manifold record is the walk
 *  entry's copy (the same words: the adjacency holds this step's manifolds). A same-chunk partner
 *  (a serial colour) and joints, whose duals read both poses from memory, come after the chunk's
 *  stores completed. */
__device__ __forceinline__ void flowDuals(const Dev &d, uint32_t i, uint32_t q0, uint32_t end, uint32_t lane, uint32_t lanes, float alpha,
                                          float4 pos, float4 rot, bool serial) {
  if (serial) flowStoresDone();
  PairBody self = flowSelf(d, i, pos, rot);
  bool joints = false;
  for (uint32_t q = q0 + lane; q < end; q += lanes) {
    uint4 it = d.items[q];
    if (!(it.w & FLOW_OWN)) continue;
    if (it.w & 2u) { joints = true; continue; }
    float4 geo = d.adjGeo[it.y];
    Contact k;
    k.rA = Coherent::ld(d.con[it.x].rA);  // the key's stick bit is the dual's
    k.rB = d.con[it.x].rB;
    k.pen = Coherent::ld(d.con[it.x].pen);
    k.lam = Coherent::ld(d.con[it.x].lam);
    PairBody other = flowPartner(d, it.z);
    bool isA = (it.w & 1u) != 0u;
    PairBody A = self, B = other;
    if (!isA) { A = other; B = self; }
    V basis[3];
    orthonormal(xyz(geo), basis);
    dualContactOf<Coherent>(d, it.x, k, basis, geo.w, A, B, alpha);
  }
  if (__any(joints)) {
    flowStoresDone();
    for (uint32_t q = q0 + lane; q < end; q += lanes) {
      uint4 it = d.items[q];
      if ((it.w & (FLOW_OWN | 2u)) == (FLOW_OWN | 2u)) dualJoint<Coherent>(d, it.x, alpha);
    }
  }
}

/** Chunk `chunk` of a colour's bucket [start, start + count) at iteration iter: one wavefront,
 *  lanes threads per body, as the colour-by-colour sweep splits it. */
template <bool EARLY = false>
__device__ __forceinline__ void flowChunk(const Dev &d, uint32_t start, uint32_t count, uint32_t lanes, bool serial, uint32_t chunk,
                                          uint32_t iter, uint32_t polls) {
  uint32_t per = 64u / lanes, slot = threadIdx.x / lanes, lane = threadIdx.x % lanes;
  uint32_t k = chunk * per + slot;
  bool live = k < count;
  uint4 item = live ? d.colorItems[start + k] : make_uint4(0u, 0u, 0u, 0u);
  uint32_t i = item.x;
  bool enabled = live && item.w != 0u;
  Acc acc;
  float4 pos = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rot = pos;
  if (live) flowAccumulate(d, i, item.y, item.z, lane, lanes, d.p.alpha, iter, polls, acc, pos, rot);
  else accZero(acc);
  reduceLanes(acc, lanes);
  if (enabled && lane == 0) finishBody<Coherent>(d, i, acc, pos, rot);
  if (enabled) {
    uint32_t first = threadIdx.x - lane;
    pos = flowShfl(pos, first);
    rot = flowShfl(rot, first);
  }
  if (EARLY) {
    // One iteration per launch: the only readers this launch releases are later bodies of the same
    // iteration, which read this body's pose and none of the constraints it owns (a constraint's
    // owner is its later body, so theirs are their own). The owned duals then run off the chain.
    flowStoresDone();
    if (live && lane == 0) __hip_atomic_store(&d.flowDone[i], iter + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (live) flowDuals(d, i, item.y, item.z, lane, lanes, d.p.alpha, pos, rot, serial);
    return;
  }
  if (live) flowDuals(d, i, item.y, item.z, lane, lanes, d.p.alpha, pos, rot, serial);
  flowStoresDone();  // everything out before the counts that release the readers
  if (live && lane == 0) __hip_atomic_store(&d.flowDone[i], iter + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/** The sweep: iterations x (colours in use x their chunks). Chunk k of colour c goes to workgroup
 *  (g_c + k) mod gridDim.x, g_c the chunks of the colours before c: the same workgroup every
 *  iteration, so a body's updates follow each other in one instruction stream. One 64-thread
 *  workgroup per wavefront. Orphan joints' duals first. */
__global__ void __launch_bounds__(64) k_flow(Dev d, uint32_t polls) {
  const uint32_t W = gridDim.x, w = blockIdx.x;
  for (uint32_t j = w * 64u + threadIdx.x; j < d.nj; j += W * 64u)
    if (d.flowOrphan[j])
      for (uint32_t t = 0; t < d.p.iterations; t++) dualJoint<Coherent>(d, j, d.p.alpha);
  uint32_t colors = d.counters[C_NUM_COLORS];
  if (colors > d.p.colorCap) colors = d.p.colorCap;
  for (uint32_t it = 0; it < d.p.iterations; it++) {
    uint32_t g = 0;  // chunks of the colours before c
    for (uint32_t c = 0; c < colors; c++) {
      uint32_t start = d.colorStart[c], count = d.colorStart[c + 1] - start;
      uint32_t lanes = lanesFor(d.p, count), per = 64u / lanes;
      uint32_t chunks = (count + per - 1u) / per;
      bool serial = (d.counters[C_CLASHING + (c >> 5)] >> (c & 31u)) & 1u;
      for (uint32_t k = (w + W - g % W) % W; k < chunks; k += W) flowChunk(d, start, count, lanes, serial, k, it, polls);
      g += chunks;
    }
  }
}

/** The later of a constraint's bodies in the sweep's order (a fixed body counts as first). */
__device__ __forceinline__ uint32_t flowOwnerKey(const Dev &d, int a, int b) {
  uint32_t ka = a >= 0 && isDynamic(d, a) ? d.bodyKey[a] : 0u, kb = b >= 0 && isDynamic(d, b) ? d.bodyKey[b] : 0u;
  return ka > kb ? ka : kb;
}
/** The dataflow tail of one iteration, after the per-colour launches of colours [0, first): the
 *  colours from `first` as the dataflow sweep runs them (each body waiting only for its neighbours
 *  ordered before it in these colours, the duals it owns after its update), and beside them the
 *  duals of the constraints whose bodies all come before `first` (final when the launch starts) and
 *  of the joints no body walks. One launch replaces the trailing colours' launches and the dual's;
 *  the reads are the colour-by-colour sweep's. Workgroups without a chunk of the trailing colours
 *  take those duals (when every workgroup has a chunk, all take them after it): they read bodies
 *  and records nothing in the launch writes, through the caches as k_dual does. Completion counts
 *  accumulate over the substep's iterations (cleared before its first). */
__global__ void __launch_bounds__(64) k_flow_tail(Dev d, uint32_t first, uint32_t iter, uint32_t polls) {
  const uint32_t W = gridDim.x, w = blockIdx.x, firstKey = flowFirstKey(first);
  uint32_t colors = d.counters[C_NUM_COLORS];
  if (colors > d.p.colorCap) colors = d.p.colorCap;
  uint32_t g = 0;
  for (uint32_t c = first; c < colors; c++) {
    uint32_t start = d.colorStart[c], count = d.colorStart[c + 1] - start;
    uint32_t lanes = lanesFor(d.p, count), per = 64u / lanes;
    uint32_t chunks = (count + per - 1u) / per;
    bool serial = (d.counters[C_CLASHING + (c >> 5)] >> (c & 31u)) & 1u;
    for (uint32_t k = (w + W - g % W) % W; k < chunks; k += W) flowChunk<true>(d, start, count, lanes, serial, k, iter, polls);
    g += chunks;
  }
  // the other duals: on the workgroups the chunks left free (chunk k went to workgroup k mod W)
  uint32_t busy = g < W ? g : 0u;
  if (w < busy) return;
  uint32_t t0 = (w - busy) * 64u + threadIdx.x, nt = (W - busy) * 64u;
  for (uint32_t j = t0; j < d.nj; j += nt) {
    int4 in = d.jinfo[j];
    if (d.flowOrphan[j] || flowOwnerKey(d, in.y, in.z) < firstKey) dualJoint(d, j, d.p.alpha);
  }
  uint32_t points = d.counters[C_CONTACTS];
  if (points > d.conCap) points = d.conCap;
  for (uint32_t c = t0; c < points; c += nt) {
    uint32_t m = d.cman[c];
    if (m >= manifoldCount(d)) continue;
    uint4 ids = d.mIds[m];
    if (c < ids.z || c >= ids.z + (ids.w & 15u)) continue;
    if (flowOwnerKey(d, (int)ids.x, (int)ids.y) >= firstKey) continue;  // the later body's, in the sweep above
    float4 geo = d.mGeo[m];
    PairBody A = pairBody(d, ids.x), B = pairBody(d, ids.y);
    V basis[3];
    orthonormal(xyz(geo), basis);
    dualContact(d, c, basis, geo.w, A, B, d.p.alpha);
  }
}

/** The leading colours [0, first) of one iteration in one launch, workgroup w taking chunk w in
 *  (colour, chunk) order (one chunk a workgroup, the grid at least the chunks). A body waits only
 *  for its partners ordered before it, whose chunks have lower indices: workgroups of a launch
 *  start in index order on each XCD, so every chunk a waiting one needs has started and waits on
 *  none after it, and the lowest unfinished chunk can always run. What it reads and nothing in the
 *  launch writes comes through the caches: its own pose, its constraints' penalties and
 *  multipliers (the duals run in the tail launch after this one), partners ordered after it (they
 *  wait for it); the poses of partners ordered before it coherently, after their counts. Waits are
 *  bounded as the tail's. */
__global__ void __launch_bounds__(64) k_flow_lead(Dev d, uint32_t first, uint32_t iter, uint32_t polls) {
  uint32_t colors = d.counters[C_NUM_COLORS];
  if (colors > d.p.colorCap) colors = d.p.colorCap;
  uint32_t last = first < colors ? first : colors, w = blockIdx.x, g = 0;
  for (uint32_t c = 0; c < last; c++) {
    uint32_t start = d.colorStart[c], count = d.colorStart[c + 1] - start;
    uint32_t lanes = lanesFor(d.p, count), per = 64u / lanes;
    uint32_t chunks = (count + per - 1u) / per;
    if (w >= g + chunks) {
      g += chunks;
      continue;
    }
    uint32_t slot = threadIdx.x / lanes, lane = threadIdx.x % lanes;
    uint32_t k = (w - g) * per + slot;
    bool live = k < count;
    uint4 item = live ? d.colorItems[start + k] : make_uint4(0u, 0u, 0u, 0u);
    uint32_t i = item.x;
    bool enabled = live && item.w != 0u;
    Acc acc;
    float4 pos = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rot = pos;
    if (live) flowAccumulate<true>(d, i, item.y, item.z, lane, lanes, d.p.alpha, iter, polls, acc, pos, rot);
    else accZero(acc);
    reduceLanes(acc, lanes);
    if (enabled && lane == 0) finishBody<Coherent>(d, i, acc, pos, rot);
    flowStoresDone();
    if (live && lane == 0) __hip_atomic_store(&d.flowDone[i], iter + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
}

}  // namespace avbd
