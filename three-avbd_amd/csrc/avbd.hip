// Host side of the MI355X AVBD step: device storage, the step's launch sequence and the C ABI
// declared in include/avbd.h. Mirrors GpuSolver3D (src/avbd3d/gpu/solver.ts:504): construction
// (:722-851), step (:2112-2293), readbacks (:2297-2352), seedFrom (:1320), fixedColors (:560),
// uploadStatics (:1263) and its large-body rule (radii.ts:47).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <array>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#include <stdexcept>
#include "../../include/avbd.h"
#include "avbd_adjoint.h"
#include "avbd_flow.h"
#include "avbd_events.h"
#include "avbd_commands.h"
#include "avbd_plan.h"
#include "avbd_sdf.h"

using namespace avbd;

struct avbd_sdf;
/** The installed SDF bridge's part of a step's collision phase (defined with the bridge below). */
static void sdfEncode(avbd_sdf *g, const Dev &d);
/** Frees a bridge and its device memory (its solver's stream is synchronised first). */
static void sdfRelease(avbd_sdf *g);
/** The bridge's checks before a step (prepare, sdf-bridge.ts:154-160); false sets why. */
static bool sdfPrepare(const avbd_sdf *g, std::string &why);

// Staging of avbd_read_bodies_noalloc: file-scope device memory, one copy on every GPU the library
// runs on (held until the process ends), room for AVBD_READ_NOALLOC_MAX body records. Host code
// never passes the array's name to a kernel: that name is a host-side stand-in. Each device's
// address comes from hipGetSymbolAddress, taken once per device with that device current.
static __device__ __align__(256) float4 scratch[10u * AVBD_READ_NOALLOC_MAX];

namespace {

thread_local std::string g_create_error;

const int SCRATCH_DEVICES = 64;
std::mutex g_scratch_mutex;
float4 *g_scratch_ptr[SCRATCH_DEVICES] = {};
bool g_scratch_tried[SCRATCH_DEVICES] = {};

/** The address of `scratch` on device `dev` (the current device), looked up on its first use. */
float4 *scratchOn(int dev) {
  if (dev < 0 || dev >= SCRATCH_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_tried[dev]) {
    g_scratch_tried[dev] = true;
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_ptr[dev] = (float4 *)ptr;
    else (void)hipGetLastError();
  }
  return g_scratch_ptr[dev];
}

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

struct KernelTimer {
  bool on = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;  // kind, (start, stop)
};

enum Kind { K_PRIMAL = 0, K_DUAL = 1, K_NARROW = 2, K_BROAD = 3, K_SOLVE = 4, K_OTHER = 5, K_TAIL = 6 };

uint32_t pow2At(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return (uint32_t)p;
}
uint32_t blocksFor(uint64_t n, uint32_t threads, uint32_t cap = 16384) {
  uint64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (uint32_t)b;
}
/** k_primal's grid: enough workgroups to take any colour of up to n bodies in one pass,
 *  at lanesFor(size) threads a body. size x lanes is largest at the top of one of the lane rule's
 *  ranges (lanes are constant between its thresholds), so only those sizes and n are tried. A
 *  colour then spreads over as many wavefronts as it has bodies' threads: none walks two chunks
 *  one after the other while other SIMDs idle. */
uint32_t primalBlocks(const avbd_params &hp, uint32_t n) {
  const uint64_t tops[5] = {n, (uint64_t)hp.primal_lanes[0] - 1, (uint64_t)hp.primal_lanes[1] - 1, (uint64_t)hp.primal_lanes[2] - 1,
                            (uint64_t)hp.primal_lanes[3] - 1};
  uint64_t most = 1;
  for (uint64_t c : tops) {
    if (c < 1 || c > n) continue;
    uint64_t lanes = c >= hp.primal_lanes[0] ? 1 : c >= hp.primal_lanes[1] ? 2 : c >= hp.primal_lanes[2] ? 4 : c >= hp.primal_lanes[3] ? 8 : 16;
    most = std::max(most, c * lanes);
  }
  return blocksFor(most, PRIMAL_THREADS);
}
float finiteF(float x) { return x > 3.0e38f ? 3.0e38f : (x < -3.0e38f ? -3.0e38f : x); }

/** The k-th smallest (Wirth's selection, radii.ts:21-43), reordering a. */
double selectK(std::vector<double> &a, size_t k) {
  size_t lo = 0, hi = a.size() - 1;
  while (lo < hi) {
    double pivot = a[k];
    size_t i = lo, j = hi;
    do {
      while (a[i] < pivot) i++;
      while (pivot < a[j]) j--;
      if (i <= j) { std::swap(a[i], a[j]); i++; if (j == 0) break; j--; }
    } while (i <= j);
    if (j < k) lo = i;
    if (k < i) hi = j;
  }
  return a[k];
}
/** largestSmall (radii.ts:47-55): the largest radius that still sizes the grid cells. */
double largestSmall(std::vector<double> radii) {
  size_t n = radii.size();
  double median = n > 0 ? selectK(radii, n >> 1) : 1.0;
  double maxSmall = 0;
  for (double r : radii) if (r <= 2 * median) maxSmall = std::max(maxSmall, r);
  if (n > 64) maxSmall = std::max(maxSmall, selectK(radii, n - 64 - 1));
  return maxSmall;
}

}  // namespace

struct avbd_solver {
  hipStream_t stream = nullptr;
  avbd_params hp{};
  uint32_t n = 0, nj = 0, bodyCap = 0, jointCap = 0, pairCap = 0, manCap = 0, conCap = 0, hullCap = 0;
  uint32_t tableSize = 0, hashSize = 0, colorGroups = 0, adjCap = 0;
  bool adjacencyHeld = false;  // the adjacency buffers hold the last substep's adjacency
  float cellSize = 1;
  float fineCell = 1, fineRadius = 0;  // the grid's fine level (uploadStatics, solver.ts:1276-1290)
  bool hasFine = false, coarseMovers = false, anyFixedSmall = false;
  std::vector<uint32_t> large;          // large bodies, grouped by world
  std::vector<uint32_t> worldStartHost;  // W + 1 body offsets
  uint32_t *worldOf = nullptr, *worldStart = nullptr, *largeStart = nullptr;
  float *largeRadiusW = nullptr;
  struct NC { uint32_t hi, lo, slot; };
  std::vector<NC> ignorePairs;  // IgnoreCollision entries
  std::vector<int32_t> jinfoHost;
  uint32_t noCollideCount = 0;
  int parity = 0;
  uint32_t stepCount = 0;
  bool fixedColors = false;
  uint32_t fixedCap = 1;
  bool manifoldsValid = false;
  bool anyDynamicLarge = false;
  bool destroyed = false;
  uint32_t coopBlocks = 0;  // grid of the cooperative solve kernel; 0: per-colour launches
  uint32_t *barrier = nullptr;  // the solve kernel's grid barrier: arrivals over one launch
  uint32_t barrierPolls = 1u << 22;  // a barrier wait gives up after this many polls
  uint32_t barrierGiveUps = 0;       // give-ups seen by the host since avbd_read_counters last reported them
  bool barrierFailed = false;        // seen during the current avbd_step/avbd_steps call
  bool coopFlow = false;        // the cooperative launch is the dataflow sweep (k_flow), not k_solve
  bool coopTail = false;        // per-colour launches up to flowTailStart, then one dataflow tail (k_flow_tail) per iteration
  bool leadFlow = false;        // with coopTail: the colours before the tail as one dataflow launch (k_flow_lead) per iteration
  uint32_t flowTailStart = MAX_COLORS;  // chosen at each adaptive-colouring readback
  uint32_t coopThreads = 512;   // threads per workgroup of the solve kernel
  uint32_t coopPerCU = 1;       // its workgroups per CU
  uint32_t coopMaxBlocks = 0;   // co-resident limit for coopThreads (0: no cooperative launch here)
  uint32_t numCUs = 0;
  uint32_t hullBlocks = 4;  // k_narrow_hull's grid: its co-resident workgroups (a wavefront per pair, grid-strided)
  bool coopSupported = false;
  uint32_t shrinkVotes = 0;
  uint32_t stepsSinceAdapt = 0;
  // per-colour schedule: colours [tailStart, cap) run fused in one workgroup (k_primal_tail); 64 = none.
  // tailAuto: chosen from the colour buckets at each adaptive-colouring readback
  uint32_t tailStart = MAX_COLORS;
  bool tailAuto = true;
  std::string err;
  std::vector<DevBuf> bufs;
  // device storage
  Bodies b{};
  float4 *refPos = nullptr, *refRot = nullptr;
  uint2 *filters = nullptr;
  float4 *joints = nullptr;
  int4 *jinfo = nullptr;
  uint32_t *flowOrphan = nullptr;  // per joint: its dual runs before the dataflow sweep (no body walks it)
  uint32_t *bodyKey = nullptr, *flowDone = nullptr;  // the dataflow sweep's order keys and counts
  uint32_t *statics = nullptr;
  uint32_t staticsCap = 0;
  uint4 *hulls = nullptr;
  uint32_t *cellStart = nullptr, *cellSorted = nullptr;
  uint32_t *sortKey[2] = {nullptr, nullptr}, *sortIdx[2] = {nullptr, nullptr}, *sortHist = nullptr;
  int4 *cellOf = nullptr;
  float4 *sortedPos = nullptr;
  int4 *sortedKey = nullptr;
  uint2 *pairs = nullptr;
  uint2 *hullPairs = nullptr;  // the narrowphase's deferred hull pairs
  uint2 *boxPairs = nullptr;   // and box-box pairs
  float4 *hullOut = nullptr;   // and their contacts for the emit pass
  uint4 *mIds[2] = {nullptr, nullptr};
  float4 *mGeo[2] = {nullptr, nullptr};
  Contact *con[2] = {nullptr, nullptr};
  uint32_t *cman = nullptr;  // per contact of the step, its manifold
  uint32_t *table = nullptr;
  uint32_t *adjCount = nullptr, *adjStart = nullptr, *adjFill = nullptr, *adjList = nullptr;
  uint32_t *adjTmp = nullptr, *adjOwner = nullptr;  // the fill's unsorted entries and their bodies
  uint2 *adjKeys = nullptr;                          // and their sort keys
  int *adjOther = nullptr;                           // per sorted entry, the dynamic neighbour or -1
  uint32_t *prio = nullptr;                          // per body, its colouring priority
  uint4 *adjIds = nullptr, *colorItems = nullptr, *items = nullptr;
  uint32_t *itemStart = nullptr, itemCap = 0;
  float4 *adjGeo = nullptr;
  uint32_t *colA = nullptr, *colB = nullptr, *colorStart = nullptr, *colorBodies = nullptr, *colorRank = nullptr, *hist = nullptr;
  uint32_t *fixedDev = nullptr;
  uint32_t *counters = nullptr;
  uint32_t *scanSums = nullptr;
  float4 *staging = nullptr;
  size_t stagingBytes = 0;
  // host mirrors for the scene edits (World.flush, world.ts:1861-1995): per body its bounding
  // radius and mass (uploadStatics reads them), its lifecycle bits (patchBody's marker), its
  // gravity scale and damping rates (the aux factors follow dt); the released joint slots
  std::vector<float> radiusHost, massHost, sizeHost;  // sizeHost: full widths xyz per body (torque arms)
  std::vector<uint8_t> lifeHost;
  std::vector<float> auxScale, auxLin, auxAng;
  float4 *aux = nullptr;  // per body: scale, exp(-lin dt), exp(-ang dt), damped (null: no aux yet)
  std::vector<uint32_t> freeSlots;
  bool staticsDirty = false;  // radii, masses or joint pairs changed: uploadStatics before the next step
  bool seeding = false;       // avbd_seed's body write (colours kept)
  int device = 0;                    // the GPU the solver was created on (current at avbd_create)
  float4 *readScratch = nullptr;     // that GPU's `scratch` (avbd_read_bodies_noalloc)
  KernelTimer timer;
  // ---- adjoint: checkpoints of the forward rollout, the recorded segment, f64 adjoints ----
  struct Checkpoint {
    float4 *streams = nullptr;  // 12 x bodyCap: the 10 body streams, refPos, refRot
    uint32_t *colA = nullptr, *counters = nullptr;
    uint4 *mIds = nullptr;
    float4 *mGeo = nullptr;
    Contact *con = nullptr;
    uint32_t stepCount = 0, shrinkVotes = 0, stepsSinceAdapt = 0;
    int parity = 0;
    bool manifoldsValid = false;
    avbd_params hp{};
  };
  /** Device and pinned host allocations of one part of the adjoint's state (freed together). */
  struct Mem { std::vector<void *> dev, host; };
  std::vector<Checkpoint> ckpts;   // segment starts, then the rollout's end (the first ckptCount)
  uint32_t ckptCount = 0;
  Mem ckptMem, adjMem;             // checkpoints + per-contact sources + remap table; AdjBuf
  // contact events (avbd_watch_*): the per-body watch words are kept, the sets and events are
  // replaced by each avbd_watch_create
  WatchBuf watchBuf{};
  Mem watchMem;
  uint32_t *watchWords = nullptr;
  uint32_t watchFlip = 0;
  int *src = nullptr;              // per contact warm-start source while recording
  // cap: colours each recorded step swept; hostStarts/hostCounts: pinned copies of the tape's
  // colour buckets and counts, filled on the solver stream after a segment is recorded
  struct TapeBuf { Tape t{}; uint32_t slots = 0; std::vector<uint32_t> cap; uint32_t *hostStarts = nullptr, *hostCounts = nullptr; Mem mem; };
  TapeBuf tapes[2];                // a segment recorded while the previous one is reversed
  TapeBuf *rec = nullptr;          // when set, steps record into slot recSlot
  uint32_t recSlot = 0;
  hipStream_t revStream = nullptr; // the reverse sweep's stream (the re-runs use the solver's)
  hipEvent_t recorded[2] = {nullptr, nullptr}, reversed[2] = {nullptr, nullptr};  // per tape buffer
  uint32_t *remapTable = nullptr;
  AdjBuf adj{};
  uint32_t adjConCap = 0;
  uint32_t fwdSteps = 0, fwdEvery = 0, fwdEndStep = 0, fwdMaxContacts = 0, fwdMaxManifolds = 0;
  bool fwdValid = false;
  bool noGrow = false;             // set while an adjoint rollout records or re-runs: storage stays
  bool fwdTail = false;            // the rollout's last segment is already on tape (recorded by the forward)
  std::string fwdWhy = "no forward rollout recorded";

  template <class T> T *alloc(size_t count) {
    DevBuf d;
    d.bytes = std::max<size_t>(count * sizeof(T), 256);
    if (hipMalloc(&d.p, d.bytes) != hipSuccess) throw std::runtime_error("hipMalloc failed");
    hipMemsetAsync(d.p, 0, d.bytes, stream);
    bufs.push_back(d);
    return (T *)d.p;
  }
  /** Frees one buffer made by alloc(). */
  void release(void *p) {
    for (size_t k = 0; k < bufs.size(); k++)
      if (bufs[k].p == p) { hipFree(p); bufs.erase(bufs.begin() + (long)k); return; }
  }
  /** Replaces buffer `p` (of `oldCount` T) by one of `count`, keeping the first min of them. */
  template <class T> T *regrow(T *p, size_t oldCount, size_t count, bool keep) {
    T *q = alloc<T>(count);
    if (keep && p) hipMemcpyAsync(q, p, sizeof(T) * std::min(oldCount, count), hipMemcpyDeviceToDevice, stream);
    if (p) { hipStreamSynchronize(stream); release(p); }
    return q;
  }
  avbd_sdf *backend = nullptr;      // the installed collision backend (setCollisionBackend)
  std::vector<avbd_sdf *> sdfs;     // every SDF bridge made for this solver (they end with it)
  ~avbd_solver() {
    for (avbd_sdf *g : sdfs) sdfRelease(g);
    sdfs.clear();
    backend = nullptr;
    dropGraphs();
    freeAdjoint();
    freeMem(watchMem);
    if (revStream) hipStreamDestroy(revStream);
    for (auto e : recorded) if (e) hipEventDestroy(e);
    for (auto e : reversed) if (e) hipEventDestroy(e);
    for (auto &d : bufs) hipFree(d.p);
    if (staging) hipFree(staging);
    for (auto &m : timer.marks) { hipEventDestroy(m.second.first); hipEventDestroy(m.second.second); }
    if (stream) hipStreamDestroy(stream);
  }

  /** Sets the solve schedule; a cooperative one needs its grid co-resident on this device. */
  /** The cooperative solve kernel for a workgroup size (nullptr: not one of 64 ... 1024). */
  static const void *solveKernel(uint32_t threads) {
    switch (threads) {
      case 64: return (const void *)k_solve<64>;
      case 128: return (const void *)k_solve<128>;
      case 256: return (const void *)k_solve<256>;
      case 512: return (const void *)k_solve<512>;
      case 1024: return (const void *)k_solve<1024>;
      default: return nullptr;
    }
  }
  int applySchedule(const avbd_schedule &sch) {
    uint32_t t = sch.threads;
    if (sch.cooperative == 2 || sch.cooperative == 3 || sch.cooperative == 4) {
      // the dataflow sweep (or tail): one wavefront per workgroup, all co-resident
      if (t != 64 || sch.blocks_per_cu < 1) return AVBD_ERR_INVALID;
      int perCU = 0;
      (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, sch.cooperative == 2 ? (const void *)k_flow : (const void *)k_flow_tail, 64, 0);
      if (!coopSupported || perCU < (int)sch.blocks_per_cu) return AVBD_ERR_INVALID;
      coopFlow = sch.cooperative == 2;
      coopTail = sch.cooperative >= 3;
      leadFlow = sch.cooperative == 4;
      coopThreads = 64;
      coopPerCU = sch.blocks_per_cu;
      coopBlocks = coopPerCU * numCUs;
    } else if (sch.cooperative == 1) {
      if (!solveKernel(t) || sch.blocks_per_cu < 1) return AVBD_ERR_INVALID;
      int perCU = 0;
      (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, solveKernel(t), (int)t, 0);
      if (!coopSupported || perCU < (int)sch.blocks_per_cu) return AVBD_ERR_INVALID;
      coopFlow = false;
      coopTail = false;
      leadFlow = false;
      coopThreads = t;
      coopPerCU = sch.blocks_per_cu;
      coopBlocks = coopPerCU * numCUs;
    } else if (sch.cooperative == 0) {
      coopBlocks = 0;
      coopFlow = false;
      coopTail = false;
      leadFlow = false;
    } else {
      return AVBD_ERR_INVALID;
    }
    return AVBD_OK;
  }
  avbd_schedule schedule() const {
    return avbd_schedule{coopBlocks ? (coopFlow ? 2u : (coopTail ? (leadFlow ? 4u : 3u) : 1u)) : 0u, coopThreads, coopPerCU};
  }
  /** The fused tail starts where every later colour held at most TAIL_BODIES bodies at this
   *  readback (the cap's empty spares included). One workgroup runs on one CU, so only small
   *  colours gain: a 55-box pyramid sweeps in one launch per iteration (1.12 -> 0.61 ms per
   *  step), while an 80-body colour already runs slower there than spread over the chip. A tail
   *  that has grown since only runs slower; the results are the same. */
  void chooseTail(const uint32_t *starts) {
#ifndef AVBD_TAIL_BODIES
#define AVBD_TAIL_BODIES 32
#endif
    const uint32_t TAIL_BODIES = AVBD_TAIL_BODIES;
    uint32_t cap = fixedColors ? fixedCap : hp.color_cap, c = std::min(cap, MAX_COLORS);
    while (c > 0 && starts[c] - starts[c - 1] <= TAIL_BODIES) c--;
    tailStart = c;
  }
  /** The dataflow tail starts where every later colour held at most FLOW_TAIL_BODIES bodies at this
   *  readback: a colour that small costs a launch for little more than one body's dependent chain,
   *  which the tail overlaps with the next colour's and with the dual. */
  void chooseFlowTail(const uint32_t *starts) {
    const uint32_t FLOW_TAIL_BODIES = 2048;
    uint32_t cap = fixedColors ? fixedCap : hp.color_cap, c = std::min(cap, MAX_COLORS);
    while (c > 0 && starts[c] - starts[c - 1] <= FLOW_TAIL_BODIES) c--;
    flowTailStart = c;
  }
  /** adaptColors (solver.ts:2379-2407). */
  void adapt(const avbd_counters &c) {
    if (fixedColors) return;
    const uint32_t SPARE = 3;
    uint32_t used = c.colors;
    if (c.clashes > 0) hp.color_rounds = std::min(32u, hp.color_rounds * 2);
    if (c.clashes > 0 || hp.color_cap < used + 2) {
      hp.color_cap = std::min(MAX_COLORS, std::max(used + SPARE + 2, hp.color_cap + 4));
      shrinkVotes = 0;
    } else if (used + SPARE + 2 < hp.color_cap || hp.color_rounds > 4) {
      if (++shrinkVotes >= 3) {
        hp.color_cap = std::min(MAX_COLORS, std::max(1u, used + SPARE));
        hp.color_rounds = std::max(4u, hp.color_rounds - 4);
        shrinkVotes = 0;
      }
    } else {
      shrinkVotes = 0;
    }
  }

  /** adapt's capacity half (solver.ts:2366-2374, GROW_AT 0.8, GROW_TO 2): a buffer whose demand
   *  overflowed or passed 0.8 of its capacity grows to max(2 demand, 1.25 capacity); manifolds
   *  never past pairs. Not with fixed storage, fixed colourings, or inside an adjoint rollout. */
  void adaptStorage(const avbd_counters &c) {
    if ((hp.flags & AVBD_FLAG_FIXED_STORAGE) || fixedColors || noGrow) return;
    const uint64_t LIMIT = 1ull << 30;
    auto want = [](uint32_t demand, uint32_t cap, uint32_t bit, uint32_t overflow) -> uint64_t {
      if ((overflow & bit) || (double)demand > 0.8 * cap) return std::max<uint64_t>(2ull * demand, (uint64_t)std::ceil(1.25 * cap));
      return cap;
    };
    uint64_t con = std::min(want(c.contacts, conCap, 2, c.overflow), LIMIT);
    uint64_t pr = std::min(want(c.pairs, pairCap, 1, c.overflow), LIMIT);
    uint64_t man = std::min(std::min(want(c.manifolds, manCap, 4, c.overflow), pr), LIMIT);
    if (con > conCap || pr > pairCap || man > manCap) growStorage((uint32_t)std::max<uint64_t>(con, conCap), (uint32_t)std::max<uint64_t>(pr, pairCap),
                                                                 (uint32_t)std::max<uint64_t>(man, manCap));
  }
  /** allocateContacts (solver.ts:1006-1040) when growing: pairs, manifolds, contacts and what is
   *  sized by them. Both ping-pong manifold and contact buffers are copied over and the counters
   *  left alone, so warm starts survive; the adjoint's buffers and any recorded rollout are dropped. */
  void growStorage(uint32_t nCon, uint32_t nPairs, uint32_t nMan) {
    hipStreamSynchronize(stream);
    dropGraphs();
    freeAdjoint();
    fwdValid = false;
    fwdWhy = "the solver's storage grew after avbd_forward";
    pairs = regrow(pairs, pairCap, nPairs, false);
    if (hullPairs) hullPairs = regrow(hullPairs, pairCap, nPairs, false);
    if (boxPairs) boxPairs = regrow(boxPairs, pairCap, nPairs, false);
    if (hullOut) hullOut = regrow(hullOut, (size_t)HULL_OUT * pairCap, (size_t)HULL_OUT * nPairs, false);
    for (int k = 0; k < 2; k++) {
      mIds[k] = regrow(mIds[k], manCap, nMan, true);
      mGeo[k] = regrow(mGeo[k], manCap, nMan, true);
      con[k] = regrow(con[k], conCap, nCon, true);
    }
    cman = regrow(cman, conCap, nCon, false);
    uint32_t hs = pow2At(2ull * nMan), ac = 2u * (jointCap + nMan);
    table = regrow(table, hashSize, hs, false);
    adjList = regrow(adjList, adjCap, ac, false);
    adjTmp = regrow(adjTmp, adjCap, ac, false);
    adjOwner = regrow(adjOwner, adjCap, ac, false);
    adjKeys = regrow(adjKeys, adjCap, ac, false);
    adjOther = regrow(adjOther, adjCap, ac, false);
    adjIds = regrow(adjIds, adjCap, ac, false);
    adjGeo = regrow(adjGeo, adjCap, ac, false);
    uint32_t ic = itemsFor(nCon, jointCap);
    items = regrow(items, itemCap, ic, false);
    itemCap = ic;
    pairCap = nPairs; manCap = nMan; conCap = nCon; hashSize = hs; adjCap = ac;
    manifoldsValid = false;
    hipStreamSynchronize(stream);
  }
  Params params() const {
    Params p;
    p.dt = hp.dt; p.gravity = hp.gravity; p.betaLin = hp.beta_lin; p.betaAng = hp.beta_ang; p.gamma = hp.gamma; p.alpha = hp.alpha;
    p.margin = hp.contact_margin; p.upx = hp.up[0]; p.upy = hp.up[1]; p.upz = hp.up[2];
    p.flags = hp.flags; p.iterations = hp.iterations; p.colorCap = fixedColors ? fixedCap : hp.color_cap; p.rounds = hp.color_rounds;
    p.lanes1 = hp.primal_lanes[0]; p.lanes2 = hp.primal_lanes[1]; p.lanes4 = hp.primal_lanes[2];
    p.lanes8 = hp.primal_lanes[3];
    return p;
  }
  Dev dev() const {
    Dev d;
    std::memset(&d, 0, sizeof d);  // every byte defined: a substep's graph is keyed by these bytes
    d.n = n; d.nj = nj; d.b = b; d.refPos = refPos; d.refRot = refRot; d.filters = filters; d.joints = joints; d.jinfo = jinfo;
    d.statics = statics; d.largeCount = (uint32_t)large.size(); d.noCollideCount = noCollideCount; d.hulls = hulls;
    d.ncFirst = statics + large.size() + 3ull * noCollideCount;
    d.worldOf = worldOf; d.worldStart = worldStart; d.largeStart = largeStart; d.largeRadius = largeRadiusW;
    d.worlds = (uint32_t)worldStartHost.size() - 1;
    d.cellSize = cellSize; d.tableMask = tableSize - 1;
    d.fineCell = fineCell; d.fineRadius = fineRadius; d.hasFine = hasFine ? 1u : 0u; d.coarseMovers = coarseMovers ? 1u : 0u;
    d.cellStart = cellStart; d.cellSorted = cellSorted;
    d.sortKey[0] = sortKey[0]; d.sortKey[1] = sortKey[1]; d.sortIdx[0] = sortIdx[0]; d.sortIdx[1] = sortIdx[1]; d.sortHist = sortHist; d.cellOf = cellOf; d.sortedPos = sortedPos; d.sortedKey = sortedKey;
    d.pairs = pairs; d.pairCap = pairCap; d.hullPairs = hullPairs; d.hullOut = hullOut; d.boxPairs = boxPairs;
    d.mIds = mIds[parity]; d.pmIds = mIds[1 - parity]; d.mGeo = mGeo[parity]; d.pmGeo = mGeo[1 - parity];
    d.con = con[parity]; d.pcon = con[1 - parity]; d.cman = cman; d.manCap = manCap; d.conCap = conCap;
    d.table = table; d.hashMask = hashSize - 1;
    d.adjCount = adjCount; d.adjStart = adjStart; d.adjFill = adjFill; d.adjList = adjList; d.adjCap = adjCap;
    d.adjTmp = adjTmp; d.adjOwner = adjOwner; d.adjKeys = adjKeys; d.adjOther = adjOther; d.prio = prio;
    d.adjIds = adjIds; d.adjGeo = adjGeo; d.colorItems = colorItems; d.itemStart = itemStart; d.items = items; d.itemCap = itemCap;
    d.colA = colA; d.colB = colB; d.colorStart = colorStart; d.colorBodies = colorBodies; d.colorRank = colorRank; d.hist = hist;
    d.colorGroups = colorGroups; d.counters = counters; d.p = params();
    d.bodyKey = bodyKey; d.flowDone = flowDone; d.flowOrphan = flowOrphan;
    d.aux = aux;
    d.src = rec ? src : nullptr;
    d.bits = nullptr;
    return d;
  }
  hipEvent_t mark(int kind) {
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, stream);
    timer.marks.push_back({kind, {e0, e1}});
    return e1;
  }
  template <class F> void timed(int kind, F launch) {
    if (!timer.on) { launch(); return; }
    hipEvent_t e1 = mark(kind);
    launch();
    hipEventRecord(e1, stream);
  }
  /** Blocks for a grid-strided kernel over up to `count` items (a capacity; the live count is on
   *  the device): enough to cover them, at most what the chip holds at once (2048 threads per
   *  CU), so a launch over a capacity dispatches no waves that find nothing to do. */
  uint32_t strideBlocks(uint64_t count, uint32_t threads) const {
    return blocksFor(count, threads, std::max(numCUs, 1u) * (2048u / threads));
  }
  void fill(uint32_t *p, uint32_t count, uint32_t v, const uint32_t *skip = nullptr) { k_fill<<<blocksFor(count, 256), 256, 0, stream>>>(p, count, v, skip); }
  /** Exclusive scan of in[0, count) into out (in may equal out); nothing when skip is given and its
   *  bit 0 is set (read on the device). */
  void scan(const uint32_t *in, uint32_t *out, uint32_t count, const uint32_t *skip = nullptr) {
    uint32_t nb = (count + 1023) / 1024;
    if (nb == 0) return;
    k_scan_local<<<nb, 256, 0, stream>>>(in, out, scanSums, count, skip);
    if (nb > 1) k_scan_add<<<nb, 256, 0, stream>>>(out, scanSums, count, skip);
  }

  /** uploadStatics (solver.ts:1263-1305): cell size, large list and sorted no-collide list. Each
   *  world classifies its own bodies (largestSmall, radii.ts:47) and lists its large ones. */
  void uploadStatics(const std::vector<float> &radius, const std::vector<float> &mass) {
    uint32_t W = (uint32_t)worldStartHost.size() - 1;
    std::vector<float> thr(W, 1e30f);
    std::vector<uint32_t> lstart(W + 1, 0);
    double maxCell = 0;
    large.clear();
    anyDynamicLarge = false;
    for (uint32_t w = 0; w < W; w++) {
      uint32_t b0 = worldStartHost[w], b1 = worldStartHost[w + 1];
      lstart[w] = (uint32_t)large.size();
      if (b1 <= b0) continue;
      std::vector<double> r(radius.begin() + b0, radius.begin() + b1);
      double maxSmall = largestSmall(r);
      double minLarge = INFINITY;
      for (double x : r) if (x > maxSmall) minLarge = std::min(minLarge, x);
      double threshold = std::isinf(minLarge) ? maxSmall * 1.5 : (maxSmall + minLarge) / 2;
      thr[w] = (float)threshold;
      maxCell = std::max(maxCell, 2 * maxSmall * (1 + 1e-4));
      for (uint32_t i = b0; i < b1; i++)
        if (radius[i] > thr[w]) { large.push_back(i); if (mass[i] > 0) anyDynamicLarge = true; }
    }
    lstart[W] = (uint32_t)large.size();
    cellSize = (float)std::max(maxCell, 1e-3);
    // A fine level for small moving bodies much smaller than the small fixed ones (held slabs must
    // not force them into slab-sized cells), as the reference's uploadStatics picks it
    // (solver.ts:1274-1290); one world only (batched worlds share one cell size).
    hasFine = coarseMovers = false;
    anyFixedSmall = false;
    for (uint32_t w = 0; w < W; w++)
      for (uint32_t i = worldStartHost[w]; i < worldStartHost[w + 1]; i++)
        if (!(mass[i] > 0) && radius[i] <= thr[w]) anyFixedSmall = true;
    fineCell = cellSize;
    fineRadius = W ? thr[0] : 0.0f;
    if (W == 1 && n > 0) {
      std::vector<double> r(radius.begin(), radius.begin() + n), moving;
      for (uint32_t i = 0; i < n; i++) if (mass[i] > 0) moving.push_back(radius[i]);
      double maxSmall = largestSmall(r);
      double target = moving.empty() ? maxSmall : std::min(maxSmall, 2.0 * largestSmall(moving));
      double fine = 0;
      for (double x : r) if (x <= target) fine = std::max(fine, x);
      double fc = cellSize, fr = thr[0];
      if (fine > 0 && fine * 2 < maxSmall) {
        double next = INFINITY;
        for (double x : r) if (x > fine) next = std::min(next, x);
        fc = std::max(2 * fine * (1 + 1e-4), 1e-3);
        fr = (fine + next) / 2;
      }
      fineCell = (float)fc;
      fineRadius = (float)fr;
      hasFine = fineCell < cellSize;
      for (uint32_t i = 0; i < n && hasFine; i++)
        if (mass[i] > 0 && radius[i] > fineRadius && radius[i] <= thr[0]) { coarseMovers = true; break; }
    }
    hipMemcpyAsync(largeRadiusW, thr.data(), W * 4ull, hipMemcpyHostToDevice, stream);
    hipMemcpyAsync(largeStart, lstart.data(), (W + 1) * 4ull, hipMemcpyHostToDevice, stream);
    std::vector<NC> nc = ignorePairs;
    for (uint32_t j = 0; j < nj; j++) {
      int a = jinfoHost[4 * j + 1], bb = jinfoHost[4 * j + 2];
      if (jinfoHost[4 * j] != (int)T_NONE && a >= 0) nc.push_back({(uint32_t)std::max(a, bb), (uint32_t)std::min(a, bb), j});
    }
    std::stable_sort(nc.begin(), nc.end(), [](const NC &x, const NC &y) { return x.hi < y.hi || (x.hi == y.hi && x.lo < y.lo); });
    noCollideCount = (uint32_t)nc.size();
    std::vector<uint32_t> words(large.begin(), large.end());
    for (auto &e : nc) { words.push_back(e.hi); words.push_back(e.lo); words.push_back(e.slot); }
    // per body, the first no-collide entry whose higher body it is (entries are sorted by hi)
    for (uint32_t body = 0, k = 0; body <= n; body++) {
      while (k < nc.size() && nc[k].hi < body) k++;
      words.push_back(k);
    }
    if (words.size() > staticsCap) {
      staticsCap = (uint32_t)words.size() * 2 + 16;
      statics = alloc<uint32_t>(staticsCap);
    }
    if (!words.empty()) hipMemcpyAsync(statics, words.data(), words.size() * 4, hipMemcpyHostToDevice, stream);
  }

  void collision(bool reuseManifolds, bool mayKeepAdjacency) {
    Dev d = dev();
    k_begin<<<1, 1, 0, stream>>>(d, reuseManifolds ? 1 : 0, mayKeepAdjacency ? 1 : 0);
    if (backend) { sdfEncode(backend, d); return; }  // the backend replaces refs, broadphase and narrowphase
    if (hp.flags & FLAG_REUSE_CONTACTS) k_update_refs<<<blocksFor(n, 256), 256, 0, stream>>>(d);
    if (reuseManifolds) {
      timed(K_NARROW, [&] { k_reuse_manifolds<<<strideBlocks(manCap, 256), 256, 0, stream>>>(d); });
      return;
    }
    fill(table, hashSize, 0);
    k_hash_insert<<<strideBlocks(manCap, 256), 256, 0, stream>>>(d);
    timed(K_BROAD, [&] {
      sortCells(d);
      k_find_pairs<<<blocksFor((uint64_t)n * PROBE_LANES, 64), 64, 0, stream>>>(d);
      if (anyDynamicLarge || anyFixedSmall) k_find_fixed<<<blocksFor(n, 256), 256, 0, stream>>>(d);
    });
    timed(K_NARROW, [&] {
      if (hullPairs) k_narrowphase<false><<<strideBlocks(pairCap, 1024), 1024, 0, stream>>>(d);
      else k_narrowphase<true><<<strideBlocks(pairCap, 1024), 1024, 0, stream>>>(d);
      if (boxPairs) k_narrow_box<<<strideBlocks(pairCap, 256), 256, 0, stream>>>(d);
      if (hullPairs) {
        k_narrow_hull<<<blocksFor(64ull * pairCap, 256, hullBlocks), 256, 0, stream>>>(d);
        k_hull_emit<<<strideBlocks(pairCap, 1024), 1024, 0, stream>>>(d);
      }
    });
  }
  /** The grid's cell order by an LSD radix sort of the bodies' cell hashes (the hash bits plus one
   *  for bodies outside the grid, in passes of at most 10 bits), then the buckets' starts and the
   *  sorted streams. */
  void sortCells(Dev &d) {
    uint32_t keyBits = 1;
    while ((1ull << keyBits) <= tableSize) keyBits++;  // keys 0 .. tableSize
    uint32_t blocks = (n + 255) / 256, maxBits = 10;
    // the histogram's scan takes at most 4096 x 1024 entries (4096 block totals)
    while (maxBits > 1 && ((uint64_t)blocks << maxBits) + 1 > 4096ull * 1024ull) maxBits--;
    uint32_t passes = (keyBits + maxBits - 1) / maxBits, bits = (keyBits + passes - 1) / passes;
    k_grid_keys<<<blocksFor(n, 256), 256, 0, stream>>>(d);
    uint32_t src = 0;
    for (uint32_t p = 0, shift = 0; p < passes && n; p++, shift += bits) {
      uint32_t b = std::min(bits, keyBits - shift);
      k_radix<false><<<blocks, 256, 0, stream>>>(sortKey[src], sortIdx[src], sortKey[1 - src], sortIdx[1 - src], sortHist, n, shift, b);
      scan(sortHist, sortHist, (blocks << b) + 1);
      k_radix<true><<<blocks, 256, 0, stream>>>(sortKey[src], sortIdx[src], sortKey[1 - src], sortIdx[1 - src], sortHist, n, shift, b);
      src = 1 - src;
    }
    k_cell_starts<<<blocksFor(tableSize + 1ull, 256), 256, 0, stream>>>(d, sortKey[src]);
    k_grid_place<<<blocksFor(n, 256), 256, 0, stream>>>(d, sortKey[src], sortIdx[src]);
  }
  void topology() {
    Dev d = dev();
    // a substep whose reuse kept every manifold keeps the adjacency too (C_TOPO bit 0, set on the
    // device): these launches then return at once
    const uint32_t *keep = counters + C_TOPO;
    fill(adjCount, n + 1, 0, keep);
    fill(adjFill, n, 0, keep);
    k_degree<<<strideBlocks(nj + manCap, 256), 256, 0, stream>>>(d);
    scan(adjCount, adjStart, n + 1, keep);
    k_fill_adj<<<strideBlocks(nj + manCap, 256), 256, 0, stream>>>(d);
    k_sort_adj<<<strideBlocks(adjCap, 256), 256, 0, stream>>>(d);
    k_item_count<<<blocksFor(n + 1, 256), 256, 0, stream>>>(d);
    scan(itemStart, itemStart, n + 1, keep);
    k_item_fill<<<blocksFor(n, 256), 256, 0, stream>>>(d);
    if (fixedColors) {
      hipMemcpyAsync(colA, fixedDev, n * 4, hipMemcpyDeviceToDevice, stream);
    } else {
      k_color_compact<<<blocksFor(n, 64), 64, 0, stream>>>(d);
      k_color_mark<<<blocksFor(n, 64), 64, 0, stream>>>(d);
      for (uint32_t r = 0; r < hp.color_rounds; r++) {
        if (r % 2 == 0) k_color_round<<<blocksFor(n, 64), 64, 0, stream>>>(d, colA, colB);
        else k_color_round<<<blocksFor(n, 64), 64, 0, stream>>>(d, colB, colA);
      }
    }
    uint32_t groups = (n + 255) / 256;
    if (groups == 0) groups = 1;
    k_color_count<<<groups, 256, 0, stream>>>(d);
    scan(hist, hist, 64 * colorGroups + 1);
    k_color_starts<<<1, 64, 0, stream>>>(d);
    k_color_scatter<<<groups, 256, 0, stream>>>(d);
    if (rec) {
      uint32_t cap = fixedColors ? fixedCap : hp.color_cap;
      k_tape_topology<<<blocksFor(std::max({conCap, manCap, adjCap, n + 1}), 256), 256, 0, stream>>>(d, rec->t, recSlot);
      k_tape_colors<<<blocksFor(n, 256), 256, 0, stream>>>(d, rec->t, recSlot, cap);
      uint32_t *off = rec->t.ptOff + (size_t)recSlot * (n + 1);
      k_tape_point_count<<<blocksFor(n + 1, 256), 256, 0, stream>>>(d, rec->t, recSlot, cap);
      scan(off, off, n + 1);
      k_tape_point_fill<<<blocksFor(n, 256), 256, 0, stream>>>(d, rec->t, recSlot, cap);
      rec->cap[recSlot] = cap;
    }
  }
  void solve() {
    Dev d = dev();
    if (coopBlocks && coopFlow && !rec) {
      // the dataflow sweep; prep writes only the walk items' flags and the orphan joints' marks, so
      // a launch refused here leaves the per-colour sweep below everything it needs
      if (nj) k_warm_joints<<<blocksFor(nj, 256), 256, 0, stream>>>(d);
      k_warm_bodies<<<blocksFor(n, 256), 256, 0, stream>>>(d);
      k_flow_prep<<<strideBlocks(std::max<uint64_t>(16ull * n, nj), 256), 256, 0, stream>>>(d, 0u, 0u);
      fill(flowDone, n, 0u);
      void *args[] = {&d, &barrierPolls};
      hipError_t e;
      timed(K_SOLVE, [&] { e = hipLaunchCooperativeKernel((const void *)k_flow, dim3(coopBlocks), dim3(64), args, 0, stream); });
      if (e == hipSuccess) {
        k_velocities<<<blocksFor(n, 256), 256, 0, stream>>>(d);
        return;
      }
      (void)hipGetLastError();
      if (capturing) { captureFailed = true; return; }
      coopBlocks = 0;
      coopFlow = false;
      sweepColors(d);
      return;
    }
    if (coopBlocks && !coopTail && !rec) {
      fill(barrier, 1, 0u);  // the arrival count starts each launch at zero
      void *args[] = {&d, &barrier, &barrierPolls};
      hipError_t e;
      timed(K_SOLVE, [&] { e = hipLaunchCooperativeKernel(solveKernel(coopThreads), dim3(coopBlocks), dim3(coopThreads), args, 0, stream); });
      if (e == hipSuccess) return;
      (void)hipGetLastError();
      if (capturing) { captureFailed = true; return; }  // not capturable: the substep runs plainly
      coopBlocks = 0;  // not co-resident here: per-colour launches from now on
    }
    if (nj) k_warm_joints<<<blocksFor(nj, 256), 256, 0, stream>>>(d);
    k_warm_bodies<<<blocksFor(n, 256), 256, 0, stream>>>(d);
    if (coopBlocks && coopTail && !rec && sweepWithTail(d)) return;
    sweepColors(d);
  }
  /** The iterations as per-colour launches of the colours before flowTailStart, then one dataflow
   *  tail per iteration (the later colours and the dual), then the velocities. False when the tail's
   *  first launch was refused before anything of the sweep ran (the caller sweeps colour by colour
   *  instead). */
  bool sweepWithTail(Dev &d) {
    uint32_t cap = fixedColors ? fixedCap : hp.color_cap, first = std::min(tailAuto ? flowTailStart : tailStart, cap);
    uint32_t pb = primalBlocks(hp, n);
    bool lead = leadFlow && first > 0;
    // one workgroup a chunk: at most a quarter of the bodies (sixteen threads a body) plus a partial
    // chunk per colour
    uint32_t leadBlocks = (uint32_t)std::min<uint64_t>((uint64_t)n / 4u + first + 1u, 0x7fffffffull);
    k_flow_prep<<<strideBlocks(std::max<uint64_t>(16ull * n, nj), 256), 256, 0, stream>>>(d, first, lead ? 1u : 0u);
    fill(flowDone, n, 0u);
    for (uint32_t it = 0; it < hp.iterations; it++) {
      if (lead) timed(K_PRIMAL, [&] { k_flow_lead<<<leadBlocks, 64, 0, stream>>>(d, first, it, barrierPolls); });
      else
        for (uint32_t c = 0; c < first; c++) timed(K_PRIMAL, [&] { k_primal<<<pb, PRIMAL_THREADS, 0, stream>>>(d, c); });
      // A plain launch: the grid is at most what the occupancy check found resident on every CU, and
      // on the solver's stream the kernel starts once the one before it has ended, so every
      // workgroup is resident together (a cooperative launch, which guarantees the same, cost ~12 us
      // of idle device before each of these 40 launches a step); a wait that still found a
      // workgroup missing would give up, be reported and drop the schedule, never hang.
      hipError_t e;
      timed(K_DUAL, [&] {
        k_flow_tail<<<coopBlocks, 64, 0, stream>>>(d, first, it, barrierPolls);
        e = hipGetLastError();
      });
      if (e != hipSuccess) {
        (void)hipGetLastError();
        if (capturing) { captureFailed = true; return true; }
        if (it == 0 && first == 0) return false;  // nothing of the sweep ran yet
        // the colours before the tail of this iteration ran: finish the iteration, then sweep on
        for (uint32_t c = first; c < cap; c++) k_primal<<<pb, PRIMAL_THREADS, 0, stream>>>(d, c);
        k_dual<<<strideBlocks(nj + conCap, 256), 256, 0, stream>>>(d);
        coopBlocks = 0;
        coopTail = false;
        leadFlow = false;
        for (uint32_t rest = it + 1; rest < hp.iterations; rest++) {
          for (uint32_t c = 0; c < cap; c++) k_primal<<<pb, PRIMAL_THREADS, 0, stream>>>(d, c);
          k_dual<<<strideBlocks(nj + conCap, 256), 256, 0, stream>>>(d);
        }
        break;
      }
    }
    k_velocities<<<blocksFor(n, 256), 256, 0, stream>>>(d);
    return true;
  }
  /** The iterations colour by colour (one launch per colour, the dual, per iteration), then the
   *  velocities; the warm starts have run. */
  void sweepColors(Dev &d) {
    uint32_t cap = fixedColors ? fixedCap : hp.color_cap;
    uint32_t pb = primalBlocks(hp, n), snapBlocks = blocksFor(std::max(n, conCap), 256);
    if (rec) k_tape_snap<<<snapBlocks, 256, 0, stream>>>(d, rec->t, recSlot, 0);
    for (uint32_t it = 0; it < hp.iterations; it++) {
      if (rec) d.bits = rec->t.bits + ((size_t)recSlot * rec->t.iters + it) * rec->t.conCap;
      for (uint32_t c = 0; c < cap; c++) {
        if (!rec && c >= tailStart) {
          timed(K_TAIL, [&] { k_primal_tail<<<1, 512, 0, stream>>>(d, c, cap); });  // apart from k_primal's launches
          break;
        }
        timed(K_PRIMAL, [&] { k_primal<<<pb, PRIMAL_THREADS, 0, stream>>>(d, c); });
      }
      timed(K_DUAL, [&] { k_dual<<<strideBlocks(nj + conCap, 256), 256, 0, stream>>>(d); });
      if (rec) k_tape_snap<<<snapBlocks, 256, 0, stream>>>(d, rec->t, recSlot, it + 1);
    }
    if (rec) k_tape_systems<<<blocksFor(4ull * n * hp.iterations, 64), 64, 0, stream>>>(d, rec->t, recSlot, rec->cap[recSlot], 4);
    k_velocities<<<blocksFor(n, 256), 256, 0, stream>>>(d);
  }
  // ---- adjoint (avbd_forward / avbd_backward) ----
  template <class T> T *adjAlloc(Mem &m, size_t count) {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(count * sizeof(T), 256)) != hipSuccess) throw std::runtime_error("hipMalloc failed (adjoint)");
    m.dev.push_back(p);
    return (T *)p;
  }
  template <class T> T *hostAlloc(Mem &m, size_t count) {
    void *p = nullptr;
    if (hipHostMalloc(&p, std::max<size_t>(count * sizeof(T), 64)) != hipSuccess) throw std::runtime_error("hipHostMalloc failed (adjoint)");
    m.host.push_back(p);
    return (T *)p;
  }
  void freeMem(Mem &m) {
    for (void *p : m.dev) hipFree(p);
    for (void *p : m.host) hipHostFree(p);
    m = Mem{};
  }
  /** Every adjoint buffer (the solver's end, or an error). Between rollouts the buffers stay and
   *  are reused: a part is reallocated only when a rollout needs more than it holds. */
  void freeAdjoint() {
    if (stream) hipStreamSynchronize(stream);
    if (revStream) hipStreamSynchronize(revStream);
    freeMem(ckptMem);
    freeMem(adjMem);
    for (auto &t : tapes) { freeMem(t.mem); t = TapeBuf{}; }
    ckpts.clear();
    ckptCount = 0;
    src = nullptr; adj = AdjBuf{}; adjConCap = 0; rec = nullptr; remapTable = nullptr;
  }
  void ensureReverseStream() {
    if (revStream) return;
    if (hipStreamCreateWithFlags(&revStream, hipStreamNonBlocking) != hipSuccess) throw std::runtime_error("hipStreamCreate failed");
    for (auto &e : recorded) hipEventCreateWithFlags(&e, hipEventDisableTiming);
    for (auto &e : reversed) hipEventCreateWithFlags(&e, hipEventDisableTiming);
  }
  /** The body state as (pointer, float4s per body): the interleaved pose block, the other six
   *  streams, the reference poses. */
  int streamPtrs(float4 *out[9], uint32_t width[9]) {
    float4 *list[6] = {b.pos.p, b.size.p, b.vel, b.angvel, refPos, refRot};
    for (int k = 0; k < 6; k++) { out[k] = list[k]; width[k] = k < 2 ? 4 : 1; }
    return 6;
  }
  void allocCheckpoint(Checkpoint &c) {
    c.streams = adjAlloc<float4>(ckptMem, 12ull * bodyCap);  // 4 (pose block) + 8 streams
    c.colA = adjAlloc<uint32_t>(ckptMem, bodyCap);
    c.counters = adjAlloc<uint32_t>(ckptMem, C_WORDS);
    c.mIds = adjAlloc<uint4>(ckptMem, manCap);
    c.mGeo = adjAlloc<float4>(ckptMem, manCap);
    c.con = adjAlloc<Contact>(ckptMem, conCap);
  }
  /** count checkpoints (their sizes are the solver's capacities: kept across rollouts). */
  void ensureCheckpoints(uint32_t count) {
    while (ckpts.size() < count) {
      ckpts.emplace_back();
      allocCheckpoint(ckpts.back());
    }
    ckptCount = count;
    if (!src) src = adjAlloc<int>(ckptMem, conCap);
    if (!remapTable) remapTable = adjAlloc<uint32_t>(ckptMem, hashSize);
  }
  /** Everything the next step reads from earlier steps, on the stream. */
  void saveCheckpoint(Checkpoint &c) {
    float4 *ps[9];
    uint32_t wd[9];
    size_t off = 0;
    for (int k = 0, m = streamPtrs(ps, wd); k < m; k++) {
      hipMemcpyAsync(c.streams + off, ps[k], 16ull * wd[k] * n, hipMemcpyDeviceToDevice, stream);
      off += (size_t)wd[k] * bodyCap;
    }
    hipMemcpyAsync(c.colA, colA, 4ull * n, hipMemcpyDeviceToDevice, stream);
    hipMemcpyAsync(c.counters, counters, 4ull * C_WORDS, hipMemcpyDeviceToDevice, stream);
    int last = 1 - parity;  // the buffers the last step wrote: the next step's previous ones
    k_copy_manifolds<<<blocksFor(manCap, 256), 256, 0, stream>>>(mIds[last], mGeo[last], c.mIds, c.mGeo, counters, C_MANIFOLDS, manCap);
    k_copy_contacts<<<blocksFor(conCap, 256), 256, 0, stream>>>(con[last], c.con, counters, C_CONTACTS, conCap);
    c.stepCount = stepCount; c.shrinkVotes = shrinkVotes; c.stepsSinceAdapt = stepsSinceAdapt;
    c.parity = parity; c.manifoldsValid = manifoldsValid; c.hp = hp;
  }
  void restoreCheckpoint(const Checkpoint &c) {
    stepCount = c.stepCount; shrinkVotes = c.shrinkVotes; stepsSinceAdapt = c.stepsSinceAdapt;
    parity = c.parity; manifoldsValid = c.manifoldsValid; hp = c.hp;
    float4 *ps[9];
    uint32_t wd[9];
    size_t off = 0;
    for (int k = 0, m = streamPtrs(ps, wd); k < m; k++) {
      hipMemcpyAsync(ps[k], c.streams + off, 16ull * wd[k] * n, hipMemcpyDeviceToDevice, stream);
      off += (size_t)wd[k] * bodyCap;
    }
    hipMemcpyAsync(colA, c.colA, 4ull * n, hipMemcpyDeviceToDevice, stream);
    hipMemcpyAsync(counters, c.counters, 4ull * C_WORDS, hipMemcpyDeviceToDevice, stream);
    int last = 1 - parity;
    k_copy_manifolds<<<blocksFor(manCap, 256), 256, 0, stream>>>(c.mIds, c.mGeo, mIds[last], mGeo[last], c.counters, C_MANIFOLDS, manCap);
    k_copy_contacts<<<blocksFor(conCap, 256), 256, 0, stream>>>(c.con, con[last], c.counters, C_CONTACTS, conCap);
  }
  /** A tape of at least `slots` steps with room for cCap contacts and mCap manifolds per step
   *  (the one held is kept when it is large enough). */
  void ensureTape(TapeBuf &tb, uint32_t slots, uint32_t cCap, uint32_t mCap) {
    cCap = std::max(cCap, 1u); mCap = std::max(mCap, 1u);
    if (tb.slots >= slots && tb.t.conCap >= cCap && tb.t.manCap >= mCap && tb.t.n == n && tb.t.iters == hp.iterations) return;
    freeMem(tb.mem);
    Mem &m = tb.mem;
    Tape t{};
    t.n = n; t.iters = hp.iterations; t.conCap = cCap; t.manCap = mCap; t.adjCap = 2 * t.manCap + 2 * nj;
    size_t S = slots, K = hp.iterations + 1;
    float4 **bodyArrays[6] = {&t.p0, &t.q0, &t.inpos, &t.inrot, &t.v0, &t.w0};
    for (auto a : bodyArrays) *a = adjAlloc<float4>(m, S * n);
    t.P = adjAlloc<float4>(m, S * K * n);
    t.R = adjAlloc<float4>(m, S * K * n);
    t.con = adjAlloc<Contact>(m, S * t.conCap);
    t.src = adjAlloc<int>(m, S * t.conCap);
    t.lam = adjAlloc<float4>(m, S * K * t.conCap);
    t.pen = adjAlloc<float4>(m, S * K * t.conCap);
    t.mIds = adjAlloc<uint4>(m, S * t.manCap);
    t.mGeo = adjAlloc<float4>(m, S * t.manCap);
    t.adjStart = adjAlloc<uint32_t>(m, S * (n + 1));
    t.adjList = adjAlloc<uint32_t>(m, S * t.adjCap);
    t.colorStart = adjAlloc<uint32_t>(m, S * 65);
    t.colorBodies = adjAlloc<uint32_t>(m, S * n);
    t.col = adjAlloc<uint32_t>(m, S * n);
    t.counts = adjAlloc<uint32_t>(m, S * 4);
    t.cman = adjAlloc<uint32_t>(m, S * t.conCap);
    t.bits = adjAlloc<uchar4>(m, S * hp.iterations * t.conCap);
    t.sys = adjAlloc<double>(m, S * hp.iterations * (size_t)n * SYS_WORDS);
    t.ptCap = 2 * t.conCap;
    t.ptOff = adjAlloc<uint32_t>(m, S * (n + 1));
    t.pts = adjAlloc<uint4>(m, S * t.ptCap);
    tb.t = t;
    tb.slots = slots;
    tb.cap.assign(slots, 0);
    tb.hostStarts = hostAlloc<uint32_t>(m, 65 * S);
    tb.hostCounts = hostAlloc<uint32_t>(m, 4 * S);
  }
  /** The reverse sweep's accumulators for up to cCap contacts per step (kept when large enough). */
  void ensureAdjoint(uint32_t cCap) {
    cCap = std::max(cCap, 1u);
    if (adj.state && adjConCap >= cCap) return;
    freeMem(adjMem);
    AdjBuf a{};
    a.state = adjAlloc<double>(adjMem, 16ull * n);
    double **bodies4[5] = {&a.ip, &a.ir, &a.np, &a.nr, &a.vs};
    for (auto p : bodies4) *p = adjAlloc<double>(adjMem, 4ull * n);
    a.mu = adjAlloc<double>(adjMem, n);
    a.lamb = adjAlloc<double>(adjMem, 4ull * cCap);
    a.lambPrev = adjAlloc<double>(adjMem, 4ull * cCap);
    a.c0b = adjAlloc<double>(adjMem, 4ull * cCap);
    a.mub = adjAlloc<double>(adjMem, cCap);
    a.dpart = adjAlloc<double>(adjMem, 32ull * cCap);
    a.nslot = adjAlloc<double>(adjMem, 32ull * cCap);
    adj = a;
    adjConCap = cCap;
  }
  void zeroF64(double *p, size_t count, hipStream_t st) { k_fill_f64<<<blocksFor(count, 256), 256, 0, st>>>(p, count, 0.0); }
  /** Recorded colour buckets and counts of slots [0, len) into the pinned host copies, on the
   *  stream after the recording (they size the reverse's launches). */
  void copySizes(TapeBuf &tb, uint32_t len, hipStream_t st) {
    hipMemcpyAsync(tb.hostStarts, tb.t.colorStart, 4ull * 65 * len, hipMemcpyDeviceToHost, st);
    hipMemcpyAsync(tb.hostCounts, tb.t.counts, 4ull * 4 * len, hipMemcpyDeviceToHost, st);
  }
  /** Reverse of the recorded slots [0, len) (global steps first + slot). The host copies of the
   *  colour buckets and counts (copySizes) must have arrived: they size each launch. */
  void reverseSlots(const TapeBuf &tb, uint32_t len, uint32_t first, hipStream_t st) {
    const Tape &tape = tb.t;
    uint32_t manBlocks = blocksFor(tape.manCap, 64), nb = blocksFor(n, 256);
    for (uint32_t t = len; t-- > 0;) {
      Dev d = dev();
      d.b.irot = Strided4{tape.q0 + (size_t)t * n, 1};  // x- of step t: the sphere contacts' rotations in evalContact
      const uint32_t *starts = tb.hostStarts + 65 * t;
      uint32_t conBlocks = blocksFor(std::max(tb.hostCounts[4 * t + 1], 1u), 64);
      k_adj_velocities<<<nb, 256, 0, st>>>(d, tape, adj, t);
      zeroF64(adj.nslot, 32ull * std::max(tb.hostCounts[4 * t + 1], 1u), st);
      for (uint32_t it = hp.iterations; it-- > 0;) {
        k_adj_dual<<<conBlocks, 64, 0, st>>>(d, tape, adj, t, it);
        for (uint32_t c = tb.cap[t]; c-- > 0;) {
          uint32_t size = starts[c + 1] - starts[c];
          if (size) k_adj_color<<<blocksFor((uint64_t)size * ADJ_LANES, 64), 64, 0, st>>>(d, tape, adj, t, it, c);
        }
      }
      k_adj_slots_final<<<nb, 256, 0, st>>>(d, tape, adj, t, tb.cap[t]);
      k_adj_contacts<<<manBlocks, 64, 0, st>>>(d, tape, adj, t, first + t > 0 ? 1 : 0);
      k_adj_bodies<<<nb, 256, 0, st>>>(d, tape, adj, t);
      std::swap(adj.lamb, adj.lambPrev);
    }
  }
  void recordSteps(TapeBuf &tb, uint32_t count) {
    rec = &tb;
    for (uint32_t t = 0; t < count; t++) { recSlot = t; step(false); }
    rec = nullptr;
  }

  bool fixedStorage() const { return (hp.flags & AVBD_FLAG_FIXED_STORAGE) != 0; }
  /** uploadStatics when a scene edit changed radii, masses, the body count or joint pairs (the
   *  reference's radiiDirty, solver.ts:2119): cell size, large lists, no-collide list and the
   *  colouring priorities of every body. */
  void refreshStatics() {
    if (!staticsDirty) return;
    std::vector<float> r(radiusHost.begin(), radiusHost.begin() + n), m(massHost.begin(), massHost.begin() + n);
    uploadStatics(r, m);
    Dev d = dev();
    k_priorities<<<blocksFor(std::max(n, 1u), 256), 256, 0, stream>>>(d);
    staticsDirty = false;
  }
  /** A body's aux record: gravity scale, the velocity decay factors per step exp(-rate dt) with
   *  the exponent rounded to f32 as the reference's exp(-aux.y * dt) takes it, and 1 when damped. */
  float4 auxRecord(uint32_t i) const {
    float lin = auxLin[i], ang = auxAng[i], dt = hp.dt;
    float xl = -(lin * dt), xa = -(ang * dt);
    return make_float4(auxScale[i], (float)std::exp((double)xl), (float)std::exp((double)xa), (lin > 0.0f || ang > 0.0f) ? 1.0f : 0.0f);
  }
  void uploadAux(uint32_t first, uint32_t count) {
    if (!aux || !count) return;
    std::vector<float4> r(count);
    for (uint32_t k = 0; k < count; k++) r[k] = auxRecord(first + k);
    hipMemcpyAsync(aux + first, r.data(), 16ull * count, hipMemcpyHostToDevice, stream);
    hipStreamSynchronize(stream);
  }
  /** The primal's walk items: one per contact point and joint on each side (both bodies). */
  static uint32_t itemsFor(uint32_t contacts, uint32_t joints) { return (uint32_t)std::min<uint64_t>(2ull * contacts + 2ull * joints, 0xffffffffu); }
  /** Joint storage for `cap` joints (the reference's allocateJoints, doubling): records, info and
   *  the adjacency sized by it. */
  void growJoints(uint32_t cap) {
    if (cap <= jointCap) return;
    hipStreamSynchronize(stream);
    joints = regrow(joints, 10ull * jointCap, 10ull * cap, true);
    jinfo = regrow(jinfo, jointCap, cap, true);
    flowOrphan = regrow(flowOrphan, jointCap, cap, false);
    uint32_t ac = 2u * (cap + manCap);
    adjList = regrow(adjList, adjCap, ac, false);
    adjTmp = regrow(adjTmp, adjCap, ac, false);
    adjOwner = regrow(adjOwner, adjCap, ac, false);
    adjKeys = regrow(adjKeys, adjCap, ac, false);
    adjOther = regrow(adjOther, adjCap, ac, false);
    adjIds = regrow(adjIds, adjCap, ac, false);
    adjGeo = regrow(adjGeo, adjCap, ac, false);
    uint32_t ic = itemsFor(conCap, cap);
    items = regrow(items, itemCap, ic, false);
    itemCap = ic;
    adjCap = ac;
    jointCap = cap;
    adjacencyHeld = false;
    hipStreamSynchronize(stream);
  }
  /** Body records [first, first + count) from host memory (rewriteBodies / addBodies' writeBodies,
   *  solver.ts:1087): the host mirrors follow; uncoloured unless `keepColours` (seedFrom). */
  void writeRecords(const float *in, uint32_t first, uint32_t count, bool keepColours) {
    for (uint32_t k = 0; k < count; k++) {
      const float *r = in + 40ull * k;
      uint32_t i = first + k;
      if (radiusHost[i] != r[23] || (massHost[i] > 0.0f) != (r[19] > 0.0f) || massHost[i] != r[19]) staticsDirty = true;
      radiusHost[i] = r[23];
      massHost[i] = r[19];
      for (int c = 0; c < 3; c++) sizeHost[3ull * i + c] = r[16 + c];
      float w = r[11];
      lifeHost[i] = (w >= 0.0f && w <= 7.0f && w == std::floor(w)) ? (uint8_t)((uint32_t)w & 6u) : 0;
    }
    hipMemcpyAsync(staging, in, 160ull * count, hipMemcpyHostToDevice, stream);
    k_records_to_bodies<<<blocksFor(count, 256), 256, 0, stream>>>(dev(), staging, first, count);
    if (!keepColours) { fill(colA + first, count, NO_COLOR); fill(colB + first, count, NO_COLOR); }
    manifoldsValid = false;
  }
  /** Joint or spring records into `slots` with their info, a copy per run of consecutive slots
   *  (appendConstraints' uploads, solver.ts:1718-1725); the no-collide list follows. */
  void putJoints(const std::vector<uint32_t> &slots, const std::vector<float> &recs, const std::vector<int> &ends) {
    std::vector<uint32_t> order(slots.size());
    for (size_t k = 0; k < slots.size(); k++) order[k] = (uint32_t)k;
    std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return slots[x] < slots[y]; });
    std::vector<float> run;
    for (size_t k = 0; k < order.size();) {
      size_t e = k + 1;
      while (e < order.size() && slots[order[e]] == slots[order[e - 1]] + 1) e++;
      run.assign(40 * (e - k), 0.0f);
      for (size_t i = k; i < e; i++) {
        uint32_t q = order[i], slot = slots[q];
        std::copy(recs.begin() + 40 * q, recs.begin() + 40 * q + 40, run.begin() + 40 * (i - k));
        int32_t *in = &jinfoHost[4ull * slot];
        in[0] = ends[3 * q]; in[1] = ends[3 * q + 1]; in[2] = ends[3 * q + 2]; in[3] = 0;
      }
      uint32_t first = slots[order[k]];
      hipMemcpyAsync(joints + 10ull * first, run.data(), 160ull * (e - k), hipMemcpyHostToDevice, stream);
      hipMemcpyAsync(jinfo + first, &jinfoHost[4ull * first], 16ull * (e - k), hipMemcpyHostToDevice, stream);
      hipStreamSynchronize(stream);
      k = e;
    }
    manifoldsValid = false;
    staticsDirty = true;
  }
  /** A cooperative launch whose grid barrier gave up waiting ran part of its solve unordered: that
   *  step's result is not the step's. From then on the solver uses per-colour launches; the
   *  give-ups are reported (bit AVBD_OVERFLOW_BARRIER of the counters' overflow word, and an error
   *  from the avbd_step/avbd_steps call that saw them). */
  void noteBarrier(const uint32_t *c) {
    if (!c[C_BARRIER]) return;
    barrierGiveUps += c[C_BARRIER];
    barrierFailed = true;
    coopBlocks = 0;
    coopFlow = false;
    coopTail = false;
    leadFlow = false;
    fill(counters + C_BARRIER, 1, 0u);
    err = "the cooperative solve's grid barrier gave up waiting (" + std::to_string(c[C_BARRIER]) +
          " waits): the solve of the steps since the last readback ran unordered; the solver now uses per-colour launches";
  }
  void substep(bool reuseManifolds, bool keepAdjacency) {
    collision(reuseManifolds, keepAdjacency);
    topology();
    solve();
  }
  // ---- per-substep hipGraphs (avbd_set_graphs) ----
  /** Everything the host decides a substep's launches by: the kernels' arguments (buffers,
   *  capacities, params) and the choices taken outside them. Equal bytes, equal launches. */
  struct GraphKey {
    Dev d;
    uint32_t reuse, keep, coopBlocks, coopThreads, tailStart, barrierPolls, hullBlocks, anyLarge, coopFlow, coopTail, flowTailStart, leadFlow;
    const uint32_t *fixedDev;
  };
  struct GraphEntry { GraphKey key; hipGraphExec_t exec; uint64_t used; };
  static const size_t GRAPH_CACHE = 8;
  bool graphs = false, capturing = false, captureFailed = false, coopUncapturable = false;
  std::vector<GraphEntry> graphCache;
  uint64_t graphClock = 0, graphCaptures = 0, graphReplays = 0;
  GraphKey graphKey(bool reuse, bool keep) const {
    GraphKey k;
    std::memset(&k, 0, sizeof k);
    k.d = dev();
    k.reuse = reuse; k.keep = keep; k.coopBlocks = coopBlocks; k.coopThreads = coopThreads; k.tailStart = tailStart;
    k.barrierPolls = barrierPolls; k.hullBlocks = hullBlocks; k.anyLarge = anyDynamicLarge || anyFixedSmall; k.fixedDev = fixedColors ? fixedDev : nullptr;
    k.coopFlow = coopFlow; k.coopTail = coopTail; k.flowTailStart = flowTailStart; k.leadFlow = leadFlow;
    return k;
  }
  void dropGraphs() {
    for (auto &g : graphCache) hipGraphExecDestroy(g.exec);
    graphCache.clear();
  }
  /** The substep as one graph launch: replayed when a graph of the same key was captured, else
   *  captured now (the capture runs nothing) and launched. False: run the launches plainly (a
   *  cooperative launch this runtime will not capture, or a failed capture). */
  bool replayGraph(bool reuse, bool keep) {
    if (coopBlocks && coopUncapturable) return false;
    GraphKey key = graphKey(reuse, keep);
    graphClock++;
    for (auto &g : graphCache)
      if (std::memcmp(&g.key, &key, sizeof key) == 0) {
        if (hipGraphLaunch(g.exec, stream) != hipSuccess) return false;
        g.used = graphClock;
        graphReplays++;
        return true;
      }
    if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) { (void)hipGetLastError(); return false; }
    capturing = true;
    captureFailed = false;
    substep(reuse, keep);
    capturing = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(stream, &g);
    hipGraphExec_t exec = nullptr;
    bool ok = e == hipSuccess && !captureFailed && g && hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) == hipSuccess;
    if (g) hipGraphDestroy(g);
    if (!ok) {
      (void)hipGetLastError();
      if (coopBlocks) coopUncapturable = true;
      return false;
    }
    if (graphCache.size() >= GRAPH_CACHE) {
      auto old = std::min_element(graphCache.begin(), graphCache.end(), [](const GraphEntry &a, const GraphEntry &b) { return a.used < b.used; });
      hipGraphExecDestroy(old->exec);
      graphCache.erase(old);
    }
    graphCache.push_back(GraphEntry{key, exec, graphClock});
    graphCaptures++;
    return hipGraphLaunch(exec, stream) == hipSuccess;
  }
  /** One substep. keepAdjacency: the caller ran the previous substep right before this one with
   *  nothing of the host's in between (the substeps of one avbd_steps call), so a reuse that keeps
   *  every manifold may keep the adjacency too; any other call rebuilds it. */
  void step(bool reuseManifolds, bool keepAdjacency = false) {
    refreshStatics();
    stepCount++;
    // a collision backend writes its own contacts each step: manifold reuse is the built-in
    // narrowphase's (solver.ts:2121), and its host-side pair uploads keep it out of graphs
    reuseManifolds = reuseManifolds && manifoldsValid && !backend;
    bool keep = reuseManifolds && keepAdjacency && adjacencyHeld;
    if (!(graphs && !rec && !timer.on && !backend && replayGraph(reuseManifolds, keep))) substep(reuseManifolds, keep);
    parity = 1 - parity;
    manifoldsValid = true;
    adjacencyHeld = true;
    if ((hp.flags & AVBD_FLAG_ADAPT_COLORS) && ++stepsSinceAdapt >= 10) {
      stepsSinceAdapt = 0;
      adjacencyHeld = false;  // storage may be regrown here
      uint32_t c[C_WORDS], starts[MAX_COLORS + 1];
      if (hipMemcpyAsync(c, counters, sizeof(c), hipMemcpyDeviceToHost, stream) == hipSuccess &&
          hipMemcpyAsync(starts, colorStart, sizeof(starts), hipMemcpyDeviceToHost, stream) == hipSuccess &&
          hipStreamSynchronize(stream) == hipSuccess) {
        avbd_counters k{};
        k.colors = c[C_NUM_COLORS];
        k.clashes = c[C_CLASHES];
        k.pairs = c[C_PAIRS]; k.contacts = c[C_CONTACTS]; k.manifolds = c[C_MANIFOLDS]; k.overflow = c[C_OVERFLOW];
        noteBarrier(c);
        adapt(k);
        adaptStorage(k);
        if (tailAuto) chooseTail(starts);
        chooseFlowTail(starts);
      }
    }
  }
};

// Fix the size of a scan's temporary: blocks of 1024 words, at most 4096 of them.
static const uint32_t MAX_SCAN = 4096u * 1024u;

/** Storage stays put while an adjoint rollout records or re-runs (its tapes index the buffers). */
struct NoGrowScope {
  avbd_solver *s;
  explicit NoGrowScope(avbd_solver *x) : s(x) { s->noGrow = true; }
  ~NoGrowScope() { s->noGrow = false; }
};
static int fail(avbd_solver *s, int code, const std::string &msg) {
  if (s) s->err = msg;
  return code;
}
static bool checkParams(const avbd_params *p, std::string &why) {
  if (!p) { why = "params is null"; return false; }
  if (!(p->dt > 0) || !std::isfinite(p->dt)) { why = "dt must be finite and > 0"; return false; }
  if (!std::isfinite(p->contact_margin) || p->contact_margin < 0) { why = "Invalid contactMargin"; return false; }
  if (p->color_cap < 1 || p->color_cap > 64) { why = "color_cap must be in [1, 64]"; return false; }
  if (p->color_rounds % 2 != 0) { why = "color_rounds must be even"; return false; }
  return true;
}

extern "C" {

const char *avbd_create_error(void) { return g_create_error.c_str(); }
const char *avbd_last_error(const avbd_solver *s) { return s ? s->err.c_str() : "null solver"; }

int avbd_create(const avbd_params *params, const float *bodies, uint32_t body_count, const float *jrec, const int32_t *jinfo,
                uint32_t joint_count, const avbd_capacity *capacity, avbd_solver **out) {
  uint32_t ws[2] = {0, body_count};
  return avbd_create_batch(params, bodies, body_count, jrec, jinfo, joint_count, ws, 1, capacity, out);
}

int avbd_create_batch(const avbd_params *params, const float *bodies, uint32_t body_count, const float *jrec, const int32_t *jinfo,
                      uint32_t joint_count, const uint32_t *world_start, uint32_t worlds, const avbd_capacity *capacity,
                      avbd_solver **out) {
  g_create_error.clear();
  if (!out) { g_create_error = "out is null"; return AVBD_ERR_INVALID; }
  *out = nullptr;
  std::string why;
  if (!checkParams(params, why)) { g_create_error = why; return AVBD_ERR_INVALID; }
  if (body_count > 0 && !bodies) { g_create_error = "bodies is null"; return AVBD_ERR_INVALID; }
  if (joint_count > 0 && (!jrec || !jinfo)) { g_create_error = "joints or joint_info is null"; return AVBD_ERR_INVALID; }
  if (worlds < 1 || !world_start || world_start[0] != 0 || world_start[worlds] != body_count) {
    g_create_error = "world_start must run from 0 to body_count over worlds + 1 entries";
    return AVBD_ERR_INVALID;
  }
  for (uint32_t w = 0; w < worlds; w++)
    if (world_start[w + 1] < world_start[w]) { g_create_error = "world_start must not decrease"; return AVBD_ERR_INVALID; }
  for (uint32_t j = 0; j < joint_count; j++) {
    int t = jinfo[4 * j], a = jinfo[4 * j + 1], b = jinfo[4 * j + 2];
    if (t < 0 || t > 2 || b < 0 || (uint32_t)b >= body_count || a < -1 || a >= (int)body_count || (t == 2 && a < 0)) {
      g_create_error = "joint " + std::to_string(j) + ": invalid type or bodies";
      return AVBD_ERR_INVALID;
    }
  }
  avbd_capacity cap = capacity ? *capacity : avbd_capacity{};
  avbd_solver *s = new avbd_solver();
  try {
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) throw std::runtime_error("hipStreamCreate failed");
    s->hp = *params;
    {
      int dev = 0, coop = 0, cus = 0;
      (void)hipGetDevice(&dev);
      s->device = dev;
      s->readScratch = scratchOn(dev);
      (void)hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev);
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      s->numCUs = (uint32_t)(cus > 0 ? cus : 0);
      s->coopSupported = coop != 0 && cus > 0;
      // as many hull workgroups as fit at once (registers and the LDS stage decide): every pair
      // of the grid-strided loop then runs in the first wave of workgroups
      int hullPerCU = 0;
      (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&hullPerCU, (const void *)k_narrow_hull, 256, 0);
      s->hullBlocks = (uint32_t)std::max(hullPerCU, 1) * std::max(s->numCUs, 1u);
      avbd_schedule sch{1u, 512u, 1u};
      if (const char *t = getenv("AVBD_SOLVE_THREADS")) sch.threads = (uint32_t)atoi(t);
      if (const char *b = getenv("AVBD_SOLVE_BLOCKS_PER_CU")) sch.blocks_per_cu = (uint32_t)atoi(b);
      if (const char *mode = getenv("AVBD_SOLVE")) {
        std::string m(mode);
        sch.cooperative = m == "colors" ? 0u : (m == "flow" ? 2u : 1u);
        if (m == "flow") {
          sch.threads = 64;
          if (!getenv("AVBD_SOLVE_BLOCKS_PER_CU")) sch.blocks_per_cu = 4;
        }
      }
      if (s->applySchedule(sch) != AVBD_OK) s->applySchedule(avbd_schedule{0u, 512u, 1u});
    }
    s->n = body_count;
    s->nj = joint_count;
    s->bodyCap = std::max({cap.bodies, body_count, 1u});
    s->jointCap = std::max({cap.joints, joint_count, 1u});
    s->pairCap = cap.pairs ? cap.pairs : std::max(4096u, 8u * s->bodyCap);
    s->manCap = cap.manifolds ? cap.manifolds : s->pairCap;
    s->conCap = cap.contacts ? cap.contacts : 4u * s->manCap;
    s->hullCap = std::max(cap.hull_words, 4u);
    s->tableSize = pow2At(2ull * s->bodyCap);
    s->hashSize = pow2At(2ull * s->manCap);
    s->colorGroups = (s->bodyCap + 255) / 256;
    s->adjCap = 2u * (s->jointCap + s->manCap);
    if (s->tableSize + 1 > MAX_SCAN || 64ull * s->colorGroups + 1 > MAX_SCAN) throw std::runtime_error("body capacity too large for one solver");
    uint32_t N = s->bodyCap;
    float4 *pose = s->alloc<float4>(4ull * N);  // pos, rot, x-, q- interleaved: 64 bytes per body
    s->b.pos = Strided4{pose, 4}; s->b.rot = Strided4{pose + 1, 4}; s->b.ipos = Strided4{pose + 2, 4}; s->b.irot = Strided4{pose + 3, 4};
    float4 *solveBlock = s->alloc<float4>(4ull * N);  // size, moment, inertial pos, inertial rot
    s->b.size = Strided4{solveBlock, 4}; s->b.moment = Strided4{solveBlock + 1, 4};
    s->b.inpos = Strided4{solveBlock + 2, 4}; s->b.inrot = Strided4{solveBlock + 3, 4};
    float4 **streams[2] = {&s->b.vel, &s->b.angvel};
    for (auto p : streams) *p = s->alloc<float4>(N);
    s->refPos = s->alloc<float4>(N);
    s->refRot = s->alloc<float4>(N);
    s->filters = s->alloc<uint2>(N);
    s->joints = s->alloc<float4>(10ull * s->jointCap);
    s->jinfo = s->alloc<int4>(s->jointCap);
    s->flowOrphan = s->alloc<uint32_t>(s->jointCap);
    s->hulls = s->alloc<uint4>(s->hullCap / 4 + 1);
    s->cellStart = s->alloc<uint32_t>(s->tableSize + 1);
    for (int k = 0; k < 2; k++) {
      s->sortKey[k] = s->alloc<uint32_t>(N);
      s->sortIdx[k] = s->alloc<uint32_t>(N);
    }
    s->sortHist = s->alloc<uint32_t>(1024ull * ((N + 255) / 256) + 1);
    s->cellSorted = s->alloc<uint32_t>(N);
    s->cellOf = s->alloc<int4>(N);
    s->sortedPos = s->alloc<float4>(N);
    s->sortedKey = s->alloc<int4>(N);
    s->pairs = s->alloc<uint2>(s->pairCap);
    if (getenv("AVBD_HULL_INLINE") == nullptr) {
      s->hullPairs = s->alloc<uint2>(s->pairCap);
      s->hullOut = s->alloc<float4>((size_t)HULL_OUT * s->pairCap);
      s->boxPairs = s->alloc<uint2>(s->pairCap);
    }
    for (int k = 0; k < 2; k++) {
      s->mIds[k] = s->alloc<uint4>(s->manCap);
      s->mGeo[k] = s->alloc<float4>(s->manCap);
      s->con[k] = s->alloc<Contact>(s->conCap);
    }
    s->cman = s->alloc<uint32_t>(s->conCap);
    s->table = s->alloc<uint32_t>(s->hashSize);
    s->adjCount = s->alloc<uint32_t>(N + 1);
    s->adjStart = s->alloc<uint32_t>(N + 1);
    s->adjFill = s->alloc<uint32_t>(N);
    s->adjList = s->alloc<uint32_t>(s->adjCap);
    s->adjTmp = s->alloc<uint32_t>(s->adjCap);
    s->adjOwner = s->alloc<uint32_t>(s->adjCap);
    s->adjKeys = s->alloc<uint2>(s->adjCap);
    s->adjOther = s->alloc<int>(s->adjCap);
    s->adjIds = s->alloc<uint4>(s->adjCap);
    s->adjGeo = s->alloc<float4>(s->adjCap);
    s->itemCap = avbd_solver::itemsFor(s->conCap, s->jointCap);
    s->items = s->alloc<uint4>(s->itemCap);
    s->itemStart = s->alloc<uint32_t>(N + 1);
    s->colA = s->alloc<uint32_t>(N);
    s->colB = s->alloc<uint32_t>(N);
    s->colorStart = s->alloc<uint32_t>(65);
    s->colorBodies = s->alloc<uint32_t>(N);
    s->colorItems = s->alloc<uint4>(N);
    s->colorRank = s->alloc<uint32_t>(N);
    s->bodyKey = s->alloc<uint32_t>(N);
    s->flowDone = s->alloc<uint32_t>(N);
    s->hist = s->alloc<uint32_t>(64ull * s->colorGroups + 1);
    s->fixedDev = s->alloc<uint32_t>(N);
    s->counters = s->alloc<uint32_t>(C_WORDS);
    s->worldStartHost.assign(world_start, world_start + worlds + 1);
    s->worldOf = s->alloc<uint32_t>(N);
    s->worldStart = s->alloc<uint32_t>(worlds + 1);
    s->largeStart = s->alloc<uint32_t>(worlds + 1);
    s->largeRadiusW = s->alloc<float>(worlds);
    {
      std::vector<uint32_t> wo(N, 0);
      for (uint32_t w = 0; w < worlds; w++) for (uint32_t i = world_start[w]; i < world_start[w + 1]; i++) wo[i] = w;
      hipMemcpyAsync(s->worldOf, wo.data(), N * 4ull, hipMemcpyHostToDevice, s->stream);
      hipMemcpyAsync(s->worldStart, world_start, (worlds + 1) * 4ull, hipMemcpyHostToDevice, s->stream);
      if (hipStreamSynchronize(s->stream) != hipSuccess) throw std::runtime_error("device error during create");
    }
    s->prio = s->alloc<uint32_t>(N);
    s->scanSums = s->alloc<uint32_t>(4096);
    s->barrier = s->alloc<uint32_t>(64);
    s->fill(s->colA, N, NO_COLOR);
    s->fill(s->colB, N, NO_COLOR);
    std::vector<uint32_t> f(2ull * N, 0xffffffffu);
    hipMemcpyAsync(s->filters, f.data(), f.size() * 4, hipMemcpyHostToDevice, s->stream);
    // bodies: records -> SoA streams; new bodies have no contact history (DIRTY, layout.ts:17)
    s->radiusHost.assign(s->bodyCap, 0.0f);
    s->massHost.assign(s->bodyCap, 0.0f);
    s->sizeHost.assign(3ull * s->bodyCap, 0.0f);
    s->lifeHost.assign(s->bodyCap, 0);
    s->jinfoHost.assign(4ull * s->jointCap, 0);
    for (uint32_t j = 0; j < s->jointCap; j++) { s->jinfoHost[4ull * j + 1] = -1; s->jinfoHost[4ull * j + 2] = -1; }
    if (body_count) {
      int rc = avbd_write_bodies(s, bodies, 0, body_count);
      if (rc) throw std::runtime_error(s->err);
    }
    if (joint_count) {
      std::vector<float> jr(jrec, jrec + 40ull * joint_count);
      for (auto &x : jr) x = finiteF(x);
      hipMemcpyAsync(s->joints, jr.data(), jr.size() * 4, hipMemcpyHostToDevice, s->stream);
      std::copy(jinfo, jinfo + 4ull * joint_count, s->jinfoHost.begin());
      hipMemcpyAsync(s->jinfo, s->jinfoHost.data(), 16ull * joint_count, hipMemcpyHostToDevice, s->stream);
    }
    std::vector<float> radius(s->radiusHost.begin(), s->radiusHost.begin() + body_count);
    std::vector<float> mass(s->massHost.begin(), s->massHost.begin() + body_count);
    s->uploadStatics(radius, mass);
    s->staticsDirty = false;
    {
      Dev d = s->dev();
      k_priorities<<<blocksFor(s->n, 256), 256, 0, s->stream>>>(d);  // the colouring's fixed priorities
    }
    if (hipStreamSynchronize(s->stream) != hipSuccess) throw std::runtime_error("device error during create");
  } catch (const std::exception &e) {
    g_create_error = e.what();
    delete s;
    return AVBD_ERR_DEVICE;
  }
  *out = s;
  return AVBD_OK;
}

void avbd_destroy(avbd_solver *s) {
  if (!s) return;
  hipStreamSynchronize(s->stream);
  delete s;
}

int avbd_set_params(avbd_solver *s, const avbd_params *p) {
  if (!s) return AVBD_ERR_INVALID;
  std::string why;
  if (!checkParams(p, why)) return fail(s, AVBD_ERR_INVALID, why);
  bool dtChanged = p->dt != s->hp.dt;
  s->hp = *p;
  if (dtChanged && s->aux) s->uploadAux(0, s->bodyCap);  // the damping factors are per step of dt
  return AVBD_OK;
}
int avbd_get_params(const avbd_solver *s, avbd_params *p) {
  if (!s || !p) return AVBD_ERR_INVALID;
  *p = s->hp;
  return AVBD_OK;
}

int avbd_step(avbd_solver *s, const avbd_step_opts *opts) {
  if (!s) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  if (s->backend) { std::string why; if (!sdfPrepare(s->backend, why)) return fail(s, AVBD_ERR_INVALID, why); }
  s->barrierFailed = false;
  s->step(opts && opts->reuse_manifolds);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(s, AVBD_ERR_DEVICE, hipGetErrorString(e));
  return s->barrierFailed ? AVBD_ERR_DEVICE : AVBD_OK;
}
int avbd_steps(avbd_solver *s, uint32_t steps, uint32_t substeps) {
  if (!s) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  if (substeps == 0) return fail(s, AVBD_ERR_INVALID, "substeps must be >= 1");
  if (s->backend) { std::string why; if (!sdfPrepare(s->backend, why)) return fail(s, AVBD_ERR_INVALID, why); }
  s->barrierFailed = false;
  for (uint32_t k = 0; k < steps; k++)
    for (uint32_t u = 0; u < substeps; u++) s->step(u > 0, u > 0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(s, AVBD_ERR_DEVICE, hipGetErrorString(e));
  return s->barrierFailed ? AVBD_ERR_DEVICE : AVBD_OK;
}
int avbd_sync(avbd_solver *s) {
  if (!s) return AVBD_ERR_INVALID;
  hipError_t e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return fail(s, AVBD_ERR_DEVICE, hipGetErrorString(e));
  return AVBD_OK;
}

/** Shape codes a record may carry: box 0, sphere 1, hull 3 + a header offset (vec4 units) inside
 *  the hull buffer's capacity. Sails (2) have no wind model here and are refused. */
static int checkShapes(avbd_solver *s, const float *in, uint32_t count) {
  for (uint32_t k = 0; k < count; k++) {
    float code = in[40ull * k + 39];
    if (code == 0.0f || code == 1.0f) continue;
    if (code == 2.0f) return fail(s, AVBD_ERR_INVALID, "body " + std::to_string(k) + ": sails (shape 2) are not supported");
    if (!(code >= 3.0f) || code != std::floor(code) || ((double)code - 3.0) * 4.0 + 4.0 > (double)s->hullCap)
      return fail(s, AVBD_ERR_INVALID, "body " + std::to_string(k) + ": shape code " + std::to_string(code) + " is no box, sphere or hull in the hull buffer");
  }
  return AVBD_OK;
}
static int ensureStaging(avbd_solver *s, size_t bytes) {
  if (s->stagingBytes >= bytes) return AVBD_OK;
  if (s->staging) hipFree(s->staging);
  s->staging = nullptr;
  s->stagingBytes = 0;
  if (hipMalloc(&s->staging, bytes) != hipSuccess) return fail(s, AVBD_ERR_DEVICE, "hipMalloc staging failed");
  s->stagingBytes = bytes;
  return AVBD_OK;
}

int avbd_read_bodies(avbd_solver *s, float *out, uint32_t first, uint32_t count) {
  if (!s || (!out && count)) return AVBD_ERR_INVALID;
  if ((uint64_t)first + count > s->n) return fail(s, AVBD_ERR_INVALID, "body range out of bounds");
  if (!count) return AVBD_OK;
  size_t bytes = 160ull * count;
  if (int rc = ensureStaging(s, bytes)) return rc;
  k_bodies_to_records<<<blocksFor(count, 256), 256, 0, s->stream>>>(s->dev(), s->staging, first, count);
  hipMemcpyAsync(out, s->staging, bytes, hipMemcpyDeviceToHost, s->stream);
  return avbd_sync(s);
}
int avbd_read_bodies_noalloc(avbd_solver *s, float *out, uint32_t first, uint32_t count) {
  if (!s || (!out && count)) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  if ((uint64_t)first + count > s->n) return fail(s, AVBD_ERR_INVALID, "body range out of bounds");
  if (count > AVBD_READ_NOALLOC_MAX)
    return fail(s, AVBD_ERR_INVALID, "avbd_read_bodies_noalloc reads at most " + std::to_string(AVBD_READ_NOALLOC_MAX) +
                                         " bodies per call (asked for " + std::to_string(count) + ")");
  if (!count) return AVBD_OK;
  if (!s->readScratch) return fail(s, AVBD_ERR_DEVICE, "the readback scratch has no address on this device");
  k_bodies_to_records<<<blocksFor(count, 256), 256, 0, s->stream>>>(s->dev(), s->readScratch, first, count);
  hipMemcpyAsync(out, s->readScratch, 160ull * count, hipMemcpyDeviceToHost, s->stream);
  return avbd_sync(s);
}
int avbd_write_bodies(avbd_solver *s, const float *in, uint32_t first, uint32_t count) {
  if (!s || (!in && count)) return AVBD_ERR_INVALID;
  if ((uint64_t)first + count > s->n) return fail(s, AVBD_ERR_INVALID, "body range out of bounds");
  for (uint64_t k = 0; k < 40ull * count; k++)
    if (!std::isfinite(in[k]) && (k % 40) != 27) return fail(s, AVBD_ERR_INVALID, "body record holds a non-finite value");
  if (!count) return AVBD_OK;
  if (int rc = checkShapes(s, in, count)) return rc;
  if (int rc = ensureStaging(s, 160ull * count)) return rc;
  s->writeRecords(in, first, count, s->seeding);
  return avbd_sync(s);
}
int avbd_read_joints(avbd_solver *s, float *out, uint32_t first, uint32_t count) {
  if (!s || (!out && count)) return AVBD_ERR_INVALID;
  if ((uint64_t)first + count > s->nj) return fail(s, AVBD_ERR_INVALID, "joint range out of bounds");
  if (!count) return AVBD_OK;
  hipMemcpyAsync(out, s->joints + 10ull * first, 160ull * count, hipMemcpyDeviceToHost, s->stream);
  return avbd_sync(s);
}
int avbd_read_counters(avbd_solver *s, avbd_counters *out) {
  if (!s || !out) return AVBD_ERR_INVALID;
  uint32_t c[C_WORDS];
  hipMemcpyAsync(c, s->counters, sizeof(c), hipMemcpyDeviceToHost, s->stream);
  if (int rc = avbd_sync(s)) return rc;
  out->pairs = c[C_PAIRS];
  out->contacts = c[C_CONTACTS];
  out->manifolds = c[C_MANIFOLDS];
  s->noteBarrier(c);
  out->overflow = c[C_OVERFLOW] & ~AVBD_OVERFLOW_BARRIER;
  if (s->barrierGiveUps) out->overflow |= AVBD_OVERFLOW_BARRIER;
  s->barrierGiveUps = 0;
  out->clashes = c[C_CLASHES];
  out->colors = c[C_NUM_COLORS];
  out->color_cap = s->fixedColors ? s->fixedCap : s->hp.color_cap;
  out->step = s->stepCount;
  out->clashing_colors[0] = c[C_CLASHING];
  out->clashing_colors[1] = c[C_CLASHING + 1];
  return AVBD_OK;
}
int avbd_read_contacts(avbd_solver *s, uint32_t *man, uint32_t mcap, uint32_t *con, uint32_t ccap) {
  if (!s) return AVBD_ERR_INVALID;
  avbd_counters c;
  if (int rc = avbd_read_counters(s, &c)) return rc;
  uint32_t nm = std::min(c.manifolds, s->manCap), nc = std::min(c.contacts, s->conCap);
  if (nm > mcap || nc > ccap) return fail(s, AVBD_ERR_CAPACITY, "output arrays too small");
  int last = 1 - s->parity;  // the buffers the last step wrote
  std::vector<uint4> ids(nm);
  std::vector<float4> geo(nm);
  if (nm) {
    hipMemcpyAsync(ids.data(), s->mIds[last], nm * 16ull, hipMemcpyDeviceToHost, s->stream);
    hipMemcpyAsync(geo.data(), s->mGeo[last], nm * 16ull, hipMemcpyDeviceToHost, s->stream);
  }
  if (nc && con) hipMemcpyAsync(con, s->con[last], nc * 64ull, hipMemcpyDeviceToHost, s->stream);
  if (int rc = avbd_sync(s)) return rc;
  for (uint32_t m = 0; m < nm && man; m++) {
    memcpy(man + 8ull * m, &ids[m], 16);
    memcpy(man + 8ull * m + 4, &geo[m], 16);
  }
  return (int)nm;
}
int avbd_read_pairs(avbd_solver *s, uint32_t *out, uint32_t cap) {
  if (!s) return AVBD_ERR_INVALID;
  avbd_counters c;
  if (int rc = avbd_read_counters(s, &c)) return rc;
  uint32_t np = std::min(c.pairs, s->pairCap);
  if (np > cap) return fail(s, AVBD_ERR_CAPACITY, "output array too small");
  if (np) hipMemcpyAsync(out, s->pairs, np * 8ull, hipMemcpyDeviceToHost, s->stream);
  if (int rc = avbd_sync(s)) return rc;
  return (int)np;
}
int avbd_read_colors(avbd_solver *s, uint32_t *state, uint32_t count, uint32_t *starts) {
  if (!s || count > s->n) return AVBD_ERR_INVALID;
  if (count && state) hipMemcpyAsync(state, s->colA, count * 4ull, hipMemcpyDeviceToHost, s->stream);
  if (starts) hipMemcpyAsync(starts, s->colorStart, 65 * 4, hipMemcpyDeviceToHost, s->stream);
  return avbd_sync(s);
}

int avbd_read_refs(avbd_solver *s, float *out, uint32_t first, uint32_t count) {
  if (!s || (!out && count)) return AVBD_ERR_INVALID;
  if ((uint64_t)first + count > s->n) return fail(s, AVBD_ERR_INVALID, "body range out of bounds");
  if (!count) return AVBD_OK;
  hipMemcpy2DAsync(out, 32, s->refPos + first, 16, 16, count, hipMemcpyDeviceToHost, s->stream);
  hipMemcpy2DAsync(out + 4, 32, s->refRot + first, 16, 16, count, hipMemcpyDeviceToHost, s->stream);
  return avbd_sync(s);
}
int avbd_seed(avbd_solver *s, const float *bodies, uint32_t body_count, const float *jrec, uint32_t joint_count, const uint32_t *man,
              uint32_t nm, const uint32_t *con, uint32_t nc) {
  if (!s) return AVBD_ERR_INVALID;
  if (body_count != s->n) return fail(s, AVBD_ERR_INVALID, "seedFrom: bodies differ from the solver's");
  if (joint_count != s->nj) return fail(s, AVBD_ERR_INVALID, "seedFrom: joints differ from the reference");
  if (nm > s->manCap || nc > s->conCap) return fail(s, AVBD_ERR_CAPACITY, "seedFrom: too many contacts");
  for (uint32_t m = 0; m < nm; m++) {
    const uint32_t *o = man + 8ull * m;
    if (o[0] >= s->n || o[1] >= s->n || (uint64_t)o[2] + (o[3] & 15u) > nc) return fail(s, AVBD_ERR_INVALID, "seedFrom: manifold out of range");
  }
  s->seeding = true;  // seedFrom keeps the colour state (solver.ts:1324)
  int wrc = avbd_write_bodies(s, bodies, 0, body_count);
  s->seeding = false;
  if (wrc) return wrc;
  if (joint_count) {
    std::vector<float> jr(jrec, jrec + 40ull * joint_count);
    for (auto &x : jr) x = finiteF(x);
    hipMemcpyAsync(s->joints, jr.data(), jr.size() * 4, hipMemcpyHostToDevice, s->stream);
  }
  int prev = 1 - s->parity;  // the next step reads last step's contacts from these
  std::vector<uint4> ids(nm);
  std::vector<float4> geo(nm);
  for (uint32_t m = 0; m < nm; m++) { memcpy(&ids[m], man + 8ull * m, 16); memcpy(&geo[m], man + 8ull * m + 4, 16); }
  if (nm) {
    hipMemcpyAsync(s->mIds[prev], ids.data(), nm * 16ull, hipMemcpyHostToDevice, s->stream);
    hipMemcpyAsync(s->mGeo[prev], geo.data(), nm * 16ull, hipMemcpyHostToDevice, s->stream);
  }
  if (nc) hipMemcpyAsync(s->con[prev], con, nc * 64ull, hipMemcpyHostToDevice, s->stream);
  uint32_t c[C_WORDS] = {0};
  c[C_CONTACTS] = nc;
  c[C_MANIFOLDS] = nm;
  c[C_STEP] = s->stepCount;
  hipMemcpyAsync(s->counters, c, sizeof(c), hipMemcpyHostToDevice, s->stream);
  s->manifoldsValid = false;
  return avbd_sync(s);
}
int avbd_set_fixed_colors(avbd_solver *s, const uint32_t *colors, uint32_t count) {
  if (!s) return AVBD_ERR_INVALID;
  if (!colors) { s->fixedColors = false; return AVBD_OK; }
  if (count != s->n) return fail(s, AVBD_ERR_INVALID, "fixedColors: one colour per body");
  uint32_t cap = 1;
  for (uint32_t i = 0; i < count; i++) {
    if (colors[i] != NO_COLOR && colors[i] >= MAX_COLORS) return fail(s, AVBD_ERR_INVALID, "fixedColors: colour >= 64");
    if (colors[i] != NO_COLOR) cap = std::max(cap, colors[i] + 1);
  }
  hipMemcpyAsync(s->fixedDev, colors, count * 4ull, hipMemcpyHostToDevice, s->stream);
  s->fixedColors = true;
  s->fixedCap = cap;
  return avbd_sync(s);
}
int avbd_set_color_state(avbd_solver *s, const uint32_t *state, uint32_t count) {
  if (!s || !state || count != s->n) return AVBD_ERR_INVALID;
  hipMemcpyAsync(s->colA, state, count * 4ull, hipMemcpyHostToDevice, s->stream);
  return avbd_sync(s);
}
int avbd_add_ignore_pairs(avbd_solver *s, const uint32_t *pairs, uint32_t count) {
  if (!s || (!pairs && count)) return AVBD_ERR_INVALID;
  for (uint32_t k = 0; k < count; k++) {
    uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
    if (a >= s->n || b >= s->n || a == b) return fail(s, AVBD_ERR_INVALID, "IgnoreCollision: bodies out of range");
  }
  for (uint32_t k = 0; k < count; k++) {
    uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
    s->ignorePairs.push_back({std::max(a, b), std::min(a, b), IGNORE_SLOT});
  }
  s->staticsDirty = true;
  s->refreshStatics();
  s->manifoldsValid = false;
  return avbd_sync(s);
}
int avbd_set_filters(avbd_solver *s, const uint32_t *groups, const uint32_t *with, uint32_t count) {
  if (!s || !groups || !with || count != s->n) return AVBD_ERR_INVALID;
  std::vector<uint32_t> f(2ull * count);
  for (uint32_t i = 0; i < count; i++) { f[2 * i] = groups[i]; f[2 * i + 1] = with[i]; }
  hipMemcpyAsync(s->filters, f.data(), f.size() * 4, hipMemcpyHostToDevice, s->stream);
  s->manifoldsValid = false;
  return avbd_sync(s);
}
int avbd_set_hulls(avbd_solver *s, const uint32_t *words, uint32_t count) {
  if (!s || (!words && count)) return AVBD_ERR_INVALID;
  if (count % 4 != 0) return fail(s, AVBD_ERR_INVALID, "hull buffer is whole vec4u words");
  if (count > s->hullCap) return fail(s, AVBD_ERR_CAPACITY, "hull storage: " + std::to_string(count) + " words > capacity " + std::to_string(s->hullCap));
  if (count) hipMemcpyAsync(s->hulls, words, count * 4ull, hipMemcpyHostToDevice, s->stream);
  return avbd_sync(s);
}

int avbd_adapt(avbd_solver *s, const avbd_counters *c) {
  if (!s || !c) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "solver destroyed");
  try {
    s->adapt(*c);
    s->adaptStorage(*c);
  } catch (const std::exception &e) {
    return fail(s, AVBD_ERR_DEVICE, e.what());
  }
  return AVBD_OK;
}
int avbd_get_capacity(const avbd_solver *s, avbd_capacity *out) {
  if (!s || !out) return AVBD_ERR_INVALID;
  out->bodies = s->bodyCap; out->joints = s->jointCap; out->pairs = s->pairCap; out->manifolds = s->manCap;
  out->contacts = s->conCap; out->hull_words = s->hullCap;
  return AVBD_OK;
}
int avbd_broadphase_info(const avbd_solver *s, float *cell_size, float *fine_cell_size, float *fine_radius, uint32_t *coarse_movers,
                         uint32_t *large_count) {
  if (!s) return AVBD_ERR_INVALID;
  if (cell_size) *cell_size = s->cellSize;
  if (fine_cell_size) *fine_cell_size = s->hasFine ? s->fineCell : s->cellSize;
  if (fine_radius) *fine_radius = s->fineRadius;
  if (coarse_movers) *coarse_movers = s->coarseMovers ? 1u : 0u;
  if (large_count) *large_count = (uint32_t)s->large.size();
  return AVBD_OK;
}
int avbd_set_schedule(avbd_solver *s, const avbd_schedule *sch) {
  if (!s || !sch) return AVBD_ERR_INVALID;
  if (s->applySchedule(*sch) != AVBD_OK) return fail(s, AVBD_ERR_INVALID, "schedule: not a co-resident cooperative grid on this device");
  return AVBD_OK;
}
int avbd_get_schedule(const avbd_solver *s, avbd_schedule *sch) {
  if (!s || !sch) return AVBD_ERR_INVALID;
  *sch = s->schedule();
  return AVBD_OK;
}
int avbd_set_barrier_polls(avbd_solver *s, uint32_t polls) {
  if (!s) return AVBD_ERR_INVALID;
  s->barrierPolls = polls ? polls : (1u << 22);
  return AVBD_OK;
}
/** Upload `count` records of `words` 32-bit words and check their body words (one per record at
 *  offset `bodyWord`): in range and, when `unique`, no body twice. */
static int stageCommands(avbd_solver *s, const void *rec, uint32_t count, uint32_t words, uint32_t bodyWord, bool unique, const char *what) {
  if (!s) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "solver destroyed");
  if (count && !rec) return AVBD_ERR_INVALID;
  const uint32_t *w = (const uint32_t *)rec;
  std::vector<uint32_t> seen;
  for (uint32_t k = 0; k < count; k++) {
    uint32_t i = w[(size_t)k * words + bodyWord];
    if (i >= s->n) return fail(s, AVBD_ERR_INVALID, std::string(what) + ": no body slot " + std::to_string(i));
    if (unique) seen.push_back(i);
  }
  std::sort(seen.begin(), seen.end());
  if (std::adjacent_find(seen.begin(), seen.end()) != seen.end()) return fail(s, AVBD_ERR_INVALID, std::string(what) + ": a body commanded twice");
  if (!count) return AVBD_OK;
  if (int rc = ensureStaging(s, 4ull * words * count)) return rc;
  hipMemcpyAsync(s->staging, rec, 4ull * words * count, hipMemcpyHostToDevice, s->stream);
  return AVBD_OK;
}
int avbd_push(avbd_solver *s, const float *pushes, uint32_t count) {
  if (int rc = stageCommands(s, pushes, count, 12, 3, true, "push")) return rc;
  if (count) k_push<<<blocksFor(count, 64), 64, 0, s->stream>>>(s->dev(), (const float4 *)s->staging, count);
  return avbd_sync(s);
}
int avbd_move(avbd_solver *s, const float *moves, uint32_t count) {
  if (int rc = stageCommands(s, moves, count, 12, 3, true, "move")) return rc;
  if (count) k_move<<<blocksFor(count, 64), 64, 0, s->stream>>>(s->dev(), (const float4 *)s->staging, count);
  return avbd_sync(s);
}
int avbd_reset(avbd_solver *s, const uint32_t *indices, uint32_t count) {
  if (int rc = stageCommands(s, indices, count, 1, 0, false, "reset")) return rc;
  if (count) k_reset<<<blocksFor(count, 64), 64, 0, s->stream>>>(s->dev(), (const uint32_t *)s->staging, count);
  return avbd_sync(s);
}
int avbd_watch_create(avbd_solver *s, uint32_t pairs, uint32_t capacity) {
  if (!s) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "solver destroyed");
  if (capacity == 0) return fail(s, AVBD_ERR_INVALID, "contact events: capacity must be > 0");
  try {
    if (hipStreamSynchronize(s->stream) != hipSuccess) throw std::runtime_error("device error before the contact watch");
    if (!s->watchWords) {
      s->watchWords = s->alloc<uint32_t>(s->bodyCap);  // zeroed: no body watched
    }
    s->freeMem(s->watchMem);
    uint64_t want = std::max<uint64_t>(64, 2ull * pairs), slots = 1;
    while (slots < want) slots <<= 1;
    if (slots > (1u << 28)) return fail(s, AVBD_ERR_INVALID, "contact events: too many pairs");
    WatchBuf w{};
    w.watch = s->watchWords;
    w.setSlots = (uint32_t)slots;
    w.capacity = capacity;
    w.table = s->adjAlloc<uint32_t>(s->watchMem, WATCH_HEAD + 4 * slots);
    w.events = s->adjAlloc<float4>(s->watchMem, 3ull * capacity);
    k_fill<<<blocksFor(WATCH_HEAD + 4 * slots, 256), 256, 0, s->stream>>>(w.table, (uint32_t)(WATCH_HEAD + 4 * slots), 0u, nullptr);
    s->watchBuf = w;
    s->watchFlip = 0;
    return avbd_sync(s);
  } catch (const std::exception &e) {
    return fail(s, AVBD_ERR_DEVICE, e.what());
  }
}
int avbd_watch_set(avbd_solver *s, const uint32_t *indices, const uint32_t *ids, const uint8_t *on, uint32_t count) {
  if (!s || (count && (!indices || !ids || !on))) return AVBD_ERR_INVALID;
  if (!s->watchBuf.table) return fail(s, AVBD_ERR_INVALID, "contact events: no watch (avbd_watch_create first)");
  std::vector<uint32_t> words(2ull * count);
  for (uint32_t k = 0; k < count; k++) {
    if (indices[k] >= s->n) return fail(s, AVBD_ERR_INVALID, "contact events: body index out of range");
    words[k] = indices[k];
    words[count + k] = (ids[k] & 0x7fffffffu) | (on[k] ? 0x80000000u : 0u);
  }
  if (!count) return AVBD_OK;
  if (int rc = ensureStaging(s, 8ull * count)) return rc;
  uint32_t *st = (uint32_t *)s->staging;
  hipMemcpyAsync(st, words.data(), 8ull * count, hipMemcpyHostToDevice, s->stream);
  k_watch_scatter<<<blocksFor(count, 256), 256, 0, s->stream>>>(s->watchWords, st, st + count, count, s->n);
  return avbd_sync(s);
}
int avbd_watch_run(avbd_solver *s, uint32_t pairs, uint32_t step, float dt) {
  if (!s) return AVBD_ERR_INVALID;
  if (!s->watchBuf.table) return fail(s, AVBD_ERR_INVALID, "contact events: no watch (avbd_watch_create first)");
  WatchBuf &w = s->watchBuf;
  s->watchFlip = 1 - s->watchFlip;
  uint32_t prev = s->watchFlip ? 0u : w.setSlots, now = s->watchFlip ? w.setSlots : 0u;
  k_fill<<<blocksFor(2ull * w.setSlots, 256), 256, 0, s->stream>>>(w.table + WATCH_HEAD + 2ull * now, 2 * w.setSlots, 0u, nullptr);
  int last = 1 - s->parity;  // the buffers the last step wrote
  uint32_t scan = std::min(pairs, s->manCap);
  Dev d = s->dev();
  k_watch_touching<<<blocksFor(std::max(scan, 1u), 64), 64, 0, s->stream>>>(d, w, s->mIds[last], s->mGeo[last], s->con[last], prev, now,
                                                                           scan, step, dt);
  k_watch_ended<<<blocksFor(w.setSlots, 64), 64, 0, s->stream>>>(w, prev, now, step);
  return avbd_sync(s);
}
int avbd_watch_read(avbd_solver *s, avbd_contact_event *out, uint32_t capacity, uint32_t *count, uint32_t *dropped, uint32_t *full) {
  if (!s || !count || !dropped || !full || (capacity && !out)) return AVBD_ERR_INVALID;
  if (!s->watchBuf.table) return fail(s, AVBD_ERR_INVALID, "contact events: no watch (avbd_watch_create first)");
  WatchBuf &w = s->watchBuf;
  uint32_t head[WATCH_HEAD];
  hipMemcpyAsync(head, w.table, sizeof(head), hipMemcpyDeviceToHost, s->stream);
  if (int rc = avbd_sync(s)) return rc;
  uint32_t n = std::min(head[0], w.capacity), copied = std::min(n, capacity);
  if (copied) hipMemcpyAsync(out, w.events, 48ull * copied, hipMemcpyDeviceToHost, s->stream);
  k_fill<<<1, 64, 0, s->stream>>>(w.table, WATCH_HEAD, 0u, nullptr);
  *count = copied;
  *dropped = head[0] - copied;
  *full = head[1] != 0 ? 1u : 0u;
  return avbd_sync(s);
}
int avbd_set_sparse_tail(avbd_solver *s, uint32_t first_colour) {
  if (!s) return AVBD_ERR_INVALID;
  if (first_colour == AVBD_TAIL_AUTO) { s->tailAuto = true; s->tailStart = MAX_COLORS; return AVBD_OK; }
  if (first_colour > MAX_COLORS) return fail(s, AVBD_ERR_INVALID, "sparse tail: first colour must be <= 64 or AVBD_TAIL_AUTO");
  s->tailAuto = false;
  s->tailStart = first_colour;
  return AVBD_OK;
}
int avbd_autotune(avbd_solver *s, uint32_t steps, uint32_t substeps, avbd_schedule *best, float *best_ms) {
  if (!s || substeps == 0 || steps == 0) return AVBD_ERR_INVALID;
  const avbd_schedule cands[] = {{0, 512, 1}, {1, 256, 1}, {1, 512, 1}, {1, 1024, 1}, {1, 256, 2}, {1, 512, 2},
                                 {2, 64, 4}, {2, 64, 8}, {2, 64, 12}, {3, 64, 4}, {3, 64, 8}, {3, 64, 12},
                                 {4, 64, 8}, {4, 64, 12}};
  const int nc = (int)(sizeof(cands) / sizeof(cands[0]));
  float bestT[sizeof(cands) / sizeof(cands[0])];
  for (int k = 0; k < nc; k++) bestT[k] = INFINITY;
  avbd_schedule keep = s->schedule();
  for (int pass = 0; pass < 2; pass++)
    for (int k = 0; k < nc; k++) {
      if (s->applySchedule(cands[k]) != AVBD_OK) continue;
      float ms = 0;
      if (int rc = avbd_time_steps(s, steps, substeps, &ms)) return rc;
      bestT[k] = std::min(bestT[k], ms / steps);
    }
  int pick = -1;
  for (int k = 0; k < nc; k++) if (bestT[k] < INFINITY && (pick < 0 || bestT[k] < bestT[pick])) pick = k;
  if (pick < 0) { s->applySchedule(keep); return fail(s, AVBD_ERR_DEVICE, "autotune: no schedule ran"); }
  // Threads per body, under the per-colour or dataflow schedule when it won: the given lane rule, eight threads
  // for every colour above the rule's last size, each of the two without sixteen-thread colours,
  // and eight threads with sixteen up to four times the last size; the fastest is kept (results
  // differ only in the order the lanes' sums are added). Sixteen threads pay where small colours
  // hold bodies with long adjacency walks, and cost an extra reduction step where they do not.
  if (cands[pick].cooperative != 1) {
    const uint32_t *g = s->hp.primal_lanes;
    const uint32_t wide = g[3] <= (1u << 29) ? 4u * g[3] : g[3];
    const uint32_t rules[5][4] = {{g[0], g[1], g[2], g[3]}, {1u << 31, 1u << 31, 1u << 31, g[3]}, {g[0], g[1], g[2], 0u}, {1u << 31, 1u << 31, 1u << 31, 0u},
                                  {1u << 31, 1u << 31, 1u << 31, wide}};
    const int nr = 5;
    uint32_t given[4] = {g[0], g[1], g[2], g[3]}, bestRule = 0;
    float t[5] = {INFINITY, INFINITY, INFINITY, INFINITY, INFINITY};
    s->applySchedule(cands[pick]);
    // every rule timed again, passes interleaved: the scene keeps changing while it is timed
    for (int pass = 0; pass < 2; pass++)
      for (int r = 0; r < nr; r++) {
        bool seen = false;
        for (int q = 0; q < r; q++) seen = seen || std::equal(rules[q], rules[q] + 4, rules[r]);
        if (seen) continue;
        std::copy(rules[r], rules[r] + 4, s->hp.primal_lanes);
        float ms = 0;
        if (int rc = avbd_time_steps(s, steps, substeps, &ms)) { std::copy(given, given + 4, s->hp.primal_lanes); return rc; }
        t[r] = std::min(t[r], ms / steps);
      }
    for (int r = 1; r < nr; r++) if (t[r] < t[bestRule]) bestRule = (uint32_t)r;
    std::copy(rules[bestRule], rules[bestRule] + 4, s->hp.primal_lanes);
    bestT[pick] = t[bestRule];
  }
  s->applySchedule(cands[pick]);
  if (best) *best = cands[pick];
  if (best_ms) *best_ms = bestT[pick];
  return AVBD_OK;
}

int avbd_profile_step(avbd_solver *s, const avbd_step_opts *opts, avbd_profile *out) {
  if (!s || !out) return AVBD_ERR_INVALID;
  hipEvent_t e[5];
  for (auto &x : e) hipEventCreate(&x);
  s->stepCount++;
  bool reuse = opts && opts->reuse_manifolds && s->manifoldsValid;
  hipEventRecord(e[0], s->stream);
  s->collision(reuse, false);  // a lone step: the adjacency is rebuilt
  hipEventRecord(e[1], s->stream);
  // adjacency and colouring share topology(); split them at the colouring launch
  s->topology();
  hipEventRecord(e[3], s->stream);
  s->solve();
  hipEventRecord(e[4], s->stream);
  s->parity = 1 - s->parity;
  s->manifoldsValid = true;
  int rc = avbd_sync(s);
  float a = 0, c = 0, d = 0;
  hipEventElapsedTime(&a, e[0], e[1]);
  hipEventElapsedTime(&c, e[1], e[3]);
  hipEventElapsedTime(&d, e[3], e[4]);
  out->collision = a;
  out->adjacency = 0;
  out->coloring = c;
  out->solve = d;
  out->total = a + c + d;
  for (auto &x : e) hipEventDestroy(x);
  return rc;
}
int avbd_time_steps(avbd_solver *s, uint32_t steps, uint32_t substeps, float *ms) {
  if (!s || !ms || substeps == 0) return AVBD_ERR_INVALID;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, s->stream);
  for (uint32_t k = 0; k < steps; k++)
    for (uint32_t u = 0; u < substeps; u++) s->step(u > 0, u > 0);
  hipEventRecord(e1, s->stream);
  int rc = avbd_sync(s);
  hipEventElapsedTime(ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return rc;
}
int avbd_profile_kernels(avbd_solver *s, uint32_t steps, uint32_t substeps, avbd_kernel_stats *out) {
  if (!s || !out || substeps == 0) return AVBD_ERR_INVALID;
  memset(out, 0, sizeof(*out));
  s->timer.on = true;
  for (uint32_t k = 0; k < steps; k++)
    for (uint32_t u = 0; u < substeps; u++) s->step(u > 0, u > 0);
  s->timer.on = false;
  int rc = avbd_sync(s);
  out->solve_blocks = s->coopBlocks;
  for (auto &m : s->timer.marks) {
    float t = 0;
    hipEventElapsedTime(&t, m.second.first, m.second.second);
    switch (m.first) {
      case K_PRIMAL: out->primal_ms += t; out->primal_launches++; break;
      case K_DUAL: out->dual_ms += t; out->dual_launches++; break;
      case K_NARROW: out->narrowphase_ms += t; out->narrowphase_launches++; break;
      case K_BROAD: out->broadphase_ms += t; out->broadphase_launches++; break;
      case K_SOLVE: out->solve_ms += t; out->solve_launches++; break;
      case K_TAIL: out->tail_ms += t; out->tail_launches++; break;
      default: out->other_ms += t; break;
    }
    hipEventDestroy(m.second.first);
    hipEventDestroy(m.second.second);
  }
  s->timer.marks.clear();
  return rc;
}

/* Differentiable rollout: `steps` detecting steps with checkpoints every `checkpoint_every`
 * steps (0: record every step's tape during the rollout itself, no re-run in the backward). */
int avbd_forward(avbd_solver *s, uint32_t steps, uint32_t checkpoint_every) {
  if (!s) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "solver destroyed");
  if (s->backend) return fail(s, AVBD_ERR_INVALID, "the adjoint covers the built-in collision only (an SDF bridge is installed)");
  for (uint32_t j = 0; j < s->nj; j++)
    if (s->jinfoHost[4 * j] != (int)T_NONE) return fail(s, AVBD_ERR_INVALID, "the adjoint covers contact-only scenes (joint " + std::to_string(j) + ")");
  if (steps == 0) return fail(s, AVBD_ERR_INVALID, "steps must be > 0");
  NoGrowScope scope(s);
  try {
    s->fwdValid = false;
    s->fwdTail = false;
    s->fwdWhy = "forward rollout incomplete";
    s->barrierFailed = false;
    uint32_t K = checkpoint_every;
    uint32_t segments = K ? (steps + K - 1) / K : 0;
    if (hipStreamSynchronize(s->stream) != hipSuccess) throw std::runtime_error("device error before the forward");
    s->ensureCheckpoints(segments + 1);
    k_fill<<<1, 64, 0, s->stream>>>(s->counters + C_STICKY, 3, 0u, nullptr);
    if (K == 0) {
      s->ensureTape(s->tapes[0], steps, s->conCap, s->manCap);
      s->recordSteps(s->tapes[0], steps);
    } else {
      // the last segment records as it runs: re-running it in the backward would overlap nothing
      uint32_t tail = (segments - 1) * K;
      avbd_solver::TapeBuf &tb = s->tapes[(segments - 1) % 2];
      s->ensureTape(tb, K, s->conCap, s->manCap);
      for (uint32_t t = 0; t < tail; t++) {
        if (t % K == 0) s->saveCheckpoint(s->ckpts[t / K]);
        s->step(false);
      }
      s->saveCheckpoint(s->ckpts[segments - 1]);
      s->recordSteps(tb, steps - tail);
      s->copySizes(tb, steps - tail, s->stream);
      s->fwdTail = true;
    }
    s->saveCheckpoint(s->ckpts[s->ckptCount - 1]);
    uint32_t c[C_WORDS];
    hipMemcpyAsync(c, s->counters, sizeof(c), hipMemcpyDeviceToHost, s->stream);
    if (int rc = avbd_sync(s)) return rc;
    uint32_t sticky = c[C_STICKY] | c[C_OVERFLOW] | (c[C_CLASHES] ? 0x10000u : 0u);
    if (c[C_BARRIER] || s->barrierFailed) {
      s->noteBarrier(c);
      s->fwdWhy = "the cooperative solve's grid barrier gave up waiting during the rollout";
      return AVBD_OK;
    }
    s->fwdMaxContacts = std::max(c[C_MAX_CONTACTS], c[C_CONTACTS]);
    s->fwdMaxManifolds = std::max(c[C_MAX_MANIFOLDS], c[C_MANIFOLDS]);
    s->fwdSteps = steps; s->fwdEvery = K; s->fwdEndStep = s->stepCount;
    if (sticky & 0xffffu) { s->fwdWhy = "capacity overflow during the rollout: raise the pair/manifold/contact capacities"; return AVBD_OK; }
    if (sticky & 0x10000u) { s->fwdWhy = "clashing colours during the rollout: the adjoint needs a proper colouring (raise color_cap)"; return AVBD_OK; }
    s->fwdValid = true;
    s->fwdWhy.clear();
    return AVBD_OK;
  } catch (const std::exception &e) {
    s->freeAdjoint();
    return fail(s, AVBD_ERR_DEVICE, e.what());
  }
}
/* Reverse sweep over the last avbd_forward: adjoint of the final state (16 doubles per body:
 * pos xyz, previous up-velocity, rot xyzw, vel xyz, 0, angvel xyz, 0) -> adjoint of the initial
 * state (same layout) and of every body's friction. Leaves the solver at the rollout's end.
 * Segments are re-run from their checkpoints on the solver's stream while the reverse sweep of
 * the segment after runs on a second stream (two tape buffers). */
int avbd_backward(avbd_solver *s, const double *d_final, double *d_initial, double *d_friction) {
  if (!s || !d_final || !d_initial) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "solver destroyed");
  if (!s->fwdValid) return fail(s, AVBD_ERR_INVALID, s->fwdWhy);
  if (s->stepCount != s->fwdEndStep) return fail(s, AVBD_ERR_INVALID, "the solver stepped after avbd_forward");
  NoGrowScope scope(s);
  try {
    uint32_t K = s->fwdEvery, steps = s->fwdSteps, n = s->n;
    s->ensureReverseStream();
    hipStream_t R = s->revStream, S = s->stream;
    s->ensureAdjoint(K ? s->fwdMaxContacts : s->conCap);
    uint32_t tailBuf = K ? ((steps + K - 1) / K - 1) % 2 : 0;
    for (uint32_t k = 0; K && k < 2; k++)
      if (!(s->fwdTail && k == tailBuf)) s->ensureTape(s->tapes[k], K, s->fwdMaxContacts, s->fwdMaxManifolds);
    AdjBuf &a = s->adj;
    if (hipStreamSynchronize(S) != hipSuccess) throw std::runtime_error("device error before the backward");
    hipMemcpyAsync(a.state, d_final, 128ull * n, hipMemcpyHostToDevice, R);
    double *zero[4] = {a.lamb, a.lambPrev, a.c0b, a.mub};
    size_t sizes[4] = {4ull * s->adjConCap, 4ull * s->adjConCap, 4ull * s->adjConCap, (size_t)s->adjConCap};
    for (int k = 0; k < 4; k++) s->zeroF64(zero[k], sizes[k], R);
    s->zeroF64(a.mu, n, R);
    if (K == 0) {
      s->copySizes(s->tapes[0], steps, S);
      if (hipStreamSynchronize(S) != hipSuccess) throw std::runtime_error("reading the tape's colour buckets failed");
      s->reverseSlots(s->tapes[0], steps, 0, R);
    } else {
      uint32_t segments = (steps + K - 1) / K;
      auto runSegment = [&](uint32_t g) {
        uint32_t len = std::min(K, steps - g * K);
        s->restoreCheckpoint(s->ckpts[g]);
        s->recordSteps(s->tapes[g % 2], len);
        s->copySizes(s->tapes[g % 2], len, S);
        hipEventRecord(s->recorded[g % 2], S);
      };
      // Segment g's reverse is enqueued on the reverse stream before segment g-1's re-run is
      // enqueued on the solver stream (whose host side may block: adaptive colouring reads the
      // counters), so the two run side by side; the re-run into a tape buffer waits for the
      // reverse that read it last.
      if (s->fwdTail) hipEventRecord(s->recorded[(segments - 1) % 2], S);  // recorded by the forward
      else runSegment(segments - 1);
      s->fwdTail = false;  // the sweep below overwrites that buffer: a second backward re-runs the segment
      for (uint32_t g = segments; g-- > 0;) {
        uint32_t first = g * K, len = std::min(K, steps - first);
        avbd_solver::TapeBuf &tb = s->tapes[g % 2];
        if (hipEventSynchronize(s->recorded[g % 2]) != hipSuccess) throw std::runtime_error("re-running a segment failed");
        if (g + 1 < segments) {
          // λ adjoint of this segment's last step arrives in the checkpoint's contact order
          const avbd_solver::Checkpoint &c = s->ckpts[g + 1];
          k_fill<<<blocksFor(s->hashSize, 256), 256, 0, R>>>(s->remapTable, s->hashSize, 0u, nullptr);
          uint32_t mb = blocksFor(s->manCap, 256);
          k_adj_remap_insert_counted<<<mb, 256, 0, R>>>(c.mIds, c.counters + C_MANIFOLDS, s->manCap, s->remapTable, s->hashSize - 1);
          s->zeroF64(a.lambPrev, 4ull * s->adjConCap, R);
          k_adj_remap_counted<<<blocksFor(tb.t.manCap, 256), 256, 0, R>>>(c.mIds, tb.t.mIds + (size_t)(len - 1) * tb.t.manCap,
                                                                         tb.t.counts + 4 * (len - 1), s->remapTable, s->hashSize - 1, a.lamb, a.lambPrev);
          std::swap(a.lamb, a.lambPrev);
          s->zeroF64(a.lambPrev, 4ull * s->adjConCap, R);
        }
        s->reverseSlots(tb, len, first, R);
        hipEventRecord(s->reversed[g % 2], R);
        if (g > 0) {
          // the other buffer was last read by the reverse of segment g+1
          if (g + 1 < segments) hipStreamWaitEvent(S, s->reversed[(g + 1) % 2], 0);
          runSegment(g - 1);
        }
      }
      hipStreamWaitEvent(S, s->reversed[0], 0);
      s->restoreCheckpoint(s->ckpts[s->ckptCount - 1]);
    }
    hipMemcpyAsync(d_initial, a.state, 128ull * n, hipMemcpyDeviceToHost, R);
    if (d_friction) hipMemcpyAsync(d_friction, a.mu, 8ull * n, hipMemcpyDeviceToHost, R);
    hipError_t e1 = hipStreamSynchronize(R), e2 = hipStreamSynchronize(S);
    if (e1 != hipSuccess || e2 != hipSuccess) return fail(s, AVBD_ERR_DEVICE, hipGetErrorString(e1 != hipSuccess ? e1 : e2));
    return AVBD_OK;
  } catch (const std::exception &e) {
    return fail(s, AVBD_ERR_DEVICE, e.what());
  }
}

}  // extern "C"

// ---- Scene edits (World.flush, world.ts:1861-1995) -------------------------------------------

namespace {
const float JBIG = 3.0e38f;
float finiteOr(float x) { return std::isinf(x) ? (x > 0 ? JBIG : -JBIG) : x; }
bool finite32(float x) { return std::isfinite(x); }
/** freshJoint (solver.ts:132-138): never breaks or yields, rest and starting rest no turn. */
void freshJoint(float *o) {
  std::fill(o, o + 40, 0.0f);
  o[11] = JBIG;  // J_LAM_LIN + 3: fracture
  o[23] = JBIG;  // J_YIELD
  o[31] = JBIG;  // J_BREAK_BEND
  o[35] = 1.0f;  // J_REST + 3
  o[39] = 1.0f;  // J_REST_START + 3
}
/** The torque arm of a joint between sizes sa and sb (the reference Joint constructor). */
float torqueArm(const float *sa, const float *sb) {
  double x = (double)sa[0] + sb[0], y = (double)sa[1] + sb[1], z = (double)sa[2] + sb[2];
  return (float)(x * x + y * y + z * z);
}
}  // namespace

/** The solver's slots for `n` new constraints (appendConstraints, solver.ts:1697-1727): released
 *  slots first, lowest first, then appended ones; grows the joint storage, or with fixed storage
 *  refuses before anything changes. */
static int takeSlots(avbd_solver *s, uint32_t n, std::vector<uint32_t> &slots) {
  if (s->fixedStorage() && (uint64_t)n > s->freeSlots.size() + (uint64_t)s->jointCap - s->nj)
    return fail(s, AVBD_ERR_CAPACITY, "StorageCapacityError: joints " + std::to_string(s->nj - s->freeSlots.size() + n) +
                                          " > capacity " + std::to_string(s->jointCap));
  std::sort(s->freeSlots.begin(), s->freeSlots.end(), [](uint32_t x, uint32_t y) { return x > y; });
  while (slots.size() < n && !s->freeSlots.empty()) { slots.push_back(s->freeSlots.back()); s->freeSlots.pop_back(); }
  uint32_t fresh = n - (uint32_t)slots.size();
  uint32_t cap = s->jointCap;
  while ((uint64_t)s->nj + fresh > cap) cap *= 2;
  try {
    if (cap > s->jointCap) {
      s->growJoints(cap);
      s->jinfoHost.resize(4ull * cap, 0);
      for (uint32_t j = s->nj; j < cap; j++) { s->jinfoHost[4ull * j + 1] = -1; s->jinfoHost[4ull * j + 2] = -1; }
    }
  } catch (const std::exception &e) {
    for (uint32_t x : slots) s->freeSlots.push_back(x);
    return fail(s, AVBD_ERR_DEVICE, e.what());
  }
  for (uint32_t k = 0; k < fresh; k++) slots.push_back(s->nj++);
  return AVBD_OK;
}
static bool isBody(const avbd_solver *s, int32_t i) { return i >= 0 && (uint32_t)i < s->n; }
static const float ZERO3[3] = {0.0f, 0.0f, 0.0f};
static const float *sizeOf(const avbd_solver *s, int32_t i) { return i < 0 ? ZERO3 : &s->sizeHost[3ull * i]; }

extern "C" {

int avbd_read_joint_info(const avbd_solver *s, int32_t *out, uint32_t first, uint32_t count) {
  if (!s || (!out && count)) return AVBD_ERR_INVALID;
  if ((uint64_t)first + count > s->nj) return AVBD_ERR_INVALID;
  std::copy(s->jinfoHost.begin() + 4ull * first, s->jinfoHost.begin() + 4ull * (first + count), out);
  return AVBD_OK;
}
int avbd_get_sizes(const avbd_solver *s, uint32_t *bodies, uint32_t *joints) {
  if (!s) return AVBD_ERR_INVALID;
  if (bodies) *bodies = s->n;
  if (joints) *joints = s->nj;
  return AVBD_OK;
}

int avbd_add_bodies(avbd_solver *s, const float *bodies, uint32_t count, int32_t *first) {
  if (!s || (!bodies && count)) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  if (first) *first = -1;
  if (s->worldStartHost.size() != 2) return fail(s, AVBD_ERR_INVALID, "addBodies: a batched solver's worlds are fixed");
  uint32_t k = std::min(count, s->bodyCap - s->n);
  if (k == 0) return 0;
  for (uint64_t i = 0; i < 40ull * k; i++)
    if (!std::isfinite(bodies[i]) && (i % 40) != 27) return fail(s, AVBD_ERR_INVALID, "body record holds a non-finite value");
  if (int rc = checkShapes(s, bodies, k)) return rc;
  if (int rc = ensureStaging(s, 160ull * k)) return rc;
  uint32_t at = s->n;
  s->n += k;
  s->writeRecords(bodies, at, k, false);
  s->worldStartHost[1] = s->n;
  hipMemcpyAsync(s->worldStart + 1, &s->worldStartHost[1], 4, hipMemcpyHostToDevice, s->stream);
  s->staticsDirty = true;
  s->adjacencyHeld = false;
  if (int rc = avbd_sync(s)) return rc;
  if (s->aux) s->uploadAux(at, k);
  if (first) *first = (int32_t)at;
  return (int)k;
}

int avbd_patch_body(avbd_solver *s, uint32_t index, uint32_t fields, const float *record, int fixed, int marker) {
  if (!s || !record) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  if (index >= s->n) return fail(s, AVBD_ERR_INVALID, "patchBody: body index out of range");
  if (fields & ~15u) return fail(s, AVBD_ERR_INVALID, "patchBody: unknown field bits");
  const int lanes[4][2] = {{0, 3}, {4, 4}, {32, 3}, {36, 3}};
  for (int f = 0; f < 4; f++)
    if (fields & (1u << f))
      for (int c = 0; c < lanes[f][1]; c++)
        if (!std::isfinite(record[lanes[f][0] + c])) return fail(s, AVBD_ERR_INVALID, "patchBody: a patched field is not finite");
  if (fixed && !(std::isfinite(record[19]) && std::isfinite(record[20]) && std::isfinite(record[21]) && std::isfinite(record[22])))
    return fail(s, AVBD_ERR_INVALID, "patchBody: mass and moments must be finite");
  PatchRecord r;
  memcpy(r.r, record, 160);
  uint32_t f = fields;
  if (fixed) {
    f |= 16u;
    if (s->massHost[index] != record[19]) s->staticsDirty = true;
    s->massHost[index] = record[19];
    s->fill(s->colA + index, 1, NO_COLOR);
    s->fill(s->colB + index, 1, NO_COLOR);
  }
  if (marker) {
    f |= 32u;
    r.r[2].w = (float)(1u | s->lifeHost[index]);  // BODY_DIRTY | the body's lifecycle bits
  }
  k_patch_body<<<1, 64, 0, s->stream>>>(s->dev(), index, f, r);
  s->manifoldsValid = false;
  return avbd_sync(s);
}

int avbd_rewrite_fixed(avbd_solver *s, const uint32_t *indices, uint32_t count, const float *positions, const float *size,
                       float friction, const float *rotations) {
  if (!s || (count && (!indices || !positions)) || !size) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  s->manifoldsValid = false;
  if (!count) return AVBD_OK;
  for (uint32_t k = 0; k < count; k++) {
    if (indices[k] >= s->n) return fail(s, AVBD_ERR_INVALID, "rewriteFixed: body index out of range");
    for (int c = 0; c < 3; c++) if (!std::isfinite(positions[3ull * k + c])) return fail(s, AVBD_ERR_INVALID, "rewriteFixed: positions must be finite");
    if (rotations) for (int c = 0; c < 4; c++) if (!std::isfinite(rotations[4ull * k + c])) return fail(s, AVBD_ERR_INVALID, "rewriteFixed: rotations must be finite");
  }
  // what a Rigid of density 0 packs to, at rest (solver.ts:1145-1162)
  float t[40] = {0};
  t[3] = friction;
  t[7] = 1.0f;
  t[11] = 1.0f;  // a new occupant: no contact history
  t[15] = 1.0f;
  t[16] = size[0]; t[17] = size[1]; t[18] = size[2];
  t[23] = (float)std::sqrt((double)size[0] * size[0] / 4 + (double)size[1] * size[1] / 4 + (double)size[2] * size[2] / 4);
  t[31] = 1.0f;
  if (int rc = ensureStaging(s, 160ull * count)) return rc;
  for (uint32_t k = 0; k < count;) {
    uint32_t e = k + 1;
    while (e < count && indices[e] == indices[e - 1] + 1) e++;
    std::vector<float> run(40ull * (e - k));
    for (uint32_t i = k; i < e; i++) {
      float *o = run.data() + 40ull * (i - k);
      std::copy(t, t + 40, o);
      for (int c = 0; c < 3; c++) o[c] = positions[3ull * i + c];
      if (rotations) for (int c = 0; c < 4; c++) o[4 + c] = rotations[4ull * i + c];
    }
    s->writeRecords(run.data(), indices[k], e - k, false);
    if (int rc = avbd_sync(s)) return rc;
    k = e;
  }
  return AVBD_OK;
}

int avbd_append_joints(avbd_solver *s, const avbd_joint_spec *specs, uint32_t count, float fracture, int linear, uint32_t *slots) {
  if (!s || (!specs && count)) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  // every spec checked before a slot is taken (solver.ts:1637-1660)
  std::vector<std::array<float, 4>> rests(count);
  for (uint32_t k = 0; k < count; k++) {
    const avbd_joint_spec &j = specs[k];
    std::string at = "appendJoints: spec " + std::to_string(k) + ": ";
    if (!isBody(s, j.a)) return fail(s, AVBD_ERR_INVALID, at + "a must be the index of a body in the solver");
    if (!isBody(s, j.b)) return fail(s, AVBD_ERR_INVALID, at + "b must be the index of a body in the solver");
    bool hasRest = j.rest[0] != 0.0f || j.rest[1] != 0.0f || j.rest[2] != 0.0f || j.rest[3] != 0.0f;
    if (hasRest) {
      double l = std::sqrt((double)j.rest[0] * j.rest[0] + (double)j.rest[1] * j.rest[1] + (double)j.rest[2] * j.rest[2] + (double)j.rest[3] * j.rest[3]);
      if (!(l > 0) || !std::isfinite(l)) return fail(s, AVBD_ERR_INVALID, at + "a rotation is four finite numbers (x, y, z, w), not all zero");
      for (int c = 0; c < 4; c++) rests[k][c] = (float)(j.rest[c] / l);
    }
    float limit = std::isnan(j.fracture) ? fracture : j.fracture;
    if (!(limit >= 0.0f)) return fail(s, AVBD_ERR_INVALID, at + "fracture is a force of at least 0");
    if (hasRest && j.angular == 0.0f) return fail(s, AVBD_ERR_INVALID, at + "rest needs an angle lock to hold it: a ball joint (angular 0) has none");
    bool hasYield = !std::isnan(j.yield);
    if (!std::isnan(j.break_bend)) {
      if (!std::isinf(j.break_bend) && !(j.break_bend >= 0.0f && (double)j.break_bend <= 3.14159265358979323846))
        return fail(s, AVBD_ERR_INVALID, at + "breakBend is an angle from 0 to pi (rad)");
      if (!std::isinf(j.break_bend) && !(hasYield && !std::isinf(j.yield)))
        return fail(s, AVBD_ERR_INVALID, at + "breakBend limits how far a plastic joint bends: it needs a yield");
    }
    if (!hasYield) continue;
    if (!(j.yield >= 0.0f)) return fail(s, AVBD_ERR_INVALID, at + "yield is a force of at least 0");
    if (!std::isnan(j.angular)) return fail(s, AVBD_ERR_INVALID, at + "yield needs a rigid angle lock (no angular stiffness given)");
  }
  std::vector<uint32_t> got;
  if (int rc = takeSlots(s, count, got)) return rc;
  std::vector<float> recs(40ull * count);
  std::vector<int> ends(3ull * count);
  for (uint32_t k = 0; k < count; k++) {
    const avbd_joint_spec &j = specs[k];
    float *o = &recs[40ull * k];
    freshJoint(o);
    o[3] = JBIG;
    o[7] = std::isnan(j.angular) ? JBIG : finiteOr(j.angular);
    float limit = finiteOr(std::isnan(j.fracture) ? fracture : j.fracture);
    bool lin = j.linear < 0 ? linear != 0 : j.linear != 0;
    o[11] = lin ? -limit : limit;
    if (!std::isnan(j.yield)) o[23] = finiteOr(j.yield);
    if (!std::isnan(j.break_bend)) o[31] = finiteOr(j.break_bend);
    bool hasRest = j.rest[0] != 0.0f || j.rest[1] != 0.0f || j.rest[2] != 0.0f || j.rest[3] != 0.0f;
    if (hasRest) for (int c = 0; c < 4; c++) { o[32 + c] = rests[k][c]; o[36 + c] = rests[k][c]; }
    for (int c = 0; c < 3; c++) { o[24 + c] = j.rA[c]; o[28 + c] = j.rB[c]; }
    o[15] = torqueArm(sizeOf(s, j.a), sizeOf(s, j.b));
    ends[3 * k] = (int)T_JOINT; ends[3 * k + 1] = j.a; ends[3 * k + 2] = j.b;
    if (slots) slots[k] = got[k];
  }
  s->putJoints(got, recs, ends);
  return avbd_sync(s);
}

int avbd_append_joint(avbd_solver *s, int32_t a, int32_t b, const float *rA, const float *rB, float stiff_lin, float stiff_ang,
                      uint32_t *slot) {
  if (!s || !rA || !rB) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  if (!(a == -1 || isBody(s, a)) || !isBody(s, b)) return fail(s, AVBD_ERR_INVALID, "appendJoint: a must be -1 (the world) or a body, b a body");
  std::vector<uint32_t> got;
  if (s->fixedStorage()) {
    if (int rc = takeSlots(s, 1, got)) return rc;
  } else {  // appendJoint appends (released slots are for appendJoints), growing the storage
    std::vector<uint32_t> keep = s->freeSlots;
    s->freeSlots.clear();
    int rc = takeSlots(s, 1, got);
    s->freeSlots = keep;
    if (rc) return rc;
  }
  std::vector<float> o(40);
  freshJoint(o.data());
  o[3] = finiteOr(stiff_lin);
  o[7] = finiteOr(stiff_ang);
  for (int c = 0; c < 3; c++) { o[24 + c] = rA[c]; o[28 + c] = rB[c]; }
  o[15] = torqueArm(sizeOf(s, a), sizeOf(s, b));
  s->putJoints(got, o, {(int)T_JOINT, a, b});
  if (slot) *slot = got[0];
  return avbd_sync(s);
}

int avbd_append_world_joint(avbd_solver *s, uint32_t b, const float *anchor_world, const float *anchor_body, uint32_t *slot) {
  if (!s || !anchor_world || !anchor_body) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  if (!isBody(s, (int32_t)b)) return fail(s, AVBD_ERR_INVALID, "appendWorldJoint: b must be a body in the solver");
  for (int c = 0; c < 3; c++)
    if (!finite32(anchor_world[c]) || !finite32(anchor_body[c])) return fail(s, AVBD_ERR_INVALID, "appendWorldJoint: anchors need three finite f32 numbers");
  const float *sb = sizeOf(s, (int32_t)b);
  double arm = (double)sb[0] * sb[0] + (double)sb[1] * sb[1] + (double)sb[2] * sb[2];
  if (!std::isfinite((float)arm)) return fail(s, AVBD_ERR_INVALID, "appendWorldJoint: torque arm must be finite as f32");
  std::vector<uint32_t> got;
  if (int rc = takeSlots(s, 1, got)) return rc;
  std::vector<float> o(40);
  freshJoint(o.data());
  o[3] = JBIG;
  for (int c = 0; c < 3; c++) { o[24 + c] = anchor_world[c]; o[28 + c] = anchor_body[c]; }
  o[15] = (float)arm;
  s->putJoints(got, o, {(int)T_JOINT, -1, (int)b});
  if (slot) *slot = got[0];
  return avbd_sync(s);
}

int avbd_append_springs(avbd_solver *s, const avbd_spring_spec *springs, uint32_t count, uint32_t *slots) {
  if (!s || (!springs && count)) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  for (uint32_t k = 0; k < count; k++)
    if (!(springs[k].a == -1 || isBody(s, springs[k].a)) || !isBody(s, springs[k].b))
      return fail(s, AVBD_ERR_INVALID, "appendSprings: bodies " + std::to_string(springs[k].a) + " and " + std::to_string(springs[k].b) +
                                           " are not both bodies in the solver (a may be -1, the world)");
  std::vector<uint32_t> got;
  if (int rc = takeSlots(s, count, got)) return rc;
  std::vector<float> recs(40ull * count, 0.0f);
  std::vector<int> ends(3ull * count);
  for (uint32_t k = 0; k < count; k++) {
    const avbd_spring_spec &sp = springs[k];
    float *o = &recs[40ull * k];
    o[3] = finiteOr(sp.stiffness);
    for (int c = 0; c < 3; c++) { o[24 + c] = sp.rA[c]; o[28 + c] = sp.rB[c]; }
    o[27] = sp.rest;
    ends[3 * k] = (int)T_SPRING; ends[3 * k + 1] = sp.a; ends[3 * k + 2] = sp.b;
    if (slots) slots[k] = got[k];
  }
  s->putJoints(got, recs, ends);
  return avbd_sync(s);
}

int avbd_release_joints(avbd_solver *s, const uint32_t *slots, uint32_t count) {
  if (!s || (!slots && count)) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  s->manifoldsValid = false;
  std::vector<uint32_t> order;
  for (uint32_t k = 0; k < count; k++)
    if (slots[k] < s->nj && s->jinfoHost[4ull * slots[k]] != (int)T_NONE) order.push_back(slots[k]);
  std::sort(order.begin(), order.end());
  order.erase(std::unique(order.begin(), order.end()), order.end());
  if (order.empty()) return AVBD_OK;
  std::vector<float> zeros(40 * order.size(), 0.0f);
  std::vector<int> ends;
  for (uint32_t slot : order) {
    ends.insert(ends.end(), {(int)T_NONE, -1, -1});
    s->freeSlots.push_back(slot);
  }
  s->putJoints(order, zeros, ends);
  return avbd_sync(s);
}

int avbd_set_world_anchor(avbd_solver *s, uint32_t slot, const float *p) {
  if (!s || !p) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  if (slot >= s->nj) return fail(s, AVBD_ERR_INVALID, "setWorldAnchor: joint slot out of range");
  hipMemcpyAsync((float *)(s->joints + 10ull * slot) + 24, p, 12, hipMemcpyHostToDevice, s->stream);
  return avbd_sync(s);
}

int avbd_disable_constraint(avbd_solver *s, uint32_t slot) {
  if (!s) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  if (slot >= s->nj) return fail(s, AVBD_ERR_INVALID, "disableConstraint: joint slot out of range");
  s->manifoldsValid = false;
  float zero[8] = {0};
  float *k = (float *)(s->joints + 10ull * slot);
  hipMemcpyAsync(k, zero, 32, hipMemcpyHostToDevice, s->stream);       // penalties and stiffnesses
  hipMemcpyAsync(k + 8, zero, 12, hipMemcpyHostToDevice, s->stream);   // lambda lin xyz
  hipMemcpyAsync(k + 12, zero, 12, hipMemcpyHostToDevice, s->stream);  // lambda ang xyz
  return avbd_sync(s);
}

int avbd_set_body_aux(avbd_solver *s, const uint32_t *indices, uint32_t count, const float *scales, const float *linear,
                      const float *angular) {
  if (!s || (count && (!indices || !scales))) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  if ((linear == nullptr) != (angular == nullptr)) return fail(s, AVBD_ERR_INVALID, "setBodyAux: pass both damping rate arrays or neither");
  bool rates = linear != nullptr;
  for (uint32_t k = 0; k < count; k++) {
    if (indices[k] >= s->bodyCap) return fail(s, AVBD_ERR_INVALID, "setBodyAux: slot " + std::to_string(indices[k]) + " is outside the body capacity");
    if (!std::isfinite(scales[k])) return fail(s, AVBD_ERR_INVALID, "setBodyAux: scales must be finite (as f32)");
    if (rates && !(std::isfinite(linear[k]) && linear[k] >= 0.0f && std::isfinite(angular[k]) && angular[k] >= 0.0f))
      return fail(s, AVBD_ERR_INVALID, "setBodyAux: damping rates must be finite (as f32) and at least 0");
  }
  if (!s->aux) {
    try {
      s->aux = s->alloc<float4>(s->bodyCap);
    } catch (const std::exception &e) {
      return fail(s, AVBD_ERR_DEVICE, e.what());
    }
    s->auxScale.assign(s->bodyCap, 1.0f);
    s->auxLin.assign(s->bodyCap, 0.0f);
    s->auxAng.assign(s->bodyCap, 0.0f);
    s->uploadAux(0, s->bodyCap);
  }
  for (uint32_t k = 0; k < count;) {
    uint32_t e = k + 1;
    while (e < count && indices[e] == indices[e - 1] + 1) e++;
    for (uint32_t i = k; i < e; i++) {
      uint32_t b = indices[i];
      s->auxScale[b] = scales[i];
      if (rates) { s->auxLin[b] = linear[i]; s->auxAng[b] = angular[i]; }
      else if (e - k > 1) { s->auxLin[b] = 0.0f; s->auxAng[b] = 0.0f; }  // a run's whole records hold rates 0
    }
    s->uploadAux(indices[k], e - k);
    k = e;
  }
  return avbd_sync(s);
}

int avbd_set_lifecycle(avbd_solver *s, uint32_t index, uint32_t bits) {
  if (!s) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  if (index >= s->n || bits > 6 || (bits & ~6u) != 0) return fail(s, AVBD_ERR_INVALID, "Invalid body lifecycle flags/index");
  if (s->backend) return fail(s, AVBD_ERR_INVALID, "Body lifecycle requires built-in standard shaders");
  s->lifeHost[index] = (uint8_t)bits;
  float w = (float)(1u | bits);  // BODY_DIRTY | bits
  hipMemcpyAsync((float *)&s->b.ipos[index] + 3, &w, 4, hipMemcpyHostToDevice, s->stream);
  s->fill(s->colA + index, 1, NO_COLOR);
  s->fill(s->colB + index, 1, NO_COLOR);
  s->manifoldsValid = false;
  return avbd_sync(s);
}

}  // extern "C"

// ---- scene planning (host only) ----
int avbd_morton_order(const float *bodies, uint32_t body_count, uint32_t *order) {
  if ((!bodies || !order) && body_count) return AVBD_ERR_INVALID;
  avbd_plan::mortonOrder(bodies, body_count, order);
  return AVBD_OK;
}
int avbd_estimate_pairs(const float *bodies, uint32_t body_count, uint64_t *pairs, uint64_t *touching, int32_t *degree) {
  if (!bodies && body_count) return AVBD_ERR_INVALID;
  avbd_plan::Estimate e = avbd_plan::estimatePairs(bodies, body_count, degree, largestSmall);
  if (pairs) *pairs = e.pairs;
  if (touching) *touching = e.touching;
  return AVBD_OK;
}
int avbd_plan_capacity(const float *bodies, uint32_t body_count, const int32_t *joint_info, uint32_t joint_count,
                       const avbd_capacity *least, uint32_t least_colors, avbd_capacity *out, uint32_t *color_cap) {
  if ((!bodies && body_count) || (!joint_info && joint_count) || !out) return AVBD_ERR_INVALID;
  avbd_capacity l = least ? *least : avbd_capacity{};
  std::vector<int32_t> degree(body_count, 0);
  avbd_plan::Estimate e = avbd_plan::estimatePairs(bodies, body_count, degree.data(), largestSmall);
  for (uint32_t k = 0; k < joint_count; k++) {
    int32_t a = joint_info[4ull * k + 1], b = joint_info[4ull * k + 2];
    if (a >= 0 && (uint32_t)a < body_count) degree[a]++;
    if (b >= 0 && (uint32_t)b < body_count) degree[b]++;
  }
  int32_t maxDegree = 0;
  for (uint32_t i = 0; i < body_count; i++)
    if (bodies[40ull * i + 19] > 0.0f) maxDegree = std::max(maxDegree, degree[i]);
  uint64_t cap = l.bodies ? std::max<uint64_t>({l.bodies, body_count, 1}) : std::max<uint64_t>(body_count + 1024ull, 1);
  auto fit = [](double v) { return (uint32_t)std::min(std::ceil(v), 4294967295.0); };
  out->bodies = (uint32_t)std::min<uint64_t>(cap, 0xffffffffull);
  out->joints = std::max(joint_count + 256u, l.joints);
  out->contacts = fit(std::max({8192.0, 4.0 * cap, 8.0 * (double)e.touching, (double)l.contacts}));
  out->pairs = fit(std::max({4096.0, 4.0 * cap, 2.0 * (double)e.pairs, (double)l.pairs}));
  out->manifolds = std::min(fit(std::max({4096.0, (double)cap, 1.6 * (double)e.touching, (double)l.manifolds})), out->pairs);
  out->hull_words = l.hull_words;
  if (color_cap) *color_cap = (uint32_t)std::min<int64_t>(64, std::max<int64_t>({8, (int64_t)maxDegree + 1 + 3, (int64_t)least_colors}));
  return AVBD_OK;
}
int avbd_set_graphs(avbd_solver *s, int on) {
  if (!s) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  if (avbd_sync(s)) return AVBD_ERR_DEVICE;
  s->graphs = on != 0;
  if (!s->graphs) s->dropGraphs();
  return AVBD_OK;
}
int avbd_graph_stats(const avbd_solver *s, uint64_t *captures, uint64_t *replays) {
  if (!s) return AVBD_ERR_INVALID;
  if (captures) *captures = s->graphCaptures;
  if (replays) *replays = s->graphReplays;
  return AVBD_OK;
}

// ---- SDF collision backend (SdfCollisionBridge, src/avbd3d/gpu/sdf-bridge.ts) ----------------
using namespace avbd::sdfk;

struct avbd_sdf {
  avbd_solver *s = nullptr;
  bool installed = false, automatic = false;
  uint32_t n = 0, points = 5, cap = 1;
  float cellSize = 1.0f;
  avbd_sdf_params prm{};
  std::string err;
  std::vector<void *> mem;
  uint4 *atlas = nullptr;
  float4 *bodies = nullptr, *configRot = nullptr, *stage = nullptr;
  uint4 *config = nullptr, *pairs = nullptr, *results[2] = {nullptr, nullptr};
  uint32_t current = 0;    // results[current]: the last step's
  uint32_t stepIndex = 0;  // the pass's step counter (seeds the per-pair streams)
  uint32_t count = 0;      // explicit: candidates of the last step
  std::vector<uint32_t> pairsHost, priorHost;  // explicit candidates (a, b, key), the last step's
  // the spatial grid
  uint32_t tableSize = 0, cacheSize = 0;
  uint32_t *bucketCount = nullptr, *bucketStart = nullptr, *bucketCursor = nullptr, *sorted = nullptr, *largeList = nullptr;
  uint32_t *largeCount = nullptr, *history = nullptr, *pairCount = nullptr, *pairOffset = nullptr, *live = nullptr;
  int4 *cellKey = nullptr;

  template <class T> T *alloc(size_t count) {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(count * sizeof(T), 256)) != hipSuccess) throw std::runtime_error("hipMalloc failed (SDF bridge)");
    hipMemsetAsync(p, 0, std::max<size_t>(count * sizeof(T), 256), s->stream);
    mem.push_back(p);
    return (T *)p;
  }
  void release(void *p) {
    for (size_t k = 0; k < mem.size(); k++) if (mem[k] == p) { hipFree(p); mem.erase(mem.begin() + (long)k); return; }
  }
  ~avbd_sdf() {
    if (s) hipStreamSynchronize(s->stream);
    for (void *p : mem) hipFree(p);
  }
  uint32_t stride() const { return RESULT_HEAD + POINT_F4 * points; }
  /** The results, the staging of points being built and (automatic) the history table for `c` pairs. */
  void allocPairs(uint32_t c) {
    cap = c;
    pairs = alloc<uint4>(cap);
    results[0] = alloc<uint4>((size_t)stride() * cap);
    results[1] = alloc<uint4>((size_t)stride() * cap);
    stage = alloc<float4>((size_t)POINT_F4 * points * cap);
    k_fill<<<blocksFor((uint64_t)cap * 4, 256), 256, 0, s->stream>>>((uint32_t *)pairs, cap * 4u, EMPTY, nullptr);
    for (int k = 0; k < 2; k++) k_sdf_copy_history<<<blocksFor(cap, 256), 256, 0, s->stream>>>(nullptr, 0, results[k], cap, stride());
    if (automatic) {
      cacheSize = pow2At(2ull * cap);
      history = alloc<uint32_t>(cacheSize);
    }
  }
  SdfDev dev() const {
    SdfDev g;
    std::memset(&g, 0, sizeof g);
    g.atlas = atlas; g.bodies = bodies; g.config = config; g.configRot = configRot; g.pairs = pairs;
    g.prev = results[current]; g.cur = results[current ^ 1u]; g.stage = stage; g.cap = cap; g.live = live;
    SdfParams &p = g.p;
    p.pairCount = automatic ? cap : count; p.step = stepIndex; p.cacheSize = prm.cache_size; p.randomSamples = prm.random_samples;
    p.dt = s->hp.dt; p.epsilon = prm.epsilon; p.alpha = prm.alpha; p.tol = prm.tolerance;
    p.margin = prm.margin; p.replacement = prm.replacement_threshold; p.signedSep = prm.signed_separation ? 1u : 0u;
    p.meshVertex = prm.mesh_vertex_contacts ? 1u : 0u; p.maxIt = prm.max_iterations; p.warmStart = prm.warm_start ? 1u : 0u;
    p.curvature = prm.curvature_sampling ? 1u : 0u; p.skipConvex = prm.skip_convex_samples ? 1u : 0u; p.optimizer = prm.optimizer;
    p.stepSize = prm.optimizer == 2 ? prm.step_size : 0.0f; p.sgdStopOnly = prm.optimizer == 2 && prm.sgd_gradient_stop_only ? 1u : 0u;
    p.points = points; p.allowedPen = prm.allowed_penetration; p.refresh = prm.refresh_deeper_contacts ? prm.tolerance : 0.0f;
    g.cellSize = cellSize; g.tableMask = tableSize ? tableSize - 1 : 0; g.cacheMask = cacheSize ? cacheSize - 1 : 0;
    g.bucketCount = bucketCount; g.bucketStart = bucketStart; g.bucketCursor = bucketCursor; g.sorted = sorted;
    g.largeList = largeList; g.largeCount = largeCount; g.cellKey = cellKey; g.history = history;
    g.pairCount = pairCount; g.pairOffset = pairOffset;
    return g;
  }
  /** prepare (sdf-bridge.ts:154-160): the checks before a step. */
  bool prepare(std::string &why) const {
    if (s->n != n) { why = "Recreate SDF bridge after changing body slots"; return false; }
    uint64_t need = (uint64_t)(automatic ? cap : count) * points;
    if (need > std::min(s->conCap, s->manCap)) { why = "Insufficient SDF solver contact capacity"; return false; }
    if (!std::isfinite(prm.allowed_penetration) || prm.allowed_penetration < 0) { why = "Invalid SDF allowed penetration"; return false; }
    return true;
  }
  /** encode (sdf-bridge.ts:162-181): the collision phase of one step. */
  void encode(const Dev &d) {
    hipStream_t st = s->stream;
    if (!automatic) {
      // previous slots: the last step's slot of the same key, if its bodies are the same
      std::vector<uint32_t> up(4ull * count);
      for (uint32_t i = 0; i < count; i++) {
        uint32_t a = pairsHost[3 * i], b = pairsHost[3 * i + 1], key = pairsHost[3 * i + 2], prevSlot = EMPTY;
        for (uint32_t j = 0; 3 * j < priorHost.size(); j++)
          if (priorHost[3 * j + 2] == key) { if (priorHost[3 * j] == a && priorHost[3 * j + 1] == b) prevSlot = j; break; }
        up[4 * i] = a; up[4 * i + 1] = b; up[4 * i + 2] = key; up[4 * i + 3] = prevSlot;
      }
      if (count) hipMemcpyAsync(pairs, up.data(), 16ull * count, hipMemcpyHostToDevice, st);
      if (count) hipStreamSynchronize(st);  // `up` is a host temporary
    }
    SdfDev g = dev();
    k_sdf_prepare<<<blocksFor(n, 256), 256, 0, st>>>(d, g);
    if (automatic) {
      s->fill(bucketCount, tableSize + 1, 0);
      s->fill(bucketCursor, tableSize, 0);
      s->fill(largeCount, 1, 0);
      s->fill(live, 4, 0);
      s->fill(history, cacheSize, EMPTY);
      k_sdf_cells<<<s->strideBlocks(n, 256), 256, 0, st>>>(d, g);
      s->scan(bucketCount, bucketStart, tableSize + 1);
      k_sdf_scatter<<<blocksFor(n, 256), 256, 0, st>>>(d, g);
      k_sdf_history<<<blocksFor(cap, 256), 256, 0, st>>>(d, g);
      k_sdf_find<false><<<blocksFor(n, 64), 64, 0, st>>>(d, g);
      s->scan(pairCount, pairOffset, n + 1);
      k_sdf_find<true><<<blocksFor(n, 64), 64, 0, st>>>(d, g);
      k_sdf_close<<<blocksFor(cap, 256), 256, 0, st>>>(d, g);
    } else if (count) {
      k_sdf_filter<<<blocksFor(count, 256), 256, 0, st>>>(d, g);
    }
    uint32_t slots = automatic ? cap : count;
    if (slots) {
      size_t lds = (size_t)points * 64 * (4 * sizeof(float) + 1);
      k_sdf_narrow<<<std::min<uint32_t>(blocksFor(slots, 64, 1u << 20), std::max(s->numCUs, 1u) * 32u), 64, lds, st>>>(d, g);
      k_sdf_export<<<s->strideBlocks(slots, 256), 256, 0, st>>>(d, g);
    }
    current ^= 1u;
    stepIndex++;
    if (!automatic) priorHost = pairsHost;
  }
};

static void sdfEncode(avbd_sdf *g, const Dev &d) { g->encode(d); }
static void sdfRelease(avbd_sdf *g) { delete g; }
static bool sdfPrepare(const avbd_sdf *g, std::string &why) { return g->prepare(why); }

namespace {
bool finite3(const float *v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }
/** packSdfParams's checks (sdf-layout.ts:100-118) and the bridge's (sdf-bridge.ts:62-64). */
bool checkSdfParams(const avbd_sdf_params &p, uint32_t points, std::string &why) {
  if (p.optimizer > 2) { why = "Unknown SDF optimizer " + std::to_string(p.optimizer); return false; }
  if (p.optimizer == 2 && !(p.step_size > 0 && std::isfinite(p.step_size))) { why = "SDF gradient descent needs a finite positive stepSize"; return false; }
  if (p.cache_size > points - 1 || p.cache_size + p.random_samples > points || p.max_iterations < 1 || p.max_iterations > 512) {
    why = "This GPU SDF pass supports at most " + std::to_string(points - 1) + " cached/" + std::to_string(points) + " total samples and 1-512 iterations";
    return false;
  }
  if (!(p.epsilon == 0 || p.epsilon > 0) || !(p.tolerance > 0 && std::isfinite(p.tolerance)) || !std::isfinite(p.epsilon) ||
      !(p.alpha >= 0 && std::isfinite(p.alpha)) || !(p.margin >= 0 && std::isfinite(p.margin)) ||
      !(p.replacement_threshold >= 0 && std::isfinite(p.replacement_threshold))) {
    why = "Invalid SDF collision distances";
    return false;
  }
  if (!std::isfinite(p.allowed_penetration) || p.allowed_penetration < 0) { why = "Invalid SDF allowed penetration"; return false; }
  return true;
}
/** packSdfAtlas (sdf-layout.ts:33-71): per field a 32-word header, its grid, seeds and supports. */
bool packAtlas(const avbd_sdf_field *fields, uint32_t nf, std::vector<uint32_t> &words, std::vector<uint32_t> &offsets, std::string &why) {
  words.clear();
  offsets.clear();
  auto putf = [&](size_t at, float v) { std::memcpy(&words[at], &v, 4); };
  for (uint32_t f = 0; f < nf; f++) {
    const avbd_sdf_field &F = fields[f];
    if (F.kind > 9) { why = "Unsupported GPU SDF kind " + std::to_string(F.kind); return false; }
    if (!finite3(F.bounds_min) || !finite3(F.bounds_max) || !finite3(F.center_of_mass)) { why = "SDF header exceeds finite float32 range"; return false; }
    for (int k = 0; k < 3; k++) if (F.bounds_min[k] > F.bounds_max[k]) { why = "Invalid SDF bounds"; return false; }
    uint64_t nodes = 0;
    if (F.kind == AVBD_SDF_VOXEL) {
      if (F.dims[0] < 2 || F.dims[1] < 2 || F.dims[2] < 2 || !(F.cell > 0) || !std::isfinite(F.cell) || !finite3(F.origin) || !F.grid) {
        why = "Invalid SDF grid geometry";
        return false;
      }
      nodes = (uint64_t)F.dims[0] * F.dims[1] * F.dims[2];
      for (uint64_t i = 0; i < 4 * nodes; i++) if (!std::isfinite(F.grid[i])) { why = "SDF grid must be finite"; return false; }
    } else {
      for (int k = 0; k < (F.kind == AVBD_SDF_WARP ? 4 : 8); k++) if (!std::isfinite(F.params[k])) { why = "SDF secondary parameters exceed finite float32 range"; return false; }
    }
    if ((F.seed_count && !F.seeds) || (F.support_count && !F.supports)) { why = "SDF seeds or supports missing"; return false; }
    size_t at = words.size();
    offsets.push_back((uint32_t)(at / 4));
    uint64_t total = at + 32 + 4 * nodes + 4ull * (F.seed_count + F.support_count);
    if (total * 4 > (1ull << 31)) { why = "SDF atlas exceeds byte/index limit"; return false; }
    words.resize((size_t)total, 0);
    size_t gridAt = at + 32, seedAt = gridAt + 4 * nodes;
    words[at] = F.kind; words[at + 1] = F.convex ? 1u : 0u; words[at + 2] = (uint32_t)(seedAt / 4); words[at + 3] = F.seed_count;
    if (F.kind == AVBD_SDF_VOXEL) {
      words[at + 4] = (uint32_t)(gridAt / 4); words[at + 5] = F.dims[0]; words[at + 6] = F.dims[1]; words[at + 7] = F.dims[2];
      for (int k = 0; k < 3; k++) putf(at + 28 + k, F.origin[k]);
      std::memcpy(&words[gridAt], F.grid, 16 * nodes);
    }
    for (int k = 0; k < 3; k++) { putf(at + 8 + k, F.bounds_min[k]); putf(at + 12 + k, F.bounds_max[k]); putf(at + 16 + k, F.center_of_mass[k]); }
    putf(at + 11, F.kind == AVBD_SDF_VOXEL ? F.cell : 0.0f);
    words[at + 15] = F.support_count;
    for (int k = 0; k < 8; k++) putf(at + 20 + k, F.params[k]);
    if (F.kind == AVBD_SDF_WARP) words[at + 24] = (uint32_t)F.params[4];
    for (uint32_t i = 0; i < F.seed_count; i++) for (int k = 0; k < 3; k++) putf(seedAt + 4 * i + k, F.seeds[3 * i + k]);
    size_t supAt = seedAt + 4ull * F.seed_count;
    for (uint32_t i = 0; i < F.support_count; i++) for (int k = 0; k < 3; k++) putf(supAt + 4 * i + k, F.supports[3 * i + k]);
  }
  if (words.empty()) words.assign(4, 0);
  return true;
}
int sdfFail(avbd_sdf *g, int code, const std::string &msg) {
  if (g) g->err = msg;
  return code;
}
}  // namespace

extern "C" {

void avbd_sdf_default_params(avbd_sdf_params *o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->optimizer = 0; o->cache_size = 4; o->random_samples = 1; o->warm_start = 1; o->curvature_sampling = 1; o->skip_convex_samples = 1;
  o->epsilon = 0.1f; o->alpha = 0.01f; o->replacement_threshold = 0.01f; o->tolerance = 0.01f; o->max_iterations = 50;
  o->step_size = 0.1f; o->margin = 0.02f; o->allowed_penetration = 0.01f;
}

int avbd_sdf_create(avbd_solver *s, const avbd_sdf_field *fields, uint32_t nf, const uint32_t *bodyField, const float *bodyScale,
                    const float *bodyRot, const avbd_sdf_params *params, int automatic, uint32_t maxPairs, uint32_t points, float cellSize,
                    avbd_sdf **out) {
  if (!out) return AVBD_ERR_INVALID;
  *out = nullptr;
  if (!s) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "SDF solver was destroyed");
  if (!s->n || !bodyField || !fields || !nf) return fail(s, AVBD_ERR_INVALID, "SDF bridge requires a field for every GPU body slot");
  if (s->worldStartHost.size() > 2) return fail(s, AVBD_ERR_INVALID, "SDF bridge: one world per solver (not a batched solver)");
  if (points == 0) points = 5;
  if (points < 2 || points > 20) return fail(s, AVBD_ERR_INVALID, "SDF point capacity must be an integer 2-20");
  avbd_sdf_params prm;
  if (params) prm = *params; else avbd_sdf_default_params(&prm);
  std::string why;
  if (!checkSdfParams(prm, points, why)) return fail(s, AVBD_ERR_INVALID, why);
  uint32_t n = s->n;
  for (uint32_t i = 0; i < n; i++) {
    if (bodyField[i] >= nf) return fail(s, AVBD_ERR_INVALID, "body " + std::to_string(i) + ": no such field");
    float sc = bodyScale ? bodyScale[i] : 1.0f;
    if (!(sc > 0) || !std::isfinite(sc)) return fail(s, AVBD_ERR_INVALID, "Invalid SDF bridge scale");
    if (bodyRot) {
      const float *q = bodyRot + 4 * i;
      double norm = std::sqrt((double)q[0] * q[0] + (double)q[1] * q[1] + (double)q[2] * q[2] + (double)q[3] * q[3]);
      if (!std::isfinite(norm) || std::fabs(norm - 1) > 1e-5) return fail(s, AVBD_ERR_INVALID, "SDF field rotation must be a unit quaternion");
    }
  }
  std::vector<uint32_t> words, offsets;
  if (!packAtlas(fields, nf, words, offsets, why)) return fail(s, AVBD_ERR_INVALID, why);
  uint64_t maxAuto = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)n * (n - 1) / 2, std::max<uint64_t>(64, 8ull * n)));
  uint64_t capWanted = maxPairs ? maxPairs : (automatic ? maxAuto : 1);
  if (capWanted > (1u << 24)) return fail(s, AVBD_ERR_INVALID, "Invalid SDF broadphase pair capacity");
  if (automatic) {
    if (!(cellSize >= 0) || !std::isfinite(cellSize)) return fail(s, AVBD_ERR_INVALID, "Invalid SDF broadphase cell size");
    if (cellSize == 0) {  // the median field bounds diagonal x scale (sdfBroadphaseOptions)
      std::vector<double> sizes(n);
      for (uint32_t i = 0; i < n; i++) {
        const avbd_sdf_field &F = fields[bodyField[i]];
        double dx = F.bounds_max[0] - F.bounds_min[0], dy = F.bounds_max[1] - F.bounds_min[1], dz = F.bounds_max[2] - F.bounds_min[2];
        sizes[i] = std::sqrt(dx * dx + dy * dy + dz * dz) * (bodyScale ? bodyScale[i] : 1.0f);
      }
      std::sort(sizes.begin(), sizes.end());
      cellSize = (float)sizes[n / 2];
    }
    if (!(cellSize > 0)) return fail(s, AVBD_ERR_INVALID, "Invalid SDF broadphase cell size");
  } else if ((uint64_t)capWanted * points > std::min(s->conCap, s->manCap)) {
    return fail(s, AVBD_ERR_CAPACITY, "SDF bridge requires capacity for " + std::to_string(points) + " contacts/manifolds per candidate pair");
  }
  if (automatic && (s->hp.flags & AVBD_FLAG_FIXED_STORAGE) && capWanted * points > std::min(s->conCap, s->manCap))
    return fail(s, AVBD_ERR_CAPACITY, "Fixed storage: the SDF pairs' contacts exceed the solver's contact capacity");
  avbd_sdf *g = new avbd_sdf();
  g->s = s; g->n = n; g->points = points; g->automatic = automatic != 0; g->cellSize = cellSize; g->prm = prm;
  try {
    hipStreamSynchronize(s->stream);
    g->atlas = g->alloc<uint4>(words.size() / 4);
    hipMemcpyAsync(g->atlas, words.data(), words.size() * 4, hipMemcpyHostToDevice, s->stream);
    std::vector<uint32_t> cfg(4ull * n);
    std::vector<float> rot(4ull * n);
    for (uint32_t i = 0; i < n; i++) {
      float sc = bodyScale ? bodyScale[i] : 1.0f;
      cfg[4 * i] = offsets[bodyField[i]];
      std::memcpy(&cfg[4 * i + 1], &sc, 4);
      for (int k = 0; k < 4; k++) rot[4 * i + k] = bodyRot ? bodyRot[4 * i + k] : (k == 3 ? 1.0f : 0.0f);
    }
    g->config = g->alloc<uint4>(n);
    g->configRot = g->alloc<float4>(n);
    g->bodies = g->alloc<float4>((size_t)BODY_F4 * n);
    hipMemcpyAsync(g->config, cfg.data(), cfg.size() * 4, hipMemcpyHostToDevice, s->stream);
    hipMemcpyAsync(g->configRot, rot.data(), rot.size() * 4, hipMemcpyHostToDevice, s->stream);
    if (g->automatic) {
      g->tableSize = pow2At(2ull * n);
      g->bucketCount = g->alloc<uint32_t>(g->tableSize + 1);
      g->bucketStart = g->alloc<uint32_t>(g->tableSize + 1);
      g->bucketCursor = g->alloc<uint32_t>(g->tableSize);
      g->sorted = g->alloc<uint32_t>(n);
      g->largeList = g->alloc<uint32_t>(n);
      g->largeCount = g->alloc<uint32_t>(1);
      g->pairCount = g->alloc<uint32_t>(n + 1);
      g->pairOffset = g->alloc<uint32_t>(n + 1);
      g->live = g->alloc<uint32_t>(4);
      g->cellKey = g->alloc<int4>(n);
      if (g->tableSize + 1 > MAX_SCAN) throw std::runtime_error("SDF spatial grid exceeds the scan's size");
    }
    g->allocPairs((uint32_t)capWanted);
    if (g->automatic) {  // reserveCollisionCapacity (solver.ts:2043-2055)
      uint64_t need = (uint64_t)g->cap * points;
      if (need > std::min(s->conCap, s->manCap)) s->growStorage(std::max<uint32_t>((uint32_t)need, s->conCap), std::max<uint32_t>((uint32_t)need, s->pairCap),
                                                               std::max<uint32_t>((uint32_t)need, s->manCap));
    }
    if (hipStreamSynchronize(s->stream) != hipSuccess) throw std::runtime_error("device error creating the SDF bridge");
  } catch (const std::exception &e) {
    delete g;
    return fail(s, AVBD_ERR_DEVICE, e.what());
  }
  s->sdfs.push_back(g);
  *out = g;
  return AVBD_OK;
}

int avbd_sdf_set_pairs(avbd_sdf *g, const uint32_t *pairs, uint32_t count) {
  if (!g) return AVBD_ERR_INVALID;
  if (g->automatic) return sdfFail(g, AVBD_ERR_INVALID, "Automatic SDF candidates are GPU generated");
  if (count > g->cap) return sdfFail(g, AVBD_ERR_CAPACITY, "SDF candidate capacity exceeded");
  if (count && !pairs) return sdfFail(g, AVBD_ERR_INVALID, "pairs is null");
  std::vector<std::pair<uint32_t, uint32_t>> ab;
  std::vector<uint32_t> keys;
  for (uint32_t i = 0; i < count; i++) {
    uint32_t a = pairs[3 * i], b = pairs[3 * i + 1], key = pairs[3 * i + 2];
    if (a >= g->n || b >= g->n || a == b) return sdfFail(g, AVBD_ERR_INVALID, "Invalid SDF pair body IDs");
    if (key == EMPTY) return sdfFail(g, AVBD_ERR_INVALID, "SDF pair keys must be unique uint32 IDs");
    if (a <= b) return sdfFail(g, AVBD_ERR_INVALID, "SDF candidates require unique body pairs with a > b");
    ab.push_back({a, b});
    keys.push_back(key);
  }
  std::sort(ab.begin(), ab.end());
  std::sort(keys.begin(), keys.end());
  if (std::adjacent_find(keys.begin(), keys.end()) != keys.end()) return sdfFail(g, AVBD_ERR_INVALID, "SDF pair keys must be unique uint32 IDs");
  if (std::adjacent_find(ab.begin(), ab.end()) != ab.end()) return sdfFail(g, AVBD_ERR_INVALID, "SDF candidates require unique body pairs with a > b");
  g->pairsHost.assign(pairs, pairs + 3ull * count);
  g->count = count;
  return AVBD_OK;
}

int avbd_sdf_set_params(avbd_sdf *g, const avbd_sdf_params *params) {
  if (!g || !params) return AVBD_ERR_INVALID;
  std::string why;
  if (!checkSdfParams(*params, g->points, why)) return sdfFail(g, AVBD_ERR_INVALID, why);
  g->prm = *params;
  return AVBD_OK;
}

int avbd_set_collision_backend(avbd_solver *s, avbd_sdf *g) {
  if (!s) return AVBD_ERR_INVALID;
  if (s->destroyed) return fail(s, AVBD_ERR_DESTROYED, "Solver is destroyed");
  if (g == s->backend) return AVBD_OK;
  if (g && g->s != s) return fail(s, AVBD_ERR_INVALID, "Invalid SDF bridge owner/lifecycle");
  if (s->hp.flags & AVBD_FLAG_FIXED_STORAGE)
    return fail(s, AVBD_ERR_INVALID, "Fixed storage: collision backends are not available (a fixed solver keeps the standard narrowphase; use adaptive storage for SDF collisions)");
  for (uint32_t i = 0; i < s->n; i++)
    if (s->lifeHost[i] != 0) return fail(s, AVBD_ERR_INVALID, "Submit an enabled standard step before switching collision backends: pending body lifecycle state");
  std::string why;
  if (g && !g->prepare(why)) return fail(s, AVBD_ERR_INVALID, why);
  hipStreamSynchronize(s->stream);
  if (s->backend) {
    s->sdfs.erase(std::remove(s->sdfs.begin(), s->sdfs.end(), s->backend), s->sdfs.end());
    sdfRelease(s->backend);
  }
  s->backend = g;
  if (g) g->installed = true;
  s->manifoldsValid = false;
  s->adjacencyHeld = false;
  // the counts of the last step's contacts go, so no later step warm-starts from them
  uint32_t zero[2] = {0, 0};
  const int words[] = {C_PAIRS, C_CONTACTS, C_MANIFOLDS, C_OVERFLOW, C_PREV_CONTACTS, C_PREV_MANIFOLDS};
  for (int w : words) hipMemcpyAsync(s->counters + w, zero, 4, hipMemcpyHostToDevice, s->stream);
  return avbd_sync(s);
}

int avbd_sdf_read_results(avbd_sdf *g, uint32_t *out, uint32_t cap, uint32_t *count) {
  if (!g) return AVBD_ERR_INVALID;
  uint32_t slots = g->automatic ? g->cap : g->count;
  if (count) *count = slots;
  if (!out || !cap) return AVBD_OK;
  uint32_t nr = std::min(cap, slots);
  if (nr && hipMemcpyAsync(out, g->results[g->current], 16ull * g->stride() * nr, hipMemcpyDeviceToHost, g->s->stream) != hipSuccess)
    return sdfFail(g, AVBD_ERR_DEVICE, "read failed");
  if (hipStreamSynchronize(g->s->stream) != hipSuccess) return sdfFail(g, AVBD_ERR_DEVICE, "read failed");
  return AVBD_OK;
}

int avbd_sdf_read_stats(avbd_sdf *g, uint32_t *pairs, uint32_t *tested, uint32_t *large, uint32_t *capacity, float *cell) {
  if (!g) return AVBD_ERR_INVALID;
  if (!g->automatic) return sdfFail(g, AVBD_ERR_INVALID, "Explicit candidates have no spatial grid");
  uint32_t w[4] = {0, 0, 0, 0};
  hipMemcpyAsync(w, g->live, 16, hipMemcpyDeviceToHost, g->s->stream);
  if (hipStreamSynchronize(g->s->stream) != hipSuccess) return sdfFail(g, AVBD_ERR_DEVICE, "read failed");
  if (pairs) *pairs = w[0];
  if (tested) *tested = w[1];
  if (large) *large = w[2];
  if (capacity) *capacity = g->cap;
  if (cell) *cell = g->cellSize;
  return AVBD_OK;
}

int avbd_sdf_grow(avbd_sdf *g, uint32_t required) {
  if (!g) return AVBD_ERR_INVALID;
  if (!g->automatic) return sdfFail(g, AVBD_ERR_INVALID, "SDF growth requires an active spatial bridge");
  if (required <= g->cap || required > (1u << 24)) return sdfFail(g, AVBD_ERR_INVALID, "Invalid SDF requested capacity");
  avbd_solver *s = g->s;
  uint32_t capacity = std::max(required, std::min(2u * g->cap, 1u << 24));
  try {
    hipStreamSynchronize(s->stream);
    uint32_t oldCap = g->cap, st = g->stride();
    uint4 *keep = g->results[g->current];
    uint4 *oldOther = g->results[g->current ^ 1u], *oldPairs = g->pairs, *oldHistory = (uint4 *)g->history;
    float4 *oldStage = g->stage;
    g->allocPairs(capacity);
    k_sdf_copy_history<<<blocksFor(capacity, 256), 256, 0, s->stream>>>(keep, oldCap, g->results[g->current], capacity, st);
    hipStreamSynchronize(s->stream);
    g->release(keep); g->release(oldOther); g->release(oldPairs); g->release(oldStage); g->release(oldHistory);
    uint64_t need = (uint64_t)g->cap * g->points;
    if (need > std::min(s->conCap, s->manCap))
      s->growStorage(std::max<uint32_t>((uint32_t)need, s->conCap), std::max<uint32_t>((uint32_t)need, s->pairCap), std::max<uint32_t>((uint32_t)need, s->manCap));
    if (hipStreamSynchronize(s->stream) != hipSuccess) throw std::runtime_error("device error growing the SDF bridge");
  } catch (const std::exception &e) {
    return sdfFail(g, AVBD_ERR_DEVICE, e.what());
  }
  return AVBD_OK;
}

void avbd_sdf_destroy(avbd_sdf *g) {
  if (!g) return;
  avbd_solver *s = g->s;
  if (s && s->backend == g) {
    s->backend = nullptr;
    s->manifoldsValid = false;
    s->adjacencyHeld = false;
  }
  if (s) s->sdfs.erase(std::remove(s->sdfs.begin(), s->sdfs.end(), g), s->sdfs.end());
  sdfRelease(g);
}

const char *avbd_sdf_last_error(const avbd_sdf *g) { return g ? g->err.c_str() : "null SDF bridge"; }

}  // extern "C"
