// Contact events (ContactWatch, src/lib/contacts.ts): after a step, the pairs it touched that have a
// reporting body go into this step's hash set of lifetime-ID pairs; a pair missing from the last
// step's set began touching (where, the normal and how hard), one of the last step's set missing
// from this one stopped. Only the changes are written, three float4 each.
#pragma once
#include "avbd_kernels.h"

namespace avbd {

/** The table's head words (events written, may pass the capacity; set full), then two sets of
 *  setSlots (a + 1, b) slots; events of three float4 (kind, a, b, step | point, impulse | normal). */
struct WatchBuf {
  uint32_t *watch;   // per body: lifetime ID in bits 0..30, reporting in bit 31
  uint32_t *table;
  float4 *events;
  uint32_t setSlots, capacity;
};
static const uint32_t WATCH_HEAD = 4, EVENT_BEGIN = 1, EVENT_END = 2;

__device__ __forceinline__ uint32_t watchHash(uint32_t x) {
  x = (x ^ (x >> 16)) * 0x7feb352du;
  x = (x ^ (x >> 15)) * 0x846ca68bu;
  return x ^ (x >> 16);
}
__device__ __forceinline__ uint32_t pairSlot(uint32_t a, uint32_t b, uint32_t mask) { return watchHash((a * 0x9e3779b1u) ^ watchHash(b)) & mask; }
/** v rotated by unit quaternion q as contacts.ts's turn: t = 2 q.xyz × v, v + w t + q.xyz × t. */
__device__ __forceinline__ V turn(float4 q, V v) {
  V u = vv(q.x, q.y, q.z);
  V t = mul(cross(u, v), 2.0f);
  return add(add(v, mul(t, q.w)), cross(u, t));
}
/** insert (contacts.ts:49-61): open addressing, the slot claimed by compare-and-swap on a + 1. */
__device__ __forceinline__ void watchInsert(const WatchBuf &w, uint32_t base, uint32_t a, uint32_t b) {
  uint32_t mask = w.setSlots - 1, h = pairSlot(a, b, mask);
  for (uint32_t probe = 0; probe <= 2 * mask + 1; probe++) {
    uint32_t *s = w.table + WATCH_HEAD + 2 * (base + h);
    uint32_t old = atomicCAS(s, 0u, a + 1u);
    if (old == 0u) { s[1] = b; return; }
    h = (h + 1) & mask;
  }
  atomicOr(w.table + 1, 1u);
}
/** has (contacts.ts:64-74), on a set no thread is writing. */
__device__ __forceinline__ bool watchHas(const WatchBuf &w, uint32_t base, uint32_t a, uint32_t b) {
  uint32_t mask = w.setSlots - 1, h = pairSlot(a, b, mask);
  for (uint32_t probe = 0; probe <= mask; probe++) {
    const uint32_t *s = w.table + WATCH_HEAD + 2 * (base + h);
    uint32_t k = s[0];
    if (k == 0u) return false;
    if (k == a + 1u && s[1] == b) return true;
    h = (h + 1) & mask;
  }
  return false;
}
__device__ __forceinline__ void watchEmit(const WatchBuf &w, uint32_t kind, uint32_t a, uint32_t b, uint32_t step, V point, float impulse, V normal) {
  uint32_t e = atomicAdd(w.table, 1u);
  if (e >= w.capacity) return;
  w.events[3 * e] = make_float4(__uint_as_float(kind), __uint_as_float(a), __uint_as_float(b), __uint_as_float(step));
  w.events[3 * e + 1] = f4(point, impulse);
  w.events[3 * e + 2] = f4(normal, 0.0f);
}
/** touching (contacts.ts:86-109): per manifold of the step, into this step's set, and a begin if
 *  it was not in the last step's; where = the points' mean on body a (world), how hard = |Σ λn| dt. */
__global__ void k_watch_touching(Dev d, WatchBuf w, const uint4 *mIds, const float4 *mGeo, const Contact *con, uint32_t prev,
                                 uint32_t now, uint32_t pairs, uint32_t step, float dt) {
  uint32_t count = d.counters[C_MANIFOLDS];
  if (count > pairs) count = pairs;
  GRID_STRIDE(m, count) {
    uint4 ids = mIds[m];
    uint32_t cnt = ids.w & 15u, a = ids.x, b = ids.y;
    uint32_t wa = w.watch[a], wb = w.watch[b], idA = wa & 0x7fffffffu, idB = wb & 0x7fffffffu;
    if (cnt == 0u || ((wa | wb) & 0x80000000u) == 0u || idA == 0u || idB == 0u) continue;
    watchInsert(w, now, idA, idB);
    if (watchHas(w, prev, idA, idB)) continue;
    V pA = xyz(d.b.pos[a]);
    float4 qA = d.b.rot[a];
    V point = vv(0.0f, 0.0f, 0.0f);
    float force = 0.0f;
    for (uint32_t i = 0; i < cnt; i++) {
      Contact k = con[ids.z + i];
      point = add(point, add(turn(qA, xyz(k.rA)), pA));
      force = force + k.lam.x;
    }
    float fc = (float)cnt;
    watchEmit(w, EVENT_BEGIN, idA, idB, step, divs(point, fc), fabsf(force) * dt, xyz(mGeo[m]));
  }
}
/** ended (contacts.ts:112-120): per slot of the last step's set, an end if this step's lacks it. */
__global__ void k_watch_ended(WatchBuf w, uint32_t prev, uint32_t now, uint32_t step) {
  GRID_STRIDE(h, w.setSlots) {
    const uint32_t *s = w.table + WATCH_HEAD + 2 * (prev + h);
    uint32_t k = s[0];
    if (k == 0u) continue;
    uint32_t b = s[1];
    if (!watchHas(w, now, k - 1u, b)) watchEmit(w, EVENT_END, k - 1u, b, step, vv(0.0f, 0.0f, 0.0f), 0.0f, vv(0.0f, 0.0f, 0.0f));
  }
}
/** setWatched (contacts.ts:193-196): watch[index[k]] = word[k]. */
__global__ void k_watch_scatter(uint32_t *watch, const uint32_t *index, const uint32_t *word, uint32_t count, uint32_t n) {
  GRID_STRIDE(k, count) {
    uint32_t i = index[k];
    if (i < n) watch[i] = word[k];
  }
}

}  // namespace avbd
