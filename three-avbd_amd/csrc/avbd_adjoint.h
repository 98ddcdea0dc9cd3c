// Adjoint of the step with its active set frozen: tape recording, checkpoint copies and the
// reverse sweep over the unrolled Gauss-Seidel iterations.
//
// The reference has no adjoint. The differentiated map is the solve of each step (warm start,
// iterations x (primal per colour, dual), velocities) with everything discrete taken from the
// forward run: the contact set with its body-local anchors, normals and warm-start sources, the
// colouring, the penalties of every iteration, the normal clamp and friction-cone branch of every
// contact evaluation, the adaptive warm-start branches and the pivot floor. Poses, velocities,
// λ, the constraint values c0 (recomputed from the step-start poses) and the bodies' friction are
// differentiated. The forward stays f32 and bit-identical to the plain step; the reverse runs in
// f64 at the forward's recorded f32 states, because contact dynamics over many steps amplify
// round-off by orders of magnitude. Branch bits are recomputed with the forward's own f32
// evalContact on the recorded inputs, so they are the forward's decisions exactly.
//
// Per step the reverse is: velocities, then for each iteration from the last the dual (per
// contact) and the colours from the last (per body: solve the adjoint of the body's 6x6 system
// with the LDLᵀ factors recorded by the forward, then each contact point's rows back through its
// evaluation), then the λ warm start, the c0 and friction terms, and the bodies' warm start. A
// contact's parts for one of its bodies wait in a per-contact slot until that body's colour comes
// up, so the sweep needs no float atomics (except for a dynamic body no colour sweeps). Within one
// colour no two bodies share a contact, so a contact's own adjoints are written by one thread.
//
// The reverse is checked against the oracle's f64 reverse to a relative tolerance, not bit for
// bit: multiply-adds in this file may be fused (the f32 forward's arithmetic is elsewhere and
// stays unfused). The pragma is undone at the end of the file.
#pragma once
#include "avbd_kernels.h"

#pragma clang fp contract(fast)
namespace avbd {

// ---- f64 vectors and quaternions -----------------------------------------------------------
struct DV { double x, y, z; };
struct DQ { double x, y, z, w; };
__device__ __forceinline__ DV dv(double x, double y, double z) { DV r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ DQ dq(double x, double y, double z, double w) { DQ r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
__device__ __forceinline__ DV dv3(float4 a) { return dv(a.x, a.y, a.z); }
__device__ __forceinline__ DQ dq4(float4 a) { return dq(a.x, a.y, a.z, a.w); }
__device__ __forceinline__ DV operator+(DV a, DV b) { return dv(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ DV operator-(DV a, DV b) { return dv(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ DV operator*(DV a, double s) { return dv(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ DV operator-(DV a) { return dv(-a.x, -a.y, -a.z); }
__device__ __forceinline__ DQ operator+(DQ a, DQ b) { return dq(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ __forceinline__ double ddot(DV a, DV b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double ddot4(DQ a, DQ b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
__device__ __forceinline__ DV dcross(DV a, DV b) { return dv(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
__device__ __forceinline__ double dcomp(DV a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
__device__ __forceinline__ DQ dqmul(DQ a, DQ b) {
  return dq(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
__device__ __forceinline__ DV dqsub(DQ a, DQ b) {
  double id = 1.0 / ddot4(b, b);
  DQ r = dqmul(a, dq(-b.x * id, -b.y * id, -b.z * id, b.w * id));
  return dv(r.x * 2.0, r.y * 2.0, r.z * 2.0);
}
__device__ __forceinline__ DQ dqadd(DQ a, DV w) {
  DQ m = dqmul(dq(w.x, w.y, w.z, 0.0), a);
  DQ s = dq(a.x + m.x * 0.5, a.y + m.y * 0.5, a.z + m.z * 0.5, a.w + m.w * 0.5);
  double il = 1.0 / sqrt(ddot4(s, s));
  return dq(s.x * il, s.y * il, s.z * il, s.w * il);
}
__device__ __forceinline__ DV dqrotate(DQ r, DV a) {
  DV u = dv(r.x, r.y, r.z);
  DV t = dcross(u, a) * 2.0;
  return a + t * r.w + dcross(u, t);
}
__device__ __forceinline__ void dorthonormal(DV n, DV *b) {
  DV t1 = fabs(n.x) > fabs(n.z) ? dv(-n.y, n.x, 0.0) : dv(0.0, -n.z, n.y);
  t1 = t1 * (1.0 / sqrt(ddot(t1, t1)));
  b[0] = n; b[1] = t1; b[2] = dcross(n, t1);
}
/** Reverse of dqadd(q, w) = normalize(q + ½(w,0)q): adds the adjoints of q and w for output adjoint g. */
__device__ __forceinline__ void qaddRev(DQ q, DV w, DQ g, DQ &qb, DV &wb) {
  DQ m = dqmul(dq(w.x, w.y, w.z, 0.0), q);
  DQ s = dq(q.x + m.x * 0.5, q.y + m.y * 0.5, q.z + m.z * 0.5, q.w + m.w * 0.5);
  double il = 1.0 / sqrt(ddot4(s, s));
  DQ o = dq(s.x * il, s.y * il, s.z * il, s.w * il);
  double og = ddot4(o, g);
  DQ mb = dq((g.x - o.x * og) * il, (g.y - o.y * og) * il, (g.z - o.z * og) * il, (g.w - o.w * og) * il);
  qb.x += mb.x + 0.5 * (w.z * mb.y - w.y * mb.z - w.x * mb.w);
  qb.y += mb.y + 0.5 * (-w.z * mb.x + w.x * mb.z - w.y * mb.w);
  qb.z += mb.z + 0.5 * (w.y * mb.x - w.x * mb.y - w.z * mb.w);
  qb.w += mb.w + 0.5 * (w.x * mb.x + w.y * mb.y + w.z * mb.z);
  wb.x += 0.5 * (q.w * mb.x - q.z * mb.y + q.y * mb.z - q.x * mb.w);
  wb.y += 0.5 * (q.z * mb.x + q.w * mb.y - q.x * mb.z - q.y * mb.w);
  wb.z += 0.5 * (-q.y * mb.x + q.x * mb.y + q.w * mb.z - q.z * mb.w);
}
/** Reverse of dqsub(a, b) = 2 vec(a b⁻¹). */
__device__ __forceinline__ void qsubRev(DQ a, DQ b, DV g, DQ &ab, DQ &bb) {
  double id = 1.0 / ddot4(b, b);
  DQ c = dq(-b.x * id, -b.y * id, -b.z * id, b.w * id);
  double gx = 2 * g.x, gy = 2 * g.y, gz = 2 * g.z;
  ab.x += gx * c.w - gy * c.z + gz * c.y;
  ab.y += gx * c.z + gy * c.w - gz * c.x;
  ab.z += -gx * c.y + gy * c.x + gz * c.w;
  ab.w += gx * c.x + gy * c.y + gz * c.z;
  DQ cb = dq(gx * a.w + gy * a.z - gz * a.y, -gx * a.z + gy * a.w + gz * a.x, gx * a.y - gy * a.x + gz * a.w, gx * a.x + gy * a.y + gz * a.z);
  double cc2 = 2 * ddot4(cb, c);
  bb.x += (-cb.x - b.x * cc2) * id;
  bb.y += (-cb.y - b.y * cc2) * id;
  bb.z += (-cb.z - b.z * cc2) * id;
  bb.w += (cb.w - b.w * cc2) * id;
}
/** Reverse of dqrotate(q, v) with respect to q. */
__device__ __forceinline__ void qrotateRev(DQ q, DV v, DV g, DQ &qb) {
  DV u = dv(q.x, q.y, q.z);
  DV t = dcross(u, v) * 2.0;
  DV tb = g * q.w + dcross(g, u);
  DV ub = dcross(t, g) + dcross(v, tb) * 2.0;
  qb.x += ub.x; qb.y += ub.y; qb.z += ub.z; qb.w += ddot(g, t);
}

// ---- tape, adjoint buffers ------------------------------------------------------------------
/** A segment of recorded steps (slot s = step within the segment). Snapshot k of a step: k = 0
 *  the warm-started state, k = it + 1 after iteration it. */
/** Per body update: L (15), 1/D (6), the step x (6), one spare. */
static const uint32_t SYS_WORDS = 28;
struct Tape {
  uint32_t n, iters, conCap, manCap, adjCap;
  float4 *p0, *q0, *inpos, *inrot, *v0, *w0;  // [s*n + i]: x-, inertial target, velocities at the step's start
  float4 *P, *R;                              // [(s*(iters+1) + k)*n + i]
  Contact *con;                               // [s*conCap + c]: as collision produced it
  int *src;                                   // [s*conCap + c]: previous step's contact it warm-started from, or -1
  float4 *lam, *pen;                          // [(s*(iters+1) + k)*conCap + c]
  uint4 *mIds;
  float4 *mGeo;                               // [s*manCap + m]
  uint32_t *adjStart;                         // [s*(n+1) + i]
  uint32_t *adjList;                          // [s*adjCap + e]
  uint32_t *colorStart;                       // [s*65 + c]
  uint32_t *colorBodies, *col;                // [s*n + k], colour of body i (255: not swept)
  uint32_t *counts;                           // [s*4]: manifolds, contacts, adjacency entries
  uint32_t *cman;                             // [s*conCap + c]: the manifold contact c belongs to
  uchar4 *bits;                               // [(s*iters + it)*conCap + c]: the forward's branch bits
  double *sys;                                // [((s*iters + it)*n + i)*SYS_WORDS]: f64 LDLᵀ factors and step of each update
  uint32_t ptCap;
  uint32_t *ptOff;                            // [s*(n+1) + k]: first point of the k-th swept body (colour-bucket order)
  uint4 *pts;                                 // [s*ptCap + p]: a body's contact point (contact, manifold, body | side << 31
                                              //   with side 0 = A / 1 = B, partner)
  __device__ __forceinline__ size_t snap(uint32_t s, uint32_t k) const { return (size_t)s * (iters + 1) + k; }
};
/** f64 adjoints. state: 16 per body (pos xyz, prev-up, rot xyzw, vel xyz, 0, angvel xyz, 0): the
 *  live pose adjoints during a step's reverse and the end-state adjoints between steps. */
struct AdjBuf {
  double *state;
  double *ip, *ir, *np, *nr, *vs;  // 4 per body: adjoints of x-, inertial target, start velocity via prev-up
  double *mu;                      // per body friction
  double *lamb, *lambPrev;         // 4 per contact: this step's λ, the previous step's
  double *c0b, *mub;               // 4 / 1 per contact of this step
  double *dpart;                   // 32 per contact: the dual's partials of the iteration for body A, then
                                   // body B (pose p, pad, q, x- p, pad, x- q)
  double *nslot;                   // 32 per contact, the same layout: a point's partials for the partner
                                   // body, written by the owner's colour, read by the partner's
};

// f64 adds into the adjoint buffers (device memory from hipMalloc, coarse-grained): the hardware
// float atomics (global_atomic_add_f64), not the compare-and-swap loop atomicAdd compiles to.
__device__ __forceinline__ void addD(double *p, double v) { (void)unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomicAddV(double *p, DV v) { addD(p, v.x); addD(p + 1, v.y); addD(p + 2, v.z); }
__device__ __forceinline__ void atomicAddQ(double *p, DQ v) { addD(p, v.x); addD(p + 1, v.y); addD(p + 2, v.z); addD(p + 3, v.w); }
__device__ __forceinline__ DV ldV(const double *p) { return dv(p[0], p[1], p[2]); }
__device__ __forceinline__ DQ ldQ(const double *p) { return dq(p[0], p[1], p[2], p[3]); }
__device__ __forceinline__ void stV(double *p, DV v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
__device__ __forceinline__ void stQ(double *p, DQ v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; p[3] = v.w; }
__device__ __forceinline__ double shx(double v, int m) { return __shfl_xor(v, m); }
__device__ __forceinline__ DV shxV(DV v, int m) { return dv(shx(v.x, m), shx(v.y, m), shx(v.z, m)); }
__device__ __forceinline__ DQ shxQ(DQ v, int m) { return dq(shx(v.x, m), shx(v.y, m), shx(v.z, m), shx(v.w, m)); }

// ---- recording ------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t contactCount(const Dev &d) { uint32_t k = d.counters[C_CONTACTS]; return k < d.conCap ? k : d.conCap; }

/** After the topology: contact set, warm-start sources, adjacency, colour buckets. */
__global__ void k_tape_topology(Dev d, Tape T, uint32_t s) {
  uint32_t M = manifoldCount(d), C = contactCount(d), E = d.adjStart[d.n], n = d.n;
  uint32_t total = max(max(max(M, C), max(E, n + 1)), MAX_COLORS + 1);
  GRID_STRIDE(i, total) {
    if (i < C) { T.con[(size_t)s * T.conCap + i] = d.con[i]; T.src[(size_t)s * T.conCap + i] = d.src[i]; }
    if (i < M) {
      uint4 ids = d.mIds[i];
      T.mIds[(size_t)s * T.manCap + i] = ids; T.mGeo[(size_t)s * T.manCap + i] = d.mGeo[i];
      for (uint32_t k = 0; k < (ids.w & 15u); k++) T.cman[(size_t)s * T.conCap + ids.z + k] = i;
    }
    if (i < E) T.adjList[(size_t)s * T.adjCap + i] = d.adjList[i];
    if (i <= n) T.adjStart[(size_t)s * (n + 1) + i] = d.adjStart[i];
    if (i < n) T.colorBodies[(size_t)s * n + i] = d.colorBodies[i];
    if (i < n) T.col[(size_t)s * n + i] = NO_COLOR;
    if (i <= MAX_COLORS) T.colorStart[s * 65 + i] = d.colorStart[i];
    if (i == 0) { T.counts[s * 4] = M; T.counts[s * 4 + 1] = C; T.counts[s * 4 + 2] = E; }
  }
}
/** Colour of every swept body (buckets below the step's cap). */
__global__ void k_tape_colors(Dev d, Tape T, uint32_t s, uint32_t cap) {
  const uint32_t *cs = d.colorStart;
  uint32_t total = cs[cap];
  GRID_STRIDE(k, total) {
    uint32_t lo = 0, hi = cap;  // cs[lo] <= k < cs[hi]
    while (hi - lo > 1) { uint32_t mid = (lo + hi) / 2; if (cs[mid] <= k) lo = mid; else hi = mid; }
    T.col[(size_t)s * d.n + d.colorBodies[k]] = lo;
  }
}
/** The contact points of every swept body (k-th in colour-bucket order) over its adjacency,
 *  counted (zero past the swept ones) for the exclusive scan into ptOff. */
__global__ void k_tape_point_count(Dev d, Tape T, uint32_t s, uint32_t cap) {
  uint32_t n = d.n, total = d.colorStart[cap];
  uint32_t *off = T.ptOff + (size_t)s * (n + 1);
  GRID_STRIDE(k, n + 1) {
    uint32_t cnt = 0;
    if (k < total) {
      uint32_t i = d.colorBodies[k];
      for (uint32_t e = d.adjStart[i], e1 = d.adjStart[i + 1]; e < e1; e++) {
        uint32_t a = d.adjList[e];
        if (a >= d.nj) cnt += d.mIds[a - d.nj].w & 15u;
      }
    }
    off[k] = cnt;
  }
}
/** The point list: each swept body's points in adjacency order at its scanned offset. */
__global__ void k_tape_point_fill(Dev d, Tape T, uint32_t s, uint32_t cap) {
  uint32_t n = d.n, total = d.colorStart[cap];
  const uint32_t *off = T.ptOff + (size_t)s * (n + 1);
  uint4 *pts = T.pts + (size_t)s * T.ptCap;
  GRID_STRIDE(k, total) {
    uint32_t i = d.colorBodies[k], o = off[k];
    for (uint32_t e = d.adjStart[i], e1 = d.adjStart[i + 1]; e < e1; e++) {
      uint32_t a = d.adjList[e];
      if (a < d.nj) continue;
      uint4 ids = d.mIds[a - d.nj];
      bool isA = ids.x == i;
      for (uint32_t q = 0; q < (ids.w & 15u); q++, o++)
        if (o < T.ptCap) pts[o] = make_uint4(ids.z + q, a - d.nj, isA ? i : i | 0x80000000u, isA ? ids.y : ids.x);
    }
  }
}
/** Snapshot k of step s: poses and the contacts' λ and penalties; k = 0 adds the step's start. */
__global__ void k_tape_snap(Dev d, Tape T, uint32_t s, uint32_t k) {
  uint32_t n = d.n, C = contactCount(d);
  size_t sn = T.snap(s, k);
  GRID_STRIDE(i, max(n, C)) {
    if (i < n) {
      T.P[sn * n + i] = d.b.pos[i];
      T.R[sn * n + i] = d.b.rot[i];
      if (k == 0) {
        size_t o = (size_t)s * n + i;
        T.p0[o] = d.b.ipos[i]; T.q0[o] = d.b.irot[i]; T.inpos[o] = d.b.inpos[i]; T.inrot[o] = d.b.inrot[i];
        T.v0[o] = d.b.vel[i]; T.w0[o] = d.b.angvel[i];
      }
    }
    if (i < C) { T.lam[sn * T.conCap + i] = d.con[i].lam; T.pen[sn * T.conCap + i] = d.con[i].pen; }
  }
}
/** Copies of counted prefixes (checkpoints of the previous step's manifolds and contacts). */
__global__ void k_copy_manifolds(const uint4 *ids, const float4 *geo, uint4 *ids2, float4 *geo2, const uint32_t *counters, uint32_t word, uint32_t cap) {
  uint32_t m = counters[word]; if (m > cap) m = cap;
  GRID_STRIDE(i, m) { ids2[i] = ids[i]; geo2[i] = geo[i]; }
}
__global__ void k_copy_contacts(const Contact *a, Contact *b, const uint32_t *counters, uint32_t word, uint32_t cap) {
  uint32_t m = counters[word]; if (m > cap) m = cap;
  GRID_STRIDE(i, m) b[i] = a[i];
}
__global__ void k_fill_f64(double *p, size_t n, double v) { for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v; }

// ---- one contact evaluation in f64 with the forward's branches --------------------------------
struct PoseView { DV p; DQ q; DV ip; DQ ir; };
struct DEv { DV rAW, rBW, dLinA, dLinB, dAngA, dAngB, C, G, F; double fs, bounds, s; };
__device__ __forceinline__ DEv evalD(const Contact &k, DV c0, const DV *l, DV pen, DV lam, uint32_t bits, double mu, double alpha,
                                     const PoseView &A, const PoseView &B) {
  DEv e;
  bool sph = (__float_as_uint(k.rA.w) & ~STICK_BIT) == SPHERE_FEATURE;
  e.rAW = dqrotate(sph ? A.ir : A.q, dv3(k.rA));
  e.rBW = dqrotate(sph ? B.ir : B.q, dv3(k.rB));
  e.dLinA = A.p - A.ip; e.dLinB = B.p - B.ip;
  e.dAngA = dqsub(A.q, A.ir); e.dAngB = dqsub(B.q, B.ir);
  double keep = 1.0 - alpha;
  double C[3], G[3];
  for (int r = 0; r < 3; r++) {
    C[r] = dcomp(c0, r) * keep + ddot(l[r], e.dLinA) - ddot(l[r], e.dLinB) + ddot(dcross(e.rAW, l[r]), e.dAngA) +
           ddot(dcross(e.rBW, -l[r]), e.dAngB);
    G[r] = dcomp(pen, r) * C[r] + dcomp(lam, r);
  }
  e.C = dv(C[0], C[1], C[2]);
  e.G = dv(G[0], G[1], G[2]);
  e.F = e.G;
  if (bits & 1u) e.F.x = 0.0;
  e.bounds = fabs(e.F.x) * mu;
  e.fs = sqrt(e.G.y * e.G.y + e.G.z * e.G.z);
  e.s = 1.0;
  if (bits & 2u) { e.s = e.fs > 0 ? e.bounds * (1.0 / e.fs) : 0.0; e.F.y = e.G.y * e.s; e.F.z = e.G.z * e.s; }
  return e;
}
/** Adjoints of one evaluation's inputs for output adjoint Fb and world-anchor adjoints rAb, rBb. */
struct EvBar { DV pA, pB, ipA, ipB, lam, c0; DQ qA, qB, irA, irB; double mu; };
__device__ __forceinline__ void evalRev(const DEv &e, const Contact &k, const DV *l, DV pen, uint32_t bits, double mu, double alpha,
                                        const PoseView &A, const PoseView &B, DV Fb, DV rAb, DV rBb, EvBar &o) {
  DV Gb = dv(0, 0, 0);
  double Fxb = Fb.x;
  if (bits & 2u) {
    if (e.fs > 0) {
      double sb = Fb.y * e.G.y + Fb.z * e.G.z, ifs = 1.0 / e.fs;
      Gb.y = e.s * Fb.y; Gb.z = e.s * Fb.z;
      double boundsb = sb * ifs, fsb = -sb * e.s * ifs;
      Gb.y += fsb * e.G.y * ifs; Gb.z += fsb * e.G.z * ifs;
      o.mu += fabs(e.F.x) * boundsb;
      Fxb += (e.F.x < 0 ? -mu : mu) * boundsb;
    }
  } else { Gb.y = Fb.y; Gb.z = Fb.z; }
  Gb.x = (bits & 1u) ? 0.0 : Fxb;
  o.lam = o.lam + Gb;
  DV Cb = dv(pen.x * Gb.x, pen.y * Gb.y, pen.z * Gb.z);
  double keep = 1.0 - alpha;
  DV dLinb = dv(0, 0, 0), dAngAb = dLinb, dAngBb = dLinb;
  for (int r = 0; r < 3; r++) {
    double cb = dcomp(Cb, r);
    if (r == 0) o.c0.x += keep * cb; else if (r == 1) o.c0.y += keep * cb; else o.c0.z += keep * cb;
    dLinb = dLinb + l[r] * cb;
    dAngAb = dAngAb + dcross(e.rAW, l[r]) * cb;
    rAb = rAb + dcross(l[r], e.dAngA) * cb;
    dAngBb = dAngBb - dcross(e.rBW, l[r]) * cb;
    rBb = rBb - dcross(l[r], e.dAngB) * cb;
  }
  o.pA = o.pA + dLinb; o.ipA = o.ipA - dLinb;
  o.pB = o.pB - dLinb; o.ipB = o.ipB + dLinb;
  qsubRev(A.q, A.ir, dAngAb, o.qA, o.irA);
  qsubRev(B.q, B.ir, dAngBb, o.qB, o.irB);
  bool sph = (__float_as_uint(k.rA.w) & ~STICK_BIT) == SPHERE_FEATURE;
  qrotateRev(sph ? A.ir : A.q, dv3(k.rA), rAb, sph ? o.irA : o.qA);
  qrotateRev(sph ? B.ir : B.q, dv3(k.rB), rBb, sph ? o.irB : o.qB);
}
__device__ __forceinline__ void zeroBar(EvBar &o) {
  o.pA = o.pB = o.ipA = o.ipB = o.lam = o.c0 = dv(0, 0, 0);
  o.qA = o.qB = o.irA = o.irB = dq(0, 0, 0, 0);
  o.mu = 0;
}
__device__ __forceinline__ bool dynamicBody(const Dev &d, uint32_t i) { return d.b.size[i].w > 0.0f; }
/** Adds pose and x- adjoints to a body other than the one a thread owns (f64 atomics). */
__device__ __forceinline__ void scatterBody(const Dev &d, const AdjBuf &A, uint32_t j, DV p, DQ q, DV ip, DQ ir) {
  if (!dynamicBody(d, j)) return;
  atomicAddV(A.state + 16 * (size_t)j, p);
  atomicAddQ(A.state + 16 * (size_t)j + 4, q);
  atomicAddV(A.ip + 4 * (size_t)j, ip);
  atomicAddQ(A.ir + 4 * (size_t)j, ir);
}

// ---- reverse kernels ------------------------------------------------------------------------
/** Velocities (BDF1) of step s in reverse; clears the step's accumulators. */
__global__ void k_adj_velocities(Dev d, Tape T, AdjBuf A, uint32_t s) {
  double dt = d.p.dt;
  DV up = dv(d.p.upx, d.p.upy, d.p.upz);
  GRID_STRIDE(i, d.n) {
    double *st = A.state + 16 * (size_t)i;
    stV(A.vs + 4 * (size_t)i, up * st[3]);  // prev-up(end) = vel(start)·up
    double *ip = A.ip + 4 * (size_t)i, *ir = A.ir + 4 * (size_t)i;
    for (int k = 0; k < 4; k++) { ip[k] = 0; ir[k] = 0; A.np[4 * (size_t)i + k] = 0; A.nr[4 * (size_t)i + k] = 0; }
    if (dynamicBody(d, i)) {
      DV vb = ldV(st + 8), wb = ldV(st + 12);
      stV(st, ldV(st) + vb * (1.0 / dt));
      stV(ip, vb * (-1.0 / dt));
      DQ qb = ldQ(st + 4), irb = dq(0, 0, 0, 0);
      qsubRev(dq4(T.R[T.snap(s, T.iters) * d.n + i]), dq4(T.q0[(size_t)s * d.n + i]), wb * (1.0 / dt), qb, irb);
      stQ(st + 4, qb);
      stQ(ir, irb);
    }
    st[3] = 0; for (int k = 8; k < 16; k++) st[k] = 0;
  }
}
/** One side's partials of a contact (16 doubles, two pads): pose p, q and x- p, q. */
__device__ __forceinline__ void stPart(double *o, DV p, DQ q, DV ip, DQ ir) {
  double2 *v = reinterpret_cast<double2 *>(o);
  v[0] = make_double2(p.x, p.y); v[1] = make_double2(p.z, 0.0);
  v[2] = make_double2(q.x, q.y); v[3] = make_double2(q.z, q.w);
  v[4] = make_double2(ip.x, ip.y); v[5] = make_double2(ip.z, 0.0);
  v[6] = make_double2(ir.x, ir.y); v[7] = make_double2(ir.z, ir.w);
}
__device__ __forceinline__ void addPart(const double *o, DV &p, DQ &q, DV &ip, DQ &ir) {
  const double2 *v = reinterpret_cast<const double2 *>(o);
  double2 v0 = v[0], v1 = v[1], v2 = v[2], v3 = v[3], v4 = v[4], v5 = v[5], v6 = v[6], v7 = v[7];
  p = p + dv(v0.x, v0.y, v1.x);
  q = q + dq(v2.x, v2.y, v3.x, v3.y);
  ip = ip + dv(v4.x, v4.y, v5.x);
  ir = ir + dq(v6.x, v6.y, v7.x, v7.y);
}
/** A contact's partials for one of its bodies, left for that body's colour in the reverse
 *  (k_adj_color reads them when it takes the body up); a dynamic body that no colour sweeps gets
 *  them added at once with f64 atomics. */
__device__ __forceinline__ void leavePart(const Dev &d, const AdjBuf &A, const uint32_t *colS, double *slot, uint32_t b, DV p, DQ q, DV ip,
                                          DQ ir) {
  if (colS[b] != NO_COLOR) stPart(slot, p, q, ip, ir);
  else scatterBody(d, A, b, p, q, ip, ir);
}
/** The dual of iteration `it` of step s in reverse, one contact per thread: each body's partials
 *  go to the contact's dpart slots. */
__global__ void __launch_bounds__(64) k_adj_dual(Dev d, Tape T, AdjBuf A, uint32_t s, uint32_t it) {
  uint32_t C = T.counts[s * 4 + 1], n = d.n;
  const float4 *P = T.P + T.snap(s, it + 1) * n, *R = T.R + T.snap(s, it + 1) * n;
  const float4 *p0 = T.p0 + (size_t)s * n, *q0 = T.q0 + (size_t)s * n;
  const float4 *lam = T.lam + T.snap(s, it) * T.conCap, *pen = T.pen + T.snap(s, it) * T.conCap;
  const uint32_t *colS = T.col + (size_t)s * n;
  double alpha = d.p.alpha;
  GRID_STRIDE(c, C) {
    size_t cc = (size_t)s * T.conCap + c;
    uint32_t m = T.cman[cc];
    if (m >= T.counts[s * 4]) continue;
    uint4 ids = T.mIds[(size_t)s * T.manCap + m];
    float4 geo = T.mGeo[(size_t)s * T.manCap + m];
    uint32_t a = ids.x, b = ids.y;
    DV l[3];
    dorthonormal(dv3(geo), l);
    double mu = sqrt((double)d.b.pos[a].w * (double)d.b.pos[b].w);
    PoseView PA{dv3(P[a]), dq4(R[a]), dv3(p0[a]), dq4(q0[a])}, PB{dv3(P[b]), dq4(R[b]), dv3(p0[b]), dq4(q0[b])};
    Contact k = T.con[cc];
    k.pen = pen[c];
    k.lam = lam[c];
    uint32_t bits = T.bits[((size_t)s * T.iters + it) * T.conCap + c].z;
    DEv e = evalD(k, dv(k.rB.w, k.pen.w, k.lam.w), l, dv3(k.pen), dv3(k.lam), bits, mu, alpha, PA, PB);
    double *lb = A.lamb + 4 * (size_t)c;
    DV Fb = ldV(lb);
    EvBar o;
    zeroBar(o);
    evalRev(e, k, l, dv3(k.pen), bits, mu, alpha, PA, PB, Fb, dv(0, 0, 0), dv(0, 0, 0), o);
    stV(lb, o.lam);
    double *cb = A.c0b + 4 * (size_t)c;
    stV(cb, ldV(cb) + o.c0);
    A.mub[c] += o.mu;
    leavePart(d, A, colS, A.dpart + 32 * (size_t)c, a, o.pA, o.qA, o.ipA, o.irA);
    leavePart(d, A, colS, A.dpart + 32 * (size_t)c + 16, b, o.pB, o.qB, o.ipB, o.irB);
  }
}
/** A body's f64 system: packed lower triangle of the 6x6 (lin 00 10 11 20 21 22, ang alike,
 *  cross rows angular / columns linear) and the right-hand side. */
struct DSys { double lin[6], ang[6], cr[9], r[6]; };
__device__ __forceinline__ void addRowD(DSys &S, DV jl, DV ja, double k, double f) {
  double L[3] = {jl.x, jl.y, jl.z}, G[3] = {ja.x, ja.y, ja.z};
  const int I[6] = {0, 1, 1, 2, 2, 2}, J[6] = {0, 0, 1, 0, 1, 2};
#pragma unroll
  for (int e = 0; e < 6; e++) { S.lin[e] += L[I[e]] * L[J[e]] * k; S.ang[e] += G[I[e]] * G[J[e]] * k; }
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) S.cr[i * 3 + j] += G[i] * L[j] * k;
#pragma unroll
  for (int i = 0; i < 3; i++) { S.r[i] += L[i] * f; S.r[3 + i] += G[i] * f; }
}
/** LDLᵀ of the system with the forward's pivot floor (1e-5 of the diagonal), in f64. */
struct DFac { double L[15], D[6], iD[6]; };
__device__ __forceinline__ double sysAt(const DSys &S, int i, int j) {  // i >= j
  const int tri[3][3] = {{0, -1, -1}, {1, 2, -1}, {3, 4, 5}};
  if (i < 3) return S.lin[tri[i][j]];
  if (j >= 3) return S.ang[tri[i - 3][j - 3]];
  return S.cr[(i - 3) * 3 + j];
}
__device__ __forceinline__ int lidx(int i, int j) { return i * (i - 1) / 2 + j; }  // i > j
__device__ __forceinline__ void factor(const DSys &S, DFac &F) {
#pragma unroll
  for (int k = 0; k < 6; k++) {
    double d = sysAt(S, k, k);
#pragma unroll
    for (int j = 0; j < k; j++) d -= F.L[lidx(k, j)] * F.L[lidx(k, j)] * F.D[j];
    double floor = sysAt(S, k, k) * 1e-5;
    F.D[k] = k == 0 ? d : (d > floor ? d : floor);
    F.iD[k] = 1.0 / F.D[k];
#pragma unroll
    for (int i = k + 1; i < 6; i++) {
      double v = sysAt(S, i, k);
#pragma unroll
      for (int j = 0; j < k; j++) v -= F.L[lidx(i, j)] * F.L[lidx(k, j)] * F.D[j];
      F.L[lidx(i, k)] = v * F.iD[k];
    }
  }
}
__device__ __forceinline__ void solveF(const DFac &F, const double *b, double *x) {
  double y[6];
#pragma unroll
  for (int i = 0; i < 6; i++) {
    y[i] = b[i];
#pragma unroll
    for (int j = 0; j < i; j++) y[i] -= F.L[lidx(i, j)] * y[j];
  }
#pragma unroll
  for (int i = 5; i >= 0; i--) {
    x[i] = y[i] * F.iD[i];
#pragma unroll
    for (int j = i + 1; j < 6; j++) x[i] -= F.L[lidx(j, i)] * x[j];
  }
}
/** The poses a contact point's evaluation read in the forward: the owner's before its update,
 *  the partner's after its update when the partner's colour came first (swept), before it
 *  otherwise; x- of both; returns the pair's friction (the geometric mean). Both versions of the
 *  partner are loaded so that no load waits on its colour. */
__device__ __forceinline__ double pointPoses(const Dev &d, const float4 *Pn, const float4 *Rn, const float4 *Po, const float4 *Ro,
                                             const float4 *p0, const float4 *q0, uint32_t i, uint32_t j, bool isA, bool swept,
                                             PoseView &PA, PoseView &PB) {
  float4 pjn = Pn[j], rjn = Rn[j], pjo = Po[j], rjo = Ro[j];
  float4 pj = swept ? pjn : pjo, rj = swept ? rjn : rjo, pi = Po[i], ri = Ro[i];
  uint32_t a = isA ? i : j, b = isA ? j : i;
  PA = PoseView{dv3(isA ? pi : pj), dq4(isA ? ri : rj), dv3(p0[a]), dq4(q0[a])};
  PB = PoseView{dv3(isA ? pj : pi), dq4(isA ? rj : ri), dv3(p0[b]), dq4(q0[b])};
  return sqrt((double)d.b.pos[i].w * (double)d.b.pos[j].w);
}
/** Recording, after the step: the f64 system of every swept body's update in every iteration at
 *  the state that update read (its own pose before the update, each partner's after its update
 *  when the partner's colour came first, before it otherwise; the recorded snapshots, λ and
 *  penalties and the forward's branch bits), factored and solved; the reverse sweep then only
 *  solves the adjoint system with these factors. Work items are (iteration, swept body); `lanes`
 *  threads per item share its recorded contact points and the system is reduced across them. */
__global__ void __launch_bounds__(64) k_tape_systems(Dev d, Tape T, uint32_t s, uint32_t cap, uint32_t lanes) {
  uint32_t n = d.n, total = T.colorStart[s * 65 + cap], items = total * T.iters;
  double dt = d.p.dt, dt2 = dt * dt, alpha = d.p.alpha;
  const float4 *p0 = T.p0 + (size_t)s * n, *q0 = T.q0 + (size_t)s * n;
  const uint32_t *off = T.ptOff + (size_t)s * (n + 1), *colS = T.col + (size_t)s * n;
  const uint4 *pts = T.pts + (size_t)s * T.ptCap;
  uint32_t slot = threadIdx.x / lanes, lane = threadIdx.x % lanes, per = blockDim.x / lanes;
  for (uint32_t base = blockIdx.x * per; base < items; base += gridDim.x * per) {
    uint32_t item = base + slot;
    bool live = item < items;
    uint32_t it = live ? item / total : 0, k = live ? item % total : 0;
    uint32_t i = live ? T.colorBodies[(size_t)s * n + k] : 0;
    DSys S;
#pragma unroll
    for (int e = 0; e < 6; e++) { S.lin[e] = 0; S.ang[e] = 0; S.r[e] = 0; }
#pragma unroll
    for (int e = 0; e < 9; e++) S.cr[e] = 0;
    const float4 *Pn = T.P + T.snap(s, it + 1) * n, *Rn = T.R + T.snap(s, it + 1) * n;
    const float4 *Po = T.P + T.snap(s, it) * n, *Ro = T.R + T.snap(s, it) * n;
    if (live) {
      uint32_t col = colS[i];
      if (lane == 0) {
        float4 mo = d.b.moment[i];
        double m = (double)d.b.size[i].w / dt2, I0 = mo.x / dt2, I1 = mo.y / dt2, I2 = mo.z / dt2;
        S.lin[0] = m; S.lin[2] = m; S.lin[5] = m;
        S.ang[0] = I0; S.ang[2] = I1; S.ang[5] = I2;
        DV dx = dv3(Po[i]) - dv3(T.inpos[(size_t)s * n + i]);
        DV dqv = dqsub(dq4(Ro[i]), dq4(T.inrot[(size_t)s * n + i]));
        S.r[0] = m * dx.x; S.r[1] = m * dx.y; S.r[2] = m * dx.z;
        S.r[3] = I0 * dqv.x; S.r[4] = I1 * dqv.y; S.r[5] = I2 * dqv.z;
      }
      const float4 *lam = T.lam + T.snap(s, it) * T.conCap, *pen = T.pen + T.snap(s, it) * T.conCap;
      const uchar4 *bitsIt = T.bits + ((size_t)s * T.iters + it) * T.conCap;
      for (uint32_t p = off[k] + lane, p1 = off[k + 1]; p < p1; p += lanes) {
        uint4 pt = pts[p];
        uint32_t c = pt.x, j = pt.w;
        bool isA = (pt.z >> 31) == 0;
        float4 geo = T.mGeo[(size_t)s * T.manCap + pt.y];
        bool swept = colS[j] < col;
        PoseView PA, PB;
        double mu = pointPoses(d, Pn, Rn, Po, Ro, p0, q0, i, j, isA, swept, PA, PB);
        DV l[3];
        dorthonormal(dv3(geo), l);
        Contact kc = T.con[(size_t)s * T.conCap + c];
        kc.pen = pen[c];
        kc.lam = lam[c];
        uchar4 bb = bitsIt[c];
        DEv ev = evalD(kc, dv(kc.rB.w, kc.pen.w, kc.lam.w), l, dv3(kc.pen), dv3(kc.lam), isA ? bb.x : bb.y, mu, alpha, PA, PB);
        DV r = isA ? ev.rAW : ev.rBW;
        double sg = isA ? 1.0 : -1.0;
        for (int row = 0; row < 3; row++) {
          DV jl = l[row] * sg;
          addRowD(S, jl, dcross(r, jl), row == 0 ? kc.pen.x : (row == 1 ? kc.pen.y : kc.pen.z), dcomp(ev.F, row));
        }
      }
    }
    for (uint32_t w = 1; w < lanes; w <<= 1) {
#pragma unroll
      for (int e = 0; e < 6; e++) { S.lin[e] += __shfl_xor(S.lin[e], w); S.ang[e] += __shfl_xor(S.ang[e], w); S.r[e] += __shfl_xor(S.r[e], w); }
#pragma unroll
      for (int e = 0; e < 9; e++) S.cr[e] += __shfl_xor(S.cr[e], w);
    }
    if (live && lane == 0) {
      DFac F;
      factor(S, F);
      double x[6];
      solveF(F, S.r, x);
      double *out = T.sys + (((size_t)s * T.iters + it) * n + i) * SYS_WORDS;
#pragma unroll
      for (int q = 0; q < 15; q++) out[q] = F.L[q];
#pragma unroll
      for (int q = 0; q < 6; q++) { out[15 + q] = F.iD[q]; out[21 + q] = x[q]; }
    }
  }
}
static const uint32_t ADJ_LANES = 8;
/** What the reverse of one contact point reads, loaded in one round. */
struct PointIn {
  uint32_t c, j;
  bool isA, swept;
  float4 geo;
  Contact kc;
  uint32_t bits;
  PoseView PA, PB;
  double mu;
  DV lamb, c0b;
  double mub;
};
__device__ __forceinline__ void pointIn(const Dev &d, const Tape &T, const AdjBuf &A, uint32_t s, uint32_t col, uint4 pt, const float4 *Pn,
                                        const float4 *Rn, const float4 *Po, const float4 *Ro, const float4 *p0, const float4 *q0,
                                        const float4 *lam, const float4 *pen, const uchar4 *bitsIt, const uint32_t *colS, PointIn &v) {
  v.c = pt.x; v.j = pt.w;
  v.isA = (pt.z >> 31) == 0;
  uint32_t i = pt.z & 0x7fffffffu;
  v.geo = T.mGeo[(size_t)s * T.manCap + pt.y];
  v.swept = colS[v.j] < col;  // already updated in this iteration
  v.kc = T.con[(size_t)s * T.conCap + v.c];
  v.kc.pen = pen[v.c];
  v.kc.lam = lam[v.c];
  uchar4 bb = bitsIt[v.c];
  v.bits = v.isA ? bb.x : bb.y;
  v.mu = pointPoses(d, Pn, Rn, Po, Ro, p0, q0, i, v.j, v.isA, v.swept, v.PA, v.PB);
  v.lamb = ldV(A.lamb + 4 * (size_t)v.c);
  v.c0b = ldV(A.c0b + 4 * (size_t)v.c);
  v.mub = A.mub[v.c];
}
/** The reverse of one contact point of the body: its three rows of the system and right-hand
 *  side (adjoint solve rb, step x) back through the f64 evaluation with the forward's branches.
 *  The body's own parts go to its accumulators, the partner's to the contact's slot. */
__device__ __forceinline__ void pointRev(const Dev &d, const AdjBuf &A, const uint32_t *colS, const PointIn &v, const double *rb, const double *x,
                                         double alpha, DV &pbar, DQ &qbar, DV &ipb, DQ &irb) {
  DV l[3];
  dorthonormal(dv3(v.geo), l);
  DV penD = dv3(v.kc.pen);
  DEv ev = evalD(v.kc, dv(v.kc.rB.w, v.kc.pen.w, v.kc.lam.w), l, penD, dv3(v.kc.lam), v.bits, v.mu, alpha, v.PA, v.PB);
  DV r = v.isA ? ev.rAW : ev.rBW;
  double sg = v.isA ? 1.0 : -1.0;
  DV Fb = dv(0, 0, 0), rW = dv(0, 0, 0);
  for (int row = 0; row < 3; row++) {
    DV jl = l[row] * sg, ja = dcross(r, jl);
    double J[6] = {jl.x, jl.y, jl.z, ja.x, ja.y, ja.z};
    double kk = dcomp(penD, row), Jx = 0, rJ = 0;
#pragma unroll
    for (int q = 0; q < 6; q++) { Jx += J[q] * x[q]; rJ += rb[q] * J[q]; }
    if (row == 0) Fb.x = rJ; else if (row == 1) Fb.y = rJ; else Fb.z = rJ;
    double f = dcomp(ev.F, row);
    DV jab = dv(f * rb[3] - kk * (Jx * rb[3] + rJ * x[3]), f * rb[4] - kk * (Jx * rb[4] + rJ * x[4]), f * rb[5] - kk * (Jx * rb[5] + rJ * x[5]));
    rW = rW + dcross(jl, jab);
  }
  EvBar ob;
  zeroBar(ob);
  evalRev(ev, v.kc, l, penD, v.bits, v.mu, alpha, v.PA, v.PB, Fb, v.isA ? rW : dv(0, 0, 0), v.isA ? dv(0, 0, 0) : rW, ob);
  stV(A.lamb + 4 * (size_t)v.c, v.lamb + ob.lam);
  stV(A.c0b + 4 * (size_t)v.c, v.c0b + ob.c0);
  A.mub[v.c] = v.mub + ob.mu;
  double *ns = A.nslot + 32 * (size_t)v.c;
  if (v.isA) {
    pbar = pbar + ob.pA; qbar = qbar + ob.qA; ipb = ipb + ob.ipA; irb = irb + ob.irA;
    if (dynamicBody(d, v.j)) leavePart(d, A, colS, ns + 16, v.j, ob.pB, ob.qB, ob.ipB, ob.irB);
  } else {
    pbar = pbar + ob.pB; qbar = qbar + ob.qB; ipb = ipb + ob.ipB; irb = irb + ob.irB;
    if (dynamicBody(d, v.j)) leavePart(d, A, colS, ns, v.j, ob.pA, ob.qA, ob.ipA, ob.irA);
  }
}
/** Colour `col` of iteration `it` of step s in reverse, eight lanes per body, each lane taking
 *  the body's recorded contact points k, k+8, ...
 *  1. The partials left for the body since its colour last came up: this iteration's dual
 *     (dpart), each partner swept after it in the forward (its colour came up earlier in this
 *     reverse iteration) and each partner swept before it (in the previous reverse iteration, or
 *     nothing yet: the slots start each step at zero). They complete the adjoint of the body's
 *     updated pose.
 *  2. The adjoint of the body's step Δ from that of its updated pose (p+ = p - Δlin,
 *     q+ = qadd(q, -Δang)) through the recorded LDLᵀ factors, and the inertial rows.
 *  3. Each point's three rows of the system and right-hand side back through the f64 evaluation
 *     with the forward's branches: the body's own parts summed across its lanes, the partner's
 *     left in the contact's slot. A contact's λ, c0 and friction adjoints are written by its one
 *     lane (within one colour no two bodies share a contact). */
__global__ void __launch_bounds__(64) k_adj_color(Dev d, Tape T, AdjBuf A, uint32_t s, uint32_t it, uint32_t col) {
  uint32_t n = d.n;
  uint32_t start = T.colorStart[s * 65 + col], count = T.colorStart[s * 65 + col + 1] - start;
  const float4 *Pn = T.P + T.snap(s, it + 1) * n, *Rn = T.R + T.snap(s, it + 1) * n;
  const float4 *Po = T.P + T.snap(s, it) * n, *Ro = T.R + T.snap(s, it) * n;
  const float4 *p0 = T.p0 + (size_t)s * n, *q0 = T.q0 + (size_t)s * n;
  const float4 *lam = T.lam + T.snap(s, it) * T.conCap, *pen = T.pen + T.snap(s, it) * T.conCap;
  const uchar4 *bitsIt = T.bits + ((size_t)s * T.iters + it) * T.conCap;
  const uint32_t *off = T.ptOff + (size_t)s * (n + 1), *colS = T.col + (size_t)s * n;
  const uint4 *pts = T.pts + (size_t)s * T.ptCap;
  double dt = d.p.dt, dt2 = dt * dt, alpha = d.p.alpha;
  uint32_t slot = threadIdx.x / ADJ_LANES, lane = threadIdx.x % ADJ_LANES, per = blockDim.x / ADJ_LANES;
  for (uint32_t base = blockIdx.x * per; base < count; base += gridDim.x * per) {  // whole waves: sums across lanes
    uint32_t k = base + slot;
    bool live = k < count;
    uint32_t i = live ? T.colorBodies[(size_t)s * n + start + k] : 0;
    uint32_t pBegin = live ? off[start + k] : 0, pEnd = live ? off[start + k + 1] : 0;
    // one round of loads: the body's recorded system and adjoint, this lane's first point and the
    // slots of all its points
    uint32_t p = pBegin + lane;
    bool first = p < pEnd;
    PointIn v;
    if (first) pointIn(d, T, A, s, col, pts[p], Pn, Rn, Po, Ro, p0, q0, lam, pen, bitsIt, colS, v);
    DV gp = dv(0, 0, 0), gip = gp;
    DQ gq = dq(0, 0, 0, 0), gir = gq;
    for (uint32_t e = p; e < pEnd; e += ADJ_LANES) {
      uint4 pt = pts[e];
      size_t side = pt.z >> 31;
      addPart(A.dpart + 32 * (size_t)pt.x + 16 * side, gp, gq, gip, gir);
      addPart(A.nslot + 32 * (size_t)pt.x + 16 * side, gp, gq, gip, gir);
    }
    double x[6] = {0, 0, 0, 0, 0, 0}, rb[6] = {0, 0, 0, 0, 0, 0};
    DFac F;
    double m = 0, I0 = 0, I1 = 0, I2 = 0;
    double *st = A.state + 16 * (size_t)i;
    DV P = dv(0, 0, 0);
    DQ Q = dq(0, 0, 0, 0), qi = Q, qin = Q;
    if (live) {
      float4 mo = d.b.moment[i];
      m = (double)d.b.size[i].w / dt2; I0 = mo.x / dt2; I1 = mo.y / dt2; I2 = mo.z / dt2;
      const double *sys = T.sys + (((size_t)s * T.iters + it) * n + i) * SYS_WORDS;
#pragma unroll
      for (int q = 0; q < 15; q++) F.L[q] = sys[q];
#pragma unroll
      for (int q = 0; q < 6; q++) { F.iD[q] = sys[15 + q]; x[q] = sys[21 + q]; F.D[q] = 0; }
      P = ldV(st); Q = ldQ(st + 4);
      qi = dq4(Ro[i]); qin = dq4(T.inrot[(size_t)s * n + i]);
    }
    for (uint32_t w = 1; w < ADJ_LANES; w <<= 1) {
      gp = gp + shxV(gp, w); gip = gip + shxV(gip, w);
      gq = gq + shxQ(gq, w); gir = gir + shxQ(gir, w);
    }
    DV pbar = dv(0, 0, 0), ipb = pbar, rl = pbar;
    DQ qbar = dq(0, 0, 0, 0), irb = qbar, nrb = qbar;
    if (live) {
      P = P + gp;
      DQ qb0 = dq(0, 0, 0, 0);
      DV wb = dv(0, 0, 0);
      qaddRev(qi, dv(-x[3], -x[4], -x[5]), Q + gq, qb0, wb);
      double xb[6] = {-P.x, -P.y, -P.z, -wb.x, -wb.y, -wb.z};
      solveF(F, xb, rb);
      if (lane == 0) {  // the inertial rows and the partials gathered above
        rl = dv(m * rb[0], m * rb[1], m * rb[2]);
        pbar = P + rl;
        qbar = qb0;
        qsubRev(qi, qin, dv(I0 * rb[3], I1 * rb[4], I2 * rb[5]), qbar, nrb);
        ipb = gip;
        irb = gir;
      }
    }
    if (first) pointRev(d, A, colS, v, rb, x, alpha, pbar, qbar, ipb, irb);
    for (uint32_t e = p + ADJ_LANES; e < pEnd; e += ADJ_LANES) {  // bodies with more than eight points
      pointIn(d, T, A, s, col, pts[e], Pn, Rn, Po, Ro, p0, q0, lam, pen, bitsIt, colS, v);
      pointRev(d, A, colS, v, rb, x, alpha, pbar, qbar, ipb, irb);
    }
    for (uint32_t w = 1; w < ADJ_LANES; w <<= 1) {
      pbar = pbar + shxV(pbar, w); ipb = ipb + shxV(ipb, w);
      qbar = qbar + shxQ(qbar, w); irb = irb + shxQ(irb, w);
    }
    if (live && lane == 0) {
      stV(st, pbar);
      stQ(st + 4, qbar);
      double *ip = A.ip + 4 * (size_t)i, *ir = A.ir + 4 * (size_t)i, *np = A.np + 4 * (size_t)i, *nr = A.nr + 4 * (size_t)i;
      stV(ip, ldV(ip) + ipb);
      stQ(ir, ldQ(ir) + irb);
      stV(np, ldV(np) - rl);
      stQ(nr, ldQ(nr) + nrb);
    }
  }
}
/** After the last reverse iteration of step s: the slots each swept body's partners of lower colour
 *  left in that iteration (its colour had already come up) into the body's adjoints. */
__global__ void k_adj_slots_final(Dev d, Tape T, AdjBuf A, uint32_t s, uint32_t cap) {
  uint32_t n = d.n, total = T.colorStart[s * 65 + cap];
  const uint32_t *off = T.ptOff + (size_t)s * (n + 1), *colS = T.col + (size_t)s * n;
  const uint4 *pts = T.pts + (size_t)s * T.ptCap;
  GRID_STRIDE(k, total) {
    uint32_t i = T.colorBodies[(size_t)s * n + k], ci = colS[i];
    DV p = dv(0, 0, 0), ip = p;
    DQ q = dq(0, 0, 0, 0), ir = q;
    bool any = false;
    for (uint32_t e = off[k], e1 = off[k + 1]; e < e1; e++) {
      uint4 pt = pts[e];
      if (colS[pt.w] >= ci) continue;
      addPart(A.nslot + 32 * (size_t)pt.x + 16 * (size_t)(pt.z >> 31), p, q, ip, ir);
      any = true;
    }
    if (!any) continue;
    double *st = A.state + 16 * (size_t)i, *ipp = A.ip + 4 * (size_t)i, *irp = A.ir + 4 * (size_t)i;
    stV(st, ldV(st) + p);
    stQ(st + 4, ldQ(st + 4) + q);
    stV(ipp, ldV(ipp) + ip);
    stQ(irp, ldQ(irp) + ir);
  }
}
/** Warm start of λ (to the previous step's contacts), c0 and friction of step s in reverse; clears
 *  this step's per-contact adjoints. */
__global__ void __launch_bounds__(64) k_adj_contacts(Dev d, Tape T, AdjBuf A, uint32_t s, int hasPrev) {
  uint32_t M = T.counts[s * 4], n = d.n;
  const float4 *q0 = T.q0 + (size_t)s * n;
  double decay = (double)d.p.alpha * (double)d.p.gamma;
  GRID_STRIDE(m, M) {
    uint4 ids = T.mIds[(size_t)s * T.manCap + m];
    float4 geo = T.mGeo[(size_t)s * T.manCap + m];
    uint32_t a = ids.x, b = ids.y;
    DV l[3];
    dorthonormal(dv3(geo), l);
    double fa = d.b.pos[a].w, fb = d.b.pos[b].w, mu = sqrt(fa * fb);
    double mub = 0;
    DV ipA = dv(0, 0, 0), ipB = ipA;
    DQ irA = dq(0, 0, 0, 0), irB = irA;
    uint32_t end = ids.z + (ids.w & 15u);
    for (uint32_t c = ids.z; c < end; c++) {
      size_t cc = (size_t)s * T.conCap + c;
      double *lb = A.lamb + 4 * (size_t)c, *cb = A.c0b + 4 * (size_t)c;
      int src = T.src[cc];
      if (hasPrev && src >= 0) atomicAddV(A.lambPrev + 4 * (size_t)src, ldV(lb) * decay);
      Contact k = T.con[cc];
      DV dd = l[0] * cb[0] + l[1] * cb[1] + l[2] * cb[2];
      ipA = ipA + dd;
      qrotateRev(dq4(q0[a]), dv3(k.rA), dd, irA);
      ipB = ipB - dd;
      qrotateRev(dq4(q0[b]), dv3(k.rB), -dd, irB);
      mub += A.mub[c];
      for (int r = 0; r < 4; r++) { lb[r] = 0; cb[r] = 0; }
      A.mub[c] = 0;
    }
    if (dynamicBody(d, a)) { atomicAddV(A.ip + 4 * (size_t)a, ipA); atomicAddQ(A.ir + 4 * (size_t)a, irA); }
    if (dynamicBody(d, b)) { atomicAddV(A.ip + 4 * (size_t)b, ipB); atomicAddQ(A.ir + 4 * (size_t)b, irB); }
    if (mu > 0) {
      addD(A.mu + a, mub * fb / (2 * mu));
      addD(A.mu + b, mub * fa / (2 * mu));
    }
  }
}
/** Warm start of the bodies of step s in reverse: the adjoints of the step's start state. */
__global__ void k_adj_bodies(Dev d, Tape T, AdjBuf A, uint32_t s) {
  double dt = d.p.dt, g = d.p.gravity;
  DV up = dv(d.p.upx, d.p.upy, d.p.upz);
  GRID_STRIDE(i, d.n) {
    double *st = A.state + 16 * (size_t)i;
    if (!dynamicBody(d, i)) { for (int k = 0; k < 16; k++) st[k] = 0; continue; }
    size_t o = (size_t)s * d.n + i;
    float4 vel4 = T.v0[o], av4 = T.w0[o];
    // the forward's f32 branch decisions (warmBodies)
    V vel = xyz(vel4), angVel = xyz(av4);
    float gf = d.p.gravity, dtf = d.p.dt;
    V upf = vv(d.p.upx, d.p.upy, d.p.upz);
    bool interior = false;
    if (fabsf(gf) > 0.0f) {
      float sgf = gf > 0.0f ? 1.0f : (gf < 0.0f ? -1.0f : 0.0f);
      float raw = (dot(vel, upf) - vel4.w) / dtf * sgf / fabsf(gf);
      interior = raw > 0.0f && raw < 1.0f;
    }
    bool slow = length(vel) + length(angVel) * d.b.moment[i].w < 0.1f;
    bool atRest = (d.p.flags & FLAG_START_AT_REST) && slow;
    DQ q0 = dq4(T.q0[o]);
    DV w0 = dv3(av4);
    DV P = ldV(st);
    DQ R = ldQ(st + 4);
    DV pb = ldV(A.ip + 4 * (size_t)i) + ldV(A.np + 4 * (size_t)i);
    DV vb = ldV(A.vs + 4 * (size_t)i) + ldV(A.np + 4 * (size_t)i) * dt;
    DQ qb = ldQ(A.ir + 4 * (size_t)i);
    DV wb = dv(0, 0, 0);
    double ub = 0;
    qaddRev(q0, w0 * dt, ldQ(A.nr + 4 * (size_t)i), qb, wb);
    if (atRest) {
      pb = pb + P;
      qb = qb + R;
    } else {
      pb = pb + P;
      vb = vb + P * dt;
      double wfb = g * dt * dt * ddot(up, P);
      if (interior) {
        double sg = g > 0 ? 1.0 : -1.0, f = sg / (dt * fabs(g));
        vb = vb + up * (wfb * f);
        ub -= wfb * f;
      }
      qaddRev(q0, w0 * dt, R, qb, wb);
    }
    stV(st, pb); st[3] = ub;
    stQ(st + 4, qb);
    stV(st + 8, vb); st[11] = 0;
    stV(st + 12, wb * dt); st[15] = 0;
  }
}
/** Remap of a λ adjoint from one ordering of a step's contacts to another (the same manifolds by
 *  (a, b), e.g. a checkpoint's copy and a re-run of the step): hash of the first ordering. */
__global__ void k_adj_remap_insert_counted(const uint4 *ids, const uint32_t *countWord, uint32_t cap, uint32_t *table, uint32_t mask) {
  uint32_t count = *countWord < cap ? *countWord : cap;
  GRID_STRIDE(m, count) {
    uint4 v = ids[m];
    uint32_t h = pairHash(v.x, v.y) & mask;
    for (uint32_t probe = 0; probe <= mask; probe++) {
      if (atomicCAS(&table[h], 0u, m + 1) == 0u) break;
      h = (h + 1) & mask;
    }
  }
}
__global__ void k_adj_remap_counted(const uint4 *from, const uint4 *to, const uint32_t *countWord, const uint32_t *table, uint32_t mask,
                                    const double *in, double *out) {
  uint32_t count = *countWord;
  GRID_STRIDE(m, count) {
    uint4 v = to[m];
    uint32_t h = pairHash(v.x, v.y) & mask;
    for (uint32_t probe = 0; probe <= mask; probe++) {
      uint32_t e = table[h];
      if (e == 0) break;
      uint4 f = from[e - 1];
      if (f.x == v.x && f.y == v.y) {
        uint32_t cnt = min(v.w & 15u, f.w & 15u);
        for (uint32_t k = 0; k < cnt; k++)
          for (int r = 0; r < 4; r++) out[4 * (size_t)(v.z + k) + r] = in[4 * (size_t)(f.z + k) + r];
        break;
      }
      h = (h + 1) & mask;
    }
  }
}

}  // namespace avbd
#pragma clang fp contract(off)
