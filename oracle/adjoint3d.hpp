// TEST INFRASTRUCTURE ONLY — the oracle's adjoint of the step with the active set frozen.
//
// The reference has no adjoint: this is the checker for the library's reverse sweep. A forward
// run of gpuo::Model (f32 or f64) records a Tape (gpu3d.hpp): per step the contact set with its
// body-local anchors, normals and friction, every contact's warm-start source, the colouring
// and adjacency, and per iteration the penalties, the clamp/cone branches of every contact
// evaluation and the pivot-floor shifts of every body update. Frozen replays that tape in f64
// as a smooth function of the initial state (pos, rot, vel, angvel, previous up-velocity) and of
// the bodies' friction: every continuous quantity the step computes is recomputed (constraint
// values c0 from the step-start poses, the 6x6 systems, λ, poses, velocities), every discrete
// one comes from the tape. Its reverse sweep is the derivative of that function, written out
// by hand; central finite differences of run() check it.
//
// The forward replayed here is the model's solve (gpu3d.hpp: warmStartBodies, accumulate,
// evalContact, finishBody, dualManifold, updateVelocities; wgsl-solve.ts in the reference) with
// each comparison replaced by its recorded outcome. The pivot floor enters as the recorded
// shift δ of the pivots: LDLᵀ with D_k = S_k + δ_k is the exact factorisation of H + diag(δ),
// so the reverse solves with the same factors. Scenes with joints or springs are refused.
#pragma once
#include "gpu3d.hpp"
#include <stdexcept>

namespace gpuo {

/** Per-body state slots of the adjoint interface: pos xyz + previous up-velocity, rot xyzw,
 *  vel xyz + 0, angvel xyz + 0. */
static const int ADJ_FLOATS = 16;

struct FrozenBody { double mass, moment[3], friction; bool dyn; };

struct Frozen {
  typedef Ops<double> O; typedef V3<double> V; typedef Q4<double> Q;
  Params p;
  Tape tape;
  std::vector<FrozenBody> bodies;
  size_t N = 0;
  uint32_t joints = 0;
  // state
  std::vector<V> pos, vel, angvel; std::vector<Q> rot; std::vector<double> prevUp;
  // per step, what the reverse needs
  struct Rec {
    std::vector<V> p0, v0, w0; std::vector<Q> q0; std::vector<double> u0;
    std::vector<V> inpos; std::vector<Q> inrot;
    std::vector<std::vector<V>> P; std::vector<std::vector<Q>> R;  // [0] initial guess, [it+1] after sweep it
    std::vector<std::vector<V>> lam;                                // [it] λ before iteration it; [iters] final
    std::vector<V> c0;
    std::vector<uint32_t> colorOf;                                  // 255: not swept
  };
  std::vector<Rec> recs;

  V up() const { return O::v(p.up[0], p.up[1], p.up[2]); }
  static void basisOf(const TapeContact &t, V b[3]) { O::orthonormal(O::v(t.n[0], t.n[1], t.n[2]), b); }
  static bool sphereKey(uint32_t key) { return (key & ~STICK_BIT) == SPHERE_FEATURE; }

  // ---------------- contact evaluation (evalContact with recorded branches) ----------------
  struct Ev {
    V rAW, rBW, dLinA, dLinB, dAngA, dAngB, C, G, F;
    double fs, bounds, s;
  };
  // ---------------- 6x6 LDLᵀ with recorded pivot shifts ----------------
  struct Sys { double H[6][6], r[6]; };
  static void solve6(const Sys &S, const double *delta, const double *rhs, double *x) {
    double L[6][6] = {{0}}, D[6];
    for (int k = 0; k < 6; k++) {
      double d = S.H[k][k];
      for (int j = 0; j < k; j++) d -= L[k][j] * L[k][j] * D[j];
      D[k] = d + delta[k];
      for (int i = k + 1; i < 6; i++) {
        double v = S.H[i][k];
        for (int j = 0; j < k; j++) v -= L[i][j] * L[k][j] * D[j];
        L[i][k] = v / D[k];
      }
    }
    double y[6];
    for (int i = 0; i < 6; i++) { y[i] = rhs[i]; for (int j = 0; j < i; j++) y[i] -= L[i][j] * y[j]; }
    for (int i = 5; i >= 0; i--) { x[i] = y[i] / D[i]; for (int j = i + 1; j < 6; j++) x[i] -= L[j][i] * x[j]; }
  }

  // ---------------- quaternion helpers and their reverses ----------------
  /** qadd(q, w) = normalize(q + 0.5 (w,0)⊗q): adds the adjoints of q and w for output adjoint g. */
  static void qaddRev(Q q, V w, Q g, Q &qb, V &wb) {
    Q m = O::qmul(O::q(w.x, w.y, w.z, 0.0), q);
    Q mm = O::q(q.x + m.x * 0.5, q.y + m.y * 0.5, q.z + m.z * 0.5, q.w + m.w * 0.5);
    double len = std::sqrt(O::dot4(mm, mm));
    Q o = O::q(mm.x / len, mm.y / len, mm.z / len, mm.w / len);
    double og = O::dot4(o, g);
    Q mb = O::q((g.x - o.x * og) / len, (g.y - o.y * og) / len, (g.z - o.z * og) / len, (g.w - o.w * og) / len);
    // m = q + 0.5 N(w) q, N(w) q = (w,0)⊗q
    qb.x += mb.x + 0.5 * (w.z * mb.y - w.y * mb.z - w.x * mb.w);
    qb.y += mb.y + 0.5 * (-w.z * mb.x + w.x * mb.z - w.y * mb.w);
    qb.z += mb.z + 0.5 * (w.y * mb.x - w.x * mb.y - w.z * mb.w);
    qb.w += mb.w + 0.5 * (w.x * mb.x + w.y * mb.y + w.z * mb.z);
    wb.x += 0.5 * (q.w * mb.x - q.z * mb.y + q.y * mb.z - q.x * mb.w);
    wb.y += 0.5 * (q.z * mb.x + q.w * mb.y - q.x * mb.z - q.y * mb.w);
    wb.z += 0.5 * (-q.y * mb.x + q.x * mb.y + q.w * mb.z - q.z * mb.w);
  }
  /** qsub(a, b) = 2 vec(a ⊗ b⁻¹): adds the adjoints of a and b for output adjoint g. */
  static void qsubRev(Q a, Q b, V g, Q &ab, Q &bb) {
    double d = O::dot4(b, b);
    Q c = O::q(-b.x / d, -b.y / d, -b.z / d, b.w / d);
    double gx = 2 * g.x, gy = 2 * g.y, gz = 2 * g.z;
    ab.x += gx * c.w - gy * c.z + gz * c.y;
    ab.y += gx * c.z + gy * c.w - gz * c.x;
    ab.z += -gx * c.y + gy * c.x + gz * c.w;
    ab.w += gx * c.x + gy * c.y + gz * c.z;
    Q cb = O::q(gx * a.w + gy * a.z - gz * a.y, -gx * a.z + gy * a.w + gz * a.x, gx * a.y - gy * a.x + gz * a.w,
                gx * a.x + gy * a.y + gz * a.z);
    double cc = O::dot4(cb, c);
    bb.x += -cb.x / d - 2 * b.x * cc / d;
    bb.y += -cb.y / d - 2 * b.y * cc / d;
    bb.z += -cb.z / d - 2 * b.z * cc / d;
    bb.w += cb.w / d - 2 * b.w * cc / d;
  }
  /** qrotate(q, v) = v + w t + u × t, t = 2 u × v: adds the adjoint of q for output adjoint g. */
  static void qrotateRev(Q q, V v, V g, Q &qb) {
    V u = O::v(q.x, q.y, q.z);
    V t = O::mul(O::cross(u, v), 2.0);
    V tb = O::add(O::mul(g, q.w), O::cross(g, u));
    V ub = O::add(O::cross(t, g), O::mul(O::cross(v, tb), 2.0));
    qb.x += ub.x; qb.y += ub.y; qb.z += ub.z; qb.w += O::dot(g, t);
  }

  // ---------------- construction ----------------
  /** bodyRecords: the 40-float records (f64) of the state the tape starts from. */
  Frozen(const Params &params, const Tape &t, const double *rec, size_t n, uint32_t jointCount) : p(params), tape(t), N(n), joints(jointCount) {
    if (jointCount) throw std::runtime_error("the adjoint covers contact-only scenes");
    bodies.resize(n);
    for (size_t i = 0; i < n; i++) {
      const double *b = rec + i * 40;
      bodies[i].mass = b[19]; bodies[i].moment[0] = b[20]; bodies[i].moment[1] = b[21]; bodies[i].moment[2] = b[22];
      bodies[i].friction = b[3]; bodies[i].dyn = b[19] > 0;
    }
    for (const TapeStep &s : tape.steps)
      if (s.clashes) throw std::runtime_error("the adjoint needs a proper colouring at every step");
  }

  // ---------------- forward replay ----------------
  void setState(const double *st, const double *mu) {
    pos.resize(N); rot.resize(N); vel.resize(N); angvel.resize(N); prevUp.resize(N);
    for (size_t i = 0; i < N; i++) {
      const double *s = st + i * ADJ_FLOATS;
      pos[i] = O::v(s[0], s[1], s[2]); prevUp[i] = s[3];
      rot[i] = O::q(s[4], s[5], s[6], s[7]);
      vel[i] = O::v(s[8], s[9], s[10]); angvel[i] = O::v(s[12], s[13], s[14]);
      if (mu) bodies[i].friction = mu[i];
    }
  }
  void getState(double *st) const {
    for (size_t i = 0; i < N; i++) {
      double *s = st + i * ADJ_FLOATS;
      s[0] = pos[i].x; s[1] = pos[i].y; s[2] = pos[i].z; s[3] = prevUp[i];
      s[4] = rot[i].x; s[5] = rot[i].y; s[6] = rot[i].z; s[7] = rot[i].w;
      s[8] = vel[i].x; s[9] = vel[i].y; s[10] = vel[i].z; s[11] = 0;
      s[12] = angvel[i].x; s[13] = angvel[i].y; s[14] = angvel[i].z; s[15] = 0;
    }
  }
  double frictionOf(const TapeContact &t) const { return std::sqrt(bodies[t.a].friction * bodies[t.b].friction); }

  struct BodyView { V p; Q q; V ip; Q ir; };
  /** The 6x6 system of body i at iteration `it` of step s, poses from (P, R) for the body and
   *  its neighbours per `version` (neighbours swept earlier in this iteration read Pn, else P). */
  void system(size_t s, uint32_t it, uint32_t i, const std::vector<V> &Pn, const std::vector<Q> &Rn, const std::vector<V> &Po,
              const std::vector<Q> &Ro, const std::vector<V> &lam, Sys &S) const {
    const TapeStep &ts = tape.steps[s];
    const Rec &R = recs[s];
    double dt = p.dt, dt2 = dt * dt;
    const FrozenBody &b = bodies[i];
    std::memset(&S, 0, sizeof(S));
    double m = b.mass / dt2;
    double I[3] = {b.moment[0] / dt2, b.moment[1] / dt2, b.moment[2] / dt2};
    uint32_t ci = R.colorOf[i];
    auto view = [&](uint32_t j) -> BodyView {
      bool swept = R.colorOf[j] < ci;  // already updated in this iteration
      BodyView v; v.p = swept ? Pn[j] : Po[j]; v.q = swept ? Rn[j] : Ro[j]; v.ip = R.p0[j]; v.ir = R.q0[j];
      return v;
    };
    BodyView me = view(i);
    for (int k = 0; k < 3; k++) S.H[k][k] = m;
    for (int k = 0; k < 3; k++) S.H[3 + k][3 + k] = I[k];
    V dx = O::sub(me.p, R.inpos[i]);
    V dq = O::qsub(me.q, R.inrot[i]);
    for (int k = 0; k < 3; k++) { S.r[k] = m * dx[k]; S.r[3 + k] = I[k] * dq[k]; }
    for (uint32_t e = ts.adjStart[i]; e < ts.adjStart[i + 1]; e++) {
      uint32_t mid = ts.adjList[e] - joints;
      for (uint32_t c = ts.manifoldFirst[mid]; c < ts.manifoldFirst[mid + 1]; c++) {
        const TapeContact &t = ts.contacts[c];
        bool isA = t.a == i;
        BodyView A = view(t.a), B = view(t.b);
        V l[3]; basisOf(t, l);
        Ev ev = evalFull(R.c0[c], l, &ts.pen[it][c * 3], lam[c], ts.branch[it][c * 3 + (isA ? 0 : 1)], frictionOf(t), t, A, B);
        double sg = isA ? 1.0 : -1.0;
        V r = isA ? ev.rAW : ev.rBW;
        for (int row = 0; row < 3; row++) {
          V jl = O::mul(l[row], sg), ja = O::cross(r, jl);
          double J[6] = {jl.x, jl.y, jl.z, ja.x, ja.y, ja.z};
          double k = ts.pen[it][c * 3 + row];
          for (int x = 0; x < 6; x++) {
            for (int y = 0; y < 6; y++) S.H[x][y] += k * J[x] * J[y];
            S.r[x] += J[x] * ev.F[row];
          }
        }
      }
    }
  }
  Ev evalFull(V c0, const V l[3], const double *pen, V lam, uint8_t bits, double mu, const TapeContact &t, const BodyView &A,
              const BodyView &B) const {
    Ev e;
    bool sph = sphereKey(t.key);
    e.rAW = O::qrotate(sph ? A.ir : A.q, O::v(t.rA[0], t.rA[1], t.rA[2]));
    e.rBW = O::qrotate(sph ? B.ir : B.q, O::v(t.rB[0], t.rB[1], t.rB[2]));
    e.dLinA = O::sub(A.p, A.ip); e.dLinB = O::sub(B.p, B.ip);
    e.dAngA = O::qsub(A.q, A.ir); e.dAngB = O::qsub(B.q, B.ir);
    double keep = 1.0 - p.alpha;
    for (int r = 0; r < 3; r++) {
      e.C[r] = c0[r] * keep + O::dot(l[r], e.dLinA) - O::dot(l[r], e.dLinB) + O::dot(O::cross(e.rAW, l[r]), e.dAngA) +
               O::dot(O::cross(e.rBW, O::neg(l[r])), e.dAngB);
      e.G[r] = pen[r] * e.C[r] + lam[r];
    }
    e.F = e.G;
    if (bits & 1) e.F.x = 0.0;
    e.bounds = std::fabs(e.F.x) * mu;
    e.fs = std::sqrt(e.G.y * e.G.y + e.G.z * e.G.z);
    e.s = 1.0;
    if (bits & 2) { e.s = e.fs > 0 ? e.bounds / e.fs : 0.0; e.F.y = e.G.y * e.s; e.F.z = e.G.z * e.s; }
    return e;
  }

  void step(size_t s) {
    const TapeStep &ts = tape.steps[s];
    Rec &R = recs[s];
    double dt = p.dt, g = p.gravity;
    V upv = up();
    R.p0 = pos; R.q0 = rot; R.v0 = vel; R.w0 = angvel; R.u0 = prevUp;
    R.inpos.resize(N); R.inrot.resize(N);
    for (size_t i = 0; i < N; i++) {
      bool dyn = bodies[i].dyn;
      V inpos = O::add(pos[i], O::mul(vel[i], dt));
      if (dyn) inpos = O::add(inpos, O::mul(upv, g * (dt * dt)));
      R.inpos[i] = inpos;
      R.inrot[i] = O::qadd(rot[i], O::mul(angvel[i], dt));
      if (dyn) {
        if (!ts.atRest[i]) {
          double w = wOf(s, i, vel[i], prevUp[i]);
          pos[i] = O::add(O::add(pos[i], O::mul(vel[i], dt)), O::mul(upv, g * (w * dt * dt)));
          rot[i] = O::qadd(rot[i], O::mul(angvel[i], dt));
        }
      } else if (vel[i].x != 0 || vel[i].y != 0 || vel[i].z != 0 || angvel[i].x != 0 || angvel[i].y != 0 || angvel[i].z != 0) {
        pos[i] = O::add(pos[i], O::mul(vel[i], dt));
        rot[i] = O::qadd(rot[i], O::mul(angvel[i], dt));
      }
    }
    size_t M = ts.contacts.size();
    R.c0.resize(M);
    std::vector<V> lam(M);
    const std::vector<V> *prevLam = s > 0 ? &recs[s - 1].lam.back() : nullptr;
    for (size_t c = 0; c < M; c++) {
      const TapeContact &t = ts.contacts[c];
      V l[3]; basisOf(t, l);
      V d = O::sub(O::add(O::qrotate(R.q0[t.a], O::v(t.rA[0], t.rA[1], t.rA[2])), R.p0[t.a]),
                   O::add(O::qrotate(R.q0[t.b], O::v(t.rB[0], t.rB[1], t.rB[2])), R.p0[t.b]));
      R.c0[c] = O::v(O::dot(l[0], d) + p.contactMargin, O::dot(l[1], d), O::dot(l[2], d));
      if (t.src >= 0 && prevLam) lam[c] = O::mul((*prevLam)[t.src], p.alpha * p.gamma);
      else if (t.src >= 0) lam[c] = O::mul(O::v(t.lam0[0], t.lam0[1], t.lam0[2]), 1.0);  // first step: the seeded warm start
      else lam[c] = O::v(0, 0, 0);
    }
    R.colorOf.assign(N, 255u);
    for (uint32_t c = 0; c < ts.colorCapUsed; c++)
      for (uint32_t k = ts.bucketStart[c]; k < ts.bucketStart[c + 1]; k++) R.colorOf[ts.bucketBodies[k]] = c;
    R.P.assign(1, pos); R.R.assign(1, rot); R.lam.clear();
    for (uint32_t it = 0; it < p.iterations; it++) {
      R.lam.push_back(lam);
      std::vector<V> Pn = pos; std::vector<Q> Rn = rot;
      const std::vector<V> &Po = R.P.back(); const std::vector<Q> &Ro = R.R.back();
      for (uint32_t c = 0; c < ts.colorCapUsed; c++)
        for (uint32_t k = ts.bucketStart[c]; k < ts.bucketStart[c + 1]; k++) {
          uint32_t i = ts.bucketBodies[k];
          Sys S; system(s, it, i, Pn, Rn, Po, Ro, lam, S);
          double x[6]; solve6(S, &ts.delta[it][(size_t)i * 6], S.r, x);
          Pn[i] = O::sub(Po[i], O::v(x[0], x[1], x[2]));
          Rn[i] = O::qadd(Ro[i], O::neg(O::v(x[3], x[4], x[5])));
        }
      R.P.push_back(Pn); R.R.push_back(Rn);
      for (size_t c = 0; c < M; c++) {
        const TapeContact &t = ts.contacts[c];
        V l[3]; basisOf(t, l);
        BodyView A{Pn[t.a], Rn[t.a], R.p0[t.a], R.q0[t.a]}, B{Pn[t.b], Rn[t.b], R.p0[t.b], R.q0[t.b]};
        lam[c] = evalFull(R.c0[c], l, &ts.pen[it][c * 3], R.lam[it][c], ts.branch[it][c * 3 + 2], frictionOf(t), t, A, B).F;
      }
    }
    R.lam.push_back(lam);
    pos = R.P.back(); rot = R.R.back();
    for (size_t i = 0; i < N; i++) {
      double pu = O::dot(vel[i], upv);
      if (bodies[i].dyn) {
        vel[i] = O::divs(O::sub(pos[i], R.p0[i]), dt);
        angvel[i] = O::divs(O::qsub(rot[i], R.q0[i]), dt);
      }
      prevUp[i] = pu;
    }
  }
  double wOf(size_t s, size_t i, V v, double u) const {
    const TapeStep &ts = tape.steps[s];
    double g = p.gravity;
    if (ts.wMode[i] == 1) return 0.0;
    if (ts.wMode[i] == 2) return 1.0;
    double sg = g > 0 ? 1.0 : -1.0;
    return (O::dot(v, up()) - u) / p.dt * sg / std::fabs(g);
  }
  /** Replays every recorded step from state `st` (ADJ_FLOATS per body) with friction `mu`
   *  (nullptr: the recorded bodies' friction); writes the final state. */
  void run(const double *st, const double *mu, double *out) {
    setState(st, mu);
    recs.assign(tape.steps.size(), Rec());
    for (size_t s = 0; s < tape.steps.size(); s++) step(s);
    if (out) getState(out);
  }

  /** Linearise at the recorded forward's own states instead of a replay: fills what backward()
   *  reads from the tape (the model's f32 states, for a reverse at the f32 trajectory). */
  void loadRecorded(const double *mu) {
    if (mu) for (size_t i = 0; i < N; i++) bodies[i].friction = mu[i];
    recs.assign(tape.steps.size(), Rec());
    auto vec = [](const std::vector<double> &x, size_t i) { return O::v(x[3 * i], x[3 * i + 1], x[3 * i + 2]); };
    auto quat = [](const std::vector<double> &x, size_t i) { return O::q(x[4 * i], x[4 * i + 1], x[4 * i + 2], x[4 * i + 3]); };
    for (size_t s = 0; s < tape.steps.size(); s++) {
      const TapeStep &ts = tape.steps[s];
      Rec &R = recs[s];
      if (ts.P.size() != p.iterations + 1) throw std::runtime_error("the tape holds no states");
      for (size_t i = 0; i < N; i++) {
        R.p0.push_back(vec(ts.p0, i)); R.q0.push_back(quat(ts.q0, i)); R.v0.push_back(vec(ts.v0, i));
        R.w0.push_back(vec(ts.w0, i)); R.u0.push_back(ts.u0[i]);
        R.inpos.push_back(vec(ts.inpos, i)); R.inrot.push_back(quat(ts.inrot, i));
      }
      R.P.resize(ts.P.size()); R.R.resize(ts.R.size());
      for (size_t k = 0; k < ts.P.size(); k++)
        for (size_t i = 0; i < N; i++) { R.P[k].push_back(vec(ts.P[k], i)); R.R[k].push_back(quat(ts.R[k], i)); }
      R.lam.resize(ts.lam.size());
      for (size_t k = 0; k < ts.lam.size(); k++)
        for (size_t c = 0; c < ts.contacts.size(); c++) R.lam[k].push_back(vec(ts.lam[k], c));
      for (const TapeContact &t : ts.contacts) R.c0.push_back(O::v(t.c0[0], t.c0[1], t.c0[2]));
      R.colorOf.assign(N, 255u);
      for (uint32_t c = 0; c < ts.colorCapUsed; c++)
        for (uint32_t k = ts.bucketStart[c]; k < ts.bucketStart[c + 1]; k++) R.colorOf[ts.bucketBodies[k]] = c;
    }
  }

  // ---------------- reverse ----------------
  struct Adj { std::vector<V> P, Ip, Np, V_, W; std::vector<Q> R, Ir, Nr; std::vector<double> U, mu; };
  /** Adds the adjoints of the inputs of one contact evaluation, for output adjoint Fb and the
   *  adjoints rAb/rBb of its world anchors. */
  void evalRev(const Ev &e, const TapeContact &t, const V l[3], const double *pen, uint8_t bits, double mu, const BodyView &A,
               const BodyView &B, V Fb, V rAb, V rBb, V &pAb, Q &qAb, V &pBb, Q &qBb, V &ipAb, Q &irAb, V &ipBb, Q &irBb, V &lamb,
               V &c0b, double &mub) const {
    V Gb = O::v(0, 0, 0);
    double Fxb = Fb.x;
    if (bits & 2) {
      if (e.fs > 0) {
        double sb = Fb.y * e.G.y + Fb.z * e.G.z;
        Gb.y = e.s * Fb.y; Gb.z = e.s * Fb.z;
        double boundsb = sb / e.fs, fsb = -sb * e.s / e.fs;
        Gb.y += fsb * e.G.y / e.fs; Gb.z += fsb * e.G.z / e.fs;
        mub += std::fabs(e.F.x) * boundsb;
        Fxb += (e.F.x < 0 ? -mu : mu) * boundsb;
      }
    } else { Gb.y = Fb.y; Gb.z = Fb.z; }
    Gb.x = (bits & 1) ? 0.0 : Fxb;
    lamb = O::add(lamb, Gb);
    V Cb = O::mulv(O::v(pen[0], pen[1], pen[2]), Gb);
    double keep = 1.0 - p.alpha;
    V dLinb = O::v(0, 0, 0), dAngAb = dLinb, dAngBb = dLinb;
    for (int r = 0; r < 3; r++) {
      c0b[r] += keep * Cb[r];
      dLinb = O::add(dLinb, O::mul(l[r], Cb[r]));
      dAngAb = O::add(dAngAb, O::mul(O::cross(e.rAW, l[r]), Cb[r]));
      rAb = O::add(rAb, O::mul(O::cross(l[r], e.dAngA), Cb[r]));
      dAngBb = O::sub(dAngBb, O::mul(O::cross(e.rBW, l[r]), Cb[r]));
      rBb = O::sub(rBb, O::mul(O::cross(l[r], e.dAngB), Cb[r]));
    }
    pAb = O::add(pAb, dLinb); ipAb = O::sub(ipAb, dLinb);
    pBb = O::sub(pBb, dLinb); ipBb = O::add(ipBb, dLinb);
    qsubRev(A.q, A.ir, dAngAb, qAb, irAb);
    qsubRev(B.q, B.ir, dAngBb, qBb, irBb);
    bool sph = sphereKey(t.key);
    qrotateRev(sph ? A.ir : A.q, O::v(t.rA[0], t.rA[1], t.rA[2]), rAb, sph ? irAb : qAb);
    qrotateRev(sph ? B.ir : B.q, O::v(t.rB[0], t.rB[1], t.rB[2]), rBb, sph ? irBb : qBb);
  }

  /** Reverse sweep over the last run(): `dFinal` is the adjoint of the final state
   *  (ADJ_FLOATS per body); writes the adjoint of the initial state and of each body's friction. */
  void backward(const double *dFinal, double *dInit, double *dMu) {
    typedef V VV;
    double dt = p.dt, g = p.gravity;
    V upv = up();
    Adj a;
    a.P.assign(N, O::v(0, 0, 0)); a.V_ = a.P; a.W = a.P; a.Ip = a.P; a.Np = a.P;
    a.R.assign(N, O::q(0, 0, 0, 0)); a.Ir = a.R; a.Nr = a.R;
    a.U.assign(N, 0.0); a.mu.assign(N, 0.0);
    for (size_t i = 0; i < N; i++) {
      const double *s = dFinal + i * ADJ_FLOATS;
      a.P[i] = O::v(s[0], s[1], s[2]); a.U[i] = s[3]; a.R[i] = O::q(s[4], s[5], s[6], s[7]);
      a.V_[i] = O::v(s[8], s[9], s[10]); a.W[i] = O::v(s[12], s[13], s[14]);
    }
    std::vector<VV> lamb;  // adjoint of the current step's final λ
    for (size_t s = tape.steps.size(); s-- > 0;) {
      const TapeStep &ts = tape.steps[s];
      const Rec &R = recs[s];
      size_t M = ts.contacts.size();
      if (lamb.size() != M) lamb.assign(M, O::v(0, 0, 0));
      std::vector<VV> c0b(M, O::v(0, 0, 0));
      std::vector<double> mub(M, 0.0);
      std::vector<V> vStart(N, O::v(0, 0, 0));
      for (size_t i = 0; i < N; i++) {
        a.Ip[i] = O::v(0, 0, 0); a.Np[i] = a.Ip[i]; a.Ir[i] = O::q(0, 0, 0, 0); a.Nr[i] = a.Ir[i];
        vStart[i] = O::mul(upv, a.U[i]);  // prevUp(end) = vel(start)·up
        if (!bodies[i].dyn) continue;
        const Q qe = R.R.back()[i];
        a.P[i] = O::add(a.P[i], O::divs(a.V_[i], dt));
        a.Ip[i] = O::sub(a.Ip[i], O::divs(a.V_[i], dt));
        qsubRev(qe, R.q0[i], O::divs(a.W[i], dt), a.R[i], a.Ir[i]);
      }
      for (uint32_t it = p.iterations; it-- > 0;) {
        const std::vector<V> &Pn = R.P[it + 1]; const std::vector<Q> &Rn = R.R[it + 1];
        const std::vector<V> &Po = R.P[it]; const std::vector<Q> &Ro = R.R[it];
        // dual of iteration it
        for (size_t c = 0; c < M; c++) {
          const TapeContact &t = ts.contacts[c];
          V l[3]; basisOf(t, l);
          BodyView A{Pn[t.a], Rn[t.a], R.p0[t.a], R.q0[t.a]}, B{Pn[t.b], Rn[t.b], R.p0[t.b], R.q0[t.b]};
          double mu = frictionOf(t);
          const double *pen = &ts.pen[it][c * 3];
          uint8_t bits = ts.branch[it][c * 3 + 2];
          Ev e = evalFull(R.c0[c], l, pen, R.lam[it][c], bits, mu, t, A, B);
          V Fb = lamb[c];
          lamb[c] = O::v(0, 0, 0);
          V pAb = O::v(0, 0, 0), pBb = pAb, ipAb = pAb, ipBb = pAb;
          Q qAb = O::q(0, 0, 0, 0), qBb = qAb, irAb = qAb, irBb = qAb;
          evalRev(e, t, l, pen, bits, mu, A, B, Fb, O::v(0, 0, 0), O::v(0, 0, 0), pAb, qAb, pBb, qBb, ipAb, irAb, ipBb, irBb, lamb[c],
                  c0b[c], mub[c]);
          addBody(a, t.a, pAb, qAb, ipAb, irAb);
          addBody(a, t.b, pBb, qBb, ipBb, irBb);
        }
        // primal sweep of iteration it, colours in reverse
        for (uint32_t col = ts.colorCapUsed; col-- > 0;)
          for (uint32_t k = ts.bucketStart[col]; k < ts.bucketStart[col + 1]; k++) {
            uint32_t i = ts.bucketBodies[k];
            Sys S; system(s, it, i, Pn, Rn, Po, Ro, R.lam[it], S);
            const double *delta = &ts.delta[it][(size_t)i * 6];
            double x[6]; solve6(S, delta, S.r, x);
            // p+ = p - Δlin ; q+ = qadd(q, -Δang)
            V pb = a.P[i];
            Q qb = O::q(0, 0, 0, 0);
            V wb = O::v(0, 0, 0);
            qaddRev(Ro[i], O::neg(O::v(x[3], x[4], x[5])), a.R[i], qb, wb);
            double xb[6] = {-a.P[i].x, -a.P[i].y, -a.P[i].z, -wb.x, -wb.y, -wb.z};
            double rb[6]; solve6(S, delta, xb, rb);
            const FrozenBody &bd = bodies[i];
            double m = bd.mass / (dt * dt);
            double I[3] = {bd.moment[0] / (dt * dt), bd.moment[1] / (dt * dt), bd.moment[2] / (dt * dt)};
            V rl = O::v(m * rb[0], m * rb[1], m * rb[2]);
            pb = O::add(pb, rl); a.Np[i] = O::sub(a.Np[i], rl);
            qsubRev(Ro[i], R.inrot[i], O::v(I[0] * rb[3], I[1] * rb[4], I[2] * rb[5]), qb, a.Nr[i]);
            V ipb = O::v(0, 0, 0); Q irb = O::q(0, 0, 0, 0);
            uint32_t ci = R.colorOf[i];
            for (uint32_t e = ts.adjStart[i]; e < ts.adjStart[i + 1]; e++) {
              uint32_t mid = ts.adjList[e] - joints;
              for (uint32_t c = ts.manifoldFirst[mid]; c < ts.manifoldFirst[mid + 1]; c++) {
                const TapeContact &t = ts.contacts[c];
                bool isA = t.a == i;
                uint32_t j = isA ? t.b : t.a;
                bool swept = R.colorOf[j] < ci;
                BodyView me{Po[i], Ro[i], R.p0[i], R.q0[i]};
                BodyView nb{swept ? Pn[j] : Po[j], swept ? Rn[j] : Ro[j], R.p0[j], R.q0[j]};
                const BodyView &A = isA ? me : nb, &B = isA ? nb : me;
                V l[3]; basisOf(t, l);
                double mu = frictionOf(t);
                const double *pen = &ts.pen[it][c * 3];
                uint8_t bits = ts.branch[it][c * 3 + (isA ? 0 : 1)];
                Ev ev = evalFull(R.c0[c], l, pen, R.lam[it][c], bits, mu, t, A, B);
                double sg = isA ? 1.0 : -1.0;
                V r = isA ? ev.rAW : ev.rBW;
                V Fb = O::v(0, 0, 0), rW = O::v(0, 0, 0);
                for (int row = 0; row < 3; row++) {
                  V jl = O::mul(l[row], sg), ja = O::cross(r, jl);
                  double J[6] = {jl.x, jl.y, jl.z, ja.x, ja.y, ja.z};
                  double k = pen[row], Jx = 0, rJ = 0;
                  for (int q = 0; q < 6; q++) { Jx += J[q] * x[q]; rJ += rb[q] * J[q]; }
                  Fb[row] = rJ;
                  V jab;
                  for (int q = 0; q < 3; q++) jab[q] = ev.F[row] * rb[3 + q] - k * (Jx * rb[3 + q] + rJ * x[3 + q]);
                  rW = O::add(rW, O::mul(O::cross(jl, jab), 1.0));  // ja = r × jl  =>  r̄ += jl × j̄a
                }
                V pAb = O::v(0, 0, 0), pBb = pAb, ipAb = pAb, ipBb = pAb;
                Q qAb = O::q(0, 0, 0, 0), qBb = qAb, irAb = qAb, irBb = qAb;
                evalRev(ev, t, l, pen, bits, mu, A, B, Fb, isA ? rW : O::v(0, 0, 0), isA ? O::v(0, 0, 0) : rW, pAb, qAb, pBb, qBb, ipAb,
                        irAb, ipBb, irBb, lamb[c], c0b[c], mub[c]);
                if (isA) {
                  pb = O::add(pb, pAb); qb = qadd4(qb, qAb); ipb = O::add(ipb, ipAb); irb = qadd4(irb, irAb);
                  addBody(a, j, pBb, qBb, ipBb, irBb);
                } else {
                  pb = O::add(pb, pBb); qb = qadd4(qb, qBb); ipb = O::add(ipb, ipBb); irb = qadd4(irb, irBb);
                  addBody(a, j, pAb, qAb, ipAb, irAb);
                }
              }
            }
            a.P[i] = pb; a.R[i] = qb;
            a.Ip[i] = O::add(a.Ip[i], ipb); a.Ir[i] = qadd4(a.Ir[i], irb);
          }
      }
      // warm start of λ, c0 and friction
      std::vector<VV> lambPrev(s > 0 ? tape.steps[s - 1].contacts.size() : 0, O::v(0, 0, 0));
      for (size_t c = 0; c < M; c++) {
        const TapeContact &t = ts.contacts[c];
        if (t.src >= 0 && s > 0) lambPrev[t.src] = O::add(lambPrev[t.src], O::mul(lamb[c], p.alpha * p.gamma));
        V l[3]; basisOf(t, l);
        V db = O::add(O::add(O::mul(l[0], c0b[c].x), O::mul(l[1], c0b[c].y)), O::mul(l[2], c0b[c].z));
        if (bodies[t.a].dyn) {
          a.Ip[t.a] = O::add(a.Ip[t.a], db);
          qrotateRev(R.q0[t.a], O::v(t.rA[0], t.rA[1], t.rA[2]), db, a.Ir[t.a]);
        }
        if (bodies[t.b].dyn) {
          a.Ip[t.b] = O::sub(a.Ip[t.b], db);
          qrotateRev(R.q0[t.b], O::v(t.rB[0], t.rB[1], t.rB[2]), O::neg(db), a.Ir[t.b]);
        }
        double fm = frictionOf(t);
        if (fm > 0) {
          a.mu[t.a] += mub[c] * bodies[t.b].friction / (2 * fm);
          a.mu[t.b] += mub[c] * bodies[t.a].friction / (2 * fm);
        }
      }
      lamb.swap(lambPrev);
      // warm start of the bodies
      for (size_t i = 0; i < N; i++) {
        if (!bodies[i].dyn) { a.P[i] = O::v(0, 0, 0); a.R[i] = O::q(0, 0, 0, 0); a.V_[i] = a.P[i]; a.W[i] = a.P[i]; a.U[i] = 0; continue; }
        V pb = O::add(a.Ip[i], a.Np[i]);
        V vb = O::add(vStart[i], O::mul(a.Np[i], dt));
        Q qb = a.Ir[i];
        V wb = O::v(0, 0, 0);
        double ub = 0;
        qaddRev(R.q0[i], O::mul(R.w0[i], dt), a.Nr[i], qb, wb);
        if (ts.atRest[i]) { pb = O::add(pb, a.P[i]); qb = qadd4(qb, a.R[i]); }
        else {
          pb = O::add(pb, a.P[i]);
          vb = O::add(vb, O::mul(a.P[i], dt));
          double wfb = g * dt * dt * O::dot(upv, a.P[i]);
          if (ts.wMode[i] == 0) {
            double sg = g > 0 ? 1.0 : -1.0, f = sg / (dt * std::fabs(g));
            vb = O::add(vb, O::mul(upv, wfb * f));
            ub -= wfb * f;
          }
          qaddRev(R.q0[i], O::mul(R.w0[i], dt), a.R[i], qb, wb);
        }
        a.P[i] = pb; a.R[i] = qb; a.V_[i] = vb; a.W[i] = O::mul(wb, dt); a.U[i] = ub;
      }
    }
    for (size_t i = 0; i < N; i++) {
      double *s = dInit + i * ADJ_FLOATS;
      s[0] = a.P[i].x; s[1] = a.P[i].y; s[2] = a.P[i].z; s[3] = a.U[i];
      s[4] = a.R[i].x; s[5] = a.R[i].y; s[6] = a.R[i].z; s[7] = a.R[i].w;
      s[8] = a.V_[i].x; s[9] = a.V_[i].y; s[10] = a.V_[i].z; s[11] = 0;
      s[12] = a.W[i].x; s[13] = a.W[i].y; s[14] = a.W[i].z; s[15] = 0;
      if (dMu) dMu[i] = a.mu[i];
    }
  }
  static Q qadd4(Q a, Q b) { return O::q(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
  void addBody(Adj &a, uint32_t j, V pb, Q qb, V ipb, Q irb) const {
    if (!bodies[j].dyn) return;
    a.P[j] = O::add(a.P[j], pb); a.R[j] = qadd4(a.R[j], qb);
    a.Ip[j] = O::add(a.Ip[j], ipb); a.Ir[j] = qadd4(a.Ir[j], irb);
  }
};

}  // namespace gpuo
