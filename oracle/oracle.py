"""TEST INFRASTRUCTURE ONLY: ctypes binding of the oracle (oracle/liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this module, as the
checker. Two restatements live behind it:

* ``RefSolver`` -- the reference's f64 CPU solver (src/avbd3d/ref/*.ts), pinned by the golden
  trajectories in tests/golden/oracle3d;
* ``Model`` -- the WebGPU step's semantics (src/avbd3d/gpu/wgsl-*.ts), f32 or f64.

``contact_events`` (plain Python, below) restates ContactWatch's two passes (src/lib/contacts.ts)
on read-back records; it is pinned by known answers in tests/test_contact_events_oracle.py.

Record layouts are the reference's (src/avbd3d/gpu/layout.ts): bodies 40 floats, joints 40
floats + info (type, a, b, 0), manifolds 8 words, contacts 16 words. Here they travel as float64
arrays with integer words (keys, ids, stamps) stored as plain numbers.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    """Compile the oracle (g++ -ffp-contract=off) into oracle/liboracle.so."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.oref_create.restype = vp
        L.omodel_create.restype = vp
        L.ofrozen_create.restype = vp
        L.ofrozen_error.restype = ctypes.c_char_p
        L.ofrozen_check_helpers.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Params(ctypes.Structure):
    """Params of the GPU-semantics model (the fields of layout.ts Params it depends on)."""

    _fields_ = [
        ("dt", ctypes.c_double), ("gravity", ctypes.c_double), ("betaLin", ctypes.c_double),
        ("betaAng", ctypes.c_double), ("gamma", ctypes.c_double), ("alpha", ctypes.c_double),
        ("contactMargin", ctypes.c_double), ("up", ctypes.c_double * 3),
        ("flags", ctypes.c_uint32), ("iterations", ctypes.c_uint32), ("colorCap", ctypes.c_uint32),
        ("colorRounds", ctypes.c_uint32), ("step", ctypes.c_uint32), ("lanesRule", ctypes.c_uint32 * 4),
    ]


FLAG_MATCH_NEAREST = 1
FLAG_FACE_BIAS = 2
FLAG_REUSE_CONTACTS = 4
FLAG_START_AT_REST = 8
FLAG_MASS_PENALTY = 16
FLAG_DEFAULTS = FLAG_MATCH_NEAREST | FLAG_FACE_BIAS | FLAG_REUSE_CONTACTS | FLAG_START_AT_REST | FLAG_MASS_PENALTY
NO_COLOR = 255


def make_params(dt=1 / 60, gravity=-10.0, iterations=10, alpha=0.99, beta_lin=1e4, beta_ang=100.0, gamma=0.999,
                flags=FLAG_DEFAULTS, color_cap=64, color_rounds=16, margin=0.01, up=(0, 0, 1), lanes=(0, 0, 0, 0)):
    p = Params()
    p.dt, p.gravity, p.betaLin, p.betaAng, p.gamma, p.alpha, p.contactMargin = dt, gravity, beta_lin, beta_ang, gamma, alpha, margin
    p.up[0], p.up[1], p.up[2] = up
    p.flags, p.iterations, p.colorCap, p.colorRounds, p.step = flags, iterations, color_cap, color_rounds, 0
    p.lanesRule[0], p.lanesRule[1], p.lanesRule[2], p.lanesRule[3] = lanes
    return p


class RefSolver:
    """The reference's f64 CPU solver (src/avbd3d/ref/solver.ts)."""

    def __init__(self, scene=None):
        self.L = lib()
        self.h = ctypes.c_void_p(self.L.oref_create())
        if scene is not None:
            self.L.oref_build_scene(self.h, scene)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.oref_destroy(self.h)
            self.h = None

    def set_params(self, dt=1 / 60, gravity=-10.0, iterations=10, alpha=0.99, beta_lin=1e4, beta_ang=100.0, gamma=0.999):
        self.L.oref_set_params(self.h, ctypes.c_double(dt), ctypes.c_double(gravity), iterations, ctypes.c_double(alpha),
                               ctypes.c_double(beta_lin), ctypes.c_double(beta_ang), ctypes.c_double(gamma))

    def add_body(self, size, density, friction, pos, rot=None, vel=None, angvel=None):
        a = lambda v: None if v is None else _p(np.ascontiguousarray(v, dtype=np.float64))
        keep = [np.ascontiguousarray(v, dtype=np.float64) for v in (size, pos) + tuple(x for x in (rot, vel, angvel) if x is not None)]
        return self.L.oref_add_body(self.h, _p(keep[0]), ctypes.c_double(density), ctypes.c_double(friction), _p(keep[1]),
                                    a(rot), a(vel), a(angvel))

    def add_joint(self, a, b, rA, rB, stiff_lin=float("inf"), stiff_ang=0.0, fracture=float("inf")):
        rA = np.ascontiguousarray(rA, dtype=np.float64)
        rB = np.ascontiguousarray(rB, dtype=np.float64)
        return self.L.oref_add_joint(self.h, a, b, _p(rA), _p(rB), ctypes.c_double(stiff_lin), ctypes.c_double(stiff_ang), ctypes.c_double(fracture))

    def add_ignore(self, a, b):
        return self.L.oref_add_ignore(self.h, a, b)

    def step(self, n=1):
        self.L.oref_step(self.h, n)

    @property
    def body_count(self):
        return self.L.oref_body_count(self.h)

    @property
    def force_count(self):
        return self.L.oref_force_count(self.h)

    def poses(self):
        out = np.zeros((self.body_count, 7))
        self.L.oref_poses(self.h, _p(out))
        return out

    def body_records(self):
        out = np.zeros((self.body_count, 40))
        self.L.oref_body_records(self.h, _p(out))
        return out

    def forces(self):
        out = np.zeros((self.force_count, 4), dtype=np.int32)
        self.L.oref_forces(self.h, _p(out))
        return out

    def joint_records(self):
        n = self.force_count
        rec = np.zeros((max(n, 1), 40))
        info = np.zeros((max(n, 1), 4), dtype=np.int32)
        k = self.L.oref_joint_records(self.h, _p(rec), _p(info))
        return rec[:k], info[:k]

    def manifold_records(self):
        forces = self.forces()
        nm = int((forces[:, 0] == 0).sum()) if len(forces) else 0
        nc = int(forces[forces[:, 0] == 0, 3].sum()) if nm else 0
        man = np.zeros((max(nm, 1), 8))
        con = np.zeros((max(nc, 1), 16))
        counts = np.zeros(2, dtype=np.int32)
        self.L.oref_manifold_records(self.h, _p(man), _p(con), _p(counts))
        return man[:counts[0]], con[:counts[1]]

    def collide(self, a, b):
        basis = np.zeros(9)
        out = np.zeros((8, 7))
        n = self.L.oref_collide(self.h, a, b, _p(basis), _p(out))
        return basis.reshape(3, 3), out[:n]


def solve6_floor(A, rhs, prec=32):
    """finishBody's floored LDLT (the oracle's): x = A^-1 rhs, the pivots D and the pivots before
    their 1e-5 A_kk floor S."""
    A = np.ascontiguousarray(A, np.float64).reshape(6, 6)
    r = np.ascontiguousarray(rhs, np.float64).reshape(6)
    x, D, S = np.zeros(6), np.zeros(6), np.zeros(6)
    lib().omodel_solve6(int(prec), _p(A), _p(r), _p(x), _p(D), _p(S))
    return x, D, S


def solve6(A, b):
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(6)
    lib().oref_solve6(_p(A), _p(b), _p(x))
    return x


class Model:
    """The WebGPU step's semantics (oracle/gpu3d.hpp) in f32 (prec=32) or f64 (prec=64)."""

    def __init__(self, prec=32):
        self.L = lib()
        self.h = ctypes.c_void_p(self.L.omodel_create(prec))
        self.n = 0
        self.nj = 0

    def __del__(self):
        if getattr(self, "h", None):
            self.L.omodel_destroy(self.h)
            self.h = None

    def set_params(self, p):
        self.L.omodel_set_params(self.h, ctypes.byref(p))

    def set_bodies(self, rec):
        rec = np.ascontiguousarray(rec, dtype=np.float64)
        self.n = len(rec)
        self.L.omodel_set_bodies(self.h, _p(rec), self.n)

    def bodies(self):
        out = np.zeros((self.n, 40))
        self.L.omodel_get_bodies(self.h, _p(out))
        return out

    def set_joints(self, rec, info):
        rec = np.ascontiguousarray(rec, dtype=np.float64).reshape(-1, 40)
        info = np.ascontiguousarray(info, dtype=np.int32).reshape(-1, 4)
        self.nj = len(rec)
        self.L.omodel_set_joints(self.h, _p(rec), _p(info), self.nj)

    def joints(self):
        out = np.zeros((self.nj, 40))
        self.L.omodel_get_joints(self.h, _p(out))
        return out

    def add_ignore(self, a, b):
        self.L.omodel_add_ignore(self.h, a, b)

    def set_hulls(self, words):
        words = np.ascontiguousarray(words, dtype=np.uint32)
        self.L.omodel_set_hulls(self.h, _p(words), len(words))

    def set_prev(self, man, con):
        man = np.ascontiguousarray(man, dtype=np.float64).reshape(-1, 8)
        con = np.ascontiguousarray(con, dtype=np.float64).reshape(-1, 16)
        self.L.omodel_set_prev(self.h, _p(man), len(man), _p(con), len(con))

    def set_fixed_colors(self, colors):
        if colors is None:
            self.L.omodel_set_fixed_colors(self.h, None)
        else:
            self._fixed = np.ascontiguousarray(colors, dtype=np.uint32)
            self.L.omodel_set_fixed_colors(self.h, _p(self._fixed))

    def set_color_state(self, state):
        state = np.ascontiguousarray(state, dtype=np.uint32)
        self.L.omodel_set_color_state(self.h, _p(state))

    def set_step_count(self, s):
        self.L.omodel_set_step_count(self.h, s)

    def hull_query(self, a, b):
        """The hull SAT's queries for bodies a, b (faces of a, faces of b, edge pairs): a dict of
        their separations and the features that give them."""
        out = np.zeros(8)
        self.L.omodel_hull_query(self.h, int(a), int(b), _p(out))
        return {"fa_sep": out[0], "fa_face": int(out[1]), "fb_sep": out[2], "fb_face": int(out[3]), "edge_sep": out[4],
                "edge_valid": bool(out[5]), "ea": int(out[6]), "eb": int(out[7])}

    def set_aux(self, aux):
        """setBodyAux's records (n x 4: gravity scale, the linear and angular per-step decay
        factors exp(-rate dt) as f32, 1 when damped)."""
        r = np.ascontiguousarray(aux, np.float64).reshape(self.n, 4)
        self.L.omodel_set_aux(self.h, _p(r))

    def set_refs(self, refs):
        """updateRefs' reference poses (n x 8: pos xyz, 0, rot xyzw; a zero rotation: none yet)."""
        r = np.ascontiguousarray(refs, np.float64).reshape(self.n, 8)
        self.L.omodel_set_refs(self.h, _p(r))

    def set_threads(self, n):
        """Host threads for the step (OpenMP); results equal the one-thread step bit for bit."""
        self.L.omodel_set_threads(self.h, int(n))

    def step(self, n=1, reuse_manifolds=False):
        self.L.omodel_step(self.h, n, int(reuse_manifolds))

    def collide(self, reuse_manifolds=False):
        self.L.omodel_collide(self.h, int(reuse_manifolds))

    def solve(self):
        self.L.omodel_solve(self.h)

    def counts(self):
        out = np.zeros(7, dtype=np.uint32)
        self.L.omodel_counts(self.h, _p(out))
        keys = ("pairs", "manifolds", "contacts", "num_colors", "clashes", "adjacency", "color_cap")
        return dict(zip(keys, (int(x) for x in out)))

    def pairs(self):
        c = self.counts()
        out = np.zeros((max(c["pairs"], 1), 2), dtype=np.uint32)
        self.L.omodel_get_pairs(self.h, _p(out))
        return out[:c["pairs"]]

    def contacts(self):
        c = self.counts()
        man = np.zeros((max(c["manifolds"], 1), 8))
        con = np.zeros((max(c["contacts"], 1), 16))
        self.L.omodel_get_contacts(self.h, _p(man), _p(con))
        return man[:c["manifolds"]], con[:c["contacts"]]

    def colors(self):
        out = np.zeros(self.n, dtype=np.uint32)
        self.L.omodel_get_colors(self.h, _p(out))
        return out

    def record_tape(self, on=True):
        """Record the adjoint's tape (the frozen decisions of every following step)."""
        self.L.omodel_tape(self.h, int(on))

    def adjacency(self):
        c = self.counts()
        start = np.zeros(self.n + 1, dtype=np.uint32)
        lst = np.zeros(max(c["adjacency"], 1), dtype=np.uint32)
        self.L.omodel_get_adjacency(self.h, _p(start), _p(lst))
        return start, lst[:c["adjacency"]]


ADJ_FLOATS = 16  # pos xyz + previous up-velocity, rot xyzw, vel xyz + 0, angvel xyz + 0


def state_from_records(rec):
    """Adjoint state slots (16 per body) from 40-float body records."""
    rec = np.asarray(rec, dtype=np.float64)
    st = np.zeros((len(rec), ADJ_FLOATS))
    st[:, 0:3] = rec[:, 0:3]
    st[:, 3] = rec[:, 35]
    st[:, 4:8] = rec[:, 4:8]
    st[:, 8:11] = rec[:, 32:35]
    st[:, 12:15] = rec[:, 36:39]
    return st


class Frozen:
    """The step with its active set frozen (oracle/adjoint3d.hpp), replayed in f64 over the tape a
    Model recorded, and its reverse sweep. ``records`` are the body records the tape started at."""

    def __init__(self, model, records):
        self.L = lib()
        rec = np.ascontiguousarray(records, dtype=np.float64)
        self.n = len(rec)
        self.records = rec
        h = self.L.ofrozen_create(model.h, _p(rec), self.n)
        if not h:
            raise RuntimeError(self.L.ofrozen_error().decode())
        self.h = ctypes.c_void_p(h)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ofrozen_destroy(self.h)
            self.h = None

    @property
    def steps(self):
        return self.L.ofrozen_steps(self.h)

    def run(self, state=None, friction=None):
        """Final state (n, 16) from the initial state (n, 16) and per-body friction."""
        st = np.ascontiguousarray(state_from_records(self.records) if state is None else state, dtype=np.float64)
        out = np.zeros((self.n, ADJ_FLOATS))
        mu = None if friction is None else np.ascontiguousarray(friction, dtype=np.float64)
        self.L.ofrozen_run(self.h, _p(st), None if mu is None else _p(mu), _p(out))
        return out

    def load_recorded(self, friction=None):
        """Linearise at the recording model's own states (its float type's trajectory)."""
        mu = None if friction is None else np.ascontiguousarray(friction, dtype=np.float64)
        if self.L.ofrozen_load_recorded(self.h, None if mu is None else _p(mu)) != 0:
            raise RuntimeError(self.L.ofrozen_error().decode())

    def backward(self, d_final):
        """Adjoints of the initial state (n, 16) and friction (n,) for the last run()."""
        d = np.ascontiguousarray(d_final, dtype=np.float64).reshape(self.n, ADJ_FLOATS)
        d_init = np.zeros((self.n, ADJ_FLOATS))
        d_mu = np.zeros(self.n)
        self.L.ofrozen_backward(self.h, _p(d), _p(d_init), _p(d_mu))
        return d_init, d_mu

    def stats(self):
        out = np.zeros(5)
        self.L.ofrozen_stats(self.h, _p(out))
        return dict(zip(("steps", "contacts", "floored_pivots", "sliding", "clamped"), out.tolist()))


def check_quaternion_reverses(q, b, w, g):
    """Largest |reverse - central difference| of the quaternion helpers' reverses at a point."""
    arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (q, b, w, g)]
    return lib().ofrozen_check_helpers(*(_p(a) for a in arr))


def sequential_colors(bodies):
    """The reference's Gauss-Seidel order as colours (solver.ts:1365-1372): newest dynamic first."""
    n = len(bodies)
    colors = np.full(n, NO_COLOR, dtype=np.uint32)
    nxt = 0
    for i in range(n - 1, -1, -1):
        if bodies[i, 19] > 0:
            colors[i] = nxt
            nxt += 1
    assert nxt <= 64
    return colors


def seed_model_from_ref(model, ref, params):
    """seedFrom (solver.ts:1320-1359): bodies, joints with warm-start state, last contacts."""
    model.set_params(params)
    model.set_bodies(ref.body_records())
    rec, info = ref.joint_records()
    model.set_joints(rec, info)
    for kind, a, b, _ in ref.forces():
        if kind == 3:
            model.add_ignore(int(a), int(b))
    man, con = ref.manifold_records()
    model.set_prev(man, con)


# --- contact events (ContactWatch, src/lib/contacts.ts) ------------------------------------------
# A restatement of the touching/ended passes with set semantics (exact whenever a set is not full;
# a full set's contents depend on the GPU's insertion order). f32 arithmetic in the passes' order.

def _cross32(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def turn32(q, v):
    """contacts.ts:38-41: v rotated by the unit quaternion q (x, y, z, w), float32 throughout."""
    f = np.float32
    q = [f(x) for x in q]
    v = [f(x) for x in v]
    u = q[:3]
    t = [c * f(2.0) for c in _cross32(u, v)]
    w = [v[i] + t[i] * q[3] for i in range(3)]
    c2 = _cross32(u, t)
    return [w[i] + c2[i] for i in range(3)]


def contact_events(prev, manifolds, contacts, bodies, watch, pairs, step, dt):
    """One ContactWatch.run (contacts.ts:86-120) on a step's read-back records: manifolds (M, 8)
    uint32 words, contacts (K, 16) uint32 words, bodies (N, 40) float32, watch (N,) uint32 words
    (lifetime ID | 0x80000000 when reporting). prev: the last run's set of (idA, idB). Returns this
    run's set and its events (dicts like GpuSolver3D.read_contact_events) in pass order: the
    begins by manifold, then the ends."""
    f = np.float32
    now, events = set(), []
    man = np.asarray(manifolds, np.uint32)
    cw = np.asarray(contacts, np.uint32)
    cf = cw.view(np.float32)
    for m in man[:pairs]:
        a, b, first, w = int(m[0]), int(m[1]), int(m[2]), int(m[3])
        cnt = w & 15
        wa, wb = int(watch[a]), int(watch[b])
        ida, idb = wa & 0x7FFFFFFF, wb & 0x7FFFFFFF
        if cnt == 0 or ((wa | wb) & 0x80000000) == 0 or ida == 0 or idb == 0:
            continue
        now.add((ida, idb))
        if (ida, idb) in prev:
            continue
        pa = [f(x) for x in bodies[a, 0:3]]
        qa = [f(x) for x in bodies[a, 4:8]]
        point, force = [f(0.0)] * 3, f(0.0)
        for c in range(first, first + cnt):
            r = turn32(qa, cf[c, 0:3])
            point = [point[i] + (r[i] + pa[i]) for i in range(3)]
            force = force + cf[c, 12]
        normal = m[4:7].view(np.float32)
        events.append({"kind": 1, "a": ida, "b": idb, "step": step, "point": [float(x / f(cnt)) for x in point],
                       "impulse": float(abs(force) * f(dt)), "normal": [float(x) for x in normal]})
    for ida, idb in sorted(prev - now):
        events.append({"kind": 2, "a": ida, "b": idb, "step": step, "point": [0.0, 0.0, 0.0], "impulse": 0.0,
                       "normal": [0.0, 0.0, 0.0]})
    return now, events


# --- World's command passes (Pusher, BodyMover, BodyReset) on 40-float body records ---------------
# Record offsets (layout.ts): pos 0, rot 4, x- 8 (w: lifecycle flags), q- 12, size 16 (w: mass),
# moment 20, inertial pos 24, inertial rot 28, vel 32, angvel 36 (w: shape).

def _lifecycle(w):
    w = np.float32(w)
    return int(w) if 0.0 <= w <= 7.0 else 0


def push_pass(records, pushes):
    """Pusher's pass (src/lib/push.ts:28-45), float32 in its order: pushes (P, 12) as avbd_push."""
    f = np.float32
    r = np.array(records, np.float32, copy=True)
    pw = np.asarray(pushes, np.float32)
    for row in pw:
        i = int(row[3:4].view(np.uint32)[0])
        if _lifecycle(r[i, 11]) & 2:
            continue
        mass = r[i, 19]
        if mass <= 0:
            continue
        for a in range(3):
            r[i, 32 + a] = r[i, 32 + a] + row[a] / mass
        pos, off, mom = [f(x) for x in r[i, 0:3]], [f(x) for x in row[4:7]], [f(x) for x in row[8:11]]
        c = _cross32(pos, off)
        L = [mom[a] - c[a] for a in range(3)]
        q = [f(x) for x in r[i, 4:8]]
        local = turn32([-q[0], -q[1], -q[2], q[3]], L)
        local = [local[a] / max(f(r[i, 20 + a]), f(1e-12)) for a in range(3)]
        dw = turn32(q, local)
        for a in range(3):
            r[i, 36 + a] = r[i, 36 + a] + dw[a]
    return r


def move_pass(records, moves):
    """BodyMover's pass (src/avbd3d/gpu/body-mover.ts:10-31): moves (M, 12) as avbd_move."""
    f = np.float32
    r = np.array(records, np.float32, copy=True)
    for row in np.asarray(moves, np.float32):
        i = int(row[3:4].view(np.uint32)[0])
        if _lifecycle(r[i, 11]) & 2:
            continue
        left, dt, has_rot = row[8], row[9], row[10]
        time = left * dt
        for a in range(3):
            r[i, 32 + a] = (row[a] - r[i, a]) / time
        omega = [f(0.0)] * 3
        if has_rot != 0:
            tq = [f(x) for x in row[4:8]]
            q = [f(x) for x in r[i, 4:8]]
            a_, b_ = tq, [-q[0], -q[1], -q[2], q[3]]
            d = [a_[3] * b_[0] + a_[0] * b_[3] + a_[1] * b_[2] - a_[2] * b_[1],
                 a_[3] * b_[1] - a_[0] * b_[2] + a_[1] * b_[3] + a_[2] * b_[0],
                 a_[3] * b_[2] + a_[0] * b_[1] - a_[1] * b_[0] + a_[2] * b_[3],
                 a_[3] * b_[3] - a_[0] * b_[0] - a_[1] * b_[1] - a_[2] * b_[2]]
            ln = np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + d[3] * d[3], dtype=np.float32)
            d = [x / ln for x in d]
            if d[3] < 0:
                d = [-x for x in d]
            sine = np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2], dtype=np.float32)
            if sine > f(1e-7):
                half = np.arctan2(sine, max(d[3], f(0.0)), dtype=np.float32) / left
                s = f(2.0) * np.tan(min(half, f(1.5706963)), dtype=np.float32) / dt
                omega = [d[a] / sine * s for a in range(3)]
        r[i, 36:39] = omega
    return r


def reset_pass(records, indices):
    """BodyReset's pass (src/avbd3d/gpu/body-reset.ts:9-17)."""
    r = np.array(records, np.float32, copy=True)
    for i in np.asarray(indices, np.uint32):
        r[i, 8:11] = r[i, 0:3]
        r[i, 24:27] = r[i, 0:3]
        r[i, 12:16] = r[i, 4:8]
        r[i, 28:32] = r[i, 4:8]
        r[i, 32:36] = 0.0
        r[i, 36:39] = 0.0
    return r


# ---- SDF collision backend (oracle/sdf3d.hpp) -------------------------------------------------
class OField(ctypes.Structure):
    """A field as the oracle takes it (kind, convex, bounds, centre of mass, the kind's 8
    parameters, voxel grid, seeds, mesh supports)."""
    _fields_ = [("kind", ctypes.c_uint32), ("convex", ctypes.c_uint32), ("lo", ctypes.c_float * 3), ("hi", ctypes.c_float * 3),
                ("com", ctypes.c_float * 3), ("prm", ctypes.c_float * 8), ("dims", ctypes.c_uint32 * 3),
                ("origin", ctypes.c_float * 3), ("cell", ctypes.c_float), ("grid", ctypes.c_void_p), ("seeds", ctypes.c_void_p),
                ("seedCount", ctypes.c_uint32), ("supports", ctypes.c_void_p), ("supportCount", ctypes.c_uint32)]


class OPrm(ctypes.Structure):
    """CollisionParams (ref/sdf/manifold.ts:13-28)."""
    _fields_ = [("optimizer", ctypes.c_uint32), ("cacheSize", ctypes.c_uint32), ("randomSamples", ctypes.c_uint32),
                ("warmStart", ctypes.c_uint32), ("curvature", ctypes.c_uint32), ("skipConvex", ctypes.c_uint32),
                ("epsilon", ctypes.c_float), ("alpha", ctypes.c_float), ("replacement", ctypes.c_float), ("tol", ctypes.c_float),
                ("maxIt", ctypes.c_uint32), ("stepSize", ctypes.c_float), ("margin", ctypes.c_float),
                ("allowedPen", ctypes.c_float), ("signedSep", ctypes.c_uint32), ("meshVertex", ctypes.c_uint32),
                ("refreshDeeper", ctypes.c_uint32), ("sgdStopOnly", ctypes.c_uint32)]


_OPRM = {"optimizer": "optimizer", "cacheSize": "cacheSize", "randomSamples": "randomSamples", "warmStart": "warmStart",
         "curvatureSampling": "curvature", "skipConvexSamples": "skipConvex", "epsilon": "epsilon", "alpha": "alpha",
         "replacementThreshold": "replacement", "tolerance": "tol", "maxIterations": "maxIt", "stepSize": "stepSize",
         "margin": "margin", "allowedPenetration": "allowedPen", "signedSeparation": "signedSep",
         "meshVertexContacts": "meshVertex", "refreshDeeperContacts": "refreshDeeper", "sgdGradientStopOnly": "sgdStopOnly"}


def sdf_params(p):
    """OPrm from a dict with the reference's CollisionParams names (missing ones: the defaults)."""
    f = np.float32
    d = {"optimizer": 0, "cacheSize": 4, "randomSamples": 1, "warmStart": True, "curvatureSampling": True,
         "skipConvexSamples": True, "epsilon": f(0.1), "alpha": f(0.01), "replacementThreshold": f(0.01),
         "tolerance": f(0.01), "maxIterations": 50, "stepSize": f(0.1), "margin": f(0.02), "allowedPenetration": f(0.01),
         "signedSeparation": False, "meshVertexContacts": False, "refreshDeeperContacts": False, "sgdGradientStopOnly": False}
    d.update(p or {})
    o = OPrm()
    for k, name in _OPRM.items():
        cur = getattr(o, name)
        setattr(o, name, int(d[k]) if isinstance(cur, int) else float(d[k]))
    return o


def sdf_field(d, keep):
    """OField from a field description dict: kind, convex, lo, hi, com, params (8), and for a voxel
    field dims, origin, cell, grid; seeds and supports (n x 3). `keep` holds the arrays."""
    o = OField()
    o.kind, o.convex = int(d["kind"]), int(d.get("convex", 0))
    z3 = np.zeros(3, np.float32)
    for i in range(3):
        o.lo[i] = d.get("lo", z3)[i]; o.hi[i] = d.get("hi", z3)[i]; o.com[i] = d.get("com", z3)[i]
        o.dims[i] = d.get("dims", (0, 0, 0))[i]; o.origin[i] = d.get("origin", z3)[i]
    prm = np.zeros(8, np.float32)
    prm[:len(d.get("params", []))] = d.get("params", [])
    for i in range(8):
        o.prm[i] = prm[i]
    o.cell = d.get("cell", 0.0)
    if d.get("grid") is not None:
        g = np.ascontiguousarray(d["grid"], np.float32)
        keep.append(g)
        o.grid = g.ctypes.data
    for name, cname, count in (("seeds", "seeds", "seedCount"), ("supports", "supports", "supportCount")):
        a = np.ascontiguousarray(d.get(name, np.zeros((0, 3))), np.float32).reshape(-1, 3)
        keep.append(a)
        setattr(o, cname, a.ctypes.data if len(a) else None)
        setattr(o, count, len(a))
    return o


def sdf_sample(field, points):
    """phi and gradient (n x 4) of a field at field-local points."""
    keep = []
    f = sdf_field(field, keep)
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    out = np.zeros((len(p), 4), np.float32)
    lib().osdf_sample(ctypes.byref(f), _p(p), len(p), _p(out))
    return out


def sdf_trig(which, x):
    """The oracle's own sin (0), cos (1) or atan2 (2, x: n x 2 of (y, x)) in float32."""
    x = np.ascontiguousarray(x, np.float32)
    n = len(x) if which < 2 else len(x) // 2 if x.ndim == 1 else len(x)
    out = np.zeros(n, np.float32)
    lib().osdf_trig(int(which), _p(x), n, _p(out))
    return out


def _pose(body):
    """pos xyz, rot xyzw, scale, static, vel xyz, omega xyz."""
    b = body
    return np.array(list(b.get("position", (0, 0, 0))) + list(b.get("orientation", (0, 0, 0, 1))) + [b.get("scale", 1.0),
                    1.0 if b.get("static", False) else 0.0] + list(b.get("velocity", (0, 0, 0))) +
                    list(b.get("angularVelocity", (0, 0, 0))), np.float32)


def sdf_optimize(kind, fa, body_a, fb, body_b, domain, repulsion, seed, params, step=None, E=None, pcg=None):
    """One optimiser run of the CPU Algorithm 1: kind 'nelder-mead', 'ellipsoid', 'descent' or
    'gradient' (the objective and its gradient at the seed). Returns a dict."""
    keep = []
    A, B = sdf_field(fa, keep), sdf_field(fb, keep)
    pa, pb = _pose(body_a), _pose(body_b)
    dom = np.ascontiguousarray(list(domain[0]) + list(domain[1]), np.float32)
    rep = np.ascontiguousarray(repulsion, np.float32).reshape(-1, 3)
    sd = np.ascontiguousarray(seed, np.float32)
    st = np.ascontiguousarray(step if step is not None else (0, 0, 0), np.float32)
    Em = np.ascontiguousarray(E if E is not None else np.eye(3), np.float32).reshape(9)
    state = ctypes.c_uint64(0x853c49e6748fea9b if pcg is None else pcg)
    out = np.zeros(4, np.float32)
    counts = np.zeros(2, np.uint32)
    k = {"nelder-mead": 0, "ellipsoid": 1, "descent": 2, "gradient": 3}[kind]
    lib().osdf_optimize(k, ctypes.byref(A), _p(pa), ctypes.byref(B), _p(pb), _p(dom), _p(rep), len(rep), _p(sd), _p(st),
                        ctypes.byref(sdf_params(params)), _p(Em), ctypes.byref(state), _p(out), _p(counts))
    if k == 3:
        return {"value": float(out[0]), "gradient": out[1:4].copy(), "queries": int(counts[1])}
    return {"x": out[:3].copy(), "objective": float(out[3]), "iterations": int(counts[0]), "queries": int(counts[1]),
            "E": Em.reshape(3, 3).copy(), "pcg": state.value}


POINT_WORDS = 28


def sdf_points_to_words(points):
    """Points as the oracle's 28-word records (x, depth, normal, separation, anchors, softmax, rank,
    active, cached, id, exported, ellipsoid)."""
    out = np.zeros((len(points), POINT_WORDS), np.float32)
    u = out.view(np.uint32)
    for i, p in enumerate(points):
        out[i, 0:3] = p["x"]; out[i, 3] = p.get("depth", 0)
        out[i, 4:7] = p.get("normal", (0, 1, 0)); out[i, 7] = p.get("separation", 0)
        out[i, 11] = p.get("softmax", 0); out[i, 15] = p.get("rank", 0)
        u[i, 16] = int(bool(p.get("active", False))); u[i, 17] = int(bool(p.get("cached", False)))
        u[i, 18] = int(p.get("id", 0))
        e = p.get("ellipsoid6", (1, 0, 0, 1, 0, 1))
        out[i, 20:24] = e[0:4]; out[i, 24:26] = e[4:6]
    return out


def sdf_words_to_points(w):
    u = w.view(np.uint32)
    return [{"x": w[i, 0:3].copy(), "depth": float(w[i, 3]), "normal": w[i, 4:7].copy(), "separation": float(w[i, 7]),
             "softmax": float(w[i, 11]), "rank": float(w[i, 15]), "active": bool(u[i, 16]), "cached": bool(u[i, 17]),
             "id": int(u[i, 18]), "ellipsoid6": np.concatenate([w[i, 20:24], w[i, 24:26]])} for i in range(len(w))]


def sdf_manifold_cpu(fa, body_a, fb, body_b, params, dt, cache, pcg):
    """The CPU reference's manifold update (manifold.ts update) with its PCG stream: returns
    (points, stats dict, pcg state after)."""
    keep = []
    A, B = sdf_field(fa, keep), sdf_field(fb, keep)
    pa, pb = _pose(body_a), _pose(body_b)
    cw = sdf_points_to_words(cache)
    out = np.zeros((20, POINT_WORDS), np.float32)
    stats = np.zeros(4, np.uint32)
    state = ctypes.c_uint64(pcg)
    n = lib().osdf_manifold_cpu(ctypes.byref(A), _p(pa), ctypes.byref(B), _p(pb), ctypes.byref(sdf_params(params)),
                                ctypes.c_float(dt), _p(cw), len(cw), ctypes.byref(state), _p(out), _p(stats))
    return sdf_words_to_points(out[:n]), dict(zip(("pairs", "contacts", "queries", "iterations"), (int(x) for x in stats))), state.value


class SdfBridge:
    """The SDF bridge installed on an f32 Model (setCollisionBackend): fields (description dicts),
    per body its field index, scale and field rotation; explicit pairs or automatic."""

    def __init__(self, model, fields, body_field, body_scale=None, body_rotation=None, params=None, automatic=False,
                 max_pairs=1, max_points=5, cell_size=1.0):
        self.L, self.model, self.points = lib(), model, max_points
        self._keep = []
        arr = (OField * len(fields))(*[sdf_field(f, self._keep) for f in fields])
        bf = np.ascontiguousarray(body_field, np.uint32)
        sc = None if body_scale is None else np.ascontiguousarray(body_scale, np.float32)
        rot = None if body_rotation is None else np.ascontiguousarray(body_rotation, np.float32).reshape(-1)
        self._keep += [bf, sc, rot]
        self.L.osdf_create.restype = ctypes.c_void_p
        self.h = ctypes.c_void_p(self.L.osdf_create(model.h, arr, len(fields), _p(bf), None if sc is None else _p(sc),
                                                    None if rot is None else _p(rot), ctypes.byref(sdf_params(params)),
                                                    int(automatic), int(max_pairs), int(max_points), ctypes.c_float(cell_size)))

    def set_pairs(self, pairs):
        w = np.ascontiguousarray([[p["a"], p["b"], p["key"]] for p in pairs], np.uint32).reshape(-1, 3)
        self.L.osdf_set_pairs(self.h, _p(w), len(w))

    def set_params(self, params):
        self.L.osdf_set_params(self.h, ctypes.byref(sdf_params(params)))

    def results_raw(self, cap=1 << 16):
        stride = 8 + POINT_WORDS * self.points
        out = np.zeros(cap * stride, np.uint32)
        ov = ctypes.c_uint32()
        self.L.osdf_results.restype = ctypes.c_uint32
        n = self.L.osdf_results(self.h, _p(out), cap, ctypes.byref(ov))
        return out[:n * stride].reshape(n, stride), ov.value

    def bodies(self):
        out = np.zeros((self.model.n, 28), np.float32)
        self.L.osdf_bodies(self.h, _p(out))
        return out

    def destroy(self):
        if self.h:
            self.L.osdf_destroy(self.h)
            self.h = None
