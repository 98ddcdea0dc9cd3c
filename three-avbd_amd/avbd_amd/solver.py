"""Host-side mirror of the reference's ``GpuSolver3D`` (src/avbd3d/gpu/solver.ts:504) over the
C ABI of libavbd_mi355x.so (include/avbd.h).

Same names, argument meaning and error behaviour as the reference's operational surface:
``step``, ``params``, ``read_bodies``/``read_joints``/``read_counters``/``read_contact_list``/
``read_pairs``, ``seed_from``, ``fixed_colors``/``sequential_colors``, ``profile_next_step``,
``destroy``. Reads return numpy arrays in the reference's record formats (layout.ts). Invalid
input raises ``ValueError`` (the reference's ``Error``), capacity refusals raise
``StorageCapacityError``; overflow during a step is counted in ``read_counters()['overflow']``,
not raised. The library must be built (``python -c "import __graft_entry__ as g; g.build()"``):
there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libavbd_mi355x.so")

FLAG_MATCH_NEAREST = 1
FLAG_FACE_BIAS = 2
FLAG_REUSE_CONTACTS = 4
FLAG_START_AT_REST = 8
FLAG_MASS_PENALTY = 16
FLAG_DEFAULTS = 31
FLAG_ADAPT_COLORS = 128
FLAG_FIXED_STORAGE = 256  # capacities stay as created (overflow counted, never grown)
OVERFLOW_BARRIER = 256  # counters' overflow bit: a cooperative solve's grid barrier gave up waiting
NO_COLOR = 255
STICK_BIT = 0x80000000
SDF_ROLLING_BIT = 0x20000000  # contact key bit of a rolling SDF surface (layout.ts:134)
BODY_FLOATS = 40
JOINT_FLOATS = 40

# record offsets (layout.ts:49-58, :84-110)
B_POS, B_ROT, B_INITIAL_POS, B_INITIAL_ROT, B_SIZE, B_MOMENT, B_INERTIAL_POS, B_INERTIAL_ROT, B_VEL, B_ANGVEL = range(0, 40, 4)
J_PEN_LIN, J_PEN_ANG, J_LAM_LIN, J_LAM_ANG, J_C0_LIN, J_C0_ANG, J_RA, J_RB, J_REST, J_REST_START = range(0, 40, 4)
SHAPE_BOX, SHAPE_SPHERE, SHAPE_SAIL, SHAPE_HULL = 0, 1, 2, 3
T_NONE, T_JOINT, T_SPRING = 0, 1, 2
BIG = 3.0e38


class StorageCapacityError(RuntimeError):
    """A capacity refused before anything changed (storage-budget.ts StorageCapacityError)."""


class Params(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_float), ("gravity", ctypes.c_float), ("beta_lin", ctypes.c_float), ("beta_ang", ctypes.c_float),
        ("gamma", ctypes.c_float), ("alpha", ctypes.c_float), ("contact_margin", ctypes.c_float), ("up", ctypes.c_float * 3),
        ("flags", ctypes.c_uint32), ("iterations", ctypes.c_uint32), ("color_cap", ctypes.c_uint32),
        ("color_rounds", ctypes.c_uint32), ("primal_lanes", ctypes.c_uint32 * 4),
    ]


class Capacity(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint32) for k in ("bodies", "joints", "pairs", "manifolds", "contacts", "hull_words")]


class Counters(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint32) for k in ("pairs", "contacts", "manifolds", "overflow", "clashes", "colors", "color_cap", "step")] + [
        ("clashing_colors", ctypes.c_uint32 * 2)]  # the last step's serial colours (bit c of word c // 32)


class Profile(ctypes.Structure):
    _fields_ = [(k, ctypes.c_float) for k in ("collision", "adjacency", "coloring", "solve", "total")]


class KernelStats(ctypes.Structure):
    _fields_ = [
        ("primal_ms", ctypes.c_float), ("dual_ms", ctypes.c_float), ("narrowphase_ms", ctypes.c_float),
        ("broadphase_ms", ctypes.c_float), ("solve_ms", ctypes.c_float), ("other_ms", ctypes.c_float),
        ("primal_launches", ctypes.c_uint32), ("dual_launches", ctypes.c_uint32),
        ("narrowphase_launches", ctypes.c_uint32), ("broadphase_launches", ctypes.c_uint32),
        ("solve_launches", ctypes.c_uint32), ("solve_blocks", ctypes.c_uint32),
        ("tail_ms", ctypes.c_float), ("tail_launches", ctypes.c_uint32),
    ]


class Schedule(ctypes.Structure):
    _fields_ = [("cooperative", ctypes.c_uint32), ("threads", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32)]


class StepOpts(ctypes.Structure):
    _fields_ = [("reuse_manifolds", ctypes.c_int)]


_lib = None

EXPORTS = [
    "avbd_create", "avbd_destroy", "avbd_last_error", "avbd_create_error", "avbd_step", "avbd_steps", "avbd_set_params",
    "avbd_get_params", "avbd_read_bodies", "avbd_write_bodies", "avbd_read_joints", "avbd_read_counters",
    "avbd_read_contacts", "avbd_read_pairs", "avbd_read_colors", "avbd_seed", "avbd_set_fixed_colors",
    "avbd_set_color_state", "avbd_add_ignore_pairs", "avbd_set_filters", "avbd_set_hulls", "avbd_profile_step",
    "avbd_profile_kernels", "avbd_time_steps", "avbd_sync", "avbd_adapt", "avbd_set_schedule", "avbd_get_schedule",
    "avbd_autotune", "avbd_create_batch", "avbd_forward", "avbd_backward", "avbd_set_sparse_tail",
    "avbd_watch_create", "avbd_watch_set", "avbd_watch_run", "avbd_watch_read", "avbd_push", "avbd_move", "avbd_reset",
    "avbd_get_capacity", "avbd_broadphase_info", "avbd_set_barrier_polls", "avbd_read_refs", "avbd_read_bodies_noalloc",
    "avbd_read_joint_info", "avbd_get_sizes", "avbd_add_bodies", "avbd_patch_body", "avbd_rewrite_fixed",
    "avbd_append_joints", "avbd_append_joint", "avbd_append_world_joint", "avbd_append_springs", "avbd_release_joints",
    "avbd_set_world_anchor", "avbd_disable_constraint", "avbd_set_body_aux", "avbd_set_lifecycle",
    "avbd_morton_order", "avbd_estimate_pairs", "avbd_plan_capacity", "avbd_set_graphs", "avbd_graph_stats",
    "avbd_sdf_default_params", "avbd_sdf_create", "avbd_sdf_set_pairs", "avbd_sdf_set_params", "avbd_set_collision_backend",
    "avbd_sdf_read_results", "avbd_sdf_read_stats", "avbd_sdf_grow", "avbd_sdf_destroy", "avbd_sdf_last_error",
]
FIELD_POSITION, FIELD_ROTATION, FIELD_VELOCITY, FIELD_ANGULAR_VELOCITY = 1, 2, 4, 8
_FIELDS = {"position": 1, "rotation": 2, "velocity": 4, "angularVelocity": 8}


class JointSpec(ctypes.Structure):
    """JointSpec (solver.ts:60-124); NaN marks a field left out."""
    _fields_ = [("a", ctypes.c_int32), ("b", ctypes.c_int32), ("rA", ctypes.c_float * 3), ("rB", ctypes.c_float * 3),
                ("angular", ctypes.c_float), ("fracture", ctypes.c_float), ("linear", ctypes.c_int32),
                ("rest", ctypes.c_float * 4), ("yield_", ctypes.c_float), ("break_bend", ctypes.c_float)]


class SpringSpec(ctypes.Structure):
    _fields_ = [("a", ctypes.c_int32), ("b", ctypes.c_int32), ("rA", ctypes.c_float * 3), ("rB", ctypes.c_float * 3),
                ("stiffness", ctypes.c_float), ("rest", ctypes.c_float)]
READ_NOALLOC_MAX = 16384  # avbd_read_bodies_noalloc: bodies per call
EVENT_BEGIN, EVENT_END = 1, 2  # contacts.ts BEGIN / END


class ContactEvent(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("a", ctypes.c_uint32), ("b", ctypes.c_uint32), ("step", ctypes.c_uint32),
                ("point", ctypes.c_float * 3), ("impulse", ctypes.c_float), ("normal", ctypes.c_float * 3), ("pad", ctypes.c_float)]
TAIL_AUTO = 0xFFFFFFFF  # avbd_set_sparse_tail: chosen at each adaptive-colouring readback
ADJ_FLOATS = 16  # adjoint state slots per body: pos xyz, previous up-velocity, rot xyzw, vel xyz, 0, angvel xyz, 0


def lib():
    """The native library; raises if it is not built (the path has no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        L.avbd_last_error.restype = ctypes.c_char_p
        L.avbd_create_error.restype = ctypes.c_char_p
        vp = ctypes.c_void_p
        L.avbd_create.argtypes = [vp, vp, ctypes.c_uint32, vp, vp, ctypes.c_uint32, vp, ctypes.POINTER(vp)]
        L.avbd_create_batch.argtypes = [vp, vp, ctypes.c_uint32, vp, vp, ctypes.c_uint32, vp, ctypes.c_uint32, vp, ctypes.POINTER(vp)]
        for name in EXPORTS[1:]:
            fn = getattr(L, name)
            if name not in ("avbd_last_error", "avbd_create_error"):
                fn.restype = ctypes.c_int
        L.avbd_destroy.restype = None
        L.avbd_sdf_destroy.restype = None
        L.avbd_sdf_default_params.restype = None
        L.avbd_sdf_last_error.restype = ctypes.c_char_p
        L.avbd_sdf_last_error.argtypes = [vp]
        L.avbd_sdf_create.argtypes = [vp, vp, ctypes.c_uint32, vp, vp, vp, vp, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_float, ctypes.POINTER(vp)]
        for name in ("avbd_sdf_set_pairs", "avbd_sdf_set_params", "avbd_sdf_read_results", "avbd_sdf_read_stats",
                     "avbd_sdf_grow", "avbd_sdf_destroy", "avbd_set_collision_backend"):
            getattr(L, name).argtypes = None
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def default_params(dt=1 / 60, gravity=-9.81, iterations=10, alpha=0.99, beta_lin=1e4, beta_ang=100.0, gamma=0.999,
                   flags=FLAG_DEFAULTS, color_cap=64, color_rounds=16, margin=0.01, up=(0.0, 0.0, 1.0),
                   primal_lanes=(2 ** 31, 2 ** 15, 2 ** 12, 2 ** 11)):
    """gpuParams3D (solver.ts:213-226) with z up, as a solver built from the reference's scenes."""
    p = Params()
    p.dt, p.gravity, p.beta_lin, p.beta_ang, p.gamma, p.alpha, p.contact_margin = dt, gravity, beta_lin, beta_ang, gamma, alpha, margin
    p.up[0], p.up[1], p.up[2] = up
    p.flags, p.iterations, p.color_cap, p.color_rounds = flags, iterations, color_cap, color_rounds
    p.primal_lanes[0], p.primal_lanes[1], p.primal_lanes[2], p.primal_lanes[3] = primal_lanes
    return p


# --- scene planning before a solver exists (host code in the library) -------------------------
def morton_order(bodies):
    """mortonOrder (solver.ts:463-485): order[g] = the index of the body placed g-th."""
    rec = np.ascontiguousarray(bodies, np.float32).reshape(-1, BODY_FLOATS)
    order = np.zeros(len(rec), np.uint32)
    if lib().avbd_morton_order(_p(rec), len(rec), _p(order)) != 0:
        raise ValueError("avbd_morton_order refused its arguments")
    return order


def estimate_pairs(bodies):
    """estimatePairs (solver.ts:350-460): (pairs, touching, degree per body)."""
    rec = np.ascontiguousarray(bodies, np.float32).reshape(-1, BODY_FLOATS)
    pairs, touching = ctypes.c_uint64(), ctypes.c_uint64()
    degree = np.zeros(len(rec), np.int32)
    if lib().avbd_estimate_pairs(_p(rec), len(rec), ctypes.byref(pairs), ctypes.byref(touching), _p(degree)) != 0:
        raise ValueError("avbd_estimate_pairs refused its arguments")
    return pairs.value, touching.value, degree


def plan_capacity(bodies, joint_info=None, least=None, colors=0):
    """The constructor's sizing (solver.ts:745, :815-843): ({bodies, joints, pairs, manifolds,
    contacts, hull_words}, starting colour cap); `least` as GpuSolverOptions.capacity."""
    rec = np.ascontiguousarray(bodies, np.float32).reshape(-1, BODY_FLOATS)
    info = np.zeros((0, 4), np.int32) if joint_info is None else np.ascontiguousarray(joint_info, np.int32).reshape(-1, 4)
    lo = Capacity()
    for k, v in (least or {}).items():
        setattr(lo, k, int(v))
    out, cap = Capacity(), ctypes.c_uint32()
    if lib().avbd_plan_capacity(_p(rec), len(rec), _p(info), len(info), ctypes.byref(lo), int(colors), ctypes.byref(out),
                                ctypes.byref(cap)) != 0:
        raise ValueError("avbd_plan_capacity refused its arguments")
    return {k: getattr(out, k) for k, _ in Capacity._fields_}, cap.value


class GpuSolver3D:
    """The MI355X AVBD step behind the C ABI, with GpuSolver3D's method names."""

    def __init__(self, bodies, joints=None, joint_info=None, params=None, capacity=None, hulls=None, ignore_pairs=None,
                 world_start=None):
        self.L = lib()
        bodies = np.ascontiguousarray(bodies, dtype=np.float32).reshape(-1, BODY_FLOATS)
        if joints is None or len(joints) == 0:
            joints = np.zeros((0, JOINT_FLOATS), np.float32)
            joint_info = np.zeros((0, 4), np.int32)
        joints = np.ascontiguousarray(joints, dtype=np.float32).reshape(-1, JOINT_FLOATS)
        joint_info = np.ascontiguousarray(joint_info, dtype=np.int32).reshape(-1, 4)
        self._params = params if params is not None else default_params()
        cap = Capacity()
        for k, v in (capacity or {}).items():
            setattr(cap, k, int(v))
        if hulls is not None and cap.hull_words < len(hulls):
            cap.hull_words = len(hulls)
        h = ctypes.c_void_p()
        if world_start is None:
            world_start = np.array([0, len(bodies)], np.uint32)
        self.world_start = np.ascontiguousarray(world_start, np.uint32)
        rc = self.L.avbd_create_batch(ctypes.byref(self._params), _p(bodies), len(bodies), _p(joints), _p(joint_info),
                                      len(joints), _p(self.world_start), len(self.world_start) - 1, ctypes.byref(cap),
                                      ctypes.byref(h))
        if rc != 0:
            msg = self.L.avbd_create_error().decode()
            raise (ValueError if rc == -1 else RuntimeError)(msg)
        self.h = h
        self.body_count = len(bodies)
        self.joint_count = len(joints)
        self.fixed_colors = None
        self.destroyed = False
        self.step_count = 0
        self._joint_info = joint_info
        if hulls is not None and len(hulls):
            self._check(self.L.avbd_set_hulls(self.h, _p(np.ascontiguousarray(hulls, np.uint32)), len(hulls)))
        if ignore_pairs is not None and len(ignore_pairs):
            pairs = np.ascontiguousarray(ignore_pairs, np.uint32).reshape(-1, 2)
            self._check(self.L.avbd_add_ignore_pairs(self.h, _p(pairs), len(pairs)))

    # --- lifecycle ---------------------------------------------------------------------------
    def _check(self, rc):
        if rc >= 0:
            return rc
        msg = self.L.avbd_last_error(self.h).decode()
        if rc == -2:
            raise StorageCapacityError(msg)
        if rc == -1:
            raise ValueError(msg)
        raise RuntimeError(msg)

    def destroy(self):
        if not self.destroyed and getattr(self, "h", None):
            self.L.avbd_destroy(self.h)  # frees every SDF bridge made for this solver too
            for b in getattr(self, "_bridges", []):
                b.disposed, b.installed, b.h = True, False, None
        self.destroyed = True
        self.h = None

    def set_collision_backend(self, backend):
        """setCollisionBackend (solver.ts:2078-2091): install an SdfCollisionBridge (None: the
        built-in collision). The solver owns it from then on: a replaced bridge is destroyed."""
        self._live()
        if backend is not None and (backend.disposed or backend.solver is not self):
            raise ValueError("Invalid SDF bridge owner/lifecycle")
        old = getattr(self, "_backend", None)
        if backend is old:
            return
        self._check(self.L.avbd_set_collision_backend(self.h, None if backend is None else backend.h))
        if old is not None:
            old.disposed, old.installed, old.h = True, False, None
        self._backend = backend
        if backend is not None:
            backend.installed = True

    @property
    def active_collision_backend(self):
        return getattr(self, "_backend", None)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def _live(self):
        if self.destroyed:
            raise RuntimeError("Solver is destroyed")

    # --- params ------------------------------------------------------------------------------
    @property
    def params(self):
        """The params the next step runs with, from the library: a copy, so edits go back through
        set_params (the adaptive colouring and the autotuner change the colour cap, the rounds
        and the lane rule inside the library)."""
        if self.destroyed or not getattr(self, "h", None):
            return self._params
        p = Params()
        self._check(self.L.avbd_get_params(self.h, ctypes.byref(p)))
        self._params = p
        return Params.from_buffer_copy(p)

    def set_params(self, params):
        self._live()
        self._check(self.L.avbd_set_params(self.h, ctypes.byref(params)))
        self._params = params

    # --- stepping ----------------------------------------------------------------------------
    def step(self, reuse_manifolds=False):
        """GpuSolver3D.step (solver.ts:2112): one substep, enqueued on the solver's stream."""
        self._live()
        self._check(self.L.avbd_step(self.h, ctypes.byref(StepOpts(int(reuse_manifolds)))))
        self.step_count += 1

    def steps(self, steps, substeps=1):
        self._live()
        self._check(self.L.avbd_steps(self.h, steps, substeps))
        self.step_count += steps * substeps

    def sync(self):
        self._check(self.L.avbd_sync(self.h))

    def time_steps(self, steps, substeps=1):
        ms = ctypes.c_float()
        self._check(self.L.avbd_time_steps(self.h, steps, substeps, ctypes.byref(ms)))
        self.step_count += steps * substeps
        return ms.value

    def profile_next_step(self, reuse_manifolds=False):
        out = Profile()
        self._check(self.L.avbd_profile_step(self.h, ctypes.byref(StepOpts(int(reuse_manifolds))), ctypes.byref(out)))
        self.step_count += 1
        return {k: getattr(out, k) for k, _ in Profile._fields_}

    def profile_kernels(self, steps, substeps=1):
        out = KernelStats()
        self._check(self.L.avbd_profile_kernels(self.h, steps, substeps, ctypes.byref(out)))
        self.step_count += steps * substeps
        return {k: getattr(out, k) for k, _ in KernelStats._fields_}

    # --- adjoint ------------------------------------------------------------------------------
    def forward(self, steps, checkpoint_every=10):
        """Differentiable rollout of `steps` detecting steps, a checkpoint every `checkpoint_every`
        (0: record every step's tape now, no re-run in backward)."""
        self._live()
        self._check(self.L.avbd_forward(self.h, steps, checkpoint_every))
        self.step_count += steps

    def backward(self, d_final):
        """Reverse sweep of the last forward(): the adjoint of the final state ((n, 16) float64,
        ADJ_FLOATS slots per body) to the adjoints of the initial state ((n, 16)) and of each
        body's friction ((n,))."""
        self._live()
        d = np.ascontiguousarray(d_final, dtype=np.float64).reshape(self.body_count, ADJ_FLOATS)
        d_init = np.zeros((self.body_count, ADJ_FLOATS))
        d_mu = np.zeros(self.body_count)
        self._check(self.L.avbd_backward(self.h, _p(d), _p(d_init), _p(d_mu)))
        return d_init, d_mu

    def capacity(self):
        """contactCapacity / pairCapacity / manifoldCapacity and the fixed capacities."""
        self._live()
        c = Capacity()
        self._check(self.L.avbd_get_capacity(self.h, ctypes.byref(c)))
        return {k: getattr(c, k) for k, _ in Capacity._fields_}

    def broadphase_info(self):
        """The grid uploadStatics chose: cell sizes of both levels, the fine radius threshold,
        whether coarse cells hold movers, and the large bodies tested against every body."""
        self._live()
        f = [ctypes.c_float() for _ in range(3)]
        u = [ctypes.c_uint32() for _ in range(2)]
        self._check(self.L.avbd_broadphase_info(self.h, *[ctypes.byref(x) for x in f + u]))
        return {"cell_size": f[0].value, "fine_cell_size": f[1].value, "fine_radius": f[2].value,
                "coarse_movers": bool(u[0].value), "large": u[1].value}

    def set_graphs(self, on=True):
        """Per-substep hipGraphs: each substep's launch sequence captured once and replayed
        (results equal the plain launches bit for bit)."""
        self._live()
        self._check(self.L.avbd_set_graphs(self.h, 1 if on else 0))

    def graph_stats(self):
        """(substeps captured, substeps replayed) since creation."""
        c, r = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self.L.avbd_graph_stats(self.h, ctypes.byref(c), ctypes.byref(r)))
        return c.value, r.value

    def adapt(self, counters=None):
        """adapt (solver.ts:2362-2407) from a counters readback: the colour cap and rounds, and
        (without FLAG_FIXED_STORAGE) pair/manifold/contact growth before they overflow."""
        c = Counters()
        for k, v in (counters or self.read_counters()).items():
            if k == "clashing_colors":
                c.clashing_colors[0], c.clashing_colors[1] = (int(x) for x in v)
            else:
                setattr(c, k, int(v))
        self._check(self.L.avbd_adapt(self.h, ctypes.byref(c)))
        p = Params()
        self._check(self.L.avbd_get_params(self.h, ctypes.byref(p)))
        self._params = p

    @property
    def schedule(self):
        s = Schedule()
        self._check(self.L.avbd_get_schedule(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in Schedule._fields_}

    def set_schedule(self, cooperative=1, threads=512, blocks_per_cu=1):
        self._check(self.L.avbd_set_schedule(self.h, ctypes.byref(Schedule(cooperative, threads, blocks_per_cu))))

    def set_barrier_polls(self, polls=0):
        """The cooperative solve's grid-barrier wait limit in polls (0: the default, 2^22). A wait
        that runs out gives up: the solver drops to per-colour launches and reports it."""
        self._live()
        self._check(self.L.avbd_set_barrier_polls(self.h, ctypes.c_uint32(polls)))

    # --- command passes (Pusher, BodyMover, BodyReset) --------------------------------------
    def push(self, bodies, lin, off_lin=None, moment=None):
        """Pusher.apply (push.ts:90-140): per body (each once) its summed linear impulse, the part
        of it applied off-centre and that part's moment about the origin plus angular impulses, as
        addImpulse (push.ts:55-64) accumulates them."""
        self._live()
        n = len(bodies)
        r = np.zeros((n, 12), np.float32)
        r[:, 0:3] = np.asarray(lin, np.float32).reshape(n, 3)
        r[:, 3] = np.asarray(bodies, np.uint32).view(np.float32)
        if off_lin is not None:
            r[:, 4:7] = np.asarray(off_lin, np.float32).reshape(n, 3)
        if moment is not None:
            r[:, 8:11] = np.asarray(moment, np.float32).reshape(n, 3)
        self._check(self.L.avbd_push(self.h, _p(r), ctypes.c_uint32(n)))

    def move(self, bodies, positions, left, dt, rotations=None):
        """BodyMover.apply (body-mover.ts:56-74): kinematic bodies toward target poses."""
        self._live()
        n = len(bodies)
        r = np.zeros((n, 12), np.float32)
        r[:, 0:3] = np.asarray(positions, np.float32).reshape(n, 3)
        r[:, 3] = np.asarray(bodies, np.uint32).view(np.float32)
        r[:, 7] = 1.0
        if rotations is not None:
            r[:, 4:8] = np.asarray(rotations, np.float32).reshape(n, 4)
            r[:, 10] = 1.0
        r[:, 8] = np.broadcast_to(np.asarray(left, np.float32), (n,))
        r[:, 9] = np.float32(dt)
        self._check(self.L.avbd_move(self.h, _p(r), ctypes.c_uint32(n)))

    def reset_bodies(self, indices):
        """BodyReset.apply (body-reset.ts:41-52)."""
        self._live()
        idx = np.ascontiguousarray(indices, np.uint32)
        self._check(self.L.avbd_reset(self.h, _p(idx), ctypes.c_uint32(len(idx))))

    # --- contact events (ContactWatch, src/lib/contacts.ts) ---------------------------------
    def watch_create(self, pairs, capacity):
        """new ContactWatch(device, bodies, maxBodies, pairs, capacity) (contacts.ts:158-177)."""
        self._live()
        self._check(self.L.avbd_watch_create(self.h, ctypes.c_uint32(int(pairs)), ctypes.c_uint32(int(capacity))))
        self._watch_capacity = int(capacity)

    def set_watched(self, indices, ids, on):
        """setWatched(indices, ids, on) (contacts.ts:193-196)."""
        self._live()
        idx = np.ascontiguousarray(indices, np.uint32)
        ii = np.ascontiguousarray(ids, np.uint32)
        oo = np.ascontiguousarray(on, np.uint8)
        if not (len(idx) == len(ii) == len(oo)):
            raise ValueError("setWatched: indices, ids and on differ in length")
        self._check(self.L.avbd_watch_set(self.h, _p(idx), _p(ii), _p(oo), ctypes.c_uint32(len(idx))))

    def watch_run(self, pairs, step, dt):
        """run(storage, pairs, step, dt) (contacts.ts:199-217), after a step."""
        self._live()
        self._check(self.L.avbd_watch_run(self.h, ctypes.c_uint32(int(pairs)), ctypes.c_uint32(int(step)), ctypes.c_float(dt)))

    def read_contact_events(self):
        """read() (contacts.ts:224-261): (events in step order, dropped, full); clears them."""
        self._live()
        cap = getattr(self, "_watch_capacity", 0)
        out = (ContactEvent * max(cap, 1))()
        n, dropped, full = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        self._check(self.L.avbd_watch_read(self.h, out, ctypes.c_uint32(cap), ctypes.byref(n), ctypes.byref(dropped), ctypes.byref(full)))
        events = [{"kind": e.kind, "a": e.a, "b": e.b, "step": e.step, "point": list(e.point), "impulse": e.impulse,
                   "normal": list(e.normal)} for e in out[:n.value]]
        events.sort(key=lambda e: e["step"])  # in step order; within a step, as the GPU wrote them
        return events, dropped.value, bool(full.value)

    def set_sparse_tail(self, first_colour=TAIL_AUTO):
        """fuseSparseColors (solver.ts:574): colours from first_colour up run in one workgroup (64: none);
        under schedule 3 they and the dual run in the dataflow tail."""
        self._live()
        self._check(self.L.avbd_set_sparse_tail(self.h, ctypes.c_uint32(int(first_colour))))

    def autotune(self, steps=4, substeps=1):
        """Times the candidate solve schedules (the simulation advances) and keeps the fastest; under
        per-colour launches also the primal lane rule (read back into params)."""
        best = Schedule()
        ms = ctypes.c_float()
        self._check(self.L.avbd_autotune(self.h, steps, substeps, ctypes.byref(best), ctypes.byref(ms)))
        self.step_count = self.read_counters()["step"]
        p = Params()
        self._check(self.L.avbd_get_params(self.h, ctypes.byref(p)))
        self._params = p
        return {k: getattr(best, k) for k, _ in Schedule._fields_}, ms.value

    # --- reads -------------------------------------------------------------------------------
    def read_bodies(self, first=0, count=None):
        self._live()
        count = self.body_count - first if count is None else count
        out = np.zeros((count, BODY_FLOATS), np.float32)
        self._check(self.L.avbd_read_bodies(self.h, _p(out), first, count))
        return out

    def write_bodies(self, records, first=0):
        """rewriteBodies (solver.ts:1774)."""
        self._live()
        rec = np.ascontiguousarray(records, np.float32).reshape(-1, BODY_FLOATS)
        self._check(self.L.avbd_write_bodies(self.h, _p(rec), first, len(rec)))

    # --- scene edits (World.flush, world.ts:1861-1995) -----------------------------------------
    def _sync_sizes(self):
        nb, nj = ctypes.c_uint32(), ctypes.c_uint32()
        self._check(self.L.avbd_get_sizes(self.h, ctypes.byref(nb), ctypes.byref(nj)))
        self.body_count, self.joint_count = nb.value, nj.value
        self._joint_info = self.joint_info()

    def joint_info(self):
        """jointInfo (solver.ts:2356): (type, a, b, 0) per joint slot."""
        out = np.zeros((self.joint_count, 4), np.int32)
        self._check(self.L.avbd_read_joint_info(self.h, _p(out), 0, self.joint_count))
        return out

    def add_bodies(self, records):
        """addBodies (solver.ts:1544): the first new index, or -1 when none fit."""
        self._live()
        rec = np.ascontiguousarray(records, np.float32).reshape(-1, BODY_FLOATS)
        first = ctypes.c_int32(-1)
        self._check(self.L.avbd_add_bodies(self.h, _p(rec), len(rec), ctypes.byref(first)))
        self._sync_sizes()
        return first.value

    def patch_body(self, index, fields, record, fixed=False, marker=True):
        """patchBody (solver.ts:1103): `fields` names from {position, rotation, velocity,
        angularVelocity} taken from the 40-float `record`."""
        self._live()
        mask = 0
        for f in fields:
            mask |= _FIELDS[f]
        rec = np.ascontiguousarray(record, np.float32).reshape(BODY_FLOATS)
        self._check(self.L.avbd_patch_body(self.h, ctypes.c_uint32(index), ctypes.c_uint32(mask), _p(rec), int(fixed), int(marker)))

    def rewrite_fixed(self, indices, positions, size, friction, rotations=None):
        """rewriteFixed (solver.ts:1140)."""
        self._live()
        idx = np.ascontiguousarray(indices, np.uint32)
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        sz = np.ascontiguousarray(size, np.float32).reshape(3)
        rot = None if rotations is None else np.ascontiguousarray(rotations, np.float32).reshape(-1, 4)
        self._check(self.L.avbd_rewrite_fixed(self.h, _p(idx), len(idx), _p(pos), _p(sz), ctypes.c_float(friction),
                                              None if rot is None else _p(rot)))

    def append_joints(self, joints, fracture=float("inf"), linear=False):
        """appendJoints (solver.ts:1637): dicts with a, b, rA, rB and optionally angular,
        fracture, linear, rest, yield, breakBend; returns their slots."""
        self._live()
        nan = float("nan")
        specs = (JointSpec * max(1, len(joints)))()
        for k, j in enumerate(joints):
            sp = specs[k]
            sp.a, sp.b = int(j["a"]), int(j["b"])
            sp.rA[:] = [float(x) for x in j["rA"]]
            sp.rB[:] = [float(x) for x in j["rB"]]
            sp.angular = j.get("angular", nan)
            sp.fracture = j.get("fracture", nan)
            sp.linear = -1 if j.get("linear") is None else int(bool(j["linear"]))
            sp.rest[:] = [float(x) for x in j.get("rest", (0, 0, 0, 0))]
            sp.yield_ = j.get("yield", nan)
            sp.break_bend = j.get("breakBend", nan)
        slots = np.zeros(max(1, len(joints)), np.uint32)
        self._check(self.L.avbd_append_joints(self.h, specs, len(joints), ctypes.c_float(fracture), int(linear), _p(slots)))
        self._sync_sizes()
        return [int(x) for x in slots[:len(joints)]]

    def append_joint(self, a, b, rA, rB, stiffness_lin, stiffness_ang):
        """appendJoint (solver.ts:1559)."""
        self._live()
        ra = np.ascontiguousarray(rA, np.float32)
        rb = np.ascontiguousarray(rB, np.float32)
        slot = ctypes.c_uint32()
        self._check(self.L.avbd_append_joint(self.h, int(a), int(b), _p(ra), _p(rb), ctypes.c_float(stiffness_lin),
                                             ctypes.c_float(stiffness_ang), ctypes.byref(slot)))
        self._sync_sizes()
        return slot.value

    def append_world_joint(self, b, anchor_world, anchor_body):
        """appendWorldJoint (solver.ts:1592), legacy policy."""
        self._live()
        aw = np.ascontiguousarray(anchor_world, np.float32)
        ab = np.ascontiguousarray(anchor_body, np.float32)
        slot = ctypes.c_uint32()
        self._check(self.L.avbd_append_world_joint(self.h, ctypes.c_uint32(b), _p(aw), _p(ab), ctypes.byref(slot)))
        self._sync_sizes()
        return slot.value

    def append_springs(self, springs):
        """appendSprings (solver.ts:1685): dicts with a, b, rA, rB, stiffness, rest."""
        self._live()
        specs = (SpringSpec * max(1, len(springs)))()
        for k, sp in enumerate(springs):
            specs[k].a, specs[k].b = int(sp["a"]), int(sp["b"])
            specs[k].rA[:] = [float(x) for x in sp["rA"]]
            specs[k].rB[:] = [float(x) for x in sp["rB"]]
            specs[k].stiffness, specs[k].rest = sp["stiffness"], sp["rest"]
        slots = np.zeros(max(1, len(springs)), np.uint32)
        self._check(self.L.avbd_append_springs(self.h, specs, len(springs), _p(slots)))
        self._sync_sizes()
        return [int(x) for x in slots[:len(springs)]]

    def release_joints(self, slots):
        """releaseJoints (solver.ts:1745)."""
        self._live()
        s = np.ascontiguousarray(slots, np.uint32)
        self._check(self.L.avbd_release_joints(self.h, _p(s), len(s)))
        self._sync_sizes()

    def set_world_anchor(self, slot, p):
        """setWorldAnchor (solver.ts:1937)."""
        self._live()
        self._check(self.L.avbd_set_world_anchor(self.h, ctypes.c_uint32(slot), _p(np.ascontiguousarray(p, np.float32))))

    def disable_constraint(self, slot):
        """disableConstraint (solver.ts:1942)."""
        self._live()
        self._check(self.L.avbd_disable_constraint(self.h, ctypes.c_uint32(slot)))

    def set_body_aux(self, indices, scales, linear=None, angular=None):
        """setBodyAux (solver.ts:1814): gravity scales and, optionally, damping rates (1/s)."""
        self._live()
        idx = np.ascontiguousarray(indices, np.uint32)
        sc = np.ascontiguousarray(scales, np.float32)
        lin = None if linear is None else np.ascontiguousarray(linear, np.float32)
        ang = None if angular is None else np.ascontiguousarray(angular, np.float32)
        self._check(self.L.avbd_set_body_aux(self.h, _p(idx), len(idx), _p(sc), None if lin is None else _p(lin),
                                             None if ang is None else _p(ang)))

    def set_lifecycle(self, index, bits):
        """setLifecycle (solver.ts:2069): 0, 2 (disabled), 4 (cold) or 6."""
        self._live()
        self._check(self.L.avbd_set_lifecycle(self.h, ctypes.c_uint32(index), ctypes.c_uint32(bits)))

    def read_joints(self, first=0, count=None):
        self._live()
        count = self.joint_count - first if count is None else count
        out = np.zeros((count, JOINT_FLOATS), np.float32)
        self._check(self.L.avbd_read_joints(self.h, _p(out), first, count))
        return out

    def read_counters(self):
        self._live()
        c = Counters()
        self._check(self.L.avbd_read_counters(self.h, ctypes.byref(c)))
        out = {k: getattr(c, k) for k, _ in Counters._fields_}
        out["clashing_colors"] = list(c.clashing_colors)
        return out

    def read_contacts_raw(self):
        """Manifold records (M, 8) u32 and contact records (K, 16) u32 of the last step."""
        c = self.read_counters()
        man = np.zeros((max(c["manifolds"], 1), 8), np.uint32)
        con = np.zeros((max(c["contacts"], 1), 16), np.uint32)
        nm = self._check(self.L.avbd_read_contacts(self.h, _p(man), len(man), _p(con), len(con)))
        return man[:nm], con[:c["contacts"]]

    def read_contact_list(self):
        """readContactList (solver.ts:2311-2346): one dict per contact point."""
        man, con = self.read_contacts_raw()
        conf = con.view(np.float32)
        out = []
        for m in man:
            a, b, first, cs = (int(x) for x in m[:4])
            for c in range(first, first + (cs & 15)):
                out.append({
                    "a": a, "b": b, "feature": int(con[c, 3]) & ~(STICK_BIT | SDF_ROLLING_BIT),
                    "stick": bool(int(con[c, 3]) & STICK_BIT), "rolling": bool(int(con[c, 3]) & SDF_ROLLING_BIT),
                    "rA": conf[c, 0:3].copy(), "rB": conf[c, 4:7].copy(), "pen": conf[c, 8:11].copy(),
                    "lam": conf[c, 12:15].copy(), "c0": np.array([conf[c, 7], conf[c, 11], conf[c, 15]]),
                    "normal": m[4:7].view(np.float32).copy(),
                })
        return out

    def read_pairs(self):
        c = self.read_counters()
        out = np.zeros((max(c["pairs"], 1), 2), np.uint32)
        n = self._check(self.L.avbd_read_pairs(self.h, _p(out), len(out)))
        return out[:n]

    def read_colors(self):
        state = np.zeros(self.body_count, np.uint32)
        starts = np.zeros(65, np.uint32)
        self._check(self.L.avbd_read_colors(self.h, _p(state), self.body_count, _p(starts)))
        return state, starts

    # --- validation hooks --------------------------------------------------------------------
    def read_refs(self):
        """updateRefs' reference poses, n x 8 (pos xyz, 0, rot xyzw; zeros: none yet)."""
        out = np.zeros((self.body_count, 8), np.float32)
        self._check(self.L.avbd_read_refs(self.h, _p(out), 0, self.body_count))
        return out

    def seed_from(self, bodies, joints, manifolds, contacts):
        """seedFrom (solver.ts:1320-1359): bodies, joints with warm-start state, last contacts."""
        self._live()
        b = np.ascontiguousarray(bodies, np.float32).reshape(-1, BODY_FLOATS)
        j = np.ascontiguousarray(joints, np.float32).reshape(-1, JOINT_FLOATS) if joints is not None else np.zeros((0, 40), np.float32)
        m = np.ascontiguousarray(manifolds, np.uint32).reshape(-1, 8)
        c = np.ascontiguousarray(contacts, np.uint32).reshape(-1, 16)
        self._check(self.L.avbd_seed(self.h, _p(b), len(b), _p(j), len(j), _p(m), len(m), _p(c), len(c)))

    def set_fixed_colors(self, colors):
        """fixedColors (solver.ts:560): colours per body; None returns to the colouring."""
        self._live()
        if colors is None:
            self._check(self.L.avbd_set_fixed_colors(self.h, None, 0))
            self.fixed_colors = None
            return
        c = np.ascontiguousarray(colors, np.uint32)
        self._check(self.L.avbd_set_fixed_colors(self.h, _p(c), len(c)))
        self.fixed_colors = c

    def sequential_colors(self):
        """sequentialColors (solver.ts:1365-1372): newest dynamic body first, one colour each."""
        rec = self.read_bodies()
        colors = np.full(self.body_count, NO_COLOR, np.uint32)
        nxt = 0
        for i in range(self.body_count - 1, -1, -1):
            if rec[i, B_SIZE + 3] > 0:
                colors[i] = nxt
                nxt += 1
        if nxt > 64:
            raise ValueError(f"sequentialColors: {nxt} dynamic bodies, at most 64")
        return colors

    def set_color_state(self, state):
        s = np.ascontiguousarray(state, np.uint32)
        self._check(self.L.avbd_set_color_state(self.h, _p(s), len(s)))

    def set_filters(self, groups, collides_with):
        g = np.ascontiguousarray(groups, np.uint32)
        w = np.ascontiguousarray(collides_with, np.uint32)
        self._check(self.L.avbd_set_filters(self.h, _p(g), _p(w), len(g)))
