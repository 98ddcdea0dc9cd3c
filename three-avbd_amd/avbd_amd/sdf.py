"""SDF collision backend: the host-side mirror of the reference's ``SdfCollisionBridge``
(src/avbd3d/gpu/sdf-bridge.ts) and its fields (src/avbd3d/ref/sdf/shapes.ts, voxel.ts) over the
C ABI (include/avbd.h, ``avbd_sdf_*``).

``AnalyticField`` / ``VoxelField`` describe a field in its local frame (bounds, convexity, centre
of mass, curvature seeds, mesh supports) exactly as the reference's constructors compute them, in
float32. ``SdfCollisionBridge.create`` (explicit candidate pairs) and ``create_automatic`` (pairs
from the GPU spatial grid) build a bridge for a ``GpuSolver3D``; ``solver.set_collision_backend``
installs it, replacing all collision detection. Errors are ``ValueError`` (the reference's
``Error``) or ``StorageCapacityError``.
"""
import ctypes

import numpy as np

from . import solver as _s

f32 = np.float32
KINDS = {"box": 0, "sphere": 1, "capsule": 2, "cylinder": 3, "cone": 4, "torus": 5, "warp": 6, "voxel": 7,
         "spike-plate": 8, "spiky-wheel": 9}
SDF_MAX_POINTS = 5
SDF_POINT_LIMIT = 20
SDF_NO_CACHE = 0xFFFFFFFF
SDF_ROLLING_BIT = 0x20000000
MESH_WITNESS_MAX_VERTICES = 4096


class SdfField(ctypes.Structure):
    """avbd_sdf_field."""
    _fields_ = [("kind", ctypes.c_uint32), ("convex", ctypes.c_uint32), ("bounds_min", ctypes.c_float * 3),
                ("bounds_max", ctypes.c_float * 3), ("center_of_mass", ctypes.c_float * 3), ("params", ctypes.c_float * 8),
                ("dims", ctypes.c_uint32 * 3), ("origin", ctypes.c_float * 3), ("cell", ctypes.c_float),
                ("grid", ctypes.c_void_p), ("seeds", ctypes.c_void_p), ("seed_count", ctypes.c_uint32),
                ("supports", ctypes.c_void_p), ("support_count", ctypes.c_uint32)]


class SdfParams(ctypes.Structure):
    """avbd_sdf_params (CollisionParams, ref/sdf/manifold.ts:13-28)."""
    _fields_ = [("optimizer", ctypes.c_uint32), ("cache_size", ctypes.c_uint32), ("random_samples", ctypes.c_uint32),
                ("warm_start", ctypes.c_uint32), ("curvature_sampling", ctypes.c_uint32),
                ("skip_convex_samples", ctypes.c_uint32), ("epsilon", ctypes.c_float), ("alpha", ctypes.c_float),
                ("replacement_threshold", ctypes.c_float), ("tolerance", ctypes.c_float),
                ("max_iterations", ctypes.c_uint32), ("step_size", ctypes.c_float), ("margin", ctypes.c_float),
                ("allowed_penetration", ctypes.c_float), ("signed_separation", ctypes.c_uint32),
                ("mesh_vertex_contacts", ctypes.c_uint32), ("refresh_deeper_contacts", ctypes.c_uint32),
                ("sgd_gradient_stop_only", ctypes.c_uint32)]


_PARAM_KEYS = {"optimizer": "optimizer", "cacheSize": "cache_size", "randomSamples": "random_samples",
               "warmStart": "warm_start", "curvatureSampling": "curvature_sampling",
               "skipConvexSamples": "skip_convex_samples", "epsilon": "epsilon", "alpha": "alpha",
               "replacementThreshold": "replacement_threshold", "tolerance": "tolerance",
               "maxIterations": "max_iterations", "stepSize": "step_size", "margin": "margin",
               "allowedPenetration": "allowed_penetration", "signedSeparation": "signed_separation",
               "meshVertexContacts": "mesh_vertex_contacts", "refreshDeeperContacts": "refresh_deeper_contacts",
               "sgdGradientStopOnly": "sgd_gradient_stop_only"}


def collision_params(**overrides):
    """collisionParams() (ref/sdf/manifold.ts:29) with overrides, by the reference's names."""
    p = {"optimizer": 0, "cacheSize": 4, "randomSamples": 1, "warmStart": True, "curvatureSampling": True,
         "skipConvexSamples": True, "epsilon": float(f32(0.1)), "alpha": float(f32(0.01)),
         "replacementThreshold": float(f32(0.01)), "tolerance": float(f32(0.01)), "maxIterations": 50,
         "stepSize": float(f32(0.1)), "margin": float(f32(0.02)), "allowedPenetration": float(f32(0.01)),
         "signedSeparation": False, "meshVertexContacts": False, "refreshDeeperContacts": False,
         "sgdGradientStopOnly": False}
    for k, v in overrides.items():
        if k not in p:
            raise ValueError(f"unknown SDF collision parameter {k}")
        p[k] = v
    return p


def to_c_params(p):
    c = SdfParams()
    for k, name in _PARAM_KEYS.items():
        v = p[k]
        setattr(c, name, int(v) if isinstance(getattr(c, name), int) else float(v))
    return c


def _vec(v):
    return np.asarray(v, np.float64).astype(f32)


class Field:
    """A field in its local frame: bounds, convexity, centre of mass and curvature seeds."""
    kind = None
    convex = False

    def __init__(self):
        self.center_of_mass = np.zeros(3, f32)
        self.seeds = np.zeros((0, 3), f32)
        self.supports = np.zeros((0, 3), f32)

    def params(self):
        return np.zeros(8, f32)

    def describe(self):
        """The field as plain data (the avbd_sdf_field contents)."""
        d = {"kind": KINDS[self.kind], "convex": int(self.convex), "lo": np.asarray(self.bounds[0], f32),
             "hi": np.asarray(self.bounds[1], f32), "com": np.asarray(self.center_of_mass, f32), "params": self.params(),
             "dims": (0, 0, 0), "origin": np.zeros(3, f32), "cell": 0.0, "grid": None,
             "seeds": np.ascontiguousarray(self.seeds, f32).reshape(-1, 3),
             "supports": np.ascontiguousarray(self.supports, f32).reshape(-1, 3)}
        return d


class AnalyticField(Field):
    """AnalyticField (ref/sdf/shapes.ts:73-138): a spec dict with the reference's keys."""
    _DIMS = {"box": [], "sphere": ["radius"], "capsule": ["radius", "halfLength"], "cylinder": ["radius", "halfHeight"],
             "cone": ["radius", "halfHeight"], "torus": ["majorRadius", "minorRadius"], "warp": [],
             "spike-plate": ["spikeRadius", "minHeight", "maxHeight"],
             "spiky-wheel": ["hubRadius", "hubHalfWidth", "spikeRadius", "spikeHalfHeight"]}

    def __init__(self, spec):
        super().__init__()
        kind = spec.get("kind")
        if kind not in self._DIMS:
            raise ValueError("Unknown field")
        for key in self._DIMS[kind]:
            v = spec.get(key)
            if not isinstance(v, (int, float)) or not np.isfinite(f32(v)) or (v < 0 if key == "halfLength" else f32(v) <= 0):
                raise ValueError(f"Invalid SDF {key}")
        if "halfExtents" in spec:
            h = spec["halfExtents"]
            if len(h) != 3 or not all(np.isfinite(f32(x)) and f32(x) > 0 for x in h):
                raise ValueError("Invalid SDF halfExtents")
        if kind == "warp" and (spec.get("type") not in (0, 1, 2) or not np.isfinite(f32(spec.get("k", np.nan)))):
            raise ValueError("Invalid SDF warp")
        if kind == "spike-plate" and (int(spec.get("count", 0)) != spec.get("count") or spec["count"] < 1 or spec["maxHeight"] < spec["minHeight"]):
            raise ValueError("Invalid SDF spike count/heights")
        if kind == "spiky-wheel" and (int(spec.get("spikes", 0)) != spec.get("spikes") or spec["spikes"] < 1):
            raise ValueError("Invalid SDF spike count")
        s = {k: (f32(v) if isinstance(v, (int, float)) and k not in ("kind", "type", "count", "spikes") else v) for k, v in spec.items()}
        if "halfExtents" in s:
            s["halfExtents"] = _vec(s["halfExtents"])
        self.spec, self.kind = s, kind
        self.convex = kind in ("box", "sphere", "capsule", "cylinder", "cone")
        if kind in ("box", "spike-plate"):
            h = s["halfExtents"].copy()
        elif kind == "sphere":
            h = np.array([s["radius"]] * 3, f32)
        elif kind == "capsule":
            h = np.array([s["radius"], f32(s["radius"] + s["halfLength"]), s["radius"]], f32)
        elif kind in ("cone", "cylinder"):
            h = np.array([s["radius"], s["halfHeight"], s["radius"]], f32)
        elif kind == "torus":
            r = f32(s["majorRadius"] + s["minorRadius"])
            h = np.array([r, s["minorRadius"], r], f32)
        elif kind == "spiky-wheel":
            r = f32(s["hubRadius"] + f32(f32(2) * s["spikeHalfHeight"]))
            h = np.array([r, s["hubHalfWidth"], r], f32)
        else:  # warp
            he = s["halfExtents"]
            r = f32(np.sqrt(f32(f32(he[0] * he[0]) + f32(he[1] * he[1]))))
            h = np.array([r, r, he[2]] if s["type"] == 0 else ([r, he[2], r] if s["type"] == 1 else he), f32)
        lo, hi = -h, h.copy()
        if kind == "warp" and s["type"] == 2:
            k = f32(abs(s["k"]))
            lo, hi = (lo - k).astype(f32), (hi + k).astype(f32)
        if kind == "spike-plate":
            hi[1] = f32(h[1] + f32(s["maxHeight"]))
        if not np.all(np.isfinite(np.concatenate([lo, hi]))):
            raise ValueError("SDF bounds exceed float32 range")
        self.bounds = (lo.astype(f32), hi.astype(f32))

    def params(self):
        s, p = self.spec, np.zeros(8, f32)
        k = self.kind
        if k == "box":
            p[0:3] = s["halfExtents"]
        elif k == "sphere":
            p[0] = s["radius"]
        elif k == "capsule":
            p[0:2] = [s["radius"], s["halfLength"]]
        elif k in ("cone", "cylinder"):
            p[0:2] = [s["radius"], s["halfHeight"]]
        elif k == "torus":
            p[0:2] = [s["majorRadius"], s["minorRadius"]]
        elif k == "spike-plate":
            if s["count"] >= 2 ** 24:
                raise ValueError("GPU SDF spike count exceeds exact float32 range")
            p[0:3] = s["halfExtents"]; p[3] = s["count"]
            p[4:7] = [s["spikeRadius"], s["minHeight"], s["maxHeight"]]
        elif k == "spiky-wheel":
            if s["spikes"] >= 2 ** 24:
                raise ValueError("GPU SDF spike count exceeds exact float32 range")
            p[0:4] = [s["hubRadius"], s["hubHalfWidth"], s["spikeRadius"], s["spikeHalfHeight"]]
            p[4] = s["spikes"]
        elif k == "warp":
            p[0:3] = s["halfExtents"]; p[3] = s["k"]; p[4] = s["type"]
        return p


class VoxelField(Field):
    """VoxelField (ref/sdf/voxel.ts): dims (each >= 2), origin, cell, grid (4 floats per node:
    phi, gradient xyz; x fastest), bounds ((min xyz), (max xyz))."""
    kind = "voxel"
    convex = False

    def __init__(self, dims, origin, cell, grid, bounds):
        super().__init__()
        dims = [int(x) for x in dims]
        if len(dims) != 3 or any(x < 2 for x in dims) or not np.isfinite(f32(cell)) or f32(cell) <= 0 or not np.all(np.isfinite(_vec(origin))):
            raise ValueError("Invalid SDF grid geometry")
        grid = np.ascontiguousarray(grid, f32).reshape(-1)
        if not np.all(np.isfinite(grid)):
            raise ValueError("SDF grid must be finite")
        if grid.size != dims[0] * dims[1] * dims[2] * 4:
            raise ValueError("SDF grid length mismatch")
        lo, hi = _vec(bounds[0]), _vec(bounds[1])
        if np.any(lo > hi) or not np.all(np.isfinite(np.concatenate([lo, hi]))):
            raise ValueError("Invalid SDF bounds")
        self.dims, self.origin, self.cell, self.grid, self.bounds = dims, _vec(origin), f32(cell), grid, (lo, hi)

    def describe(self):
        d = super().describe()
        d.update({"dims": tuple(self.dims), "origin": self.origin, "cell": float(self.cell), "grid": self.grid})
        return d


def mesh_surface_vertices(triangles):
    """meshSurfaceVertices (sdf-supports.ts:31-38): every unique vertex, sorted by x, y, z."""
    t = np.ascontiguousarray(triangles, f32).reshape(-1, 3)
    if len(t) == 0:
        return np.zeros((0, 3), f32)
    u = np.unique(t, axis=0)
    if len(u) > MESH_WITNESS_MAX_VERTICES:
        raise ValueError(f"meshVertexContacts supports at most {MESH_WITNESS_MAX_VERTICES} unique surface vertices per field "
                         f"(got {len(u)}); decimate the source mesh")
    return u[np.lexsort((u[:, 2], u[:, 1], u[:, 0]))]


def field_struct(d, keep):
    """avbd_sdf_field for a field description (`keep` holds the arrays it points into)."""
    c = SdfField()
    c.kind, c.convex = d["kind"], d["convex"]
    for i in range(3):
        c.bounds_min[i], c.bounds_max[i], c.center_of_mass[i] = d["lo"][i], d["hi"][i], d["com"][i]
        c.dims[i], c.origin[i] = d["dims"][i], d["origin"][i]
    for i in range(8):
        c.params[i] = d["params"][i]
    c.cell = d["cell"]
    if d["grid"] is not None:
        g = np.ascontiguousarray(d["grid"], f32)
        keep.append(g)
        c.grid = g.ctypes.data
    seeds, sup = np.ascontiguousarray(d["seeds"], f32), np.ascontiguousarray(d["supports"], f32)
    keep += [seeds, sup]
    c.seeds, c.seed_count = (seeds.ctypes.data if len(seeds) else None), len(seeds)
    c.supports, c.support_count = (sup.ctypes.data if len(sup) else None), len(sup)
    return c


RESULT_HEAD = 8
POINT_WORDS = 28


def decode_results(words, points):
    """decodeSdfResults (sdf-layout.ts:121-133) over avbd_sdf_read_results records."""
    u = np.ascontiguousarray(words, np.uint32)
    f = u.view(np.float32)
    stride = RESULT_HEAD + POINT_WORDS * points
    out = []
    for i in range(len(u) // stride):
        at = i * stride
        n = int(u[at + 3])
        pts = []
        for j in range(min(n, points)):
            o = at + RESULT_HEAD + j * POINT_WORDS
            pts.append({"x": f[o:o + 3].tolist(), "depth": float(f[o + 3]), "normal": f[o + 4:o + 7].tolist(),
                        "separation": float(f[o + 7]), "anchorA": f[o + 8:o + 11].tolist(), "softmax": float(f[o + 11]),
                        "anchorB": f[o + 12:o + 15].tolist(), "rank": float(f[o + 15]), "active": bool(u[o + 16]),
                        "cached": bool(u[o + 17]), "id": int(u[o + 18]), "exported": int(u[o + 19]),
                        "ellipsoid": f[o + 20:o + 26].tolist()})
        out.append({"a": int(u[at]), "b": int(u[at + 1]), "key": int(u[at + 2]), "queries": int(u[at + 4]),
                    "iterations": int(u[at + 5]), "idCounter": int(u[at + 6]), "rolling": bool(u[at + 7]), "points": pts})
    return out


class SdfCollisionBridge:
    """SdfCollisionBridge (sdf-bridge.ts:25-227) for one solver."""

    def __init__(self, solver, bodies, pairs, params, automatic, max_pairs, max_points, cell_size):
        L = _s.lib()
        self.L, self.solver = L, solver
        solver._live()
        if not bodies or len(bodies) != solver.body_count:
            raise ValueError("SDF bridge requires a field for every GPU body slot")
        fields, index, keep = [], {}, []
        body_field = np.zeros(len(bodies), np.uint32)
        scale = np.ones(len(bodies), f32)
        rot = np.tile(np.array([0, 0, 0, 1], f32), (len(bodies), 1))
        for i, b in enumerate(bodies):
            fld = b["field"]
            if id(fld) not in index:
                index[id(fld)] = len(fields)
                fields.append(fld)
            body_field[i] = index[id(fld)]
            sc = b.get("scale", 1.0)
            if not (f32(sc) > 0 and np.isfinite(f32(sc))):
                raise ValueError("Invalid SDF bridge scale")
            scale[i] = sc
            if b.get("rotation") is not None:
                q = np.asarray(b["rotation"], np.float64)
                if q.size != 4 or not np.all(np.isfinite(q)) or abs(np.linalg.norm(q) - 1) > 1e-5:
                    raise ValueError("SDF field rotation must be a unit quaternion")
                rot[i] = q
        self.fields = fields
        arr = (SdfField * len(fields))(*[field_struct(f.describe(), keep) for f in fields])
        self.params = collision_params(**(params or {}))
        cp = to_c_params(self.params)
        h = ctypes.c_void_p()
        rc = L.avbd_sdf_create(solver.h, arr, len(fields), _s._p(body_field), _s._p(scale), _s._p(rot.reshape(-1)),
                               ctypes.byref(cp), int(automatic), int(max_pairs), int(max_points), ctypes.c_float(cell_size),
                               ctypes.byref(h))
        if rc != 0:
            solver._check(rc)
        self.h, self.automatic, self.points, self.installed, self.disposed = h, automatic, max_points or SDF_MAX_POINTS, False, False
        self._keep = keep
        solver.__dict__.setdefault("_bridges", []).append(self)
        if not automatic:
            self.set_pairs(pairs)

    @classmethod
    def create(cls, solver, bodies, pairs, params=None, max_points=SDF_MAX_POINTS):
        """create (sdf-bridge.ts:56-82): explicit candidates {a, b, key} (a > b)."""
        return cls(solver, bodies, pairs, params, False, max(1, len(pairs)), max_points, 0.0)

    @classmethod
    def create_automatic(cls, solver, bodies, params=None, options=None):
        """createAutomatic (sdf-bridge.ts:48-54) with SdfBroadphaseOptions (cellSize, maxPairs,
        maxPointsPerPair)."""
        o = options or {}
        return cls(solver, bodies, [], params, True, o.get("maxPairs", 0), o.get("maxPointsPerPair", SDF_MAX_POINTS),
                   o.get("cellSize", 0.0))

    def _check(self, rc):
        if rc >= 0:
            return rc
        msg = self.L.avbd_sdf_last_error(self.h).decode()
        if rc == -2:
            raise _s.StorageCapacityError(msg)
        raise (ValueError if rc == -1 else RuntimeError)(msg)

    def _live(self):
        if self.disposed:
            raise ValueError("SDF bridge is disposed")

    def set_pairs(self, pairs):
        """setPairs (sdf-bridge.ts:133-145)."""
        self._live()
        w = np.array([[p["a"], p["b"], p["key"]] for p in pairs], np.uint32).reshape(-1, 3)
        self._check(self.L.avbd_sdf_set_pairs(self.h, _s._p(w), len(w)))

    def set_params(self, **overrides):
        """The mutable bridge.params: validated here and again before each step."""
        self._live()
        p = dict(self.params)
        p.update(overrides)
        self._check(self.L.avbd_sdf_set_params(self.h, ctypes.byref(to_c_params(p))))
        self.params = p

    @property
    def capacity(self):
        n = ctypes.c_uint32()
        self.L.avbd_sdf_read_results(self.h, None, 0, ctypes.byref(n))
        return n.value

    def read_raw(self):
        self._live()
        n = ctypes.c_uint32()
        self._check(self.L.avbd_sdf_read_results(self.h, None, 0, ctypes.byref(n)))
        out = np.zeros(max(1, n.value) * (RESULT_HEAD + POINT_WORDS * self.points), np.uint32)
        self._check(self.L.avbd_sdf_read_results(self.h, _s._p(out), n.value, ctypes.byref(n)))
        return out[:n.value * (RESULT_HEAD + POINT_WORDS * self.points)]

    def read_results(self):
        """readResults (sdf-bridge.ts:216): automatic drops the empty slots."""
        res = decode_results(self.read_raw(), self.points)
        return [r for r in res if r["a"] != SDF_NO_CACHE] if self.automatic else res

    def read_broadphase_stats(self):
        """readBroadphaseStats (sdf-bridge.ts:217)."""
        self._live()
        p, t, l, c, cell = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_float()
        self._check(self.L.avbd_sdf_read_stats(self.h, ctypes.byref(p), ctypes.byref(t), ctypes.byref(l), ctypes.byref(c),
                                               ctypes.byref(cell)))
        return {"pairs": p.value, "testedPairs": t.value, "largeBodies": l.value, "capacity": c.value, "cellSize": cell.value}

    def grown(self, required):
        """grown (sdf-bridge.ts:183-201), in place: returns this bridge with the larger capacity."""
        self._live()
        self._check(self.L.avbd_sdf_grow(self.h, int(required)))
        return self

    def destroy(self):
        if self.disposed:
            return
        self.disposed = True
        if self.installed and not self.solver.destroyed:
            self.solver._backend = None
        if self.solver.destroyed:
            return  # the solver's end destroyed its installed bridge
        self.L.avbd_sdf_destroy(self.h)
        self.h = None
