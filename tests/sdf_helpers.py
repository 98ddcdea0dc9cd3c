"""SDF test helpers: the reference test suite's field specs (tests/sdf-reference.test.ts:18-28)
and the golden fixtures in tests/golden/sdf (made by tests/golden/sdf/make_fixtures.py)."""
import json
import os

import numpy as np

from avbd_amd.sdf import AnalyticField

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sdf")

SPECS = {
    "box": {"kind": "box", "halfExtents": [.5, .5, .5]}, "sphere": {"kind": "sphere", "radius": .5},
    "capsule": {"kind": "capsule", "radius": .3, "halfLength": .4},
    "cylinder": {"kind": "cylinder", "radius": .4, "halfHeight": .4},
    "cone": {"kind": "cone", "radius": .5, "halfHeight": .5},
    "torus": {"kind": "torus", "majorRadius": .5, "minorRadius": .2},
    "spike-plate": {"kind": "spike-plate", "halfExtents": [.7, .08, .7], "count": 10, "spikeRadius": .045,
                    "minHeight": .12, "maxHeight": .26},
    "spiky-wheel": {"kind": "spiky-wheel", "hubRadius": .35, "hubHalfWidth": .48, "spikeRadius": .08,
                    "spikeHalfHeight": .7, "spikes": 16},
}
for _name, _type, _values in (("bend", 0, (.95, .15, .55, .95)), ("twist", 1, (1.25, .5, 2.5, 5)), ("displace", 2, (.1, .05, .1, .2))):
    for _i, _k in enumerate(_values):
        SPECS[_name if _i == 0 else f"{_name}-{_k:.2f}"] = {"kind": "warp", "type": _type, "k": _k,
                                                          "halfExtents": [1.15, .55, .28] if _type == 1 else [1.15, .28, .55]}

_cache = {}


def golden(name):
    if name not in _cache:
        with open(os.path.join(GOLDEN, name)) as f:
            _cache[name] = json.load(f)
    return _cache[name]


def seeded_field(name):
    """The field of a spec with the upstream oracle's centre of mass and curvature seeds (the
    reference test's field(name))."""
    f = AnalyticField(SPECS[name])
    s = golden("seeds.json")[name]
    f.center_of_mass = np.float32(s["centerOfMass"])
    f.seeds = np.float32(s["seeds"]).reshape(-1, 3)
    return f


def body_desc(b):
    return {k: b[k] for k in ("position", "orientation", "scale", "static", "velocity", "angularVelocity")}


# --- scenes -----------------------------------------------------------------------------------
import math

from avbd_amd import records
from avbd_amd.sdf import VoxelField

Z_UP = np.float32([math.sin(math.pi / 4), 0, 0, math.cos(math.pi / 4)])  # field +y -> body +z


def rand_quat(rng):
    q = rng.normal(size=4)
    return (q / np.linalg.norm(q)).astype(np.float32)


def baked_voxel(field_desc, n=12, pad=.15):
    """A voxel field sampled from an analytic field at the grid's nodes (phi and gradient), with
    that field's bounds: the same surface through the trilinear path."""
    import oracle
    lo, hi = np.asarray(field_desc["lo"], np.float64) - pad, np.asarray(field_desc["hi"], np.float64) + pad
    cell = float(np.float32((hi - lo).max() / (n - 1)))
    dims = [int(math.ceil((hi[i] - lo[i]) / cell)) + 1 for i in range(3)]
    origin = lo.astype(np.float32)
    xs = [origin[i] + np.float32(cell) * np.arange(dims[i], dtype=np.float32) for i in range(3)]
    Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    pts = np.stack([X, Y, Z], -1).reshape(-1, 3).astype(np.float32)
    grid = oracle.sdf_sample(field_desc, pts)
    return VoxelField(dims, origin, cell, grid, (field_desc["lo"], field_desc["hi"]))


def field_zoo():
    """One field of every kind, sized about a unit body, with curvature seeds from the golden
    fixtures where the kind has them (else the field's box corners pulled onto its surface)."""
    from avbd_amd.sdf import AnalyticField
    zoo = []
    for name in ("box", "sphere", "capsule", "cylinder", "cone", "torus", "bend", "twist", "displace", "spike-plate", "spiky-wheel"):
        zoo.append((name, seeded_field(name)))
    vox = baked_voxel(AnalyticField({"kind": "torus", "majorRadius": .45, "minorRadius": .18}).describe())
    vox.seeds = np.float32([[.63, 0, 0], [-.63, 0, 0], [0, 0, .63], [0, 0, -.63], [0, .18, .45], [0, -.18, -.45]])
    zoo.append(("voxel", vox))
    return zoo


def body_for_field(field, pos, rot, density=1.0, friction=0.6, vel=(0, 0, 0), angvel=(0, 0, 0)):
    """A solver body carrying a field: box inertia of the field's bounds (in the body frame with
    the Z_UP field rotation), density `density` (0: static)."""
    lo, hi = np.asarray(field.bounds[0], np.float64), np.asarray(field.bounds[1], np.float64)
    ext = hi - lo
    size = (ext[0], ext[2], ext[1])  # field y is the body's z
    r = records.box_record(pos, size, density, friction, rot=rot, vel=vel, angvel=angvel).astype(np.float32)
    r[11] = 1.0
    return r


def floor_body(friction=0.6, half=(6, .2, 6)):
    from avbd_amd.sdf import AnalyticField
    f = AnalyticField({"kind": "box", "halfExtents": list(half)})
    r = records.box_record((0, 0, -half[1]), (2 * half[0], 2 * half[2], 2 * half[1]), 0.0, friction).astype(np.float32)
    r[11] = 1.0
    return f, r


def zoo_scene(seed=3, layers=1, spacing=2.2, lift=0.05):
    """A static floor (body 0) and one body per field kind (times `layers`) resting on or just above
    it, randomly turned; returns (records, fields per body, explicit pairs)."""
    rng = np.random.default_rng(seed)
    ff, fr = floor_body()
    recs, fields = [fr], [ff]
    zoo = field_zoo()
    k = 0
    for layer in range(layers):
        for name, f in zoo:
            x = (k % 4 - 1.5) * spacing
            y = ((k // 4) % 4 - 1.5) * spacing
            z = float(max(f.bounds[1]) + lift + layer * 2.4)
            recs.append(body_for_field(f, (x, y, z), rand_quat(rng)))
            fields.append(f)
            k += 1
    pairs = [{"a": i, "b": 0, "key": 1000 + i} for i in range(1, len(recs))]
    n = len(recs)
    for i in range(1, n):
        for j in range(1, i):
            if np.linalg.norm(recs[i][0:3] - recs[j][0:3]) < 2.6:
                pairs.append({"a": i, "b": j, "key": 5000 + i * 97 + j})
    return np.array(recs, np.float32), fields, pairs


def pile_scene(n=240, seed=11, pitch=1.25):
    """A static floor and n bodies of mixed analytic fields on a jittered grid above it."""
    from avbd_amd.sdf import AnalyticField
    rng = np.random.default_rng(seed)
    ff, fr = floor_body(half=(8, .2, 8))
    kinds = [AnalyticField({"kind": "sphere", "radius": .35}), AnalyticField({"kind": "box", "halfExtents": [.3, .25, .35]}),
             AnalyticField({"kind": "capsule", "radius": .2, "halfLength": .25}),
             AnalyticField({"kind": "cylinder", "radius": .3, "halfHeight": .25}),
             AnalyticField({"kind": "torus", "majorRadius": .3, "minorRadius": .12})]
    kinds[3].seeds = np.float32([[.3, .25, 0], [-.3, .25, 0], [0, .25, .3], [0, -.25, -.3]])
    kinds[4].seeds = np.float32([[.42, 0, 0], [-.42, 0, 0], [0, 0, .42], [0, 0, -.42]])
    recs, fields = [fr], [ff]
    side = int(math.ceil(math.sqrt(n / 3)))
    for i in range(n):
        f = kinds[i % len(kinds)]
        layer, cell = divmod(i, side * side)
        x = (cell % side - side / 2) * pitch + rng.uniform(-.1, .1)
        y = (cell // side - side / 2) * pitch + rng.uniform(-.1, .1)
        z = 0.45 + layer * pitch
        recs.append(body_for_field(f, (x, y, z), rand_quat(rng), density=rng.uniform(.5, 2)))
        fields.append(f)
    return np.array(recs, np.float32), fields
