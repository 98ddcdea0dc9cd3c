"""The oracle's SDF backend (oracle/sdf3d.hpp) pinned before it is trusted as the HIP backend's
checker (CPU only):

* its own sin, cos and atan2 against the C library's to a few ulp;
* every field kind against the 3,500 samples of the upstream C++ oracle that the reference's
  tests hold (tests/sdf-reference.test.ts:41-50, tolerance 2e-6), and the host-side field
  bounds against the fixtures' bounds exactly;
* the voxel field against closed forms (a linear field interpolates exactly; outside the grid the
  distance to the grid box is added) and the reference's boundary cases;
* Nelder-Mead, the ellipsoid method (cold and warm) and gradient descent against the upstream
  optimiser runs (optimizer.json, optimizers.json: iterations and queries exactly, points within
  the tolerance, the final ellipsoid to 1e-4), as the reference's own tests check them;
* whole manifold updates (cached points advected and re-optimised, fresh samples from the
  curvature seeds, dedup, rank order, convex skip, margins, scaled and rotated bodies) against
  manifolds.json and the chained optimizer-1/2 cases of manifolds-optimizers.json, with the
  reference's PCG stream: the stream's state afterwards, the point count, flags and stats exactly.
The CPU update and the GPU one share every piece but the random stream and the seed filter; the
GPU's stream is the lowbias32 hash (pinned in test_colouring_greedy.py).
"""
import math

import numpy as np
import pytest

import oracle
from avbd_amd.sdf import AnalyticField, VoxelField, collision_params
from sdf_helpers import SPECS, golden, seeded_field, body_desc

f32 = np.float32


def ulps(got, want):
    want = np.asarray(want, np.float64)
    return np.abs(np.asarray(got, np.float64) - want) / np.spacing(np.maximum(np.abs(want), 1e-30).astype(np.float32)).astype(np.float64)


def test_trig_is_within_a_few_ulp_of_the_c_library():
    x = np.concatenate([np.linspace(-40, 40, 200001), [0.0, -0.0, math.pi / 2, math.pi, 1e-8, -1e-8, 7 * 1.3]]).astype(f32)
    s, c = oracle.sdf_trig(0, x), oracle.sdf_trig(1, x)
    xd = x.astype(np.float64)
    big = np.abs(np.sin(xd)) > 1e-3
    assert ulps(s, np.sin(xd))[big].max() <= 2.0
    assert np.abs(s - np.sin(xd)).max() < 1.5e-7
    big = np.abs(np.cos(xd)) > 1e-3
    assert ulps(c, np.cos(xd))[big].max() <= 2.0
    assert np.abs(c - np.cos(xd)).max() < 1.5e-7
    yx = np.random.default_rng(3).normal(size=(200000, 2)).astype(f32)
    yx = np.vstack([yx, [[0, 1], [0, -1], [1, 0], [-1, 0], [1, 1], [-1, -1], [0, 0]]]).astype(f32)
    a = oracle.sdf_trig(2, yx.reshape(-1))
    want = np.arctan2(yx[:, 0].astype(np.float64), yx[:, 1].astype(np.float64))
    assert ulps(a[:-1], want[:-1]).max() <= 4.0
    assert a[-1] == 0.0


@pytest.mark.parametrize("name", sorted(SPECS))
def test_field_matches_upstream_samples_and_bounds(name):
    fx = golden("fields.json")[name]
    f = AnalyticField(SPECS[name])
    d = f.describe()
    assert np.array_equal(d["lo"], f32(fx["bounds"]["min"])) and np.array_equal(d["hi"], f32(fx["bounds"]["max"]))
    P = np.array([s[0] for s in fx["samples"]], f32)
    want = np.array([[s[1]] + s[2] for s in fx["samples"]])
    got = oracle.sdf_sample(d, P)
    assert len(P) == 175
    assert np.all(np.abs(got - want) <= 2e-6 * np.maximum(1, np.abs(want))), name


def test_field_kinds_against_closed_forms():
    rng = np.random.default_rng(5)
    P = rng.uniform(-1.5, 1.5, size=(2000, 3)).astype(f32)
    sph = oracle.sdf_sample(AnalyticField({"kind": "sphere", "radius": .7}).describe(), P)
    r = np.linalg.norm(P.astype(np.float64), axis=1)
    assert np.abs(sph[:, 0] - (r - .7)).max() < 1e-6
    assert np.abs(sph[:, 1:] - P / r[:, None]).max() < 1e-6
    box = oracle.sdf_sample(AnalyticField({"kind": "box", "halfExtents": [.3, .5, .7]}).describe(), P)
    q = np.abs(P.astype(np.float64)) - [.3, .5, .7]
    want = np.linalg.norm(np.maximum(q, 0), axis=1) + np.minimum(q.max(axis=1), 0)
    assert np.abs(box[:, 0] - want).max() < 1e-6
    cap = oracle.sdf_sample(AnalyticField({"kind": "capsule", "radius": .25, "halfLength": .6}).describe(), P)
    seg = P.astype(np.float64).copy()
    seg[:, 1] -= np.clip(seg[:, 1], -.6, .6)
    assert np.abs(cap[:, 0] - (np.linalg.norm(seg, axis=1) - .25)).max() < 1e-6
    tor = oracle.sdf_sample(AnalyticField({"kind": "torus", "majorRadius": .6, "minorRadius": .2}).describe(), P)
    rr = np.hypot(P[:, 0].astype(np.float64), P[:, 2].astype(np.float64))
    assert np.abs(tor[:, 0] - (np.hypot(rr - .6, P[:, 1]) - .2)).max() < 1e-6
    # twist and bend are distance-preserving maps of a box: phi at q = phi_box at the mapped point
    for typ in (0, 1):
        spec = {"kind": "warp", "type": typ, "k": 1.3, "halfExtents": [.5, .2, .3]}
        got = oracle.sdf_sample(AnalyticField(spec).describe(), P)
        x, y, z = (P[:, i].astype(np.float64) for i in range(3))
        ang = 1.3 * (x if typ == 0 else y)
        c, s = np.cos(ang), np.sin(ang)
        m = np.stack([c * x - s * y, s * x + c * y, z], 1) if typ == 0 else np.stack([c * x - s * z, s * x + c * z, y], 1)
        q = np.abs(m) - [.5, .2, .3]
        want = np.linalg.norm(np.maximum(q, 0), axis=1) + np.minimum(q.max(axis=1), 0)
        assert np.abs(got[:, 0] - want).max() < 2e-6


def _linear_grid(n, dims, origin, cell, c):
    xs = [origin[i] + cell * np.arange(dims[i]) for i in range(3)]
    Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    pts = np.stack([X, Y, Z], -1).reshape(-1, 3)
    g = np.zeros((len(pts), 4), f32)
    g[:, 0] = pts @ n - c
    g[:, 1:] = n
    return g


def test_voxel_interpolation_and_outside_extension_by_closed_forms():
    n = np.array([.48, .6, .64])
    dims, origin, cell = (5, 6, 7), np.array([-1, -1.25, -1.5], f32), f32(.5)
    grid = _linear_grid(n, dims, origin, cell, .1)
    hi = origin + cell * (np.array(dims) - 1)
    v = VoxelField(dims, origin, cell, grid, (origin, hi))
    rng = np.random.default_rng(9)
    inside = rng.uniform(origin, hi, size=(3000, 3)).astype(f32)
    got = oracle.sdf_sample(v.describe(), inside)
    assert np.abs(got[:, 0] - (inside.astype(np.float64) @ n - .1)).max() < 2e-6  # trilinear reproduces a plane
    assert np.abs(got[:, 1:] - n).max() < 1e-6
    out = rng.uniform(origin - 2, hi + 2, size=(3000, 3)).astype(f32)
    got = oracle.sdf_sample(v.describe(), out)
    q = np.clip(out.astype(np.float64), origin, hi)
    d = np.linalg.norm(out - q, axis=1)
    assert np.abs(got[:, 0] - (q @ n - .1 + d)).max() < 3e-6
    # the reference's boundary cases (tests/sdf-reference.test.ts:180-190)
    data = np.zeros(32, f32)
    for i in range(8):
        data[i * 4] = i & 1
        data[i * 4 + 1] = 1
    v = VoxelField((2, 2, 2), (0, 0, 0), 1, data, ((0, 0, 0), (1, 1, 1)))
    for x in (0, .25, 1, 2):
        assert abs(oracle.sdf_sample(v.describe(), [[x, .5, .5]])[0, 0] - x) < 2e-6
    assert abs(oracle.sdf_sample(v.describe(), [[-1, .5, .5]])[0, 0] - 1) < 2e-6
    for bad in (((1, 2, 2), (0, 0, 0), 1, data), ((2, 2, 2), (0, 0, 0), 0, data), ((2, 2, 2), (0, 0, 0), 1, data[:4])):
        with pytest.raises(ValueError):
            VoxelField(*bad, ((0, 0, 0), (1, 1, 1)))


def _pair_fields(c):
    return AnalyticField(SPECS[c["shapeA"]]).describe(), AnalyticField(SPECS[c["shapeB"]]).describe()


def test_nelder_mead_matches_the_upstream_optimisations():
    cases = golden("optimizer.json")["cases"]
    assert len(cases) == 16 and any(c["repulsionPoints"] for c in cases)
    for c in cases:
        A, B = _pair_fields(c)
        prm = {"epsilon": f32(c["epsilon"]), "alpha": f32(c["alpha"]), "tolerance": f32(c["tolerance"]),
               "maxIterations": c["maxIterations"]}
        r = oracle.sdf_optimize("nelder-mead", A, {}, B, {"position": c["originB"], "orientation": c["rotationB"]},
                                (c["domain"]["min"], c["domain"]["max"]), c["repulsionPoints"], c["seed"], prm, step=c["initialStep"])
        assert np.linalg.norm(r["x"] - c["x"]) <= c["tolerance"]
        assert abs(r["objective"] - c["objective"]) <= 1e-5 * max(1, abs(c["objective"]))
        assert (r["iterations"], r["queries"]) == (c["iterations"], c["queries"])


def test_ellipsoid_and_gradient_descent_match_the_upstream_optimisations():
    cases = golden("optimizers.json")["cases"]
    assert len(cases) == 16
    for c in cases:
        A, B = _pair_fields(c)
        bb = {"position": c["originB"], "orientation": c["rotationB"]}
        dom = (c["domain"]["min"], c["domain"]["max"])
        prm = {"epsilon": f32(c["epsilon"]), "alpha": f32(c["alpha"]), "tolerance": f32(c["tolerance"]),
               "maxIterations": c["maxIterations"], "stepSize": f32(c["stepSize"]), "optimizer": 2}
        g = oracle.sdf_optimize("gradient", A, {}, B, bb, dom, c["repulsionPoints"], c["seed"], prm)
        assert abs(g["value"] - c["seedValue"]) <= 2e-6 * max(1, abs(c["seedValue"]))
        assert np.all(np.abs(g["gradient"] - c["seedGradient"]) <= 2e-5 * np.maximum(1, np.abs(c["seedGradient"])))
        for run, seed in (("ellipsoid", c["seed"]), ("warmEllipsoid", c["warmEllipsoid"]["seed"])):
            e = c[run]
            r = oracle.sdf_optimize("ellipsoid", A, {}, B, bb, dom, c["repulsionPoints"], seed, prm, E=np.array(e["initial"]))
            assert np.linalg.norm(r["x"] - e["x"]) <= c["tolerance"]
            assert (r["iterations"], r["queries"]) == (e["iterations"], e["queries"])
            fin = np.array(e["final"])
            assert np.all(np.abs(r["E"] - fin) <= 1e-4 * np.maximum(1, np.abs(fin)))
        r = oracle.sdf_optimize("descent", A, {}, B, bb, dom, c["repulsionPoints"], c["seed"], prm)
        gd = c["gradientDescent"]
        assert np.linalg.norm(r["x"] - gd["x"]) <= c["tolerance"]
        assert (r["iterations"], r["queries"]) == (gd["iterations"], gd["queries"])


def test_hardmax_midpoint_subgradient():
    """tests/sdf-reference.test.ts:284-292: epsilon 0 gives the midpoint subgradient at a tie."""
    sph = AnalyticField({"kind": "sphere", "radius": 1}).describe()
    prm = {"epsilon": 0.0, "alpha": 0.0}
    dom = ((-5, -5, -5), (5, 5, 5))
    g = oracle.sdf_optimize("gradient", sph, {"position": (-.5, 0, 0)}, sph, {"position": (.5, 0, 0)}, dom, [], (0, 0, 0), prm)
    assert g["value"] == -.5 and np.array_equal(g["gradient"], [0, 0, 0])
    g = oracle.sdf_optimize("gradient", sph, {"position": (-.5, 0, 0)}, sph, {"position": (.5, 0, 0)}, dom, [], (.1, 0, 0), prm)
    assert np.array_equal(g["gradient"], [1, 0, 0])


def _field_desc(name):
    d = seeded_field(name).describe()
    return d


def _params(c):
    p = dict(c["params"])
    for k in ("epsilon", "alpha", "replacementThreshold", "tolerance", "stepSize", "margin", "allowedPenetration"):
        if k in p:
            p[k] = f32(p[k])
    return p


def _check_points(label, pts, want, tol):
    assert len(pts) == len(want), label
    for p, q in zip(pts, want):
        assert np.linalg.norm(p["x"] - np.array(q["x"])) <= tol, label
        assert (p["active"], p["cached"]) == (q["active"], q["cached"]), label
        assert np.all(np.abs(p["normal"] - q["normal"]) <= 2e-5 * np.maximum(1, np.abs(q["normal"]))), label
        for k in ("depth", "separation", "rank"):
            assert abs(p[k] - q[k]) <= 2e-5 * max(1, abs(q[k])), (label, k)


def test_manifold_updates_match_upstream_with_its_random_stream():
    cases = golden("manifolds.json")["cases"]
    assert len(cases) == 21
    ran = 0
    for c in cases:
        if c["shapeA"] not in SPECS or c["shapeB"] not in SPECS:
            continue  # voxel-seeds: the bunny's 15 MB grid is not carried here
        pts, stats, after = oracle.sdf_manifold_cpu(_field_desc(c["shapeA"]), body_desc(c["bodyA"]), _field_desc(c["shapeB"]),
                                                   body_desc(c["bodyB"]), _params(c), float(f32(c["dt"])), c["initialCache"],
                                                   int(c["rng"]["before"]))
        label = c["name"]
        assert after == int(c["rng"]["after"]), label
        _check_points(label, pts, c["points"], c["params"]["tolerance"])
        assert stats == {k: c["stats"][k] for k in ("pairs", "contacts", "queries", "iterations")}, label
        ran += 1
    assert ran == 20


def test_chained_ellipsoid_and_descent_manifolds_match_upstream():
    chain = None
    cases = golden("manifolds-optimizers.json")["cases"]
    assert {c["params"].get("optimizer") for c in cases} == {1, 2}
    for c in cases:
        if c["step"] == 0:
            cache, pcg = c["initialCache"], int(c["rng"]["before"])
        else:
            cache, pcg = chain
        assert pcg == int(c["rng"]["before"])
        pts, stats, after = oracle.sdf_manifold_cpu(_field_desc(c["shapeA"]), body_desc(c["bodyA"]), _field_desc(c["shapeB"]),
                                                   body_desc(c["bodyB"]), _params(c), float(f32(c["dt"])), cache, pcg)
        assert after == int(c["rng"]["after"])
        _check_points(f"{c['name']} {c['step']}", pts, c["points"], c["params"]["tolerance"])
        assert stats == {k: c["stats"][k] for k in ("pairs", "contacts", "queries", "iterations")}
        chain = (pts, after)


def test_disjoint_bounds_consume_no_randomness():
    """tests/sdf-reference.test.ts:193-198."""
    box = _field_desc("box")
    state = 0x853c49e6748fea9b
    pts, stats, after = oracle.sdf_manifold_cpu(box, {}, box, {"position": (10, 0, 0)}, {}, float(f32(1 / 60)), [], state)
    assert pts == [] and after == state and stats == {"pairs": 0, "contacts": 0, "queries": 0, "iterations": 0}


def test_signed_separation_detects_an_embedded_surface():
    """tests/sdf-reference.test.ts:233-248: where the sample is outside the sphere but inside a static
    box, the unsigned gap calls the pair separated; the signed projection onto the box's normal sees
    the overlap. Same optimiser sample either way."""
    sphere = AnalyticField({"kind": "sphere", "radius": .5})
    sphere.seeds = np.float32([[0, -.51, 0]])
    floor = AnalyticField({"kind": "box", "halfExtents": [4, .2, 4]})
    mixed = 0
    for h in np.linspace(.36, .48, 25):
        def run(signed):
            return oracle.sdf_manifold_cpu(sphere.describe(), {"position": (0, float(h), 0)}, floor.describe(),
                                           {"position": (0, -.2, 0), "static": True},
                                           {"signedSeparation": signed, "maxIterations": 1, "cacheSize": 0, "randomSamples": 1},
                                           0.0, [], 0x853c49e6748fea9b)[0]
        legacy, signed = run(False), run(True)
        if not legacy:
            continue
        legacy, signed = legacy[0], signed[0]
        assert np.array_equal(legacy["x"], signed["x"])
        x = legacy["x"].astype(np.float32)
        pa = oracle.sdf_sample(sphere.describe(), [x - np.float32([0, h, 0])])[0, 0]
        pb = oracle.sdf_sample(floor.describe(), [x - np.float32([0, -.2, 0])])[0, 0]
        if pa > 0 and pb < 0:
            mixed += 1
            assert legacy["separation"] > 0 and not legacy["active"]
            assert signed["separation"] < 0 and signed["active"]
            assert abs(signed["separation"] - (pa + pb)) < 1e-6  # dot(gap, +y) = phi_a + phi_b for these gradients
        elif pa <= 0 and pb <= 0:
            assert legacy["separation"] == signed["separation"]
    assert mixed > 0


def test_collision_params_defaults_and_names():
    p = collision_params()
    assert (p["optimizer"], p["cacheSize"], p["randomSamples"], p["maxIterations"]) == (0, 4, 1, 50)
    assert p["margin"] == float(f32(.02)) and p["allowedPenetration"] == float(f32(.01))
    with pytest.raises(ValueError):
        collision_params(bogus=1)
    for bad in ({"kind": "sphere", "radius": 0}, {"kind": "capsule", "radius": .1, "halfLength": -1},
                {"kind": "warp", "type": 3, "k": 1, "halfExtents": [1, 1, 1]}, {"kind": "nope"},
                {"kind": "spike-plate", "halfExtents": [1, 1, 1], "count": 0, "spikeRadius": .1, "minHeight": .1, "maxHeight": .2}):
        with pytest.raises(ValueError):
            AnalyticField(bad)
