"""Scene planning before a solver exists (avbd_morton_order, avbd_estimate_pairs,
avbd_plan_capacity; the reference's mortonOrder, estimatePairs and constructor sizing,
src/avbd3d/gpu/solver.ts:350-485, :745, :815-843). Host code in the library, so these run on the
CPU. Pinned by known answers, a bit-by-bit Z-order interleave, brute force over all pairs and
the true distance between the shapes (a quadratic program), not by restating the formulas."""
import numpy as np
import pytest
from scipy.optimize import minimize

import avbd_amd
from avbd_amd import records


def quat_z(angle):
    return (0.0, 0.0, np.sin(angle / 2), np.cos(angle / 2))


def rand_quat(rng):
    q = rng.normal(size=4)
    return tuple(q / np.linalg.norm(q))


def rot_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


# --- mortonOrder -------------------------------------------------------------------------------
def test_morton_order_of_cube_corners_counts_in_binary():
    corners = [(x, y, z) for z in (0, 1) for y in (0, 1) for x in (0, 1)]  # binary counting, x fastest
    perm = np.array([5, 2, 7, 0, 3, 6, 1, 4])
    recs = np.array([records.box_record(corners[k], (0.1, 0.1, 0.1), 1.0) for k in perm])
    order = avbd_amd.morton_order(recs)
    assert [corners[perm[g]] for g in order] == corners
    # equal positions keep their input order
    same = np.array([records.box_record((0.5, 0.5, 0.5), (1, 1, 1), 1.0)] * 3 + [records.box_record((0, 0, 0), (1, 1, 1), 1.0)])
    assert avbd_amd.morton_order(same).tolist() == [3, 0, 1, 2]


def test_morton_order_matches_a_bitwise_interleave():
    rng = np.random.default_rng(3)
    pos = rng.uniform(-50, 80, (500, 3))
    pos[:40] = pos[40:80]  # ties
    recs = np.array([records.box_record(p, (1, 1, 1), 1.0) for p in pos]).astype(np.float32)
    p32 = recs[:, 0:3].astype(np.float64)
    lo, hi = p32.min(0), p32.max(0)
    q = np.minimum(1023, np.floor((p32 - lo) / np.maximum(hi - lo, 1e-9) * 1024)).astype(np.int64)
    code = np.zeros(len(q), np.int64)
    for bit in range(10):  # bit b of x, y, z -> bits 3b, 3b + 1, 3b + 2
        for axis in range(3):
            code |= ((q[:, axis] >> bit) & 1) << (3 * bit + axis)
    want = np.lexsort((np.arange(len(q)), code))
    assert np.array_equal(avbd_amd.morton_order(recs), want)


# --- estimatePairs -----------------------------------------------------------------------------
def test_estimate_pairs_known_answers():
    box = records.box_record
    sph = records.sphere_record
    cases = [
        ("stacked boxes touching", [box((0, 0, 0), (1, 1, 1), 1), box((0, 0, 1.0), (1, 1, 1), 1)], 1, 1),
        ("both static", [box((0, 0, 0), (1, 1, 1), 0), box((0, 0, 1.0), (1, 1, 1), 0)], 0, 0),
        ("axis-aligned gap: AABBs apart", [box((0, 0, 0), (1, 1, 1), 1), box((1.0005, 0, 0), (1, 1, 1), 1)], 0, 0),
        ("rotated, AABBs overlap, separated", [box((0, 0, 0), (1, 1, 1), 1), box((1.1, 1.1, 0), (1, 1, 1), 1, rot=quat_z(np.pi / 4))], 1, 0),
        ("sphere on a box", [box((0, 0, 0), (1, 1, 1), 0), sph((0.1, -0.2, 1.0), 0.5, 1)], 1, 1),
        ("sphere off a box edge", [box((0, 0, 0), (1, 1, 1), 1), sph((0.86, 0.86, 0), 0.5, 1)], 1, 0),
        ("sphere 5e-4 above a box: AABBs apart, so no pair", [box((0, 0, 0), (1, 1, 1), 1), sph((0, 0, 1.0005), 0.5, 1)], 0, 0),
        ("sphere 5e-4 off a tilted box: within the tolerance", [box((0, 0, 0), (1, 1, 1), 1, rot=quat_z(0.3)),
                                                               sph((0.5 * np.cos(0.3) + 0.5005, 0.5 * np.sin(0.3), 0), 0.5, 1)], 1, 1),
        ("spheres overlapping", [sph((0, 0, 0), 0.5, 1), sph((0, 0.5, 0.75), 0.5, 1)], 1, 1),
        ("spheres 2e-3 apart: bounding spheres apart", [sph((0, 0, 0), 0.5, 1), sph((0, 0.6016, 0.8012), 0.5, 1)], 0, 0),
    ]
    for name, bodies, pairs, touching in cases:
        got = avbd_amd.estimate_pairs(np.array(bodies))
        assert (got[0], got[1]) == (pairs, touching), name
        assert got[2].tolist() == [touching, touching], name


def box_halfspaces(rec):
    """A x <= b for a box record (local axes, half extents)."""
    R = rot_matrix(rec[4:8])
    h = np.asarray(rec[16:19]) * 0.5
    A = np.vstack([R.T, -R.T])
    b = np.concatenate([h + R.T @ rec[0:3], h - R.T @ rec[0:3]])
    return A, b


def distance(a, b):
    """True distance between two shapes (sphere: centre and radius; box: a quadratic program over a
    point in each)."""
    sa, sb = a[39] == 1, b[39] == 1
    if sa and sb:
        return max(np.linalg.norm(a[0:3] - b[0:3]) - a[23] - b[23], 0.0)
    if sa or sb:
        s, bx = (a, b) if sa else (b, a)
        R = rot_matrix(bx[4:8])
        local = R.T @ (s[0:3] - bx[0:3])
        h = np.asarray(bx[16:19]) * 0.5
        return max(np.linalg.norm(local - np.clip(local, -h, h)) - s[23], 0.0)
    Aa, ba = box_halfspaces(a)
    Ab, bb = box_halfspaces(b)
    cons = [{"type": "ineq", "fun": lambda z, A=Aa, c=ba: c - A @ z[:3], "jac": lambda z, A=Aa: np.hstack([-A, np.zeros_like(A)])},
            {"type": "ineq", "fun": lambda z, A=Ab, c=bb: c - A @ z[3:], "jac": lambda z, A=Ab: np.hstack([np.zeros_like(A), -A])}]
    z0 = np.concatenate([a[0:3], b[0:3]])
    r = minimize(lambda z: np.sum((z[:3] - z[3:]) ** 2), z0, jac=lambda z: np.concatenate([2 * (z[:3] - z[3:]), -2 * (z[:3] - z[3:])]),
                 constraints=cons, method="SLSQP", options={"ftol": 1e-14, "maxiter": 500})
    return float(np.sqrt(max(r.fun, 0.0)))


def aabb_pair(a, b):
    """The broadphase predicate (bounding spheres and world AABBs overlap, not both static)."""
    if a[19] <= 0 and b[19] <= 0:
        return False
    d = a[0:3] - b[0:3]
    if d @ d > (a[23] + b[23]) ** 2:
        return False

    def half(r):
        if r[39] == 1:
            return np.full(3, r[23])
        return np.abs(rot_matrix(r[4:8])) @ (np.asarray(r[16:19]) * 0.5)
    return bool((np.abs(d) <= half(a) + half(b)).all())


def test_estimate_pairs_match_brute_force_and_true_distances():
    """Random boxes and spheres, placed so that every broadphase pair is either interpenetrating or
    at least 0.05 apart: the pair count equals the brute force over all pairs, and a pair touches
    exactly when the shapes intersect."""
    rng = np.random.default_rng(11)
    # away from the origin: the grid's xor hash maps some cells near it onto one bucket, whose
    # pairs the estimate then counts twice (as the reference's does)
    at = np.array([1000.0, 2000.0, 3000.0])
    bodies = [records.box_record(at + (3.5, 3.5, 2.0), (6.0, 5.0, 0.6), 0.0, rot=quat_z(0.2)),  # large: tested against all
              records.box_record(at + (2.0, 5.0, 5.0), (5.0, 0.8, 4.0), 1.0, rot=quat_z(-0.4))]
    bodies = [b.astype(np.float32).astype(np.float64) for b in bodies]
    while len(bodies) < 70:
        p = rng.uniform(0, 7, 3) + at
        if rng.random() < 0.3:
            rec = records.sphere_record(p, rng.uniform(0.3, 0.7), 1.0)
        else:
            rec = records.box_record(p, rng.uniform(0.4, 1.5, 3), 0.0 if rng.random() < 0.2 else 1.0, rot=rand_quat(rng))
        rec = rec.astype(np.float32).astype(np.float64)
        ok = True
        for other in bodies:
            if aabb_pair(rec, other) or aabb_pair(other, rec):
                dist = distance(rec, other)
                if 1e-7 < dist < 0.05:
                    ok = False
                    break
        if ok:
            bodies.append(rec)
    recs = np.array(bodies)
    n = len(recs)
    pairs, touching, degree = [], 0, np.zeros(n, int)
    for i in range(n):
        for j in range(i):
            if aabb_pair(recs[i], recs[j]):
                pairs.append((i, j))
                if distance(recs[i], recs[j]) <= 1e-7:
                    touching += 1
                    degree[i] += 1
                    degree[j] += 1
    got = avbd_amd.estimate_pairs(recs)
    assert len(pairs) > 20 and touching > 5
    assert got[0] == len(pairs)
    assert got[1] == touching
    assert np.array_equal(got[2], degree)


# --- the constructor's sizing -----------------------------------------------------------------
def test_plan_capacity_known_answers():
    # 100 bodies far apart: the floors
    recs = np.array([records.box_record((10.0 * k, 0, 0), (1, 1, 1), 1.0) for k in range(100)])
    cap, colors = avbd_amd.plan_capacity(recs)
    assert cap == {"bodies": 1124, "joints": 256, "pairs": 4496, "manifolds": 4096, "contacts": 8192, "hull_words": 0}
    assert colors == 8
    # a body jointed to ten others: max degree 10 -> 10 + 1 + 3 colours
    info = np.array([[0, 0, k, 0] for k in range(1, 11)], np.int32)
    _, colors = avbd_amd.plan_capacity(recs, info)
    assert colors == 14
    # a world joint (a = -1) counts for its body only
    _, colors = avbd_amd.plan_capacity(recs, np.array([[0, -1, 5, 0]] * 6, np.int32))
    assert colors == 10
    # caller's least capacities and colours
    cap, colors = avbd_amd.plan_capacity(recs, least={"bodies": 50, "pairs": 10000, "manifolds": 20000, "contacts": 9000, "joints": 300},
                                         colors=70)
    assert cap["bodies"] == 100 and cap["pairs"] == 10000 and cap["manifolds"] == 10000  # never past the pairs
    assert cap["contacts"] == 9000 and cap["joints"] == 300 and colors == 64
    # a touching column of 2000 boxes: 8 contacts per touching pair, 1.6 manifolds
    col = np.array([records.box_record((0, 0, 1.0 * k), (1, 1, 1), 1.0) for k in range(2000)])
    p, t, _ = avbd_amd.estimate_pairs(col)
    assert (p, t) == (1999, 1999)
    cap, colors = avbd_amd.plan_capacity(col)
    assert cap["contacts"] == max(8192, 4 * 3024, 8 * 1999) and cap["manifolds"] == max(4096, 3024, int(np.ceil(1.6 * 1999)))
    assert cap["pairs"] == max(4096, 4 * 3024, 2 * 1999) and colors == 8


def test_planning_refuses_null_arguments():
    L = avbd_amd.lib()
    assert L.avbd_morton_order(None, 3, None) != 0
    assert L.avbd_estimate_pairs(None, 3, None, None, None) != 0
    assert L.avbd_plan_capacity(None, 0, None, 0, None, 0, None, None) != 0
    assert L.avbd_morton_order(None, 0, None) == 0
