"""The oracle's hull narrowphase (restating src/avbd3d/gpu/wgsl-hull.ts) pinned by brute force, on
the CPU, in f64 on the hull buffer's own f32 data:

- the face queries (queryFaces, wgsl-hull.ts:127-158) equal the largest separation of the other
  hull's vertices over each face plane;
- the SAT as a whole (faces of both, edge pairs through the Gauss-map test, :160-210) equals the
  separation over every candidate axis, every face normal and the cross product of every edge
  pair, with proper projections of both hulls: a wrong Gauss-map test or sign misses an axis;
- contacts exist exactly when that separation is not positive;
- an edge-edge contact (hullEdgeContact, :293-316) sits at the closest points of its two edges,
  computed here as a bounded least-squares problem;
- sphere against hull (sphereHull, :350-395) touches at the closest point of the polyhedron to
  the centre, whether that lies in a face, on an edge or at a vertex."""
import math

import numpy as np

import oracle
from avbd_amd import records, scenes
from avbd_amd.hull import HullBuffer, convex_hull


def qrotate(q, v):
    """v turned by q as the reference's qrotate (layout.ts PRELUDE_3D): v + 2w (u x v) + u x 2(u x v),
    exact for the stored (f32-rounded, not exactly unit) quaternions the oracle uses."""
    u, w = np.asarray(q[0:3], np.float64), float(q[3])
    v = np.asarray(v, np.float64)
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def to_world(c, q, r):
    """The world point whose body-local anchor the oracle stored as r = qrotate(conj(q), x - c):
    that map is linear in x - c, so it is inverted exactly as a 3x3 system."""
    qc = np.array([-q[0], -q[1], -q[2], q[3]], np.float64)
    M = np.column_stack([qrotate(qc, e) for e in np.eye(3)])
    return c + np.linalg.solve(M, np.asarray(r, np.float64))


def point_on_segment(a, b, p):
    t = min(1.0, max(0.0, ((p - a) @ (b - a)) / ((b - a) @ (b - a))))
    return a + t * (b - a)


class WorldHull:
    """A hull body's vertices, face planes and edges in the world frame, from the f32 values the
    hull buffer and the body record hold."""

    def __init__(self, rec, shape):
        self.c = rec[0:3].astype(np.float64)
        self.q = rec[4:8].astype(np.float64)
        self.v = np.array([qrotate(self.q, v) for v in shape["vertices"].astype(np.float32).astype(np.float64)]) + self.c
        self.n = np.array([qrotate(self.q, np.float32(f[0]).astype(np.float64)) for f in shape["faces"]])
        self.d = np.array([np.float64(np.float32(f[1])) for f in shape["faces"]]) + self.n @ self.c
        self.edges = [(int(e[0]), int(e[1])) for e in shape["edges"]]
        self.loops = [f[2] for f in shape["faces"]]


def brute_face_sep(A, B):
    """Over A's face planes, the largest of (B's lowest vertex along the normal - the plane)."""
    s = (B.v @ A.n.T).min(0) - A.d
    return s.max(), int(s.argmax()), np.sort(s)[-2] if len(s) > 1 else -np.inf


def brute_sat(A, B):
    """Separation over every candidate axis with proper projections of both polyhedra."""
    axes = list(A.n) + list(B.n)
    for a0, a1 in A.edges:
        ea = A.v[a1] - A.v[a0]
        for b0, b1 in B.edges:
            eb = B.v[b1] - B.v[b0]
            n = np.cross(ea, eb)
            l = np.linalg.norm(n)
            if l < 1e-5 * np.linalg.norm(ea) * np.linalg.norm(eb):
                continue
            axes.append(n / l)
    best = -np.inf
    for n in axes:
        pa, pb = A.v @ n, B.v @ n
        best = max(best, pb.min() - pa.max(), pa.min() - pb.max())
    return best


def segment_closest(p0, p1, q0, q1):
    """Closest points of segments [p0, p1] and [q0, q1]: the least-squares problem over
    (s, t) in [0, 1]^2, its minimum at the free optimum or on one of the four sides."""
    d1, d2, r = p1 - p0, q1 - q0, p0 - q0
    cands = []
    M = np.array([[d1 @ d1, -d1 @ d2], [-d1 @ d2, d2 @ d2]])
    if abs(np.linalg.det(M)) > 1e-14:
        s, t = np.linalg.solve(M, [-(d1 @ r), d2 @ r])
        if 0 <= s <= 1 and 0 <= t <= 1:
            cands.append((s, t))
    for s in (0.0, 1.0):
        cands.append((s, min(1.0, max(0.0, (d2 @ (r + s * d1)) / (d2 @ d2)))))
    for t in (0.0, 1.0):
        cands.append((min(1.0, max(0.0, -(d1 @ (r - t * d2)) / (d1 @ d1))), t))
    s, t = min(cands, key=lambda st: np.linalg.norm(p0 + st[0] * d1 - q0 - st[1] * d2))
    return p0 + s * d1, q0 + t * d2


def closest_on_polyhedron(H, p):
    """The closest point to an outside point p: over every face polygon, the projection when it
    falls inside the polygon, else the closest point of the polygon's edges."""
    best, bd = None, np.inf
    for n, d, loop in zip(H.n, H.d, H.loops):
        q = p - n * (n @ p - d)
        inside = True
        for j in range(len(loop)):
            a, b = H.v[loop[j]], H.v[loop[(j + 1) % len(loop)]]
            if np.cross(b - a, q - a) @ n < -1e-12:
                inside = False
        if not inside:
            cands = [point_on_segment(H.v[loop[j]], H.v[loop[(j + 1) % len(loop)]], p) for j in range(len(loop))]
            q = min(cands, key=lambda x: np.linalg.norm(x - p))
        dist = np.linalg.norm(q - p)
        if dist < bd:
            best, bd = q, dist
    return best, bd


def rand_quat(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def shapes(rng):
    """Hull shapes: points on ellipsoids, tetrahedra, wedges and box corners."""
    out = []
    for k in range(12):
        kind = k % 4
        if kind == 0:
            g = rng.normal(size=(int(rng.integers(16, 33)), 3))
            pts = g / np.linalg.norm(g, axis=1, keepdims=True) * rng.uniform(0.3, 1.2, 3)
        elif kind == 1:
            pts = rng.uniform(-1, 1, (4, 3))
        elif kind == 2:
            pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 0.4], [1, 0, 0.4], [0, 1, 0.4]]) * rng.uniform(0.5, 1.5)
        else:
            h = rng.uniform(0.2, 1.0, 3)
            pts = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)]) * h
        s = convex_hull(pts)
        if s is not None:
            out.append(s)
    return out


def pair_model(shape_a, shape_b, pos_a, pos_b, q_a, q_b, face_bias=True):
    hb = HullBuffer()
    at_a, at_b = hb.add(shape_a), hb.add(shape_b)
    ra = scenes.hull_record(pos_a, shape_a, at_a, 1.0, rot=tuple(q_a)).astype(np.float32)
    rb = scenes.hull_record(pos_b, shape_b, at_b, 1.0, rot=tuple(q_b)).astype(np.float32)
    m = oracle.Model(64)
    m.set_params(oracle.make_params(gravity=0.0, flags=oracle.FLAG_FACE_BIAS if face_bias else 0, iterations=0))
    m.set_bodies(records.bodies_to_oracle(np.vstack([rb, ra])))  # a = 1 (higher index), b = 0
    m.set_hulls(hb.words())
    return m, ra, rb


def test_hull_sat_queries_equal_the_brute_force():
    rng = np.random.default_rng(2026)
    sh = shapes(rng)
    checked = edge_axes = separated = 0
    for trial in range(240):
        A, B = sh[rng.integers(len(sh))], sh[rng.integers(len(sh))]
        qa, qb = rand_quat(rng), rand_quat(rng)
        dirn = rng.normal(size=3)
        dirn /= np.linalg.norm(dirn)
        dist = rng.uniform(0.2, 1.1) * (A["radius"] + B["radius"])
        m, ra, rb = pair_model(A, B, (0.0, 0.0, 0.0), tuple(dirn * dist), qa, qb)
        WA, WB = WorldHull(ra, A), WorldHull(rb, B)
        q = m.hull_query(1, 0)
        fa, fa_face, fa_second = brute_face_sep(WA, WB)
        fb, fb_face, fb_second = brute_face_sep(WB, WA)
        assert abs(q["fa_sep"] - fa) < 1e-9 and abs(q["fb_sep"] - fb) < 1e-9, trial
        if fa - fa_second > 1e-9:
            assert q["fa_face"] == fa_face, trial
        if fb - fb_second > 1e-9:
            assert q["fb_face"] == fb_face, trial
        sat = brute_sat(WA, WB)
        got = max(q["fa_sep"], q["fb_sep"], q["edge_sep"] if q["edge_valid"] else -np.inf)
        if q["fa_sep"] > 0:  # the reference stops at the first separating query
            assert sat > 0, trial
            separated += 1
            continue
        if q["fb_sep"] > 0:
            assert sat > 0, trial
            separated += 1
            continue
        assert abs(got - sat) < 1e-6, (trial, got, sat)
        edge_axes += q["edge_valid"] and q["edge_sep"] > max(q["fa_sep"], q["fb_sep"])
        checked += 1
        # contacts exist exactly when the separation is not positive
        m.collide()
        man, _ = m.contacts()
        if abs(sat) > 1e-4:
            assert (len(man) > 0) == (sat < 0), (trial, sat, len(man))
    assert checked > 60 and separated > 20 and edge_axes > 10, (checked, separated, edge_axes)


def test_edge_contacts_sit_at_the_closest_points_of_their_edges():
    rng = np.random.default_rng(7)
    sh = [s for s in shapes(rng) if len(s["vertices"]) <= 8]  # tetrahedra, wedges, boxes
    found = unique = 0
    for trial in range(600):
        A, B = sh[rng.integers(len(sh))], sh[rng.integers(len(sh))]
        qa, qb = rand_quat(rng), rand_quat(rng)
        dirn = rng.normal(size=3)
        dirn /= np.linalg.norm(dirn)
        m, ra, rb = pair_model(A, B, (0.0, 0.0, 0.0), tuple(dirn * rng.uniform(0.3, 0.9) * (A["radius"] + B["radius"])), qa, qb)
        q = m.hull_query(1, 0)
        m.collide()
        man, con = m.contacts()
        if len(man) != 1 or (int(con[0, 3]) >> 24) & 0x7F != 2:
            continue
        WA, WB = WorldHull(ra, A), WorldHull(rb, B)
        a0, a1 = WA.edges[q["ea"]]
        b0, b1 = WB.edges[q["eb"]]
        pa, pb = segment_closest(WA.v[a0], WA.v[a1], WB.v[b0], WB.v[b1])
        xa = to_world(WA.c, WA.q, con[0, 0:3])
        xb = to_world(WB.c, WB.q, con[0, 4:7])
        # closest points: each on its edge, as far apart as the edges' minimum distance (when the
        # edges are nearly parallel the pair is not unique: its distance is)
        on_a = point_on_segment(WA.v[a0], WA.v[a1], xa)
        on_b = point_on_segment(WB.v[b0], WB.v[b1], xb)
        assert np.linalg.norm(on_a - xa) < 1e-9 and np.linalg.norm(on_b - xb) < 1e-9, trial
        assert abs(np.linalg.norm(xa - xb) - np.linalg.norm(pa - pb)) < 1e-9, trial
        if np.abs(xa - pa).max() < 1e-9 and np.abs(xb - pb).max() < 1e-9:
            unique += 1
        found += 1
    assert found >= 10 and unique >= found // 2, (found, unique)


def test_sphere_touches_a_hull_at_its_closest_point():
    rng = np.random.default_rng(11)
    sh = shapes(rng)
    regions = {"face": 0, "edge": 0, "vertex": 0}
    for trial in range(300):
        H = sh[rng.integers(len(sh))]
        qh = rand_quat(rng)
        hb = HullBuffer()
        at = hb.add(H)
        rh = scenes.hull_record((0.0, 0.0, 0.0), H, at, 1.0, rot=tuple(qh)).astype(np.float32)
        WH = WorldHull(rh, H)
        # a centre just outside: a surface point pushed out along a random direction
        dirn = rng.normal(size=3)
        dirn /= np.linalg.norm(dirn)
        far = dirn * (H["radius"] + 2.0)
        p0, _ = closest_on_polyhedron(WH, far)
        r = rng.uniform(0.1, 0.4)
        c = p0 + (far - p0) / np.linalg.norm(far - p0) * rng.uniform(0.3, 1.5) * r
        rs = records.sphere_record(tuple(c), r, 1.0).astype(np.float32)
        m = oracle.Model(64)
        m.set_params(oracle.make_params(gravity=0.0, flags=0, iterations=0))
        m.set_bodies(records.bodies_to_oracle(np.vstack([rh, rs])))
        m.set_hulls(hb.words())
        m.collide()
        man, con = m.contacts()
        cs = rs[0:3].astype(np.float64)
        best, dist = closest_on_polyhedron(WH, cs)
        if abs(dist - r) < 1e-6:
            continue
        assert (len(man) == 1) == (dist < r), (trial, dist, r)
        if dist >= r:
            continue
        # a = the sphere (index 1), b = the hull (index 0): the hull's point is rB
        on_hull = to_world(WH.c, WH.q, con[0, 4:7])
        assert np.abs(on_hull - best).max() < 1e-9, trial
        on_sphere = to_world(cs, rs[4:8].astype(np.float64), con[0, 0:3])
        assert abs(np.linalg.norm(on_sphere - cs) - np.float32(r)) < 1e-6
        assert np.abs(np.cross(on_sphere - cs, best - cs)).max() < 1e-6  # along the normal
        hits = [np.linalg.norm(best - WH.v[k]) < 1e-9 for k in range(len(WH.v))]
        regions["vertex" if any(hits) else ("face" if any(abs(WH.n @ best - WH.d) < 1e-9) and
                                            sum(abs(WH.n @ best - WH.d) < 1e-9) == 1 else "edge")] += 1
    assert all(v >= 5 for v in regions.values()), regions
