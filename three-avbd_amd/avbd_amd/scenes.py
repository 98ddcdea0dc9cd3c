"""Synthetic scenes of the measured workloads, as body/joint records (layout.ts formats).

Deterministic from a seed with the reference bench's generator (src/avbd3d/bench-scenes.ts:24-27:
s = s*1103515245 + 12345 mod 2^32). z is up, as in the reference's scenes.
"""
import math

import numpy as np

from .records import box_record, sphere_record

T_JOINT = 1


class Lcg:
    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFF

    def __call__(self):
        self.s = (self.s * 1103515245 + 12345) & 0xFFFFFFFF
        return self.s / 4294967296.0

    def uniform(self, lo, hi):
        return lo + (hi - lo) * self()

    def gauss(self):
        u1 = max(self(), 1e-12)
        u2 = self()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def quat(self):
        """Uniformly random rotation: a normalised Gaussian 4-vector."""
        while True:
            q = [self.gauss() for _ in range(4)]
            n = math.sqrt(sum(x * x for x in q))
            if n > 1e-6:
                return [x / n for x in q]


def ground(z=-0.5, size=(100.0, 100.0, 1.0), friction=0.5):
    """The reference's ground box (ref/scenes.ts:8-10) with its top at z + 0.5."""
    return box_record((0.0, 0.0, z), size, 0.0, friction)


def pyramid(rows=10, friction=0.5):
    """Box pyramid: a static ground plane and `rows` rows of unit boxes (density 1), row y holding
    rows - y boxes (the reference's scenePyramid layout, ref/scenes.ts:50-58, with unit boxes)."""
    recs = [ground(-0.5, friction=friction)]
    for y in range(rows):
        for x in range(rows - y):
            recs.append(box_record((x * 1.01 + y * 0.5 - rows / 2.0, 0.0, y * 1.0 + 0.5), (1.0, 1.0, 1.0), 1.0, friction))
    return np.array(recs, np.float32), np.zeros((0, 40), np.float32), np.zeros((0, 4), np.int32)


def walled_bin(size=40.0, height=20.0, friction=0.6):
    half = size / 2.0
    return [
        ground(-0.5, (size + 2.0, size + 2.0, 1.0), friction),
        box_record((half + 0.5, 0.0, height / 2), (1.0, size + 2.0, height), 0.0, friction),
        box_record((-half - 0.5, 0.0, height / 2), (1.0, size + 2.0, height), 0.0, friction),
        box_record((0.0, half + 0.5, height / 2), (size + 2.0, 1.0, height), 0.0, friction),
        box_record((0.0, -half - 0.5, height / 2), (size + 2.0, 1.0, height), 0.0, friction),
    ]


def random_pile(n=10000, seed=12345, bin_size=40.0, pitch=2.5, jitter=0.1, friction=0.6):
    """Random pile: n bodies, half boxes with half-extents U[0.25, 1] m and half spheres with
    radius U[0.2, 0.6] m, density U[0.5, 2], on a jittered 3D grid (pitch 2.5 m, jitter 0.1 m)
    with uniformly random orientations, dropped into a walled bin; friction 0.6."""
    rng = Lcg(seed)
    recs = walled_bin(bin_size, friction=friction)
    per_row = max(1, int(bin_size // pitch))
    origin = -bin_size / 2.0 + pitch / 2.0
    for k in range(n):
        ix, iy, iz = k % per_row, (k // per_row) % per_row, k // (per_row * per_row)
        pos = (origin + ix * pitch + rng.uniform(-jitter, jitter), origin + iy * pitch + rng.uniform(-jitter, jitter),
               pitch / 2.0 + iz * pitch + rng.uniform(-jitter, jitter))
        density = rng.uniform(0.5, 2.0)
        q = rng.quat()
        if k % 2 == 0:
            he = [rng.uniform(0.25, 1.0) for _ in range(3)]
            recs.append(box_record(pos, [2 * h for h in he], density, friction, rot=q))
        else:
            recs.append(sphere_record(pos, rng.uniform(0.2, 0.6), density, friction, rot=q))
    return np.array(recs, np.float32), np.zeros((0, 40), np.float32), np.zeros((0, 4), np.int32)


def joint_record(rA, rB, stiff_lin=3.0e38, stiff_ang=0.0, fracture=3.0e38, torque_arm=0.0):
    """A joint record as GpuSolver3D.writeJoint packs it (solver.ts:1222-1246), before any step."""
    r = np.zeros(40, np.float32)
    r[3] = stiff_lin
    r[7] = stiff_ang
    r[11] = fracture
    r[15] = torque_arm
    r[23] = 3.0e38
    r[24:27] = rA
    r[28:31] = rB
    r[31] = 3.0e38
    r[35] = 1.0
    r[39] = 1.0
    return r


def hull_record(pos, shape, offset, density, friction=0.5, rot=(0, 0, 0, 1), vel=(0, 0, 0), angvel=(0, 0, 0)):
    """A hull body (shapes.ts:50-70): its principal frame at pos, size = the vertices' bounds,
    mass = volume x density, principal moments, bounding radius; shape code SHAPE_HULL + offset."""
    r = box_record(pos, shape["size"], density, friction, rot, vel, angvel)
    r[19] = shape["volume"] * density
    r[20:23] = np.asarray(shape["moments"]) * density
    r[23] = shape["radius"]
    r[39] = 3.0 + offset
    return r


def ellipsoid_points(rng, lo=0.3, hi=1.2, points=(16, 32)):
    """16-32 points uniform on an ellipsoid with axes U[lo, hi] m."""
    axes = [rng.uniform(lo, hi) for _ in range(3)]
    k = min(points[0] + int(rng() * (points[1] - points[0] + 1)), points[1])
    pts = []
    for _ in range(k):
        g = [rng.gauss() for _ in range(3)]
        nrm = math.sqrt(sum(x * x for x in g)) or 1.0
        pts.append([axes[c] * g[c] / nrm for c in range(3)])
    return pts


def ellipsoid_hull(rng, lo=0.3, hi=1.2, points=(16, 32)):
    """The convex hull of ellipsoid_points (drawn again when they span no volume)."""
    from .hull import convex_hull
    while True:
        shape = convex_hull(ellipsoid_points(rng, lo, hi, points))
        if shape is not None:
            return shape


def _hull_or_none(pts):
    from .hull import convex_hull
    return convex_hull(pts)


def build_hulls(point_sets, workers=8, context="fork"):
    """convex_hull of each point set; a pool of workers for many. Forked workers start fastest but
    need a process that has not initialised the GPU yet; "spawn" is safe after it."""
    if len(point_sets) < 2000 or workers <= 1:
        return [_hull_or_none(p) for p in point_sets]
    import multiprocessing as mp
    with mp.get_context(context).Pool(workers) as pool:
        return pool.map(_hull_or_none, point_sets, chunksize=256)


def large_scene(n=100000, seed=2026, chains=200, links=20, pitch=2.5, friction=0.5, workers=8, context="fork"):
    """Large scene: n bodies, boxes (half-extents U[0.3, 1.2] m), spheres (radius U[0.3, 1.2] m)
    and convex hulls (16-32 points on ellipsoids with axes U[0.3, 1.2] m) 1:1:1, density 1, on a
    jittered grid over a square walled bin; plus `chains` chains of `links` unit boxes joined by
    ball joints (linear rows hard, angular free) hanging from static anchors beside the bin.
    Returns (bodies, joints, joint_info, hull_words)."""
    from .hull import HullBuffer
    rng = Lcg(seed)
    per_layer = max(1, int(math.sqrt(n / 16.0)))
    bin_size = per_layer * pitch
    recs = walled_bin(bin_size, height=30.0, friction=friction)
    # one ground under the bin and the chains beside it
    recs[0] = ground(-0.5, (bin_size + 80.0, bin_size + 80.0, 1.0), friction)
    hulls = HullBuffer()
    origin = -bin_size / 2.0 + pitch / 2.0
    draws = []
    for k in range(n):
        ix, iy, iz = k % per_layer, (k // per_layer) % per_layer, k // (per_layer * per_layer)
        pos = (origin + ix * pitch + rng.uniform(-0.1, 0.1), origin + iy * pitch + rng.uniform(-0.1, 0.1),
               pitch / 2.0 + iz * pitch + rng.uniform(-0.1, 0.1))
        q = rng.quat()
        kind = k % 3
        if kind == 0:
            draws.append((0, pos, q, [2 * rng.uniform(0.3, 1.2) for _ in range(3)]))
        elif kind == 1:
            draws.append((1, pos, q, rng.uniform(0.3, 1.2)))
        else:
            draws.append((2, pos, q, ellipsoid_points(rng)))
    shapes = iter(build_hulls([d[3] for d in draws if d[0] == 2], workers, context))
    for kind, pos, q, arg in draws:
        if kind == 0:
            recs.append(box_record(pos, arg, 1.0, friction, rot=q))
        elif kind == 1:
            recs.append(sphere_record(pos, arg, 1.0, friction, rot=q))
        else:
            shape = next(shapes)
            if shape is None:  # degenerate draw (no volume): a box of the same extents instead
                ext = np.ptp(np.asarray(arg), axis=0)
                recs.append(box_record(pos, np.maximum(ext, 0.1), 1.0, friction, rot=q))
                continue
            at = hulls.add(shape)
            recs.append(hull_record(pos, shape, at, 1.0, friction, rot=q))
    joints, info = [], []
    side = bin_size / 2.0 + 10.0
    top = links + 3.0
    for c in range(chains):
        x = side + 4.0 * (c % 8)
        y = -bin_size / 2.0 + (c // 8) * (bin_size / max(1, (chains + 7) // 8))
        anchor = len(recs)
        recs.append(box_record((x, y, top), (1.0, 1.0, 1.0), 0.0, friction))
        prev = anchor
        for l in range(links):
            cur = len(recs)
            recs.append(box_record((x, y, top - 1.0 - l), (1.0, 1.0, 1.0), 1.0, friction))
            arm = (1.0 + 1.0) ** 2 * 3  # torqueArm = |size_a + size_b|^2 (forces.ts:146)
            joints.append(joint_record((0.0, 0.0, -0.5), (0.0, 0.0, 0.5), 3.0e38, 0.0, 3.0e38, arm))
            info.append((T_JOINT, prev, cur, 0))
            prev = cur
    return (np.array(recs, np.float32), np.array(joints, np.float32).reshape(-1, 40), np.array(info, np.int32).reshape(-1, 4),
            hulls.words())


def batched_worlds(worlds=4096, lo=8, hi=512, seed_xor=0x9E3779B9):
    """Batched: `worlds` independent worlds, world w seeded with w ^ 0x9e3779b9. Each holds a
    static ground and N_w = round(exp(U[ln lo, ln hi])) bodies, half boxes (half-extents U[0.2, 0.5]
    m) and half spheres (radius U[0.2, 0.5] m), densities log-uniform in [1, 100] (mass ratios up
    to 100:1), friction U[0.2, 1.0] for the whole world, arranged as a pyramid, a random pile or
    stacks (chosen uniformly). Returns (bodies, world_start) with world w's bodies at
    [world_start[w], world_start[w + 1]), its ground first."""
    recs = []
    starts = [0]
    for w in range(worlds):
        rng = Lcg((w ^ seed_xor) & 0xFFFFFFFF)
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        mu = rng.uniform(0.2, 1.0)
        kind = min(2, int(rng() * 3))
        recs.append(ground(-0.5, (60.0, 60.0, 1.0), mu))
        if kind == 0:  # pyramid: rows of unit-pitch slots, bottom row wide enough for n
            rows = 1
            while rows * (rows + 1) // 2 < n:
                rows += 1
            slots = [(x * 1.05 + y * 0.525 - rows / 2.0, 0.0, y * 1.0 + 0.5) for y in range(rows) for x in range(rows - y)][:n]
        elif kind == 1:  # pile: jittered grid, 1.3 m pitch
            side = max(1, int(math.ceil(n ** (1.0 / 3.0))))
            slots = [((k % side) * 1.3 - side * 0.65 + rng.uniform(-0.1, 0.1), ((k // side) % side) * 1.3 - side * 0.65 + rng.uniform(-0.1, 0.1),
                      0.65 + (k // (side * side)) * 1.3 + rng.uniform(-0.1, 0.1)) for k in range(n)]
        else:  # stacks: columns of at most 16
            cols = max(1, int(math.ceil(n / 16.0)))
            slots = [((k % cols) * 1.2 - cols * 0.6, 0.0, 0.5 + (k // cols) * 1.02) for k in range(n)]
        for k, pos in enumerate(slots):
            density = math.exp(rng.uniform(0.0, math.log(100.0)))
            if (k + w) % 2 == 0:
                he = [rng.uniform(0.2, 0.5) for _ in range(3)] if kind != 0 else [0.5, 0.5, 0.5]
                recs.append(box_record(pos, [2 * h for h in he], density, mu))
            else:
                recs.append(sphere_record(pos, rng.uniform(0.2, 0.5) if kind != 0 else 0.5, density, mu))
        starts.append(len(recs))
    return np.array(recs, np.float32), np.array(starts, np.uint32)


def thrown_at_stack(worlds=256, seed=5, stack=4, thrown=64, steps_hint=120):
    """Boxes thrown at a box stack, one copy per world: a stack of stack^3 boxes (half-extent
    0.5 m, density 1, 0.01 m gaps) on a static ground, and `thrown` boxes of the same size on a
    lattice above it (pitch 1.6 m from z = 6 m) with linear velocity U[-5, 5]^3 m/s and angular
    velocity U[-2, 2]^3 rad/s; each world's friction (all its bodies) is U[0.3, 0.8].
    Returns (bodies, world_start, friction per world)."""
    rng = Lcg(seed)
    recs, starts, mus = [], [], []
    side = int(round(thrown ** (1.0 / 3.0)))
    for w in range(worlds):
        starts.append(len(recs))
        mu = rng.uniform(0.3, 0.8)
        mus.append(mu)
        recs.append(ground(-0.5, (40.0, 40.0, 1.0), mu))
        off = (stack - 1) / 2.0
        for k in range(stack):
            for j in range(stack):
                for i in range(stack):
                    recs.append(box_record(((i - off) * 1.01, (j - off) * 1.01, 0.5 + k * 1.0), (1.0, 1.0, 1.0), 1.0, mu))
        toff = (side - 1) / 2.0
        for t in range(thrown):
            i, j, k = t % side, (t // side) % side, t // (side * side)
            v = (rng.uniform(-5, 5), rng.uniform(-5, 5), rng.uniform(-5, 5))
            om = (rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(-2, 2))
            recs.append(box_record(((i - toff) * 1.6, (j - toff) * 1.6, 6.0 + k * 1.6), (1.0, 1.0, 1.0), 1.0, mu,
                                   rng.quat(), v, om))
    starts.append(len(recs))
    return np.array(recs, np.float32), np.array(starts, np.uint32), np.array(mus)

