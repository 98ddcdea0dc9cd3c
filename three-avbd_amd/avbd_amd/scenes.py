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
