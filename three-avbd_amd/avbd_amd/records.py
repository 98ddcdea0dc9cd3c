"""Record conversions between the reference's f32 record ABI (src/avbd3d/gpu/layout.ts) and
float64 arrays whose integer words (stamps, keys, ids) are plain numbers.

Body record words that are integers on the GPU: initialPos.w (lifecycle flags, a small float),
inertialPos.w (step stamp, u32 bits), angVel.w (shape code, a float). Manifold words 0-3 and a
contact's key (word 3) are u32; everything else is f32.
"""
import numpy as np

BIG = 3.0e38


def bodies_from_oracle(rec):
    """f64 body records (stamp as a number) -> f32 records (stamp as u32 bits)."""
    rec = np.asarray(rec, np.float64).reshape(-1, 40)
    out = rec.astype(np.float32)
    out.view(np.uint32)[:, 27] = rec[:, 27].astype(np.uint32)
    return out


def bodies_to_oracle(rec):
    rec = np.ascontiguousarray(rec, np.float32).reshape(-1, 40)
    out = rec.astype(np.float64)
    out[:, 27] = rec.view(np.uint32)[:, 27]
    return out


def joints_from_oracle(rec):
    rec = np.asarray(rec, np.float64).reshape(-1, 40)
    return np.clip(rec, -BIG, BIG).astype(np.float32)


def contacts_from_oracle(man, con):
    """f64 manifolds (8) / contacts (16) -> u32 records."""
    man = np.asarray(man, np.float64).reshape(-1, 8)
    con = np.asarray(con, np.float64).reshape(-1, 16)
    m = np.zeros((len(man), 8), np.uint32)
    m[:, 0:4] = man[:, 0:4].astype(np.uint32)
    m[:, 4:8] = man[:, 4:8].astype(np.float32).view(np.uint32)
    c = con.astype(np.float32).view(np.uint32).copy()
    c[:, 3] = con[:, 3].astype(np.uint64).astype(np.uint32)
    return m, c


def contacts_to_oracle(man, con):
    man = np.ascontiguousarray(man, np.uint32).reshape(-1, 8)
    con = np.ascontiguousarray(con, np.uint32).reshape(-1, 16)
    m = np.zeros((len(man), 8))
    m[:, 0:4] = man[:, 0:4]
    m[:, 4:8] = man[:, 4:8].view(np.float32)
    c = con.view(np.float32).astype(np.float64)
    c[:, 3] = con[:, 3]
    return m, c


def box_record(pos, size, density, friction=0.5, rot=(0, 0, 0, 1), vel=(0, 0, 0), angvel=(0, 0, 0)):
    """A box body record as Rigid's constructor (ref/body.ts:34-56) and pack (solver.ts:1189)."""
    sx, sy, sz = (float(v) for v in size)
    mass = sx * sy * sz * density
    r = np.zeros(40)
    r[0:3] = pos
    r[3] = friction
    r[4:8] = rot
    r[8:11] = pos
    r[12:16] = rot
    r[16:19] = size
    r[19] = mass
    r[20:23] = [(sy * sy + sz * sz) / 12 * mass, (sx * sx + sz * sz) / 12 * mass, (sx * sx + sy * sy) / 12 * mass]
    r[23] = np.sqrt((sx * 0.5) ** 2 + (sy * 0.5) ** 2 + (sz * 0.5) ** 2)
    r[24:27] = pos
    r[28:32] = rot
    r[32:35] = vel
    r[35] = vel[2]
    r[36:39] = angvel
    r[39] = 0
    return r


def sphere_record(pos, radius, density, friction=0.5, rot=(0, 0, 0, 1), vel=(0, 0, 0), angvel=(0, 0, 0)):
    """A sphere (shapes.ts:28-48): size 2r, mass 4/3 pi r^3 density, moments 2/5 m r^2."""
    r = box_record(pos, (2 * radius,) * 3, density, friction, rot, vel, angvel)
    mass = 4.0 / 3.0 * np.pi * radius ** 3 * density
    r[19] = mass
    r[20:23] = 0.4 * mass * radius * radius
    r[23] = radius
    r[39] = 1
    return r
