"""The oracle's restatements of World's command passes (oracle.push_pass / move_pass / reset_pass:
Pusher, BodyMover, BodyReset) against closed forms: an impulse through the centre changes only
the velocity, by J/m; an off-centre one adds I^-1 (r x J) in the body's principal frame; a move
reaches its target in `left` steps, the angular velocity integrated by the solver's qadd landing
on the target rotation; a reset copies the pose and clears the velocities."""
import numpy as np
from scipy.spatial.transform import Rotation

import oracle


def body(pos=(0, 0, 0), rot=(0, 0, 0, 1), mass=2.0, moment=(1.0, 2.0, 3.0), vel=(0, 0, 0), angvel=(0, 0, 0), flags=0.0):
    r = np.zeros(40, np.float32)
    r[0:3] = pos
    r[4:8] = rot
    r[8:11] = pos
    r[11] = flags
    r[12:16] = rot
    r[16:19] = 1.0
    r[19] = mass
    r[20:23] = moment
    r[24:27] = pos
    r[28:32] = rot
    r[32:35] = vel
    r[36:39] = angvel
    r[39] = 0.0
    return r


def push_row(i, lin, off=(0, 0, 0), moment=(0, 0, 0)):
    p = np.zeros(12, np.float32)
    p[0:3] = lin
    p[3:4] = np.array([i], np.uint32).view(np.float32)
    p[4:7] = off
    p[8:11] = moment
    return p


def test_impulse_through_the_centre():
    rec = np.array([body(vel=(1, 0, 0))])
    out = oracle.push_pass(rec, [push_row(0, (2, 4, -6))])
    assert out[0, 32:35].tolist() == [2.0, 2.0, -3.0]
    assert out[0, 36:39].tolist() == [0.0, 0.0, 0.0]


def test_off_centre_impulse_in_the_principal_frame():
    pos = np.array([1.0, 2.0, 3.0], np.float32)
    at = pos + np.array([0.0, 0.5, 0.0], np.float32)  # half a metre off the centre along y
    j = np.array([3.0, 0.0, 0.0], np.float32)
    rec = np.array([body(pos=pos)])
    out = oracle.push_pass(rec, [push_row(0, j, j, np.cross(at, j))])
    assert np.allclose(out[0, 32:35], j / 2.0)
    # r x J = (0, 0.5, 0) x (3, 0, 0) = (0, 0, -1.5); I = (1, 2, 3)
    assert np.allclose(out[0, 36:39], [0.0, 0.0, -0.5], atol=1e-6)
    # a rotated body: I^-1 applied in its frame
    rng = np.random.default_rng(2)
    for _ in range(10):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        r = rng.normal(size=3).astype(np.float32)
        jj = rng.normal(size=3).astype(np.float32)
        rec = np.array([body(pos=pos, rot=q)])
        out = oracle.push_pass(rec, [push_row(0, jj, jj, np.cross(pos + r, jj))])
        R = Rotation.from_quat(q).as_matrix()
        want = R @ np.diag(1 / np.array([1.0, 2.0, 3.0])) @ R.T @ np.cross(r.astype(np.float64), jj)
        assert np.abs(out[0, 36:39] - want).max() <= 1e-5 * max(1.0, np.abs(want).max())


def test_static_and_disabled_bodies_are_left_alone():
    rec = np.array([body(mass=0.0), body(flags=2.0), body(flags=1.0)])
    out = oracle.push_pass(rec, [push_row(0, (1, 1, 1)), push_row(1, (1, 1, 1)), push_row(2, (1, 1, 1))])
    assert (out[0] == rec[0]).all() and (out[1] == rec[1]).all()
    assert out[2, 32:35].tolist() == [0.5, 0.5, 0.5]  # flag 1 (rewritten slot) still takes pushes


def move_row(i, target, left, dt, rot=None):
    m = np.zeros(12, np.float32)
    m[0:3] = target
    m[3:4] = np.array([i], np.uint32).view(np.float32)
    m[4:8] = rot if rot is not None else (0, 0, 0, 1)
    m[8:11] = (left, dt, 1.0 if rot is not None else 0.0)
    return m


def qadd(q, w, dt):
    """The solver's rotation step: q + 0.5 (w, 0) q dt, normalised (f64)."""
    wq = np.array([w[0], w[1], w[2], 0.0])
    x = wq
    prod = np.array([x[3] * q[0] + x[0] * q[3] + x[1] * q[2] - x[2] * q[1], x[3] * q[1] - x[0] * q[2] + x[1] * q[3] + x[2] * q[0],
                     x[3] * q[2] + x[0] * q[1] - x[1] * q[0] + x[2] * q[3], x[3] * q[3] - x[0] * q[0] - x[1] * q[1] - x[2] * q[2]])
    out = np.asarray(q, np.float64) + 0.5 * dt * prod
    return out / np.linalg.norm(out)


def test_move_reaches_the_target():
    dt = np.float32(1 / 60)
    rec = np.array([body(pos=(1, 2, 3), vel=(9, 9, 9), angvel=(1, 1, 1))])
    theta = 0.8
    target = np.array([0, 0, np.sin(theta / 2), np.cos(theta / 2)])
    out = oracle.move_pass(rec, [move_row(0, (4, 2, 0), 3, dt, target)])
    assert np.allclose(out[0, 32:35], np.array([3, 0, -3]) / (3 * np.float64(dt)), rtol=1e-6)
    # one step of qadd with 1/3 of the arc: 2 tan(theta / 6) / dt about z
    assert np.allclose(out[0, 36:39], [0, 0, 2 * np.tan(theta / 6) / np.float64(dt)], rtol=1e-5)
    # left = 1: one qadd step lands on the target rotation
    out = oracle.move_pass(rec, [move_row(0, (4, 2, 0), 1, dt, target)])
    q1 = qadd(np.array([0, 0, 0, 1.0]), out[0, 36:39].astype(np.float64), np.float64(dt))
    assert np.allclose(q1, target, atol=1e-6)
    # no rotation given: the angular velocity is cleared
    out = oracle.move_pass(rec, [move_row(0, (1, 2, 3), 1, dt)])
    assert out[0, 36:39].tolist() == [0.0, 0.0, 0.0] and out[0, 32:35].tolist() == [0.0, 0.0, 0.0]


def test_reset_copies_the_pose_and_clears_the_velocities():
    q = np.array([0.5, 0.5, 0.5, 0.5], np.float32)
    rec = np.array([body(pos=(1, 2, 3), rot=q, vel=(4, 5, 6), angvel=(7, 8, 9), flags=1.0)])
    rec[0, 8:11] = (0, 0, 0)
    rec[0, 24:27] = (0, 0, 0)
    rec[0, 35] = 0.25
    rec[0, 39] = 3.0
    out = oracle.reset_pass(rec, [0])
    assert out[0, 8:12].tolist() == [1, 2, 3, 1.0] and out[0, 24:27].tolist() == [1, 2, 3]
    assert (out[0, 12:16] == q).all() and (out[0, 28:32] == q).all()
    assert out[0, 32:36].tolist() == [0, 0, 0, 0] and out[0, 36:40].tolist() == [0, 0, 0, 3.0]
