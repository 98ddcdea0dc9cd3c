"""The oracle's adjoint (oracle/adjoint3d.hpp) against the mathematics: central finite differences.

The oracle's reverse sweep is hand-written; what pins it is that it is the derivative of the
frozen map it claims to differentiate (Frozen.run: the recorded step with its discrete decisions
held), checked by f64 central differences of that map, down to the quaternion helpers. A dropped
term, a wrong sign or a transposed index in the reverse fails these checks by orders of
magnitude (the agreement below is 1e-6 or better where the differences themselves are exact to
that level)."""
import numpy as np
import pytest

import oracle
from avbd_amd import scenes


def world(stack, thrown, seed=5):
    bodies, _, _ = scenes.thrown_at_stack(1, seed=seed, stack=stack, thrown=thrown)
    return bodies.astype(np.float64)


def recorded(rec, steps, prec=64):
    m = oracle.Model(prec)
    m.set_params(oracle.make_params(dt=1 / 60, gravity=-9.81))
    m.set_bodies(rec)
    m.record_tape(True)
    m.step(steps)
    return m


def centroid_cotangent(rec, out):
    dyn = rec[:, 19] > 0
    d = np.zeros_like(out)
    d[dyn, 0:3] = 1.0 / dyn.sum()
    return d, lambda o: o[dyn, 0:3].sum() / dyn.sum()


def test_quaternion_reverses_match_central_differences():
    rng = np.random.default_rng(3)
    for _ in range(20):
        q = rng.normal(size=4); q /= np.linalg.norm(q)
        b = rng.normal(size=4); b /= np.linalg.norm(b)
        err = oracle.check_quaternion_reverses(q, b, rng.normal(size=3) * 0.5, rng.normal(size=4))
        assert err < 1e-8, err


def test_frozen_replay_reproduces_the_model():
    rec = world(2, 8)
    m = recorded(rec, 30)
    fz = oracle.Frozen(m, rec)
    out = fz.run()
    ref = oracle.state_from_records(m.bodies())
    dyn = rec[:, 19] > 0
    assert np.abs(out[dyn] - ref[dyn]).max() < 1e-11


def test_adjoint_matches_central_differences_short_rollout():
    """2x2x2 stack + 8 thrown boxes, 20 steps: every input direction at h = 1e-4."""
    rec = world(2, 8)
    m = recorded(rec, 20)
    fz = oracle.Frozen(m, rec)
    stats = fz.stats()
    assert stats["contacts"] > 0 and stats["sliding"] > 0 and stats["clamped"] > 0
    st0 = oracle.state_from_records(rec)
    out = fz.run(st0)
    d_final, loss = centroid_cotangent(rec, out)
    d_init, d_mu = fz.backward(d_final)
    dyn = rec[:, 19] > 0
    rng = np.random.default_rng(0)
    h = 1e-4
    for slots in ((slice(8, 11),), (slice(12, 15),), (slice(0, 3),), (slice(4, 8),), (slice(8, 11), slice(12, 15))):
        d = np.zeros_like(st0)
        for sl in slots:
            d[dyn, sl] = rng.normal(size=(dyn.sum(), sl.stop - sl.start))
        fd = (loss(fz.run(st0 + h * d)) - loss(fz.run(st0 - h * d))) / (2 * h)
        an = float((d_init * d).sum())
        assert abs(fd - an) <= 1e-6 * max(abs(fd), 1e-3), (slots, fd, an)
    mu0 = rec[:, 3].copy()
    fd = (loss(fz.run(st0, mu0 + h)) - loss(fz.run(st0, mu0 - h))) / (2 * h)
    assert abs(fd - d_mu.sum()) <= 1e-6 * max(abs(fd), 1e-3), (fd, d_mu.sum())


def test_adjoint_is_the_limit_of_central_differences_on_a_full_world():
    """One world of the differentiation workload (4x4x4 stack, 64 thrown boxes, 120 steps). The
    frozen map is strongly curved here, so the differences converge as h shrinks: single
    components at h = 1e-6 and the shared friction at h = 1e-8 agree with the adjoint."""
    rec = world(4, 64)
    m = recorded(rec, 120)
    fz = oracle.Frozen(m, rec)
    st0 = oracle.state_from_records(rec)
    out = fz.run(st0)
    d_final, loss = centroid_cotangent(rec, out)
    d_init, d_mu = fz.backward(d_final)
    idx = np.nonzero(rec[:, 19] > 0)[0]
    checked = 0
    for i, slot in ((int(idx[-31]), 14), (int(idx[-42]), 14), (int(idx[-3]), 9), (int(idx[-20]), 10)):
        an = d_init[i, slot]
        if abs(an) < 1e-3:
            continue
        d = np.zeros_like(st0)
        d[i, slot] = 1.0
        h = 1e-6
        fd = (loss(fz.run(st0 + h * d)) - loss(fz.run(st0 - h * d))) / (2 * h)
        assert abs(fd - an) <= 1e-3 * abs(an), (i, slot, fd, an)
        checked += 1
    assert checked >= 2
    mu0 = rec[:, 3].copy()
    h = 1e-8
    fd = (loss(fz.run(st0, mu0 + h)) - loss(fz.run(st0, mu0 - h))) / (2 * h)
    assert abs(fd - d_mu.sum()) <= 1e-3 * abs(d_mu.sum()), (fd, d_mu.sum())


def test_recorded_states_linearisation_equals_the_replay_in_f64():
    """load_recorded (the reverse at the recording model's own states) is the replay's reverse
    when the model ran in f64: the path the GPU adjoint is compared through."""
    rec = world(3, 27)
    m = recorded(rec, 40)
    fz = oracle.Frozen(m, rec)
    st0 = oracle.state_from_records(rec)
    out = fz.run(st0)
    d_final, _ = centroid_cotangent(rec, out)
    g1, m1 = fz.backward(d_final)
    fz.load_recorded()
    g2, m2 = fz.backward(d_final)
    assert np.abs(g1 - g2).max() <= 1e-9 * np.abs(g1).max()
    assert np.abs(m1 - m2).max() <= 1e-9 * np.abs(m1).max()


def test_joints_are_refused():
    rec = world(2, 8)
    m = recorded(rec, 1)
    j = np.zeros((1, 40))
    j[0, 3] = 3e38
    m.set_joints(j, np.array([[1, 1, 2, 0]], np.int32))
    with pytest.raises(RuntimeError, match="contact-only"):
        oracle.Frozen(m, rec)
