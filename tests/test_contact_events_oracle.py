"""The oracle's contact-event restatement (oracle.contact_events, ContactWatch's passes in
src/lib/contacts.ts) against known answers: hand-built manifolds whose events follow from the
definitions (a pair begins the first run it touches, ends the first run it does not), and the
quaternion rotation against scipy's. A swapped body, a dropped term of the mean point, a wrong
lambda component or a missing dt fails one of these."""
import numpy as np
from scipy.spatial.transform import Rotation

import oracle

REPORT = 0x80000000


def manifold(a, b, first, count, normal):
    m = np.zeros(8, np.uint32)
    m[0:4] = [a, b, first, count | (7 << 4)]
    m[4:7] = np.asarray(normal, np.float32).view(np.uint32)
    return m


def contact(rA, lam_n):
    c = np.zeros(16, np.float32)
    c[0:3] = rA
    c[12] = lam_n
    return c.view(np.uint32)


def bodies_at(positions, rotations=None):
    b = np.zeros((len(positions), 40), np.float32)
    b[:, 0:3] = positions
    b[:, 7] = 1.0
    if rotations is not None:
        b[:, 4:8] = rotations
    return b


def test_turn_is_the_quaternion_rotation():
    rng = np.random.default_rng(1)
    for _ in range(50):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        v = rng.normal(size=3)
        want = Rotation.from_quat(q).apply(v)  # scipy: (x, y, z, w)
        got = np.array(oracle.turn32(q, v), np.float64)
        assert np.abs(got - want).max() <= 1e-5 * max(1.0, np.abs(want).max())


def test_begin_stay_end_with_point_impulse_and_normal():
    dt = np.float32(1 / 60)
    bodies = bodies_at([[0, 0, 1], [0, 0, 0], [5, 0, 0]])
    watch = np.array([11 | REPORT, 22, 33], np.uint32)
    con = np.array([contact([0.5, 0.5, -0.5], -2.0), contact([-0.5, 0.5, -0.5], -4.0)])
    man = np.array([manifold(1, 0, 0, 2, [0, 0, -1])])
    # run 1: the pair (slot 1 = id 22 first, slot 0 = id 11, which reports) begins
    now, ev = oracle.contact_events(set(), man, con, bodies, watch, 10, 1, dt)
    assert now == {(22, 11)}
    assert len(ev) == 1 and ev[0]["kind"] == 1 and (ev[0]["a"], ev[0]["b"], ev[0]["step"]) == (22, 11, 1)
    # mean of the points on body a (slot 1 at the origin, identity rotation)
    assert ev[0]["point"] == [0.0, 0.5, -0.5]
    assert ev[0]["impulse"] == float(np.float32(6.0) * dt)
    assert ev[0]["normal"] == [0.0, 0.0, -1.0]
    # run 2: still touching -> nothing; run 3: gone -> an end
    now2, ev2 = oracle.contact_events(now, man, con, bodies, watch, 10, 2, dt)
    assert now2 == now and ev2 == []
    now3, ev3 = oracle.contact_events(now2, man[:0], con, bodies, watch, 10, 3, dt)
    assert now3 == set() and ev3 == [{"kind": 2, "a": 22, "b": 11, "step": 3, "point": [0.0, 0.0, 0.0], "impulse": 0.0,
                                      "normal": [0.0, 0.0, 0.0]}]


def test_the_point_turns_with_body_a():
    q = np.array([0, 0, np.sin(np.pi / 4), np.cos(np.pi / 4)], np.float32)  # 90 degrees about z
    bodies = bodies_at([[1, 2, 3], [0, 0, 0]], rotations=[q, [0, 0, 0, 1]])
    watch = np.array([1 | REPORT, 2], np.uint32)
    con = np.array([contact([1, 0, 0], 1.0)])
    _, ev = oracle.contact_events(set(), np.array([manifold(0, 1, 0, 1, [1, 0, 0])]), con, bodies, watch, 4, 0, 1.0)
    assert np.allclose(ev[0]["point"], [1, 3, 3], atol=1e-6)  # (1,0,0) turned to (0,1,0), plus (1,2,3)


def test_only_reporting_pairs_with_ids_count():
    bodies = bodies_at([[0, 0, 0]] * 4)
    con = np.array([contact([0, 0, 0], 1.0)] * 4)
    man = np.array([manifold(1, 0, 0, 1, [0, 0, 1]), manifold(3, 2, 1, 1, [0, 0, 1]), manifold(2, 0, 2, 0, [0, 0, 1]),
                    manifold(3, 0, 3, 1, [0, 0, 1])])
    # 0 reports; 1 has an ID but does not report; 2 has no ID; 3 reports
    watch = np.array([5 | REPORT, 6, 0 | REPORT, 8 | REPORT], np.uint32)
    now, ev = oracle.contact_events(set(), man, con, bodies, watch, 10, 0, 1.0)
    # (1, 0): 0 reports -> in; (3, 2): 2 has no ID -> out; (2, 0): no points -> out; (3, 0) in
    assert now == {(6, 5), (8, 5)}
    assert sorted((e["a"], e["b"]) for e in ev) == [(6, 5), (8, 5)]
    # only the first `pairs` manifolds are scanned
    now, _ = oracle.contact_events(set(), man, con, bodies, watch, 1, 0, 1.0)
    assert now == {(6, 5)}
