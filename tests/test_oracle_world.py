"""The oracle's per-body aux (setBodyAux: gravity scale, velocity damping; wgsl-solve.ts:14-20,
:293-305) and plastic joints (yield and bend limit, wgsl-solve.ts:562-635), pinned by closed forms
and by the invariants the reference's comments state, on the CPU."""
import math

import numpy as np

import oracle
from avbd_amd import records
from helpers import aux_factors, new_box, spec_record

DT = 1 / 60
G = -9.81


def _free_body(vel=(0.0, 0.0, 0.0), angvel=(0.0, 0.0, 0.0), prec=64):
    rec = new_box((0.0, 0.0, 100.0), (1.0, 1.0, 1.0), 1.0).astype(np.float64)
    rec[32:35] = vel
    rec[35] = vel[2]  # previous up-velocity: steady, so the adaptive warm start is settled
    rec[36:39] = angvel
    m = oracle.Model(prec)
    m.set_params(oracle.make_params(dt=DT, gravity=G, flags=0))
    m.set_bodies(rec[None])
    return m


def test_scaled_gravity_free_fall_is_the_closed_form():
    """No contacts: each step lands on the inertial target x + v dt + s g dt^2, so after N steps
    v = N s g dt and z = z0 + s g dt^2 N (N + 1) / 2."""
    for scale in (0.5, -0.25, 2.0):
        m = _free_body()
        m.set_aux(np.array([[scale, 1.0, 1.0, 0.0]]))
        n = 40
        m.step(n)
        b = m.bodies()[0]
        g = scale * G
        assert abs(b[34] - n * g * DT) < 1e-9, scale
        assert abs(b[2] - (100.0 + g * DT * DT * n * (n + 1) / 2)) < 1e-9, scale


def test_damping_decays_velocities_before_the_step_reads_them():
    """v_k = f v_(k-1) + g dt with f = exp(-rate dt) (the decayed velocity is stored, then the
    step adds gravity): the vertical velocity is a geometric series, the horizontal one decays as
    f^N, the angular one as f_ang^N (to the qadd/qsub round trip's O((w dt)^2))."""
    rate, arate = 0.8, 1.5
    f = float(aux_factors(rate, DT))
    fa = float(aux_factors(arate, DT))
    assert abs(f - math.exp(-rate * DT)) < 1e-7 and abs(fa - math.exp(-arate * DT)) < 1e-7
    m = _free_body(vel=(3.0, 0.0, 0.0), angvel=(0.0, 0.0, 0.1))
    m.set_aux(np.array([[1.0, f, fa, 1.0]]))
    n = 50
    m.step(n)
    b = m.bodies()[0]
    assert abs(b[32] - 3.0 * f ** n) < 1e-9
    assert abs(b[34] - G * DT * (1 - f ** n) / (1 - f)) < 1e-9
    assert abs(b[38] - 0.1 * fa ** n) < 1e-4 * 0.1 * fa ** n
    # undamped (flag clear): the factors are not applied even when they are not 1
    m = _free_body(vel=(3.0, 0.0, 0.0))
    m.set_aux(np.array([[1.0, f, fa, 0.0]]))
    m.step(n)
    assert abs(m.bodies()[0, 32] - 3.0) < 1e-12


def _cantilever(yield_force, break_bend=None, fracture=math.inf, iterations=10):
    """A static anchor and one 2 kg box held out sideways by a rigid joint: the angle lock carries
    m g L / torqueArm = 2 * 9.81 * 1.0 / 13.5 = 1.45 at rest."""
    bodies = np.vstack([new_box((0.0, 0.0, 5.0), (1.0, 1.0, 1.0), 0.0), new_box((1.5, 0.0, 5.0), (2.0, 0.5, 0.5), 4.0)])
    spec = {"a": 0, "b": 1, "rA": (0.5, 0.0, 0.0), "rB": (-1.0, 0.0, 0.0), "yield": yield_force, "fracture": fracture}
    if break_bend is not None:
        spec["breakBend"] = break_bend
    joint = spec_record(spec, bodies[:, 16:19])
    m = oracle.Model(64)
    m.set_params(oracle.make_params(dt=DT, gravity=G, flags=0, color_cap=8, iterations=iterations))
    m.set_bodies(records.bodies_to_oracle(bodies))
    m.set_joints(joint[None].astype(np.float64), np.array([[1, 0, 1, 0]], np.int32))
    return m, joint


def _bend(j):
    """The angle of the turn restStart^-1 rest (layout.ts J_REST_START)."""
    a, b = j[36:40], j[32:36]
    w = a[3] * b[3] + a[0] * b[0] + a[1] * b[1] + a[2] * b[2]
    v = np.linalg.norm(a[3] * b[0:3] - b[3] * a[0:3] - np.cross(a[0:3], b[0:3]))
    return 2.0 * math.atan2(v, abs(w))


def test_a_plastic_joint_holds_its_yield_force_and_keeps_its_bend():
    """One iteration per step, so a step's one dual pass either yielded or not."""
    m, joint = _cantilever(yield_force=1.0, iterations=1)
    bends, yielded = [], 0
    for _ in range(60):
        before = m.joints()[0].copy()
        m.step(1)
        j = m.joints()[0]
        if not np.array_equal(j[32:36], before[32:36]):  # the rest moved: it yielded this step
            yielded += 1
            assert abs(np.linalg.norm(j[12:15]) - 1.0) < 1e-9  # cut back to yield exactly
            assert np.array_equal(j[20:23], [0.0, 0.0, 0.0])  # its angular error goes with the bend
            decayed = np.minimum(np.maximum(before[4:7] * 0.999, 1.0), 1e10)  # the warm start's decay alone
            assert np.allclose(j[4:7], decayed, rtol=1e-15, atol=0)  # a yielding row's penalty is not ramped
        assert abs(np.linalg.norm(j[32:36]) - 1.0) < 1e-12  # the rest stays a unit rotation
        assert np.array_equal(j[36:40], joint[36:40])  # the starting rest never moves
        assert j[3] > 1e30 and j[23] == 1.0  # still whole, yield kept
        bends.append(_bend(j))
    assert yielded >= 5
    assert bends[-1] > 0.02 and all(b2 >= b1 - 1e-12 for b1, b2 in zip(bends, bends[1:]))  # it droops and stays drooped
    # the box hangs below the anchor's axis: the bend is about -y (x down towards -z)
    assert m.bodies()[1, 2] < 5.0 - 0.05


def test_a_plastic_joint_breaks_once_bent_past_its_limit():
    m, _ = _cantilever(yield_force=1.0, break_bend=0.1)
    last = 0.0
    for k in range(200):
        m.step(1)
        j = m.joints()[0]
        if j[3] == 0.0:  # broken: stiffnesses, penalties and multipliers zeroed
            assert np.array_equal(j[0:3], [0, 0, 0]) and j[7] == 0.0 and np.array_equal(j[12:15], [0, 0, 0])
            assert last <= 0.1 < _bend(j)  # the bend past the limit is what broke it
            break
        last = _bend(j)
    else:
        raise AssertionError("never broke")


def test_fracture_beats_yield():
    """Yield only acts short of fracture: a joint whose load (1.45) passes both its yield (0.5)
    and its fracture (1.0) breaks. If yield cut every force back, the force could never pass
    fracture and the joint (with no bend limit) would never break."""
    m, _ = _cantilever(yield_force=0.5, fracture=1.0)
    for k in range(60):
        m.step(1)
        if m.joints()[0, 3] == 0.0:
            break
    else:
        raise AssertionError("never broke")
    m, _ = _cantilever(yield_force=0.5)  # the same joint that cannot fracture only bends
    m.step(60)
    j = m.joints()[0]
    assert j[3] > 1e30 and _bend(j) > 0.1
