"""The pivot floor of the per-body LDLT (finishBody, src/avbd3d/gpu/wgsl-solve.ts:504-558: each
pivot at least 1e-5 of its diagonal) on the reference's own case for it
(tests-gpu/pivot.gpu.test.ts:15-41): a 0.6 g sliver resting on its edge under contacts at the
penalty cap, rocking about the edge. Rotation about the edge is held only by the sliver's
inertia, ~12 orders of magnitude under the contact rows, so in f32 that pivot cancels to noise;
without the floor it can go negative and the body flies off to Inf and NaN. The oracle's f32 step
stays finite and on the ground (CPU), and the HIP step equals it bit for bit (GPU)."""
import math

import numpy as np
import pytest

import oracle
from avbd_amd import records

FLAGS = oracle.FLAG_START_AT_REST | oracle.FLAG_MASS_PENALTY  # the reference turns the other three off


def sliver_state():
    ref = oracle.RefSolver()
    ref.add_body([100, 100, 1], 0, 0.8, [0, 0, -0.5])
    s = math.sin(math.pi / 8)
    ref.add_body([0.05, 0.004, 0.004], 750, 0.8, [0, 0, 0.004 * math.sqrt(0.5)], rot=[s, 0, 0, math.cos(math.pi / 8)])
    ref.step(5)
    man, con = ref.manifold_records()
    assert any(int(m[3]) == 2 for m in man)  # resting on an edge: 2 contacts
    con[:, 8:11] = 1e10  # every contact row at the penalty cap, as under a slab's weight
    bodies = ref.body_records()
    bodies[1, 36:39] = (3.0, 0.0, 0.0)  # rocking about the edge
    return bodies, man, con


def oracle_run(steps=20):
    bodies, man, con = sliver_state()
    m = oracle.Model(32)
    m.set_params(oracle.make_params(gravity=-10.0, flags=FLAGS))
    m.set_bodies(bodies)
    m.set_prev(man, con)
    m.step(steps)
    return bodies, man, con, m


def test_the_sliver_stays_finite_and_on_the_ground():
    _, _, _, m = oracle_run()
    chip = m.bodies()[1]
    assert np.isfinite(chip).all(), chip
    assert abs(chip[2]) < 0.05 and math.hypot(chip[0], chip[1]) < 0.05, chip[0:3]


@pytest.mark.gpu
def test_the_hip_step_floors_the_pivot_as_the_oracle(gpu):
    import avbd_amd
    bodies, man, con, m = oracle_run()
    b32 = records.bodies_from_oracle(bodies)
    p = avbd_amd.default_params(gravity=-10.0, flags=FLAGS, color_cap=64, primal_lanes=(0, 0, 0, 0))
    s = avbd_amd.GpuSolver3D(b32, params=p)
    mu, cu = records.contacts_from_oracle(man, con)
    s.seed_from(b32, None, mu, cu)
    for _ in range(20):
        s.step()
    got = records.bodies_to_oracle(s.read_bodies())
    assert np.isfinite(got).all()
    assert np.array_equal(got, m.bodies())
    s.destroy()


def _spd(rng, scale_rows=1e10):
    """A body's system as the primal builds it: masses and moments over dt^2 on the diagonal plus
    contact rows k J J^T (J = (l, r x l))."""
    dt2 = (1 / 60) ** 2
    m, I = 0.6e-3, rng.uniform(1e-9, 1e-8, 3)
    A = np.diag(np.concatenate([[m, m, m], I]) / dt2)
    for _ in range(6):
        l = rng.normal(size=3)
        l /= np.linalg.norm(l)
        r = rng.normal(size=3) * 0.02
        J = np.concatenate([l, np.cross(r, l)])
        A += scale_rows * np.outer(J, J)
    return A


def test_the_floored_ldlt_solves_well_conditioned_systems():
    """f64, no pivot near its floor: the solve of a dense system (numpy)."""
    rng = np.random.default_rng(3)
    for _ in range(50):
        A = _spd(rng, scale_rows=rng.uniform(1, 1e4))
        rhs = rng.normal(size=6)
        x, D, S = oracle.solve6_floor(A, rhs, prec=64)
        assert np.array_equal(D, S)  # no floor taken
        assert np.allclose(x, np.linalg.solve(A, rhs), rtol=1e-9, atol=0)


def test_the_floor_keeps_every_pivot_at_least_1e_5_of_its_diagonal():
    """f32, one direction u held only by inertia under penalty-cap rows (the rows span u's
    complement): the pivot where the leading block first holds u cancels to noise. With u in the
    first k coordinates that is pivot k, for each k in turn: the floor is taken exactly where a
    pivot fell under 1e-5 A_kk, and the solve stays finite."""
    rng = np.random.default_rng(4)
    engaged = set()
    for trial in range(200):
        k = 2 + trial % 5
        u = np.zeros(6)
        u[:k] = rng.normal(size=k)
        u /= np.linalg.norm(u)
        A = np.diag(rng.uniform(1e-3, 1.0, 6))
        for _ in range(5):
            r = rng.normal(size=6)
            r -= (r @ u) * u
            A += 1e10 * np.outer(r, r) / (r @ r)
        rhs = rng.normal(size=6)
        x, D, S = oracle.solve6_floor(A, rhs, prec=32)
        diag = np.diag(A).astype(np.float32).astype(np.float64)
        floor = (diag * np.float32(1e-5)).astype(np.float32).astype(np.float64)
        assert np.isfinite(x).all() and np.isfinite(D).all(), trial
        assert D[0] == np.float32(A[0, 0])
        for j in range(1, 6):
            assert D[j] == max(S[j], floor[j]), (trial, j)
            if S[j] < floor[j]:
                engaged.add(j)
    assert engaged == {1, 2, 3, 4, 5}, engaged
