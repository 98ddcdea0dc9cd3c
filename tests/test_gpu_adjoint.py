"""The HIP adjoint (avbd_forward / avbd_backward through the C ABI) against the oracle's.

The forward rollout must be the plain step bit for bit (it is the same code), so each world's
f32 states equal the oracle model's; the reverse runs in f64 at those states, as the oracle's
Frozen.load_recorded does, so the two adjoints are the same linear map evaluated in a different
summation order. The oracle's reverse is itself pinned by central differences
(test_adjoint_oracle.py)."""
import sys

import numpy as np
import pytest

import avbd_amd
import oracle
from avbd_amd import records, scenes

pytestmark = pytest.mark.gpu

F32 = lambda x: float(np.float32(x))


def say(*a):
    print(*a, flush=True)


def gpu_params():
    return avbd_amd.default_params(dt=1 / 60, gravity=-9.81, primal_lanes=(0, 0, 0, 0))


def oracle_params():
    return oracle.make_params(dt=F32(1 / 60), gravity=F32(-9.81), alpha=F32(0.99), gamma=F32(0.999), margin=F32(0.01),
                              beta_lin=F32(1e4), beta_ang=F32(100.0))


def centroid_cotangent(bodies, world_start):
    """Adjoint of L = sum over worlds of (1, 1, 1) . (mean position of the world's dynamic bodies)."""
    d = np.zeros((len(bodies), avbd_amd.solver.ADJ_FLOATS))
    for w in range(len(world_start) - 1):
        sl = slice(int(world_start[w]), int(world_start[w + 1]))
        dyn = np.nonzero(bodies[sl, 19] > 0)[0] + world_start[w]
        d[dyn, 0:3] = 1.0 / len(dyn)
    return d


def oracle_world(bodies, world_start, w, steps, d_final):
    sl = slice(int(world_start[w]), int(world_start[w + 1]))
    rec = records.bodies_to_oracle(bodies[sl])
    m = oracle.Model(32)
    m.set_params(oracle_params())
    m.set_bodies(rec)
    m.record_tape(True)
    m.step(steps)
    fz = oracle.Frozen(m, rec)
    fz.load_recorded()
    g, gmu = fz.backward(d_final[sl])
    return m.bodies(), g, gmu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("every", [10, 0, 7])
def test_gpu_adjoint_matches_the_oracle_at_the_recorded_states(gpu, every):
    worlds, steps = 3, 30
    bodies, ws, _ = scenes.thrown_at_stack(worlds, seed=11, stack=3, thrown=27)
    say(f"adjoint parity: {worlds} worlds x {ws[1]} bodies, {steps} steps, checkpoint_every={every}")
    s = avbd_amd.GpuSolver3D(bodies, params=gpu_params(), world_start=ws)
    s.forward(steps, every)
    final = s.read_bodies()
    d_final = centroid_cotangent(bodies, ws)
    d_init, d_mu = s.backward(d_final)
    assert np.array_equal(s.read_bodies(), final), "backward must leave the solver at the rollout's end"
    for w in range(worlds):
        sl = slice(int(ws[w]), int(ws[w + 1]))
        ref_final, g, gmu = oracle_world(bodies, ws, w, steps, d_final)
        # the rollout is the plain step: bit-exact poses against the f32 model
        assert np.array_equal(final[sl, 0:8].astype(np.float32), ref_final[:, 0:8].astype(np.float32)), f"world {w}: forward differs"
        dyn = bodies[sl, 19] > 0
        for name, cols in (("pos", slice(0, 3)), ("rot", slice(4, 8)), ("vel", slice(8, 11)), ("angvel", slice(12, 15))):
            e = rel(d_init[sl][dyn][:, cols], g[dyn][:, cols])
            say(f"  world {w} {name}: rel {e:.2e}")
            assert e < 1e-7, (w, name, e)
        e = abs(d_mu[sl].sum() - gmu.sum()) / max(abs(gmu.sum()), 1e-30)
        say(f"  world {w} friction: gpu {d_mu[sl].sum():.9e} oracle {gmu.sum():.9e} rel {e:.2e}")
        assert e < 1e-7
    s.destroy()


def test_gpu_adjoint_reuses_its_buffers_across_rollouts(gpu):
    """The solver keeps the adjoint's checkpoints, tapes and accumulators between rollouts (and
    grows them when a rollout needs more). Rollouts chained on one solver (each backward leaves it
    at its rollout's end) against a fresh solver brought to the same state by plain steps and then
    running the same rollout with newly allocated buffers."""
    bodies, ws, _ = scenes.thrown_at_stack(2, seed=4, stack=2, thrown=8)
    d_final = centroid_cotangent(bodies, ws)
    s = avbd_amd.GpuSolver3D(bodies, params=gpu_params(), world_start=ws)
    done = 0
    for steps, every in ((12, 4), (30, 0), (30, 10)):  # growing: more checkpoints, a whole-rollout tape, then segments
        say(f"reuse: {steps} steps from step {done}, checkpoint_every={every}")
        s.forward(steps, every)
        final = s.read_bodies()
        g, gmu = s.backward(d_final)
        ref = avbd_amd.GpuSolver3D(bodies, params=gpu_params(), world_start=ws)
        if done:
            ref.steps(done)
        ref.forward(steps, every)
        assert np.array_equal(final, ref.read_bodies()), "the rollouts differ"
        g_ref, gmu_ref = ref.backward(d_final)
        ref.destroy()
        e = rel(g, g_ref)
        say(f"  gradient vs a fresh solver: rel {e:.2e}")
        assert e < 1e-12 and abs(gmu.sum() - gmu_ref.sum()) <= 1e-12 * abs(gmu_ref.sum())
        done += steps
    s.destroy()


def test_gpu_adjoint_refuses_what_it_does_not_cover(gpu):
    bodies, ws, _ = scenes.thrown_at_stack(1, seed=2, stack=2, thrown=8)
    s = avbd_amd.GpuSolver3D(bodies, params=gpu_params(), world_start=ws)
    with pytest.raises(ValueError, match="forward"):
        s.backward(np.zeros((len(bodies), 16)))
    s.forward(3, 2)
    s.step()
    with pytest.raises(ValueError, match="stepped"):
        s.backward(np.zeros((len(bodies), 16)))
    s.destroy()
    joints = np.zeros((1, 40), np.float32)
    joints[0, 3] = 3e38
    s = avbd_amd.GpuSolver3D(bodies, joints, np.array([[1, 1, 2, 0]], np.int32), params=gpu_params())
    with pytest.raises(ValueError, match="contact-only"):
        s.forward(2, 1)
    s.destroy()
