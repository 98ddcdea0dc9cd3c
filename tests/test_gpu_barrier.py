"""The cooperative solve's grid barrier gives up after a bounded wait instead of hanging the
device. Forced from the host with a tiny poll limit (every kernel still ends normally): the
give-up is reported, the solver drops to per-colour launches, the barrier's count does not stay
spoiled for later launches, and the following steps equal the oracle's bit for bit."""
import sys

import numpy as np
import pytest

import avbd_amd
from avbd_amd import records, scenes
from helpers import oracle_from_gpu

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pile_solver(flags=avbd_amd.FLAG_DEFAULTS):
    bodies = scenes.random_pile(400, seed=6, bin_size=10.0)[0]
    p = avbd_amd.default_params(color_cap=24, flags=flags, primal_lanes=(1 << 20, 1 << 19, 1 << 18, 0))
    s = avbd_amd.GpuSolver3D(bodies, params=p)
    s.set_schedule(cooperative=1, threads=256, blocks_per_cu=1)
    return s


def _same_as_oracle(s, steps):
    m = oracle_from_gpu(s)
    for _ in range(steps):
        s.step()
        m.step(1)
    assert np.array_equal(records.bodies_to_oracle(s.read_bodies()), m.bodies())


def test_barrier_give_up_is_reported_and_the_solver_drops_to_per_colour_launches(gpu):
    say("case: cooperative solve, 400-body pile, barrier poll limit 1 for one step")
    s = _pile_solver()
    s.steps(40)
    _same_as_oracle(s, 2)  # the cooperative schedule is exact while its barriers hold
    assert s.schedule["cooperative"] == 1
    assert s.read_counters()["overflow"] & avbd_amd.solver.OVERFLOW_BARRIER == 0
    s.set_barrier_polls(1)
    s.step()  # no readback inside this step: the call itself cannot know yet
    c = s.read_counters()
    assert c["overflow"] & avbd_amd.solver.OVERFLOW_BARRIER
    assert s.schedule["cooperative"] == 0
    assert "barrier" in avbd_amd.lib().avbd_last_error(s.h).decode()
    assert s.read_counters()["overflow"] & avbd_amd.solver.OVERFLOW_BARRIER == 0  # reported once
    _same_as_oracle(s, 3)  # per-colour launches from the state the unordered step left
    # the arrival count starts every launch at zero: the cooperative schedule is exact again
    s.set_barrier_polls(0)
    s.set_schedule(cooperative=1, threads=256, blocks_per_cu=1)
    _same_as_oracle(s, 3)
    assert s.schedule["cooperative"] == 1
    assert s.read_counters()["overflow"] & avbd_amd.solver.OVERFLOW_BARRIER == 0
    s.destroy()


def test_steps_with_adaptive_readbacks_fail_loudly_on_a_give_up(gpu):
    say("case: cooperative solve with adaptive readbacks, barrier poll limit 1")
    s = _pile_solver(avbd_amd.FLAG_DEFAULTS | avbd_amd.solver.FLAG_ADAPT_COLORS)
    s.steps(20)
    s.set_barrier_polls(1)
    with pytest.raises(RuntimeError, match="barrier"):
        s.steps(10)
    assert s.schedule["cooperative"] == 0
    s.set_barrier_polls(0)
    s.steps(10)  # per-colour launches: no error
    s.destroy()
