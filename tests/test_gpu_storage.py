"""adapt's capacity half (solver.ts:2362-2374) and fixed storage on the GPU: pairs, manifolds and
contacts grow before they overflow, keeping the warm-start contacts (a solver that grew steps
exactly as one that was large from the start, and as the oracle seeded from its state after the
growth), and a fixed-storage solver keeps its capacities and counts what did not fit."""
import numpy as np
import pytest

import avbd_amd
from avbd_amd import records, scenes
from helpers import oracle_from_gpu

pytestmark = pytest.mark.gpu

FLAGS = avbd_amd.FLAG_DEFAULTS | avbd_amd.solver.FLAG_ADAPT_COLORS


def say(*a):
    print(*a, flush=True)


def pile(n=300):
    return scenes.random_pile(n, seed=21, bin_size=6.0)[0]


def test_storage_grows_before_it_overflows(gpu):
    bodies = scenes.pyramid(14)[0]  # 105 boxes touching from the first step
    small = {"pairs": 64, "manifolds": 64, "contacts": 128}
    say("storage: 105-box pyramid, capacities 64/64/128, adaptive readback every 10 steps")
    s = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(gravity=-9.81, flags=FLAGS), capacity=small)
    f = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(gravity=-9.81, flags=FLAGS | avbd_amd.solver.FLAG_FIXED_STORAGE),
                             capacity=small)
    for _ in range(40):
        s.step()
        f.step()
    cap, c = s.capacity(), s.read_counters()
    say(f"  grown to {cap['pairs']} pairs, {cap['manifolds']} manifolds, {cap['contacts']} contacts; last step {c}")
    assert cap["pairs"] > 64 and cap["contacts"] > 128 and cap["manifolds"] <= cap["pairs"]
    assert c["overflow"] == 0  # grown at the readbacks, with room to spare since
    assert f.capacity()["pairs"] == 64 and f.capacity()["contacts"] == 128
    assert f.read_counters()["overflow"] != 0
    s.destroy()
    f.destroy()


def test_growth_keeps_the_warm_starts(gpu):
    bodies = pile(200)
    params = avbd_amd.default_params(gravity=-9.81)
    a = avbd_amd.GpuSolver3D(bodies, params=params, capacity={"pairs": 4000, "manifolds": 4000, "contacts": 12000})
    b = avbd_amd.GpuSolver3D(bodies, params=params, capacity={"pairs": 40000, "manifolds": 40000, "contacts": 120000})
    for _ in range(30):
        a.step()
        b.step()
    c = a.read_counters()
    assert c["overflow"] == 0 and c["contacts"] > 0
    # a readback that says the contacts are near a's capacity: a grows, b (large) does not
    forced = dict(c, contacts=int(0.9 * a.capacity()["contacts"]))
    before = a.capacity()
    a.adapt(forced)
    b.adapt(forced)
    assert a.capacity()["contacts"] > before["contacts"] and b.capacity()["contacts"] == 120000
    m = oracle_from_gpu(a)  # the grown buffers' contents are the warm starts the next step reads
    for _ in range(15):
        a.step()
        b.step()
        m.step(1)
    assert np.array_equal(a.read_bodies().view(np.uint32), b.read_bodies().view(np.uint32))
    assert np.array_equal(records.bodies_to_oracle(a.read_bodies()), m.bodies())
    a.destroy()
    b.destroy()
