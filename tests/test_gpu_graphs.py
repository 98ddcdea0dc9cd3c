"""Per-substep hipGraphs (avbd_set_graphs): a solver replaying captured substeps steps exactly as
one launching them plainly, bit for bit, over many replays with the host reading results (and
so synchronising) between calls; a changed launch sequence (adapted colour cap, new params, a
host edit that invalidates the manifolds, another schedule) is captured anew, never replayed
stale."""
import numpy as np
import pytest

import avbd_amd
from avbd_amd import records, scenes
from helpers import canonical_contacts

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def same(a, b, what):
    assert np.array_equal(a.read_bodies(), b.read_bodies()), what
    am, ac = canonical_contacts(*a.read_contacts_raw())  # emission order follows atomics: compare sorted
    bm, bc = canonical_contacts(*b.read_contacts_raw())
    assert np.array_equal(am, bm) and np.array_equal(ac, bc), what
    ca, cb = a.read_counters(), b.read_counters()
    assert ca == cb, (what, ca, cb)


def pair(bodies, params, **kw):
    g = avbd_amd.GpuSolver3D(bodies, params=params, **kw)
    e = avbd_amd.GpuSolver3D(bodies, params=params, **kw)
    g.set_graphs(True)
    return g, e


@pytest.mark.parametrize("schedule", ["colors", "coop"])
def test_batched_worlds_replay_bit_for_bit(gpu, schedule):
    say(f"case: 48 batched worlds, graphs vs plain launches, {schedule} schedule, 60 steps")
    bodies, starts = scenes.batched_worlds(48, lo=8, hi=200)
    params = avbd_amd.default_params(color_cap=24, flags=avbd_amd.FLAG_DEFAULTS | avbd_amd.solver.FLAG_ADAPT_COLORS)
    g, e = pair(bodies, params, world_start=starts)
    for s in (g, e):
        s.set_schedule(cooperative=1 if schedule == "coop" else 0)
    for k in range(12):
        g.steps(5)
        e.steps(5)
        same(g, e, f"after {5 * (k + 1)} steps")
    captures, replays = g.graph_stats()
    say(f"  captures {captures}, replays {replays}")
    if schedule == "colors":
        assert replays > 40 and captures < 20  # the adaptive cap changes the sequence now and then
    assert e.graph_stats() == (0, 0)
    g.destroy()
    e.destroy()


def test_substeps_with_reuse_and_host_changes(gpu):
    say("case: hulls, chains and substeps with manifold reuse; params and bodies changed between calls")
    bodies, joints, info, hw = scenes.large_scene(300, seed=9, chains=2, links=4, workers=1)
    params = avbd_amd.default_params(dt=1 / 240, color_cap=32)
    g, e = pair(bodies, params, joints=joints, joint_info=info, hulls=hw)
    for s in (g, e):
        s.set_schedule(cooperative=0)
    for k in range(6):
        g.steps(2, 4)
        e.steps(2, 4)
        same(g, e, f"call {k}")
    c0 = g.graph_stats()[0]
    # new params: the same substeps need new graphs
    for s in (g, e):
        p = s.params
        p.iterations = 7
        s.set_params(p)
    g.steps(2, 4)
    e.steps(2, 4)
    same(g, e, "after new params")
    assert g.graph_stats()[0] > c0
    # a host write invalidates the manifolds: the first substep detects although reuse was asked
    rec = g.read_bodies(10, 1)
    rec[0, 0:3] += (0.02, 0.0, 0.01)
    for s in (g, e):
        s.write_bodies(rec, first=10)
        s.step(reuse_manifolds=True)
        s.step(reuse_manifolds=True)
    same(g, e, "after a write")
    for _ in range(3):
        g.steps(1, 4)
        e.steps(1, 4)
    same(g, e, "after the write, more substeps")
    # graphs off: plain launches from the same state
    g.set_graphs(False)
    g.steps(1, 4)
    e.steps(1, 4)
    same(g, e, "graphs dropped")
    g.destroy()
    e.destroy()
