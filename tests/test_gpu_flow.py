"""The dataflow sweep (schedule 2: one cooperative launch per substep, bodies ordered by per-body
completion counts) against the oracle, bit for bit: it reads what the colour-by-colour sweep reads,
so every pose, contact, colour and joint must equal the oracle's restatement of that sweep. Cases
cover one to sixteen threads per body, hulls and chains, manifold reuse with joints breaking,
clashing (serial) colours, plastic joints, batched worlds, and a forced give-up of its bounded
waits (the host reports it and drops to per-colour launches; every kernel still ends normally)."""
import sys

import numpy as np
import pytest

import avbd_amd
import oracle
from avbd_amd import records, scenes
from helpers import oracle_from_gpu, sorted_contacts

pytestmark = pytest.mark.gpu

LANES1 = (0, 0, 0, 0)
FLOW = dict(cooperative=2, threads=64, blocks_per_cu=4)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def model_for(bodies, params_kw, joints=None, info=None):
    params_kw = dict(params_kw)
    params_kw.setdefault("gravity", -9.81)
    m = oracle.Model(32)
    m.set_params(oracle.make_params(**params_kw))
    m.set_bodies(records.bodies_to_oracle(bodies))
    if joints is not None and len(joints):
        m.set_joints(np.asarray(joints, np.float64), info)
    return m


def flow_for(bodies, params_kw, joints=None, info=None, **kw):
    pk = dict(params_kw)
    pk.setdefault("gravity", -9.81)
    pk["primal_lanes"] = pk.pop("lanes", LANES1)
    s = avbd_amd.GpuSolver3D(bodies, joints, info, params=avbd_amd.default_params(**pk), **kw)
    s.set_schedule(**FLOW)
    assert s.schedule["cooperative"] == 2
    return s


def same_contacts(gpu, model):
    gm, gc = gpu.read_contacts_raw()
    om, oc = model.contacts()
    G = sorted_contacts(gm, gc)
    O = sorted_contacts(*records.contacts_from_oracle(om, oc))
    assert len(G) == len(O)
    for (ma, ca), (mb, cb) in zip(G, O):
        assert (ma[:2] == mb[:2]).all() and (ma[3] & 15) == (mb[3] & 15) and (ma[4:8] == mb[4:8]).all(), (ma, mb)
        assert (ca == cb).all(), (ma[:2], ca, cb)


def same_bodies(s, m):
    got, want = records.bodies_to_oracle(s.read_bodies()), m.bodies()
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert len(bad) == 0, (len(bad), bad[:10], np.abs(got - want).max())


def no_give_up(s):
    c = s.read_counters()
    assert c["overflow"] & avbd_amd.solver.OVERFLOW_BARRIER == 0
    assert s.schedule["cooperative"] == 2


def test_flow_pile_one_lane(gpu):
    say("case: flow sweep, pile 400 x 40 steps, one thread per body")
    bodies = scenes.random_pile(400, seed=3, bin_size=10.0)[0]
    kw = dict(lanes=LANES1, color_cap=64)
    s, m = flow_for(bodies, kw), model_for(bodies, kw)
    for _ in range(40):
        s.step()
        m.step(1)
    same_bodies(s, m)
    same_contacts(s, m)
    assert np.array_equal(s.read_colors()[0], m.colors())
    no_give_up(s)
    s.destroy()


@pytest.mark.parametrize("rule,pitch", [((1 << 20, 1 << 19, 1 << 18, 0), 2.5), ((1 << 20, 1 << 19, 1 << 18, 1 << 17), 2.5),
                                        ((2 ** 31, 200, 50, 10), 1.6)], ids=["8", "16", "mixed"])
def test_flow_lane_splits(gpu, rule, pitch):
    say(f"case: flow sweep, pile 400 (pitch {pitch}) x 10 steps, lane rule {rule}")
    bodies = scenes.random_pile(400, seed=4, bin_size=10.0 if pitch == 2.5 else 8.0, pitch=pitch)[0]
    kw = dict(lanes=rule, color_cap=64)
    s, m = flow_for(bodies, kw), model_for(bodies, kw)
    for _ in range(10):
        s.step()
        m.step(1)
    same_bodies(s, m)
    no_give_up(s)
    s.destroy()


def test_flow_hulls_spheres_boxes_chains(gpu):
    say("case: flow sweep, hull/sphere/box pile + chains, 600 bodies x 20 steps")
    bodies, joints, info, hw = scenes.large_scene(600, seed=7, chains=4, links=5, workers=1)
    kw = dict(lanes=(2 ** 31, 2 ** 15, 2 ** 12, 2 ** 11), color_cap=64)
    s, m = flow_for(bodies, kw, joints, info, hulls=hw), model_for(bodies, kw, joints, info)
    m.set_hulls(hw)
    for _ in range(20):
        s.step()
        m.step(1)
    same_bodies(s, m)
    same_contacts(s, m)
    assert np.array_equal(s.read_joints().astype(np.float64), m.joints().astype(np.float32).astype(np.float64))
    no_give_up(s)
    s.destroy()


def test_flow_reuse_substeps_with_joints_breaking(gpu):
    say("case: flow sweep, 4 substeps with reuse, breakable chains, 4 steps")
    bodies, joints, info, hw = scenes.large_scene(300, seed=9, chains=2, links=4, workers=1)
    joints = joints.copy()
    joints[1::4, 11] = -4.0
    kw = dict(lanes=(2 ** 31, 2 ** 15, 2 ** 12, 2 ** 11), color_cap=32, dt=1 / 240)
    s, m = flow_for(bodies, kw, joints, info, hulls=hw), model_for(bodies, kw, joints, info)
    m.set_hulls(hw)
    for _ in range(4):
        s.steps(1, 4)
        for u in range(4):
            m.step(1, reuse_manifolds=u > 0)
    assert (m.joints()[:, 3] != 0).sum() == len(joints) - 2
    same_bodies(s, m)
    assert np.array_equal(s.read_joints().astype(np.float64), m.joints().astype(np.float32).astype(np.float64))
    same_contacts(s, m)
    no_give_up(s)
    s.destroy()


@pytest.mark.parametrize("lanes", [1, 8])
def test_flow_clashing_colours_run_in_serial_chunks(gpu, lanes):
    """At colour cap 2 the pyramid's colouring clashes: the last colour runs serially, chunk after
    chunk, neighbours in one chunk reading each other's previous poses."""
    say(f"case: flow sweep, pyramid at colour cap 2 (clashing), {lanes} lanes, then cap 16")
    bodies = scenes.pyramid(10)[0]
    rule = LANES1 if lanes == 1 else (1 << 20, 1 << 19, 1 << 18, 0)
    kw = dict(lanes=rule, color_cap=2)
    s, m = flow_for(bodies, kw), model_for(bodies, kw)
    for _ in range(6):
        s.step()
        m.step(1)
    c = s.read_counters()
    assert c["clashes"] > 0 and c["clashing_colors"] != [0, 0]
    same_bodies(s, m)
    p = s.params
    p.color_cap = 16
    s.set_params(p)
    m.set_params(oracle.make_params(gravity=-9.81, color_cap=16, lanes=rule))
    for _ in range(4):
        s.step()
        m.step(1)
    same_bodies(s, m)
    no_give_up(s)
    s.destroy()


def test_flow_batched_worlds(gpu):
    """24 batched worlds: the sweep equals the per-colour launches of the same batch (each world
    steps as in a solver of its own under those: test_gpu_parity)."""
    say("case: flow sweep, 24 batched worlds x 15 steps vs per-colour launches")
    bodies, starts = scenes.batched_worlds(24, lo=8, hi=160)
    pk = dict(gravity=-9.81, color_cap=32, primal_lanes=(2 ** 31, 2 ** 15, 2 ** 12, 2 ** 11))
    a = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(**pk), world_start=starts)
    b = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(**pk), world_start=starts)
    a.set_schedule(**FLOW)
    b.set_schedule(cooperative=0, threads=512, blocks_per_cu=1)
    for _ in range(15):
        a.step()
        b.step()
    assert np.array_equal(a.read_bodies().view(np.uint32), b.read_bodies().view(np.uint32))
    A, B = sorted_contacts(*a.read_contacts_raw()), sorted_contacts(*b.read_contacts_raw())
    assert len(A) == len(B) > 0
    for (ma, ca), (mb, cb) in zip(A, B):
        assert (ma[:2] == mb[:2]).all() and (ma[4:8] == mb[4:8]).all() and (ca == cb).all()
    no_give_up(a)
    a.destroy()
    b.destroy()


def test_flow_give_up_is_reported_and_the_solver_drops_to_per_colour_launches(gpu):
    say("case: flow sweep, 400-body pile, wait poll limit 1 for one step")
    bodies = scenes.random_pile(400, seed=6, bin_size=10.0)[0]
    kw = dict(lanes=(1 << 20, 1 << 19, 1 << 18, 0), color_cap=24)
    s = flow_for(bodies, kw)
    s.steps(20)
    m = oracle_from_gpu(s)
    s.step()
    m.step(1)
    same_bodies(s, m)
    s.set_barrier_polls(1)
    s.step()  # a wait gives up within a few polls; the launch still ends
    c = s.read_counters()
    assert c["overflow"] & avbd_amd.solver.OVERFLOW_BARRIER
    assert s.schedule["cooperative"] == 0
    s.set_barrier_polls(0)
    m = oracle_from_gpu(s)
    for _ in range(3):
        s.step()
        m.step(1)
    same_bodies(s, m)  # per-colour launches from the state the unordered step left
    s.set_schedule(**FLOW)  # and the sweep is exact again
    for _ in range(3):
        s.step()
        m.step(1)
    same_bodies(s, m)
    no_give_up(s)
    s.destroy()


TAIL = dict(cooperative=3, threads=64, blocks_per_cu=4)
LEAD = dict(cooperative=4, threads=64, blocks_per_cu=4)


@pytest.mark.parametrize("sched", [3, 4])
@pytest.mark.parametrize("first", [0, 1, 3, 64])
def test_tail_pile_matches(gpu, first, sched):
    """Schedule 3: colours before `first` launched one by one, the rest and the dual in the
    dataflow tail (0: the whole sweep; 64: only the dual). Schedule 4: the colours before `first`
    as one dataflow launch (k_flow_lead) instead."""
    say(f"case: schedule {sched}, dataflow tail from colour {first}, pile 400 x 12 steps, eight threads")
    bodies = scenes.random_pile(400, seed=4, bin_size=10.0)[0]
    kw = dict(lanes=(1 << 20, 1 << 19, 1 << 18, 0), color_cap=64)
    pk = dict(kw, gravity=-9.81)
    pk["primal_lanes"] = pk.pop("lanes")
    s = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(**pk))
    s.set_schedule(**(TAIL if sched == 3 else LEAD))
    s.set_sparse_tail(first)
    m = model_for(bodies, kw)
    for _ in range(12):
        s.step()
        m.step(1)
    same_bodies(s, m)
    same_contacts(s, m)
    assert s.schedule["cooperative"] == sched
    assert s.read_counters()["overflow"] & avbd_amd.solver.OVERFLOW_BARRIER == 0
    s.destroy()


@pytest.mark.parametrize("sched", [3, 4])
@pytest.mark.parametrize("first", [1, 2])
def test_tail_chains_reuse_and_breaking_joints(gpu, first, sched):
    say(f"case: schedule {sched}, dataflow tail from colour {first}, 4 substeps with reuse, breakable chains, 4 steps")
    bodies, joints, info, hw = scenes.large_scene(300, seed=9, chains=2, links=4, workers=1)
    joints = joints.copy()
    joints[1::4, 11] = -4.0
    kw = dict(lanes=(2 ** 31, 2 ** 15, 2 ** 12, 2 ** 11), color_cap=32, dt=1 / 240)
    pk = dict(kw, gravity=-9.81)
    pk["primal_lanes"] = pk.pop("lanes")
    s = avbd_amd.GpuSolver3D(bodies, joints, info, params=avbd_amd.default_params(**pk), hulls=hw)
    s.set_schedule(**(TAIL if sched == 3 else LEAD))
    s.set_sparse_tail(first)
    m = model_for(bodies, kw, joints, info)
    m.set_hulls(hw)
    for _ in range(4):
        s.steps(1, 4)
        for u in range(4):
            m.step(1, reuse_manifolds=u > 0)
    assert (m.joints()[:, 3] != 0).sum() == len(joints) - 2
    same_bodies(s, m)
    assert np.array_equal(s.read_joints().astype(np.float64), m.joints().astype(np.float32).astype(np.float64))
    same_contacts(s, m)
    s.destroy()


@pytest.mark.parametrize("sched", [3, 4])
def test_tail_clashing_colours_and_adaptive_start(gpu, sched):
    """A clashing last colour inside the tail (pyramid at cap 2, tail from colour 1), then the
    automatic start under adaptive colouring (readbacks every 10 steps) on a pile."""
    say(f"case: schedule {sched}, dataflow tail, pyramid at colour cap 2 from colour 1; then a pile with the automatic start")
    bodies = scenes.pyramid(10)[0]
    rule = (1 << 20, 1 << 19, 1 << 18, 0)
    kw = dict(lanes=rule, color_cap=2)
    pk = dict(kw, gravity=-9.81)
    pk["primal_lanes"] = pk.pop("lanes")
    s = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(**pk))
    s.set_schedule(**(TAIL if sched == 3 else LEAD))
    s.set_sparse_tail(1)
    m = model_for(bodies, kw)
    for _ in range(6):
        s.step()
        m.step(1)
    assert s.read_counters()["clashes"] > 0
    same_bodies(s, m)
    s.destroy()
    bodies = scenes.random_pile(1000, seed=9, bin_size=14.0)[0]
    pk = dict(gravity=-9.81, color_cap=16, primal_lanes=(2 ** 31, 2 ** 15, 2 ** 12, 2 ** 11),
              flags=avbd_amd.FLAG_DEFAULTS | avbd_amd.solver.FLAG_ADAPT_COLORS)
    a = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(**pk))
    b = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(**pk))
    a.set_schedule(**(TAIL if sched == 3 else LEAD))
    b.set_schedule(cooperative=0, threads=512, blocks_per_cu=1)
    for _ in range(3):
        a.steps(10)
        b.steps(10)
    assert np.array_equal(a.read_bodies().view(np.uint32), b.read_bodies().view(np.uint32))
    assert a.schedule["cooperative"] == sched
    a.destroy()
    b.destroy()


def test_lead_give_up_is_reported_and_the_solver_drops_to_per_colour_launches(gpu):
    """Schedule 4 with a wait limit of one poll: a body of the leading launch (or the tail) gives
    up, the launch still ends, the step is reported and the solver sweeps colour by colour from the
    state that step left, exactly; schedule 4 is exact again once set."""
    say("case: schedule 4 from colour 3, 400-body pile, wait poll limit 1 for one step")
    bodies = scenes.random_pile(400, seed=6, bin_size=10.0)[0]
    kw = dict(lanes=(1 << 20, 1 << 19, 1 << 18, 0), color_cap=24)
    pk = dict(kw, gravity=-9.81)
    pk["primal_lanes"] = pk.pop("lanes")
    s = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(**pk))
    s.set_schedule(**LEAD)
    s.set_sparse_tail(3)
    s.steps(20)
    assert s.read_counters()["overflow"] & avbd_amd.solver.OVERFLOW_BARRIER == 0
    assert s.schedule["cooperative"] == 4
    s.set_barrier_polls(1)
    s.step()
    c = s.read_counters()
    assert c["overflow"] & avbd_amd.solver.OVERFLOW_BARRIER
    assert s.schedule["cooperative"] == 0
    s.set_barrier_polls(0)
    m = oracle_from_gpu(s)
    for _ in range(3):
        s.step()
        m.step(1)
    same_bodies(s, m)
    s.set_schedule(**LEAD)
    for _ in range(3):
        s.step()
        m.step(1)
    same_bodies(s, m)
    assert s.schedule["cooperative"] == 4
    assert s.read_counters()["overflow"] & avbd_amd.solver.OVERFLOW_BARRIER == 0
    s.destroy()
