"""The oracle's OpenMP step (the all-core CPU baseline) equals its one-thread step bit for bit:
bodies of a proper colour, contacts and pairs are independent, and a worker's manifolds are
appended in pair order. Clashing colours keep the serial chunked order."""
import numpy as np

import oracle
from avbd_amd import records, scenes


def _model(bodies, joints, info, hulls, threads, color_cap=32, rounds=16, dt=1 / 240):
    m = oracle.Model(32)
    m.set_params(oracle.make_params(dt=dt, gravity=-9.81, color_cap=color_cap, color_rounds=rounds,
                                    lanes=(2 ** 31, 2 ** 15, 2 ** 12, 2 ** 11)))
    m.set_bodies(records.bodies_to_oracle(bodies))
    if len(joints):
        m.set_joints(joints.astype(np.float64), info)
    if hulls is not None:
        m.set_hulls(hulls)
    m.set_threads(threads)
    return m


def _run(m, steps, substeps):
    for _ in range(steps):
        for u in range(substeps):
            m.step(1, reuse_manifolds=u > 0)


def test_threaded_step_equals_one_thread_on_hulls_spheres_boxes_chains():
    bodies, joints, info, hulls = scenes.large_scene(600, seed=3, chains=4, links=5, workers=1)
    one = _model(bodies, joints, info, hulls, 1)
    many = _model(bodies, joints, info, hulls, 4)
    _run(one, 80, 2)
    _run(many, 80, 2)
    assert one.counts()["contacts"] > 100
    np.testing.assert_array_equal(one.bodies(), many.bodies())
    np.testing.assert_array_equal(one.joints(), many.joints())
    for x, y in zip(one.contacts(), many.contacts()):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(one.colors(), many.colors())


def test_threaded_step_equals_one_thread_with_clashing_colours():
    """A colour cap of 2 forces clashes: those colours run in the serial chunked order."""
    bodies, joints, info = scenes.random_pile(400, seed=9, bin_size=12.0)
    one = _model(bodies, joints, info, None, 1, color_cap=2, rounds=4, dt=1 / 60)
    many = _model(bodies, joints, info, None, 4, color_cap=2, rounds=4, dt=1 / 60)
    _run(one, 60, 1)
    _run(many, 60, 1)
    assert one.counts()["clashes"] > 0
    np.testing.assert_array_equal(one.bodies(), many.bodies())
