"""The broadphase's grid against the predicate over every pair (tests-gpu/avbd3d-gpu.gpu.test.ts:
144-207's method) and the oracle: a moving body probes its cell and half of its neighbours for
movers, a fixed small body its 27 cells for movers, and a scene whose moving bodies are much
smaller than its small fixed ones runs a fine level beside the coarse one (uploadStatics,
solver.ts:1274-1290), pairs across the levels found from the finer body. Pair sets must be exact;
whole steps bit-exact with the oracle."""
import sys

import numpy as np
import pytest

import avbd_amd
import oracle
from avbd_amd import records, scenes
from helpers import brute_force_pairs, lcg, oracle_from_gpu

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def slab_scene(slabs=150, small=300, medium=20, seed=21):
    """Fixed slabs (3 x 3 x 0.5 m: radius ~2.1) on a 14 x 14 grid of posts, more of them than the 64
    the large list takes; small moving boxes and spheres (radius ~0.2-0.4) falling onto them; a few
    medium moving boxes (radius ~1.9) that sit on the coarse level with the slabs; a ground."""
    rnd = lcg(seed)
    out = [scenes.ground(z=-0.5, size=(200.0, 200.0, 1.0))]
    for k in range(slabs):
        x, y = (k % 14) * 3.4 - 22.0, (k // 14) * 3.4 - 22.0
        out.append(records.box_record((x, y, 1.0 + 0.3 * (k % 3)), (3.0, 3.0, 0.5), 0.0))
    for k in range(small):
        x, y = (rnd() - 0.5) * 40.0, (rnd() - 0.5) * 40.0
        z = 3.0 + (k % 10) * 0.9
        if k % 2:
            out.append(records.sphere_record((x, y, z), 0.2 + 0.2 * rnd(), 1.0))
        else:
            s = 0.25 + 0.2 * rnd()
            out.append(records.box_record((x, y, z), (2 * s, 2 * s, 2 * s), 1.0))
    for k in range(medium):
        x, y = (rnd() - 0.5) * 36.0, (rnd() - 0.5) * 36.0
        out.append(records.box_record((x, y, 13.0 + 2.5 * (k % 4)), (2.2, 2.2, 2.2), 1.0))
    return np.asarray(out, np.float32)


def test_fine_level_pairs_and_steps_match(gpu):
    say("case: slab scene (fine + coarse grid), 80 steps vs oracle, pairs vs brute force")
    bodies = slab_scene()
    kw = dict(gravity=-9.81, color_cap=32, primal_lanes=(2 ** 31, 2 ** 15, 2 ** 12, 2 ** 11))
    s = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(**kw))
    info = s.broadphase_info()
    say(f"  grid: {info}")
    assert info["fine_cell_size"] < info["cell_size"] and info["coarse_movers"], info
    assert 0 < info["large"] <= 64
    m = oracle.Model(32)
    m.set_params(oracle.make_params(gravity=-9.81, color_cap=32, lanes=(2 ** 31, 2 ** 15, 2 ** 12, 2 ** 11)))
    m.set_bodies(records.bodies_to_oracle(bodies))
    checked = 0
    for k in range(80):
        s.step()
        m.step(1)
        if k in (0, 20, 40, 79):
            gp = s.read_pairs()
            got = {(int(a), int(b)) for a, b in gp}
            assert len(got) == len(gp)
            op = m.pairs()
            assert got == {(int(a), int(b)) for a, b in op}, (len(got), len(op))
            checked += len(got)
    assert checked > 200
    got, want = records.bodies_to_oracle(s.read_bodies()), m.bodies()
    assert np.array_equal(got, want)
    s.destroy()


def settled_pairs_match(s, steps):
    """After `steps` steps, one step's pairs: exactly the oracle's (the f32 predicate over every pair
    it finds by its own grid), and the f64 predicate over all pairs but for pairs within round-off of
    a threshold (resting bodies touch: their AABB gaps and face separations sit near the pads)."""
    s.steps(steps)
    before = s.read_bodies().copy()
    m = oracle_from_gpu(s)
    s.step()
    m.step(1)
    got = {(int(a), int(b)) for a, b in s.read_pairs()}
    assert got == {(int(a), int(b)) for a, b in m.pairs()}
    f64 = brute_force_pairs(before)
    assert len(got ^ f64) <= max(5, len(f64) // 50), (len(got - f64), len(f64 - got))
    return got


def test_fine_level_pairs_equal_the_predicate_over_all_pairs(gpu):
    say("case: slab scene settled 60 steps, then one step's pairs vs the predicate over all pairs")
    bodies = slab_scene(seed=23)
    kw = dict(gravity=-9.81, color_cap=32, primal_lanes=(2 ** 31, 2 ** 15, 2 ** 12, 2 ** 11))
    s = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(**kw))
    assert len(settled_pairs_match(s, 60)) > 100
    s.destroy()


def test_small_fixed_bodies_find_their_movers(gpu):
    """Fixed small bodies (posts under a pile) are in cells of their own: each probes the movers
    around it, so every fixed-moving pair is found from the fixed side."""
    say("case: pile over 120 small fixed posts, pairs vs the predicate over all pairs")
    rnd = lcg(5)
    out = [scenes.ground(z=-0.5, size=(60.0, 60.0, 1.0))]
    for k in range(120):
        out.append(records.box_record(((k % 12) * 1.5 - 9.0, (k // 12) * 1.5 - 7.0, 0.5), (0.6, 0.6, 1.0), 0.0))
    for k in range(400):
        x, y = (rnd() - 0.5) * 18.0, (rnd() - 0.5) * 15.0
        out.append(records.box_record((x, y, 1.4 + 0.7 * (k % 8)), (0.5, 0.5, 0.5), 1.0))
    bodies = np.asarray(out, np.float32)
    s = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(gravity=-9.81, color_cap=32))
    assert s.broadphase_info()["fine_cell_size"] == s.broadphase_info()["cell_size"]
    got = settled_pairs_match(s, 40)
    fixed_moving = {p for p in got if bodies[p[0], 19] <= 0 or bodies[p[1], 19] <= 0}
    assert len(fixed_moving) > 50
    s.destroy()
