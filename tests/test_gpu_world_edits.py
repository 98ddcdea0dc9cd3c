"""World's scene edits through the C ABI (World.flush, src/lib/world.ts:1861-1995): an empty solver
filled and edited the way World drives GpuSolver3D (addBodies, appendJoints/appendSprings/
appendWorldJoint, releaseJoints and slot reuse, setWorldAnchor, disableConstraint, setBodyAux,
setLifecycle, patchBody, rewriteFixed). The records it holds afterwards are checked against the
reference's packing rules (solver.ts:132-138, :1559-1745, :1103-1165) restated here, then it
steps bit for bit with the oracle given those records: plastic joints yielding and breaking,
springs, a world joint, scaled gravity and damping, disabled and cold bodies."""
import math
import sys

import numpy as np
import pytest

import avbd_amd
import oracle
from avbd_amd import records, scenes
from helpers import BIG, aux_factors, canonical_contacts, fresh_joint, new_box, oracle_from_gpu, spec_record

pytestmark = pytest.mark.gpu
NO_COLOR = 255


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("solve", ["coop", "flow"])
def test_world_fills_an_empty_solver_and_steps_like_the_oracle(gpu, monkeypatch, solve):
    say(f"case: World-style edits on an empty solver (pyramid, plastic chain, spring, world joint, aux, lifecycle), 90 steps, {solve} solve")
    monkeypatch.setenv("AVBD_SOLVE", solve)
    dt = 1 / 60
    p = avbd_amd.default_params(dt=dt, gravity=-9.81, color_cap=32, primal_lanes=(0, 0, 0, 0))
    s = avbd_amd.GpuSolver3D(np.zeros((0, 40), np.float32), params=p, capacity={"bodies": 64, "joints": 2})
    assert (s.body_count, s.joint_count) == (0, 0)

    # addBodies: ground + a 4-row pyramid, then an anchor, a cantilevered chain and a sphere
    pyr = scenes.pyramid(4)[0].astype(np.float32)
    pyr[:, 11] = 1.0
    assert s.add_bodies(pyr) == 0
    chain = [new_box((6.0, 0.0, 4.0), (1.0, 1.0, 1.0), 0.0)]
    chain += [new_box((7.0 + k, 0.0, 4.0), (1.0, 0.6, 0.6), 2.0) for k in range(5)]
    sph = records.sphere_record((-1.0, 3.0, 3.0), 0.4, 3.0, 0.5).astype(np.float32)
    sph[11] = 1.0
    base = s.add_bodies(np.vstack(chain + [sph]))
    assert base == len(pyr) and s.body_count == len(pyr) + 7
    anchor, links, ball = base, list(range(base + 1, base + 6)), base + 6
    bodies = np.vstack([pyr] + chain + [sph]).copy()
    sizes = bodies[:, 16:19]

    # appendJoints: a plastic joint at the anchor that bends past its yield and breaks once bent
    # past 0.25 rad, a ball joint, a joint holding a rest turn, one breaking on its linear force
    specs = [
        {"a": anchor, "b": links[0], "rA": (0.5, 0, 0), "rB": (-0.5, 0, 0), "yield": 1.0, "breakBend": 0.25},
        {"a": links[0], "b": links[1], "rA": (0.5, 0, 0), "rB": (-0.5, 0, 0), "angular": 0.0},
        {"a": links[1], "b": links[2], "rA": (0.5, 0, 0), "rB": (-0.5, 0, 0), "rest": (0, 0, 0.05, 1)},
        {"a": links[2], "b": links[3], "rA": (0.5, 0, 0), "rB": (-0.5, 0, 0), "fracture": 1e6, "linear": True},
        {"a": links[3], "b": links[4], "rA": (0.5, 0, 0), "rB": (-0.5, 0, 0)},
    ]
    slots = s.append_joints(specs)
    assert slots == [0, 1, 2, 3, 4] and s.joint_count == 5  # the joint storage grew from 2
    spring = s.append_springs([{"a": ball, "b": 5, "rA": (0, 0, 0), "rB": (0, 0, 0.5), "stiffness": 50.0, "rest": 1.5}])[0]
    drag = s.append_world_joint(9, (0.0, 0.5, 4.5), (0.0, 0.0, 0.5))
    s.set_world_anchor(drag, (0.3, 0.8, 4.6))
    # release the last chain joint; the next appendJoints takes its slot back
    s.release_joints([4, 4, 99])
    assert s.joint_info()[4].tolist() == [0, -1, -1, 0]
    again = s.append_joints([specs[4]])
    assert again == [4]
    s.disable_constraint(3)

    # per-body aux, lifecycle, patches, a fixed rewrite
    s.set_body_aux([ball], [0.5])
    s.set_body_aux([3, 4], [1.0, 2.0], [0.8, 0.0], [0.0, 1.5])
    s.set_lifecycle(8, 2)   # disabled
    s.set_lifecycle(links[4], 4)  # cold
    moved = bodies[2].copy()
    moved[32:35] = (0.0, 1.5, 0.0)
    s.patch_body(2, ["velocity"], moved)
    lifted = bodies[7].copy()
    lifted[0:3] += (0.0, 0.0, 0.02)
    s.patch_body(7, ["position"], lifted)
    s.rewrite_fixed([6], [(2.0, -3.0, 0.5)], (1.0, 2.0, 1.0), 0.7)

    # the records the edits wrote, as the reference packs them
    want_b = bodies.copy()
    want_b[2, 32:35] = (0.0, 1.5, 0.0)
    want_b[2, 11] = 1.0
    for o in (0, 8, 24):
        want_b[7, o:o + 3] = lifted[0:3]
    want_b[7, 11] = 1.0
    want_b[8, 11] = 3.0
    want_b[links[4], 11] = 5.0
    fixed = np.zeros(40, np.float32)
    fixed[0:3] = (2.0, -3.0, 0.5)
    fixed[3] = 0.7
    fixed[7] = fixed[11] = fixed[15] = fixed[31] = 1.0
    fixed[16:19] = (1.0, 2.0, 1.0)
    fixed[23] = np.float32(math.sqrt(0.25 + 1.0 + 0.25))
    want_b[6] = fixed
    got_b = s.read_bodies()
    assert np.array_equal(got_b.view(np.uint32), want_b.view(np.uint32)), np.nonzero((got_b != want_b).any(axis=1))[0]
    want_j = np.zeros((7, 40), np.float32)
    for k, sp in enumerate(specs):
        want_j[k] = spec_record(sp, sizes)
    want_j[3, 0:8] = 0.0
    want_j[3, 8:11] = 0.0
    want_j[3, 12:15] = 0.0
    want_j[spring, 3] = 50.0
    want_j[spring, 28:31] = (0, 0, 0.5)
    want_j[spring, 27] = 1.5
    want_j[drag] = fresh_joint()
    want_j[drag, 3] = BIG
    want_j[drag, 24:27] = (0.3, 0.8, 4.6)
    want_j[drag, 28:31] = (0, 0, 0.5)
    want_j[drag, 15] = np.float32(sum(float(x) ** 2 for x in sizes[9]))
    got_j = s.read_joints()
    assert np.array_equal(got_j.view(np.uint32), want_j.view(np.uint32)), np.nonzero((got_j != want_j).any(axis=1))[0]
    info = np.array([[1, anchor, links[0], 0], [1, links[0], links[1], 0], [1, links[1], links[2], 0], [1, links[2], links[3], 0],
                     [1, links[3], links[4], 0], [2, ball, 5, 0], [1, -1, 9, 0]], np.int32)
    assert np.array_equal(s.joint_info(), info)
    assert np.array_equal(s.read_colors()[0], np.full(s.body_count, NO_COLOR, np.uint32))

    # the oracle from those records
    aux = np.zeros((s.body_count, 4))
    aux[:, 0:3] = 1.0
    aux[ball, 0] = 0.5
    aux[3] = (1.0, aux_factors(0.8, dt), aux_factors(0.0, dt), 1.0)
    aux[4] = (2.0, aux_factors(0.0, dt), aux_factors(1.5, dt), 1.0)
    m = oracle.Model(32)
    m.set_params(oracle.make_params(dt=p.dt, gravity=p.gravity, color_cap=32, lanes=(0, 0, 0, 0)))
    m.set_bodies(records.bodies_to_oracle(want_b))
    m.set_joints(want_j.astype(np.float64), info)
    m.set_aux(aux)
    yielded = broke = False
    for k in range(90):
        s.step()
        m.step(1)
        j = s.read_joints()
        yielded |= not np.array_equal(j[0, 32:36], want_j[0, 32:36])  # the plastic joint's rest moved
        broke |= j[0, 3] == 0.0
    assert yielded and broke
    assert np.array_equal(records.bodies_to_oracle(s.read_bodies()), m.bodies())
    assert np.array_equal(s.read_joints().astype(np.float64), m.joints().astype(np.float32).astype(np.float64))
    gm, gc = s.read_contacts_raw()
    om, oc = records.contacts_from_oracle(*m.contacts())
    GM, GC = canonical_contacts(np.asarray(gm, np.int64), gc)
    OM, OC = canonical_contacts(np.asarray(om, np.int64), oc)
    assert np.array_equal(GM, OM) and np.array_equal(GC, OC)
    b = s.read_bodies()
    assert np.array_equal(b[8, 0:3], want_b[8, 0:3])  # disabled: untouched

    # a second round of edits mid-run, then stepping on from the GPU's state
    s.set_lifecycle(8, 0)
    s.release_joints([spring])
    more = s.add_bodies(np.vstack([new_box((0.0, -2.0, 2.0 + k), (0.8, 0.8, 0.8), 1.0) for k in range(3)]))
    assert more == ball + 1
    s.set_body_aux([more], [0.25])
    m2 = oracle_from_gpu(s)
    aux2 = np.zeros((s.body_count, 4))
    aux2[:, 0:3] = 1.0
    aux2[:len(aux)] = aux
    aux2[more, 0] = 0.25
    m2.set_aux(aux2)
    for _ in range(30):
        s.step()
        m2.step(1)
    assert np.array_equal(records.bodies_to_oracle(s.read_bodies()), m2.bodies())
    s.destroy()


def test_edits_refuse_what_the_reference_refuses(gpu):
    say("case: scene edits refuse bad input before anything changes")
    p = avbd_amd.default_params(flags=avbd_amd.FLAG_DEFAULTS | avbd_amd.solver.FLAG_FIXED_STORAGE)
    s = avbd_amd.GpuSolver3D(scenes.pyramid(2)[0], params=p, capacity={"bodies": 5, "joints": 1})
    n = s.body_count
    j = {"a": 1, "b": 2, "rA": (0, 0, 0.5), "rB": (0, 0, -0.5)}
    for bad, msg in [({**j, "b": 9}, "index of a body"), ({**j, "fracture": -1.0}, "fracture"),
                     ({**j, "rest": (0, 0, 0.1, 1), "angular": 0.0}, "rest needs an angle lock"),
                     ({**j, "breakBend": 0.3}, "needs a yield"), ({**j, "yield": 5.0, "breakBend": 4.0}, "0 to pi"),
                     ({**j, "yield": 5.0, "angular": 3.0}, "rigid angle lock"), ({**j, "yield": -1.0}, "at least 0")]:
        with pytest.raises(ValueError, match=msg):
            s.append_joints([bad])
    assert s.joint_count == 0
    assert s.append_joints([j]) == [0]
    with pytest.raises(avbd_amd.StorageCapacityError):
        s.append_joints([j])  # fixed storage: one joint slot
    s.release_joints([0])
    assert s.append_joints([j]) == [0]  # the released slot comes back
    assert s.add_bodies(np.vstack([new_box((5, 5, 1), (1, 1, 1), 1.0)] * 3)) == n  # as many as fit: one
    assert s.body_count == 5
    assert s.add_bodies(np.vstack([new_box((5, 5, 1), (1, 1, 1), 1.0)])) == -1
    with pytest.raises(ValueError, match="lifecycle"):
        s.set_lifecycle(0, 1)
    with pytest.raises(ValueError, match="finite"):
        s.set_body_aux([0], [float("nan")])
    with pytest.raises(ValueError, match="both damping"):
        s.set_body_aux([0], [1.0], [0.1], None)
    sail = new_box((0, 0, 3), (1, 1, 0.01), 1.0)
    sail[39] = 2.0
    with pytest.raises(ValueError, match="sails"):
        s.write_bodies(sail[None], 1)
    s.destroy()
