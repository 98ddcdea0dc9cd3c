"""Known answers for the shapes the reference's CPU oracle does not have (spheres, hulls):
the hull builder on a box's corners, a box-shaped hull colliding as the box, sphere contacts
against closed forms, and a sphere rolling without slip (5/7 of its launch speed,
tests-gpu/avbd3d-gpu.gpu.test.ts:346-381), run through the oracle's restatement of the step."""
import math

import numpy as np

import oracle
from avbd_amd import records, scenes
from avbd_amd.hull import HullBuffer, convex_hull


def test_box_corner_hull_has_box_topology_and_mass():
    pts = np.array([[x, y, z] for x in (-0.5, 0.5) for y in (-1, 1) for z in (-1.5, 1.5)])
    h = convex_hull(pts)
    assert len(h["vertices"]) == 8
    assert sorted(len(f[2]) for f in h["faces"]) == [4] * 6
    assert len(h["edges"]) == 12
    assert abs(h["volume"] - 6.0) < 1e-12
    # principal moments per unit density of a 1 x 2 x 3 box: m/12 (b^2 + c^2), m = 6
    assert np.allclose(sorted(h["moments"]), sorted([6 / 12 * (4 + 9), 6 / 12 * (1 + 9), 6 / 12 * (1 + 4)]))
    assert np.allclose(sorted(h["size"]), [1, 2, 3])
    for n, d, loop in h["faces"]:
        assert abs(np.linalg.norm(n) - 1) < 1e-12
        for v in loop:
            assert abs(np.dot(n, h["vertices"][v]) - d) < 1e-6
    # every edge separates two faces that both hold it
    for a, b, f0, f1 in h["edges"]:
        for f in (f0, f1):
            assert a in h["faces"][f][2] and b in h["faces"][f][2]


def _contacts_world(model, bodies):
    man, con = model.contacts()
    out = []
    for m in man:
        a, b, first, cnt = int(m[0]), int(m[1]), int(m[2]), int(m[3]) & 15
        for k in range(first, first + cnt):
            out.append((a, b, con[k, 0:3].copy()))
    return out


def test_box_shaped_hull_rests_like_the_box():
    """A hull of a box's corners on the ground gets the box's four corner contacts."""
    pts = np.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)])
    shape = convex_hull(pts)
    hb = HullBuffer()
    at = hb.add(shape)
    results = []
    for as_hull in (False, True):
        g = scenes.ground(-0.5)
        # rotate the principal frame back onto the box's axes so both bodies sit identically
        rot = np.array([0, 0, 0, 1.0])
        if as_hull:
            body = scenes.hull_record((0, 0, 0.495), shape, at, 1.0, rot=tuple(shape["rotation"]))
        else:
            body = records.box_record((0, 0, 0.495), (1, 1, 1), 1.0, rot=tuple(rot))
        m = oracle.Model(64)
        m.set_params(oracle.make_params(iterations=0, gravity=0.0, flags=0))
        m.set_bodies(np.array([g, body]))
        m.set_hulls(hb.words())
        m.collide()
        man, con = m.contacts()
        assert len(man) == 1 and int(man[0][3]) & 15 == 4
        q = body[4:8]
        world = []
        for k in range(4):
            rA = con[k, 0:3]
            u = q[:3]
            t = 2 * np.cross(u, rA)
            world.append(np.round(rA + q[3] * t + np.cross(u, t) + body[0:3], 6))
        results.append(sorted(map(tuple, world)))
        assert np.allclose(man[0][4:7], [0, 0, 1], atol=1e-9)
    assert np.allclose(results[0], results[1], atol=1e-6)


def test_sphere_contacts_closed_form():
    a = records.sphere_record((0, 0, 0.9), 0.5, 1.0)
    b = records.sphere_record((0.3, 0, 0), 0.6, 1.0)
    g = scenes.ground(-0.5)
    m = oracle.Model(64)
    m.set_params(oracle.make_params(iterations=0, gravity=0.0, flags=0))
    m.set_bodies(np.array([g, b, a]))
    m.collide()
    man, con = m.contacts()
    by = {(int(x[0]), int(x[1])): (x, con[int(x[2])]) for x in man}
    # sphere-sphere: normal along the centre line (B to A), anchors at the radii
    mm, c = by[(2, 1)]
    d = a[0:3] - b[0:3]
    n = d / np.linalg.norm(d)
    assert np.allclose(mm[4:7], n, atol=1e-12)
    assert np.allclose(c[0:3], -n * 0.5, atol=1e-12) and np.allclose(c[4:7], n * 0.6, atol=1e-12)
    # C0 normal = separation + margin
    assert abs(c[7] - ((np.linalg.norm(d) - 1.1) + 0.01)) < 1e-12
    # sphere on the ground plane: contact below its centre, depth = overlap
    mm, c = by[(1, 0)]
    assert np.allclose(mm[4:7], [0, 0, 1], atol=1e-12)
    assert abs(c[7] - ((0.0 - 0.6) + 0.01)) < 1e-12


def test_sphere_rolls_at_five_sevenths():
    """A sphere launched at 3 m/s on the ground ends rolling at 5/7 of it, without slip."""
    ball = records.sphere_record((0, 0, 0.495), 0.5, 1.0, friction=0.5, vel=(3, 0, 0))
    m = oracle.Model(32)
    m.set_params(oracle.make_params(gravity=-10.0, color_cap=8))
    m.set_bodies(np.array([scenes.ground(-0.5), ball]))
    m.step(120)
    b = m.bodies()[1]
    v = b[32]
    w = b[37]
    assert abs(v - 5 / 7 * 3) < 0.05, v
    assert abs(v - w * 0.5) < 0.02, (v, w)
