"""The adapt policy (GpuSolver3D.adapt / adaptColors, src/avbd3d/gpu/solver.ts:2362-2407) as a
known-answer table: counters readbacks fed through avbd_adapt one after another, and the colour
cap, Jones-Plassmann rounds and capacities the policy must leave, worked out by hand from the
reference's rules (COLOR_SPARE 3, shrink after three quiet reads, rounds doubled on a clash and
cut by 4 to no fewer than 4; GROW_AT 0.8, GROW_TO 2, at least 1.25x, manifolds never past pairs)."""
import numpy as np
import pytest

import avbd_amd
from avbd_amd import records

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def counters(colors=0, clashes=0, overflow=0, pairs=0, manifolds=0, contacts=0):
    return {"pairs": pairs, "contacts": contacts, "manifolds": manifolds, "overflow": overflow, "clashes": clashes,
            "colors": colors, "color_cap": 0, "step": 0, "clashing_colors": (0, 0)}


def small_scene():
    return np.array([records.box_record((3.0 * k, 0.0, 0.5), (1, 1, 1), 1.0) for k in range(8)] +
                    [records.box_record((10.0, 0.0, -0.5), (40, 10, 1), 0.0)])


# (counters, colour cap, rounds) after each readback, from cap 12 and 16 rounds
COLOUR_TABLE = [
    (dict(colors=5), 12, 16),            # 5 + 3 + 2 < 12: a quiet read (vote 1)
    (dict(colors=5), 12, 16),            # vote 2
    (dict(colors=5), 8, 12),             # vote 3: cap 5 + 3, rounds 16 - 4
    (dict(colors=5), 8, 12),             # 10 < 8 no, but rounds 12 > 4: vote 1
    (dict(colors=5), 8, 12),
    (dict(colors=5), 8, 8),              # vote 3: rounds 8 (the cap stays 5 + 3)
    (dict(colors=5), 8, 8),
    (dict(colors=5), 8, 8),
    (dict(colors=5), 8, 4),
    (dict(colors=5), 8, 4),              # neither: votes reset
    (dict(colors=7), 12, 4),             # a spare short (8 - 7 < 2): max(7 + 5, 8 + 4)
    (dict(colors=7, clashes=1), 16, 8),  # a clash: rounds doubled, max(12, 12 + 4)
    (dict(colors=62), 64, 8),            # min(64, max(67, 20))
    (dict(colors=64, clashes=3), 64, 16),
    (dict(colors=5), 64, 16),            # vote 1
    (dict(colors=61), 64, 16),           # 64 - 61 = 3 spares and 66 > 64, but rounds 16 > 4: vote 2
    (dict(colors=63), 64, 16),           # a spare short: grow (capped at 64), votes reset
    (dict(colors=5), 64, 16),
    (dict(colors=5), 64, 16),
    (dict(colors=5), 8, 12),             # three quiet reads since the reset
    (dict(colors=40, clashes=2), 45, 24),  # max(40 + 5, 8 + 4)
    (dict(colors=40, clashes=2), 49, 32),  # rounds capped at 32: max(45, 45 + 4)
]


def test_adapt_colours_known_answers(gpu):
    say("case: adaptColors known-answer table, 22 readbacks")
    s = avbd_amd.GpuSolver3D(small_scene(), params=avbd_amd.default_params(color_cap=12, color_rounds=16))
    for k, (c, cap, rounds) in enumerate(COLOUR_TABLE):
        s.adapt(counters(**c))
        p = s.params
        assert (p.color_cap, p.color_rounds) == (cap, rounds), (k, c, p.color_cap, p.color_rounds)
    s.destroy()


def test_adapt_storage_known_answers(gpu):
    say("case: adapt's capacity half, known-answer table")
    small = {"pairs": 1000, "manifolds": 800, "contacts": 2000}
    s = avbd_amd.GpuSolver3D(small_scene(), capacity=small)
    steps = [
        # (counters, pairs, manifolds, contacts)
        (dict(pairs=700, manifolds=600, contacts=1500), 1000, 800, 2000),   # all at or under 0.8
        (dict(pairs=700, manifolds=600, contacts=1700), 1000, 800, 3400),   # contacts > 0.8: 2 x 1700
        (dict(pairs=100, manifolds=700, contacts=100), 1000, 1000, 3400),   # manifolds: max(1400, 1000), at most the pairs
        (dict(pairs=100, overflow=1), 1250, 1000, 3400),                     # overflowed pairs: 1.25 x 1000
        (dict(pairs=1100, manifolds=1100, contacts=100), 2200, 2200, 3400),  # 1100 > 1000: 2 x 1100 each
        (dict(contacts=10, overflow=2), 2200, 2200, 4250),                   # overflowed contacts: 1.25 x 3400
    ]
    for k, (c, pairs, manifolds, contacts) in enumerate(steps):
        s.adapt(counters(colors=3, **c))
        cap = s.capacity()
        assert (cap["pairs"], cap["manifolds"], cap["contacts"]) == (pairs, manifolds, contacts), (k, c, cap)
    s.step()  # the grown buffers step
    assert s.read_counters()["overflow"] == 0
    s.destroy()
    # fixed storage keeps its capacities; the colour half still adapts
    f = avbd_amd.GpuSolver3D(small_scene(), capacity=small,
                             params=avbd_amd.default_params(color_cap=12, flags=avbd_amd.FLAG_DEFAULTS | avbd_amd.solver.FLAG_FIXED_STORAGE))
    f.adapt(counters(colors=11, pairs=5000, manifolds=5000, contacts=5000, overflow=7))
    cap = f.capacity()
    assert (cap["pairs"], cap["manifolds"], cap["contacts"]) == (1000, 800, 2000)
    assert f.params.color_cap == 16
    f.destroy()
