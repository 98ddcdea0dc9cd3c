"""World's command passes on the GPU (avbd_push / avbd_move / avbd_reset through the C ABI:
Pusher, BodyMover, BodyReset) against the oracle's restatements applied to the same records read
back just before. Push and reset are float32 arithmetic in the same order: bit for bit. The move's
angular velocity goes through atan2 and tan, whose last bit may differ between the device's and
the host's libraries: 1e-6 relative."""
import numpy as np
import pytest

import avbd_amd
import oracle
from avbd_amd import scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def moving_pile(n=200, steps=20):
    bodies, _, _ = scenes.random_pile(n, seed=8, bin_size=8.0)
    s = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(gravity=-9.81))
    s.steps(steps)
    return s


def test_push_matches_the_oracle(gpu):
    say("commands: push on 60 bodies of a moving pile")
    s = moving_pile()
    rec = s.read_bodies()
    rng = np.random.default_rng(4)
    idx = rng.choice(len(rec), size=60, replace=False).astype(np.uint32)
    lin = rng.normal(size=(60, 3)).astype(np.float32) * 5
    off = lin * (rng.random((60, 1)) < 0.5)  # half of them pushed off-centre
    at = rec[idx, 0:3] + rng.normal(size=(60, 3)).astype(np.float32) * 0.3
    mom = np.cross(at, off).astype(np.float32)
    s.push(idx, lin, off, mom)
    got = s.read_bodies()
    rows = np.zeros((60, 12), np.float32)
    rows[:, 0:3], rows[:, 4:7], rows[:, 8:11] = lin, off, mom
    rows[:, 3] = idx.view(np.float32)
    want = oracle.push_pass(rec, rows)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    s.destroy()


def test_move_matches_the_oracle(gpu):
    say("commands: move 30 bodies toward target poses")
    s = moving_pile()
    rec = s.read_bodies()
    rng = np.random.default_rng(5)
    idx = rng.choice(len(rec), size=30, replace=False).astype(np.uint32)
    target = rec[idx, 0:3] + rng.normal(size=(30, 3)).astype(np.float32)
    q = rng.normal(size=(30, 4)).astype(np.float32)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    left = rng.integers(1, 6, size=30).astype(np.float32)
    dt = np.float32(1 / 60)
    s.move(idx, target, left, dt, rotations=q)
    got = s.read_bodies()
    rows = np.zeros((30, 12), np.float32)
    rows[:, 0:3], rows[:, 4:8], rows[:, 8], rows[:, 9], rows[:, 10] = target, q, left, dt, 1.0
    rows[:, 3] = idx.view(np.float32)
    want = oracle.move_pass(rec, rows)
    untouched = np.setdiff1d(np.arange(len(rec)), idx)
    assert np.array_equal(got[untouched].view(np.uint32), rec[untouched].view(np.uint32))
    assert np.array_equal(got[idx, 32:35].view(np.uint32), want[idx, 32:35].view(np.uint32))  # plain division
    w_got, w_want = got[idx, 36:39].astype(np.float64), want[idx, 36:39].astype(np.float64)
    assert np.abs(w_got - w_want).max() <= 1e-6 * max(1.0, np.abs(w_want).max())
    s.destroy()


def test_reset_matches_the_oracle(gpu):
    say("commands: reset 50 bodies")
    s = moving_pile()
    rec = s.read_bodies()
    idx = np.arange(0, len(rec), 4, dtype=np.uint32)
    s.reset_bodies(idx)
    assert np.array_equal(s.read_bodies().view(np.uint32), oracle.reset_pass(rec, idx).view(np.uint32))
    s.destroy()


def test_commands_refuse_what_would_race_or_miss(gpu):
    s = moving_pile(n=20, steps=1)
    with pytest.raises(ValueError, match="twice"):
        s.push([3, 3], np.ones((2, 3)))
    with pytest.raises(ValueError, match="no body slot"):
        s.move([len(s.read_bodies())], np.zeros((1, 3)), 1, 1 / 60)
    with pytest.raises(ValueError, match="no body slot"):
        s.reset_bodies([10 ** 6])
    s.destroy()
