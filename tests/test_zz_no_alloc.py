"""avbd_read_bodies_noalloc, the readback that allocates nothing, called through the C ABI: the
same records as avbd_read_bodies on two calls in a row (the second smaller), ranges past its
fixed maximum refused before anything is read, and two solvers on one GPU sharing its array."""
import ctypes
import sys

import numpy as np
import pytest

import avbd_amd
from avbd_amd import scenes

pytestmark = pytest.mark.gpu
MAX = avbd_amd.solver.READ_NOALLOC_MAX


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc(s, first, count, out=None):
    L = avbd_amd.lib()
    out = np.full((count, 40), np.nan, np.float32) if out is None else out
    rc = L.avbd_read_bodies_noalloc(s.h, out.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint32(first), ctypes.c_uint32(count))
    return rc, out


def test_noalloc_readback_equals_the_staged_readback(gpu):
    say(f"case: read_bodies_noalloc vs read_bodies, pile of 20,000 bodies, {MAX} then 1000 records")
    bodies, _, _ = scenes.random_pile(20000, seed=4, bin_size=60.0)
    s = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(color_cap=32))
    s.steps(5)
    rc, a = _noalloc(s, 100, MAX)  # before the solver's own staging buffer exists
    assert rc == 0, avbd_amd.lib().avbd_last_error(s.h)
    b = s.read_bodies(100, MAX)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    s.step()
    rc, a2 = _noalloc(s, 7, 1000)  # smaller, after a step: the array still holds the last call's records
    assert rc == 0
    b2 = s.read_bodies(7, 1000)
    assert np.array_equal(a2.view(np.uint32), b2.view(np.uint32))
    assert not np.array_equal(a2[:10].view(np.uint32), a[:10].view(np.uint32))  # the step moved them
    s.destroy()


def test_noalloc_readback_refuses_past_its_maximum(gpu):
    say("case: read_bodies_noalloc refuses more than its maximum and ranges past the bodies")
    bodies, _, _ = scenes.random_pile(MAX + 10, seed=5, bin_size=60.0)
    s = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(color_cap=32))
    out = np.zeros((MAX + 1, 40), np.float32)
    rc, _ = _noalloc(s, 0, MAX + 1, out)
    assert rc == -1
    assert "at most" in avbd_amd.lib().avbd_last_error(s.h).decode()
    assert not out.any()
    rc, _ = _noalloc(s, len(bodies) - 5, 10, np.zeros((10, 40), np.float32))
    assert rc == -1
    rc, last = _noalloc(s, len(bodies) - 5, 5)
    assert rc == 0 and np.array_equal(last.view(np.uint32), s.read_bodies(len(bodies) - 5, 5).view(np.uint32))
    s.destroy()


def test_two_solvers_on_one_gpu_share_the_array(gpu):
    say("case: read_bodies_noalloc from two solvers in turn on one GPU")
    one = avbd_amd.GpuSolver3D(scenes.random_pile(3000, seed=1, bin_size=30.0)[0], params=avbd_amd.default_params(color_cap=32))
    two = avbd_amd.GpuSolver3D(scenes.pyramid(10)[0], params=avbd_amd.default_params(color_cap=32))
    for _ in range(3):
        one.step()
        two.step()
        rc1, a = _noalloc(one, 0, one.body_count)
        rc2, b = _noalloc(two, 0, two.body_count)
        assert rc1 == 0 and rc2 == 0
        assert np.array_equal(a.view(np.uint32), one.read_bodies().view(np.uint32))
        assert np.array_equal(b.view(np.uint32), two.read_bodies().view(np.uint32))
    one.destroy()
    two.destroy()
