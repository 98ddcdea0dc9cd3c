"""Contact events on the GPU (avbd_watch_* through the C ABI: ContactWatch, src/lib/contacts.ts)
against the oracle's restatement of its two passes run on the same step's read-back records. Same
events bit for bit (point, impulse and normal as float32 words), compared as sorted lists: the GPU
appends with atomics, so its order within a step is the order lanes reached the counter."""
import numpy as np
import pytest

import avbd_amd
import oracle
from avbd_amd import scenes

pytestmark = pytest.mark.gpu

REPORT = 0x80000000
ALL = 1 << 30


def say(*a):
    print(*a, flush=True)


def key(e):
    return (e["step"], e["kind"], e["a"], e["b"])


def expected(s, prev, watch, step, dt):
    man, con = s.read_contacts_raw()
    return oracle.contact_events(prev, man, con, s.read_bodies(), watch, ALL, step, np.float32(dt))


def watched(n, every=3):
    idx = np.arange(n, dtype=np.uint32)
    ids = idx + 1
    on = (idx % every == 0).astype(np.uint8)
    return idx, ids, on, (ids | np.where(on > 0, REPORT, 0)).astype(np.uint32)


def test_events_match_the_oracle_over_a_falling_pile(gpu):
    bodies, _, _ = scenes.random_pile(120, seed=3, bin_size=8.0)
    dt = 1 / 60
    say(f"contact events: pile of {len(bodies)} bodies, 120 steps, every third body reporting")
    s = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(dt=dt, gravity=-9.81))
    s.watch_create(pairs=4096, capacity=8192)
    idx, ids, on, watch = watched(len(bodies))
    s.set_watched(idx, ids, on)
    prev, begins, ends = set(), 0, 0
    for step in range(1, 121):
        s.step()
        s.watch_run(ALL, step, dt)
        prev, want = expected(s, prev, watch, step, dt)
        got, dropped, full = s.read_contact_events()
        assert dropped == 0 and not full
        assert sorted(got, key=key) == sorted(want, key=key), f"step {step}"
        begins += sum(e["kind"] == avbd_amd.solver.EVENT_BEGIN for e in got)
        ends += sum(e["kind"] == avbd_amd.solver.EVENT_END for e in got)
    say(f"  {begins} begins, {ends} ends")
    assert begins > 20 and ends > 0
    s.destroy()


def test_events_accumulate_between_reads_and_end_when_bodies_part(gpu):
    bodies, _, _ = scenes.pyramid(4)
    dt = 1 / 60
    s = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(dt=dt, gravity=-9.81))
    s.watch_create(pairs=64, capacity=256)
    idx, ids, on, watch = watched(len(bodies), every=1)
    s.set_watched(idx, ids, on)
    prev, want = set(), []
    for step in range(1, 4):  # three runs, one read: the events of all three, in step order
        s.step()
        s.watch_run(ALL, step, dt)
        prev, w = expected(s, prev, watch, step, dt)
        want += w
    got, dropped, full = s.read_contact_events()
    assert dropped == 0 and not full and len(got) > 0
    assert sorted(got, key=key) == sorted(want, key=key)
    assert [e["step"] for e in got] == sorted(e["step"] for e in got)
    # the top box is lifted away: its pairs end
    rec = s.read_bodies()
    top = int(np.argmax(rec[:, 2]))
    rec[top, 2] += 10.0
    s.write_bodies(rec)
    s.step()
    s.watch_run(ALL, 4, dt)
    prev, want = expected(s, prev, watch, 4, dt)
    got, _, _ = s.read_contact_events()
    assert sorted(got, key=key) == sorted(want, key=key)
    assert any(e["kind"] == avbd_amd.solver.EVENT_END and top + 1 in (e["a"], e["b"]) for e in got)
    # a new watch starts with empty sets: every touching pair begins again
    s.watch_create(pairs=64, capacity=256)
    s.step()
    s.watch_run(ALL, 5, dt)
    _, want = expected(s, set(), watch, 5, dt)
    got, _, _ = s.read_contact_events()
    assert sorted(got, key=key) == sorted(want, key=key) and all(e["kind"] == 1 for e in got)
    s.destroy()


def test_a_full_set_and_a_full_event_buffer_are_reported(gpu):
    bodies, _, _ = scenes.pyramid(12)  # 78 boxes: more touching pairs than a 64-slot set holds
    s = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params(gravity=-9.81))
    s.watch_create(pairs=1, capacity=16)
    idx, ids, on, _ = watched(len(bodies), every=1)
    s.set_watched(idx, ids, on)
    s.step()
    s.watch_run(ALL, 1, 1 / 60)
    got, dropped, full = s.read_contact_events()
    assert full and len(got) == 16 and dropped > 0
    got, dropped, full = s.read_contact_events()  # read clears
    assert got == [] and dropped == 0 and not full
    s.destroy()


def test_watch_calls_refuse_bad_arguments(gpu):
    bodies, _, _ = scenes.pyramid(2)
    s = avbd_amd.GpuSolver3D(bodies, params=avbd_amd.default_params())
    with pytest.raises(ValueError, match="no watch"):
        s.watch_run(ALL, 1, 1 / 60)
    s.watch_create(pairs=8, capacity=8)
    with pytest.raises(ValueError, match="out of range"):
        s.set_watched([len(bodies)], [1], [1])
    s.destroy()
