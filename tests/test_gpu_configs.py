"""Parity with the oracle at the measured workloads' full sizes, through the C ABI.

- pile-10k: 10k boxes and spheres in a walled bin, 300 steps (the 39-layer pile lands after
  ~4.4 s of fall) with the bench's lane rule and adaptive colouring: poses, contacts and colours bit-exact with the oracle stepped in lockstep
  (its params follow the GPU's adapted colour cap after every step);
- large-100k: the 100k-body scene settled for its 300 steps on the GPU (autotuned schedule and
  lane rule, adaptive colouring and storage), then one detecting and one reuse substep from the
  settled state: pairs, contacts, colours and poses bit-exact with the oracle seeded from that
  state;
- batched-4096: 4096 worlds, one launch per pass: sampled worlds against the oracle stepping each
  world alone (bodies and contacts bit for bit), no clash or overflow anywhere in the batch;
- adjoint-256: 256 worlds x 120 steps, checkpoints every 10: two worlds' gradients against the
  oracle's f64 reverse at the same f32 states (<= 1e-7 relative), their forwards bit-exact;
- pyramid-55: 600 adaptive steps bit-exact with the oracle (poses and colours), and the settled
  top box against the reference's f64 CPU solver within 0.05 m, as the reference's own GPU test
  checks its pyramid (tests-gpu/avbd3d-gpu.gpu.test.ts:222-229), every box at rest, no clash.
"""
import sys
import time

import numpy as np
import pytest

import avbd_amd
import oracle
from avbd_amd import records, scenes
from helpers import canonical_contacts, oracle_from_gpu, oracle_params_of

pytestmark = pytest.mark.gpu

BENCH_LANES = (2 ** 31, 2 ** 15, 2 ** 12, 2 ** 11)
ADAPT = avbd_amd.FLAG_DEFAULTS | avbd_amd.solver.FLAG_ADAPT_COLORS
THREADS = 16


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_contacts(s, m, offset=0, first_body=None, last_body=None):
    """The GPU's manifolds/contacts (optionally those of bodies [first_body, last_body), shifted by
    -offset) equal the oracle's, order-free by (a, b). Returns the number of manifolds."""
    gm, gc = s.read_contacts_raw() if not isinstance(s, tuple) else s
    gm = np.asarray(gm, np.int64).copy()
    if first_body is not None:
        keep = (gm[:, 0] >= first_body) & (gm[:, 0] < last_body)
        gm = gm[keep]
        gm[:, 0:2] -= offset
    om, oc = m.contacts()
    om_u, oc_u = records.contacts_from_oracle(om, oc)
    GM, GC = canonical_contacts(gm, np.asarray(gc))
    OM, OC = canonical_contacts(np.asarray(om_u, np.int64), np.asarray(oc_u))
    assert GM.shape == OM.shape, (GM.shape, OM.shape)
    assert np.array_equal(GM[:, 0:4], OM[:, 0:4])
    assert np.array_equal(GM[:, 4:8].astype(np.uint32), OM[:, 4:8].astype(np.uint32))
    assert np.array_equal(GC.view(np.uint32) if GC.dtype == np.float32 else GC.astype(np.uint32), OC.astype(np.uint32))
    return len(GM)


def lockstep(s, m, steps):
    """One GPU step then the oracle's, the oracle taking the GPU's params (adapted every 10 steps
    on the GPU after its step) before each of its steps."""
    for _ in range(steps):
        m.set_params(oracle_params_of(s))
        s.step()
        m.step(1)


def test_pile_10k_300_steps_bit_exact(gpu):
    say("case: pile-10k, 10,000 boxes/spheres + bin, 300 steps, bench lane rule, adaptive colours")
    bodies, _, _ = scenes.random_pile(10000, seed=12345)
    p = avbd_amd.default_params(dt=1 / 60, color_cap=32, color_rounds=16, flags=ADAPT, primal_lanes=BENCH_LANES)
    s = avbd_amd.GpuSolver3D(bodies, params=p)
    m = oracle.Model(32)
    m.set_bodies(records.bodies_to_oracle(bodies))
    m.set_threads(THREADS)
    t = time.perf_counter()
    lockstep(s, m, 300)
    c = s.read_counters()
    say(f"  300 steps in {time.perf_counter() - t:.1f} s: {c['contacts']} contacts, {c['colors']} colours, cap {c['color_cap']}")
    assert c["contacts"] > 10000 and c["overflow"] == 0
    assert np.array_equal(records.bodies_to_oracle(s.read_bodies()), m.bodies())
    assert same_contacts(s, m) == c["manifolds"]
    assert np.array_equal(s.read_colors()[0], m.colors())
    s.destroy()


def test_large_100k_settled_substeps_bit_exact(gpu):
    say("case: large-100k, 100k boxes/spheres/hulls + 200 chains, settled 300 steps x 4 substeps, then 1 detecting + 1 reuse substep")
    t = time.perf_counter()
    bodies, joints, info, hulls = scenes.large_scene(100000, seed=2026, workers=8, context="spawn")
    say(f"  scene built in {time.perf_counter() - t:.1f} s")
    n = len(bodies)
    p = avbd_amd.default_params(dt=1 / 240, iterations=10, color_cap=32, color_rounds=16, flags=ADAPT, primal_lanes=BENCH_LANES)
    s = avbd_amd.GpuSolver3D(bodies, joints, info, params=p, hulls=hulls,
                             capacity={"pairs": 12 * n, "manifolds": 12 * n, "contacts": 40 * n})
    s.steps(272, 4)
    schedule, _ = s.autotune(2, 4)  # 28 more steps: the bench's lead-in exactly
    c = s.read_counters()
    say(f"  settled: {c['contacts']} contacts, {c['manifolds']} manifolds, {c['colors']} colours, "
        f"schedule {schedule}, lanes {list(s.params.primal_lanes)}")
    assert c["contacts"] > 300000 and c["overflow"] == 0
    q = s.params
    q.flags &= ~avbd_amd.solver.FLAG_ADAPT_COLORS  # no readback between the two substeps below
    s.set_params(q)
    t = time.perf_counter()
    m = oracle_from_gpu(s, hulls, threads=THREADS)
    s.steps(1, 2)  # one avbd_steps call: a detecting substep, then a reuse that may keep the adjacency
    m.step(1, reuse_manifolds=False)
    m.step(1, reuse_manifolds=True)
    say(f"  oracle seeded and stepped in {time.perf_counter() - t:.1f} s")
    gp = s.read_pairs()
    op = m.pairs()
    assert len(gp) == len(op) > 200000
    assert np.array_equal(gp[np.lexsort((gp[:, 1], gp[:, 0]))], op)
    assert np.array_equal(s.read_colors()[0], m.colors())
    got, want = records.bodies_to_oracle(s.read_bodies()), m.bodies()
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert len(bad) == 0, (len(bad), bad[:10])
    assert np.array_equal(s.read_joints().astype(np.float64), m.joints().astype(np.float32).astype(np.float64))
    assert same_contacts(s, m) == s.read_counters()["manifolds"]
    s.destroy()


def test_batched_4096_worlds_match_each_world_alone(gpu):
    say("case: batched-4096, 4096 worlds (log-U[8,512] bodies), 30 steps; 32 sampled worlds vs the oracle alone")
    t = time.perf_counter()
    bodies, starts = scenes.batched_worlds(4096)
    n = len(bodies)
    p = avbd_amd.default_params(color_cap=24, color_rounds=16, primal_lanes=(0, 0, 0, 0))
    batch = avbd_amd.GpuSolver3D(bodies, params=p, world_start=starts,
                                 capacity={"pairs": 8 * n, "manifolds": 8 * n, "contacts": 24 * n})
    for _ in range(6):
        batch.steps(5)
        c = batch.read_counters()
        assert c["clashes"] == 0 and c["overflow"] == 0, c
    say(f"  {n} bodies, 30 steps in {time.perf_counter() - t:.1f} s: {c['contacts']} contacts")
    got = records.bodies_to_oracle(batch.read_bodies())
    gm, gc = batch.read_contacts_raw()
    rng = np.random.default_rng(7)
    sample = sorted(set(rng.choice(4096, 30, replace=False).tolist()) | {0, 4095})
    sizes = np.diff(starts)
    sample = sorted(set(sample) | {int(np.argmax(sizes))})  # the largest world too
    checked = 0
    for w in sample:
        lo, hi = int(starts[w]), int(starts[w + 1])
        m = oracle.Model(32)
        m.set_params(oracle.make_params(dt=1 / 60, gravity=-9.81, color_cap=24, color_rounds=16))
        m.set_bodies(records.bodies_to_oracle(bodies[lo:hi]))
        m.step(30)
        assert np.array_equal(got[lo:hi], m.bodies()), w
        checked += same_contacts((gm, gc), m, offset=lo, first_body=lo, last_body=hi)
    say(f"  {len(sample)} worlds, {checked} manifolds bit-exact")
    assert checked > 1000
    batch.destroy()


def test_adjoint_256_worlds_120_steps_match_the_oracle_reverse(gpu):
    say("case: adjoint-256, 256 worlds x 129 bodies, 120 steps, checkpoints every 10; worlds 0 and 1 vs the oracle's f64 reverse")
    worlds, steps = 256, 120
    bodies, ws, _ = scenes.thrown_at_stack(worlds, seed=5)
    n = len(bodies)
    p = avbd_amd.default_params(dt=1 / 60, gravity=-9.81, color_cap=16, primal_lanes=(0, 0, 0, 0))
    d_final = np.zeros((n, 16))
    for w in range(worlds):
        idx = np.arange(ws[w], ws[w + 1])
        dyn = idx[bodies[idx, 19] > 0]
        d_final[dyn, 0:3] = 1.0 / len(dyn)
    s = avbd_amd.GpuSolver3D(bodies, params=p, world_start=ws)
    t = time.perf_counter()
    s.forward(steps, 10)
    final = records.bodies_to_oracle(s.read_bodies())
    gi, gm = s.backward(d_final)
    say(f"  forward + backward in {time.perf_counter() - t:.2f} s")
    assert np.isfinite(gi).all() and np.isfinite(gm).all()
    F32 = lambda x: float(np.float32(x))
    op = oracle.make_params(dt=F32(p.dt), gravity=F32(p.gravity), alpha=F32(p.alpha), gamma=F32(p.gamma), margin=F32(p.contact_margin),
                            beta_lin=F32(p.beta_lin), beta_ang=F32(p.beta_ang), color_cap=16, lanes=(0, 0, 0, 0))
    for w in (0, 1):
        sl = slice(int(ws[w]), int(ws[w + 1]))
        rec = records.bodies_to_oracle(bodies[sl])
        m = oracle.Model(32)
        m.set_params(op)
        m.set_bodies(rec)
        m.record_tape(True)
        m.step(steps)
        assert np.array_equal(final[sl], m.bodies()), w  # the batch's world w steps as the oracle's world alone
        fz = oracle.Frozen(m, rec)
        fz.load_recorded()
        g, gmu = fz.backward(d_final[sl])
        dyn = bodies[sl, 19] > 0
        for cols in (slice(0, 3), slice(4, 8), slice(8, 11), slice(12, 15)):
            a, b = gi[sl][dyn][:, cols], g[dyn][:, cols]
            err = float(np.abs(a - b).max() / np.abs(b).max())
            say(f"  world {w} slots {cols.start}-{cols.stop - 1}: {err:.2e}")
            assert err <= 1e-7, (w, cols, err)
        assert abs(gm[sl].sum() - gmu.sum()) <= 1e-7 * abs(gmu.sum()), w
    s.destroy()


def test_pyramid_55_600_adaptive_steps(gpu):
    say("case: pyramid-55, 10 rows of unit boxes, 600 steps with adaptive colours vs the oracle and the f64 reference")
    bodies, _, _ = scenes.pyramid(10)
    p = avbd_amd.default_params(dt=1 / 60, gravity=-9.81, color_cap=32, color_rounds=16, flags=ADAPT, primal_lanes=BENCH_LANES)
    s = avbd_amd.GpuSolver3D(bodies, params=p)
    m = oracle.Model(32)
    m.set_bodies(records.bodies_to_oracle(bodies))
    clashes = 0
    for _ in range(60):
        lockstep(s, m, 10)
        clashes += s.read_counters()["clashes"]
    assert clashes == 0
    got = records.bodies_to_oracle(s.read_bodies())
    assert np.array_equal(got, m.bodies())
    assert np.array_equal(s.read_colors()[0], m.colors())
    same_contacts(s, m)
    # the reference's f64 CPU solver on the same 55 boxes
    ref = oracle.RefSolver()
    ref.set_params(dt=1 / 60, gravity=-9.81, iterations=10)
    for r in records.bodies_to_oracle(bodies):
        ref.add_body(r[16:19], 0.0 if r[19] <= 0 else 1.0, r[3], r[0:3], rot=r[4:8])
    ref.step(600)
    top_ref = ref.poses()[-1, 2]
    top = got[-1, 2]
    say(f"  top box z: GPU {top:.4f}, f64 reference {top_ref:.4f}")
    assert abs(top - top_ref) < 0.05
    assert abs(top - 9.5) < 0.3  # ten courses of unit boxes, still standing
    speeds = np.linalg.norm(got[1:, 32:35], axis=1)
    assert speeds.max() < 0.1, speeds.max()
    s.destroy()
