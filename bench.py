#!/usr/bin/env python3
"""Throughput of the MI355X AVBD step on one GPU.

Default line (the headline): the large scene, 100,000 bodies (boxes, spheres, convex hulls 1:1:1)
in a walled bin plus 200 chains of 20 ball-jointed boxes hanging from static anchors, 4 substeps x
10 iterations per step, f32, in body-substeps/s. The scene first runs the workload's 300 steps
(whatever --warmup says; a larger --warmup runs longer), so the timed window sees the settled
pile; then exactly `--steps` steps (each 4 substeps: the first detects contacts, the rest reuse
manifolds) are timed between device synchronisations, bodies resident in HBM, no readback.

Also printed on the same JSON line:
- roofline: the dominant kernel (k_primal, one colour of the primal sweep, under the per-colour
  schedule the autotuner picks; the cooperative solve kernel when it picks that one) --
  algorithmic bytes per launch (the unique-touch model below, from the counters read back after
  the timed run) / its average launch duration, timed with hipEvents on the solver's stream,
  against the MI355X HBM peak of 8 TB/s; traffic from the committed rocprofv3 PMC summary taken
  at the same settled state;
- cpu_baseline: the oracle's f32 restatement of the same step, seeded with the GPU's state after
  the timed run and stepped for a bounded sample of substeps on the host: on every host thread
  (OpenMP over the bodies of each colour, the contacts, the pairs and the colouring rounds; the
  result equals the one-thread run bit for bit) and, for context, on one thread;
- batched: 4096 independent worlds stepped by one solver, in worlds*steps/s, against looping the
  plain step over a sample of the same worlds;
- adjoint: 256 worlds x 120 steps, backward time / forward time, and the gradient against the
  oracle's f64 reverse at the same f32 states.
`--workload batched-4096` / `adjoint-256` print those alone (the adjoint with its
finite-difference checks by horizon); `pile-10k` and `pyramid-55` time the smaller scenes.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "three-avbd_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0


def primal_sweep_bytes(n, m, k, j):
    """Algorithmic bytes of one primal sweep over every colour (see solve_bytes)."""
    return 160 * n + 4 * (2 * n + 2 * (m + j)) + 4 * n + 32 * m + 64 * k + 176 * j


def solve_bytes(n, m, k, j, iters):
    """Algorithmic bytes of one solve launch (unique-touch model of the record fields each pass
    reads or writes; N bodies, M manifolds, K contact points, J joints):
      primal sweep  160N + 4(2N + 2(M+J)) + 4N + 32M + 64K + 176J
      dual          64N + 32M + 92K + 240J
      warm starts + velocities  368N + 304J"""
    primal = 160 * n + 4 * (2 * n + 2 * (m + j)) + 4 * n + 32 * m + 64 * k + 176 * j
    dual = 64 * n + 32 * m + 92 * k + 240 * j
    return iters * (primal + dual) + 368 * n + 304 * j


def measured_traffic(workload):
    """HBM bytes per launch of the roofline kernel from the committed rocprofv3 PMC summary
    (profiles/<workload>_hbm_traffic.txt: FETCH_SIZE and WRITE_SIZE runs scaled by a copy
    kernel's reading), or None when that workload has none."""
    path = os.path.join(ROOT, "profiles", f"{workload}_hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def build_workload(name, workers):
    from avbd_amd import scenes
    if name == "large-100k":
        bodies, joints, info, hulls = scenes.large_scene(100000, seed=2026, workers=workers)
        return bodies, joints, info, hulls, 4, 10
    if name == "pile-10k":
        bodies, joints, info = scenes.random_pile(10000, seed=12345)
        return bodies, joints, info, None, 1, 10
    if name == "pyramid-55":
        bodies, joints, info = scenes.pyramid(10)
        return bodies, joints, info, None, 1, 10
    raise SystemExit(f"unknown workload {name}")


SETTLE_STEPS = {"large-100k": 300, "pile-10k": 300, "pyramid-55": 600}


def host_info():
    """Host cores this process may use and the CPU model (from /proc/cpuinfo)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", usable) or usable)
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable, "threads": max(1, min(threads, usable))}


def cpu_baseline(solver, hulls, substeps, threads, budget_s=10.0, max_substeps=40):
    """The oracle's f32 step from the GPU's current state on `threads` host threads; stops after
    the first whole step (all `substeps`) that takes the elapsed time past budget_s."""
    import numpy as np
    import oracle
    from avbd_amd import records
    m = oracle.Model(32)
    p = solver.params
    m.set_params(oracle.make_params(dt=p.dt, gravity=p.gravity, iterations=p.iterations, alpha=p.alpha, beta_lin=p.beta_lin,
                                    beta_ang=p.beta_ang, gamma=p.gamma, flags=p.flags, color_cap=p.color_cap,
                                    color_rounds=p.color_rounds, margin=p.contact_margin, up=tuple(p.up),
                                    lanes=tuple(p.primal_lanes)))
    m.set_bodies(records.bodies_to_oracle(solver.read_bodies()))
    if solver.joint_count:
        m.set_joints(solver.read_joints().astype(np.float64), solver._joint_info)
    if hulls is not None:
        m.set_hulls(hulls)
    man, con = solver.read_contacts_raw()
    m.set_prev(*records.contacts_to_oracle(man, con))
    m.set_color_state(solver.read_colors()[0])
    m.set_step_count(solver.read_counters()["step"])
    m.set_threads(threads)
    done = 0
    t0 = time.perf_counter()
    while done < max_substeps:
        for u in range(substeps):
            m.step(1, reuse_manifolds=u > 0)
            done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return done, dt


def f64_reference_baseline(bodies, budget_s=6.0, max_steps=2000):
    """The reference's f64 sequential CPU solver (src/avbd3d/ref/solver.ts, restated in the oracle)
    on one host thread, for scenes of boxes only (it has no other shape), from the scene's initial
    state: bodies by size, density, friction and pose, the reference's default parameters with
    gravity -9.81. Returns (steps, seconds)."""
    import numpy as np
    import oracle
    ref = oracle.RefSolver()
    ref.set_params(dt=1 / 60, gravity=-9.81, iterations=10)
    for b in np.asarray(bodies, np.float64):
        size = b[16:19]
        density = b[19] / float(np.prod(size)) if b[19] > 0 else 0.0
        ref.add_body(size, density, b[3], b[0:3], rot=b[4:8], vel=b[32:35], angvel=b[36:39])
    done = 0
    t0 = time.perf_counter()
    while done < max_steps and time.perf_counter() - t0 < budget_s:
        ref.step(10)
        done += 10
    return done, time.perf_counter() - t0


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def bench_batched(args, worlds=4096, loop_sample=256, settle=60, steps=None):
    """worlds x steps per second of 4096 independent worlds stepped by one batched solver (every
    launch covers all worlds), against looping the plain step (one solver per world) over a
    bounded sample of the same worlds, same parameters and iterations. Both run `settle` steps
    first (the autotuner's steps among them), whatever --warmup says."""
    import numpy as np
    import torch
    import avbd_amd
    from avbd_amd import scenes
    steps = steps or args.steps
    t = time.perf_counter()
    bodies, starts = scenes.batched_worlds(worlds)
    log(f"{worlds} worlds, {len(bodies)} bodies, built in {time.perf_counter() - t:.1f} s")
    params = avbd_amd.default_params(color_cap=24, color_rounds=16, flags=avbd_amd.FLAG_DEFAULTS | avbd_amd.solver.FLAG_ADAPT_COLORS)
    n = len(bodies)
    batch = avbd_amd.GpuSolver3D(bodies, params=params, world_start=starts,
                                 capacity={"pairs": 8 * n, "manifolds": 8 * n, "contacts": 24 * n})
    lead = max(settle, args.warmup) - 24
    batch.steps(lead)
    schedule, _ = batch.autotune(2)
    batch.sync()
    log(f"batched: lead-in {lead + 24} steps (autotune {schedule})")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.steps(steps)
    torch.cuda.synchronize()
    plain_wall = time.perf_counter() - t0
    # then each substep as a hipGraph: captured in 10 untimed steps, replayed in the timed ones
    batch.set_graphs(True)
    batch.steps(10)
    batch.sync()
    cap0, rep0 = batch.graph_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.steps(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cap1, rep1 = batch.graph_stats()
    c = batch.read_counters()
    batch_lanes = list(batch.params.primal_lanes)
    graph_rate = worlds * steps / wall
    plain_rate = worlds * steps / plain_wall
    log(f"batched: graphs {graph_rate:.0f} worlds*steps/s ({rep1 - rep0} substeps replayed, {cap1 - cap0} captured "
        f"in the timed steps), plain launches {plain_rate:.0f}")
    mode = "graphs" if graph_rate >= plain_rate else "plain launches"  # the line reports the faster
    batched_rate, wall = (graph_rate, wall) if mode == "graphs" else (plain_rate, plain_wall)
    batch.destroy()
    # the plain step looped over a sample of the same worlds (one solver each), same lead-in
    rng = np.random.default_rng(1)
    sample = np.sort(rng.choice(worlds, size=min(loop_sample, worlds), replace=False))
    singles = []
    for w in sample:
        s = avbd_amd.GpuSolver3D(bodies[starts[w]:starts[w + 1]], params=avbd_amd.default_params(
            color_cap=24, color_rounds=16, flags=avbd_amd.FLAG_DEFAULTS | avbd_amd.solver.FLAG_ADAPT_COLORS))
        s.steps(lead + 24)
        singles.append(s)
    loop_steps = max(2, steps // 4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(loop_steps):
        for s in singles:
            s.step()
    torch.cuda.synchronize()
    loop_wall = time.perf_counter() - t0
    loop_rate = len(singles) * loop_steps / loop_wall
    for s in singles:
        s.destroy()
    log(f"batched: {batched_rate:.0f} worlds*steps/s, looped plain step {loop_rate:.0f}")
    return {
        "metric": "batched worlds*steps/s (4096 worlds)", "value": batched_rate, "unit": "worlds*steps/s",
        "n_gpus": 1, "steps": steps, "warmup": lead + 24, "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched-4096", "worlds": worlds, "bodies": n, "substeps": 1, "iterations": 10, "value_from": mode,
                   "lead_in_steps": lead + 24, "contacts": c["contacts"], "colors": c["colors"], "clashes": c["clashes"],
                   "overflow": c["overflow"], "schedule": schedule, "color_cap": c["color_cap"], "primal_lanes": batch_lanes},
        "looped_plain_step": {"value": loop_rate, "unit": "worlds*steps/s", "sample_worlds": len(singles),
                              "steps": loop_steps, "speedup": batched_rate / loop_rate},
        "graphs": {"value": graph_rate, "unit": "worlds*steps/s", "replayed_substeps": rep1 - rep0, "captured_substeps": cap1 - cap0},
        "plain_launches": {"value": plain_rate, "unit": "worlds*steps/s", "ms_per_step": plain_wall / steps * 1e3},
    }


def adjoint_oracle_check(bodies, ws, params, d_final, steps, every, worlds=(0, 1)):
    """The GPU's gradient against the oracle's f64 reverse at the same f32 states, for sample
    worlds each on a solver of its own (its colour sizes pick the lane split, as the oracle's).
    Returns (max relative error, CPU seconds, per-world results)."""
    import numpy as np
    import avbd_amd
    import oracle
    from avbd_amd import records
    F32 = lambda x: float(np.float32(x))
    op = oracle.make_params(dt=F32(params.dt), gravity=F32(params.gravity), alpha=F32(params.alpha), gamma=F32(params.gamma),
                            margin=F32(params.contact_margin), beta_lin=F32(params.beta_lin), beta_ang=F32(params.beta_ang),
                            lanes=tuple(params.primal_lanes))
    worst, t_cpu, out = 0.0, 0.0, {}
    for w in worlds:
        sl = slice(int(ws[w]), int(ws[w + 1]))
        one = avbd_amd.GpuSolver3D(bodies[sl], params=params)
        one.forward(steps, every)
        final = one.read_bodies()
        gi, gm = one.backward(d_final[sl])
        one.destroy()
        rec = records.bodies_to_oracle(bodies[sl])
        tc = time.perf_counter()
        m = oracle.Model(32)
        m.set_params(op)
        m.set_bodies(rec)
        m.record_tape(True)
        m.step(steps)
        fz = oracle.Frozen(m, rec)
        fz.load_recorded()
        g, gmu = fz.backward(d_final[sl])
        t_cpu += time.perf_counter() - tc
        assert np.array_equal(m.bodies()[:, 0:3].astype(np.float32), final[:, 0:3]), "forward differs from the oracle"
        dyn = bodies[sl, 19] > 0
        err = 0.0
        for cols in (slice(0, 3), slice(4, 8), slice(8, 11), slice(12, 15)):
            a, b = gi[dyn][:, cols], g[dyn][:, cols]
            err = max(err, float(np.abs(a - b).max() / np.abs(b).max()))
        err = max(err, abs(gm.sum() - gmu.sum()) / abs(gmu.sum()))
        out[w] = {"gi": gi, "gm": gm, "g": g, "gmu": gmu, "fz": fz, "rec": rec, "err": err}
        worst = max(worst, err)
    return worst, t_cpu, out


def bench_adjoint(args, worlds=256, steps=120, every=10, seed=5, light=False):
    """Differentiation: 256 worlds, each 64 boxes thrown at a 4x4x4 box stack, 120 steps at dt
    1/60. Forward: the plain step with a checkpoint every 10 steps (avbd_forward); backward: the
    reverse sweep with the active set frozen, re-running each 10-step segment from its checkpoint
    (avbd_backward), for the gradient of sum over worlds of (1,1,1) . (final mean centroid) with
    respect to every body's initial state and friction. The value is backward time / forward
    time. Gradient checks on sample worlds: the oracle's f64 reverse at the same f32 states and
    (unless `light`) f64 central differences of the oracle's frozen replay by horizon."""
    import numpy as np
    import torch
    import avbd_amd
    import oracle
    from avbd_amd import records, scenes
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    t = time.perf_counter()
    bodies, ws, mus = scenes.thrown_at_stack(worlds, seed=seed)
    n = len(bodies)
    log(f"adjoint: {worlds} worlds, {n} bodies, built in {time.perf_counter() - t:.1f} s")
    params = avbd_amd.default_params(dt=1 / 60, gravity=-9.81, color_cap=args.color_cap_adjoint)
    d_final = np.zeros((n, 16))
    for w in range(worlds):
        idx = np.arange(ws[w], ws[w + 1])
        dyn = idx[bodies[idx, 19] > 0]
        d_final[dyn, 0:3] = 1.0 / len(dyn)

    s = avbd_amd.GpuSolver3D(bodies, params=params, world_start=ws)  # load every kernel once, pick the schedule
    s.forward(12, 10)
    s.backward(d_final)
    schedule, _ = s.autotune(2)
    s.destroy()

    def fresh():
        solver = avbd_amd.GpuSolver3D(bodies, params=params, world_start=ws)
        solver.set_schedule(**schedule)
        return solver

    if args.adjoint_pair_only:
        # for kernel-trace runs: two identical pairs (the first loads the kernels and allocates)
        for rep in range(2):
            s = fresh()
            s.sync()
            t0 = time.perf_counter()
            s.forward(steps, every)
            t1 = time.perf_counter()
            s.backward(d_final)
            t2 = time.perf_counter()
            s.destroy()
            log(f"pair {rep}: forward {1e3 * (t1 - t0):.1f} ms, backward {1e3 * (t2 - t1):.1f} ms ({(t2 - t1) / (t1 - t0):.2f}x)")
        return None
    s = fresh()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.steps(steps)
    s.sync()
    plain = time.perf_counter() - t0
    s.destroy()
    timings, reps = {}, 3
    for mode in ((every,) if light else (0, every)):
        runs = []
        for _ in range(reps):  # independent solvers: each pair allocates its checkpoints, tapes and accumulators
            s = fresh()
            s.sync()
            t0 = time.perf_counter()
            s.forward(steps, mode)
            t1 = time.perf_counter()
            s.backward(d_final)
            t2 = time.perf_counter()
            runs.append((t1 - t0, t2 - t1))
            s.destroy()
        runs.sort(key=lambda r: r[1] / r[0])
        timings[mode] = runs[len(runs) // 2]  # the median backward/forward ratio
        log(f"adjoint checkpoint_every={mode}: " + ", ".join(f"forward {1e3 * f:.1f} ms backward {1e3 * b:.1f} ms ({b / f:.2f}x)" for f, b in runs))
        timings[f"runs{mode}"] = runs
    fwd, bwd = timings[every]

    vs_oracle, t_cpu, per_world = adjoint_oracle_check(bodies, ws, params, d_final, steps, every)
    log(f"adjoint: gradient vs the oracle's f64 reverse at the same states {vs_oracle:.2e}")
    out = {
        "metric": f"adjoint/forward time ({worlds} worlds, {steps} steps)",
        "value": bwd / fwd,
        "unit": "x",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": 1e3 * fwd / steps,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 forward, f64 reverse",
        "data": "synthetic",
        "config": {"workload": "adjoint-256", "worlds": worlds, "bodies": n, "steps": steps, "checkpoint_every": every,
                   "color_cap": params.color_cap, "schedule": schedule,
                   "timing": "median backward/forward ratio of 3 forward/backward pairs, each on a new solver (allocation included)",
                   "loss": "sum over worlds of (1,1,1) . mean final position of the world's dynamic bodies"},
        "adjoint": {
            "forward_ms": 1e3 * fwd, "backward_ms": 1e3 * bwd, "plain_forward_ms": 1e3 * plain,
            "forward_plus_backward_over_plain": (fwd + bwd) / plain,
            "pairs_ms": {f"checkpoint_every={m}": [[1e3 * f, 1e3 * b] for f, b in timings[f"runs{m}"]]
                         for m in ((every,) if light else (0, every))},
            "grad_rel_err_vs_oracle_same_states": vs_oracle,
            "grad_target_vs_oracle_1e-3": "pass" if vs_oracle <= 1e-3 else "fail",
        },
        "cpu_baseline": {"value": 2 * steps / t_cpu, "unit": "worlds*steps/s (forward + backward)", "cores": 1, "kind": "port",
                         "sample": f"2 worlds: the oracle's f32 step recording its tape, then its f64 reverse at those states, {t_cpu:.1f} s"},
        "gpu_worlds_steps_per_s": worlds * steps / (fwd + bwd),
    }
    if not light:
        out["adjoint"]["record_all"] = {"forward_ms": 1e3 * timings[0][0], "backward_ms": 1e3 * timings[0][1],
                                        "backward_over_forward": timings[0][1] / timings[0][0]}
        out["adjoint"].update(adjoint_fd_checks(bodies, ws, params, d_final, steps, every, per_world[0]))
    return out


def adjoint_fd_checks(bodies, ws, params, d_final, steps, every, w0):
    """f64 central differences of the oracle's frozen replay (world 0 on its own solver): at the
    configured h = 1e-4 and at h = 1e-7 for the six largest velocity components and friction, and
    by horizon at h = 1e-6 (where the differences have converged), with pass/fail against the
    1e-3 target per horizon."""
    import numpy as np
    import avbd_amd
    import oracle
    from avbd_amd import records
    F32 = lambda x: float(np.float32(x))
    op = oracle.make_params(dt=F32(params.dt), gravity=F32(params.gravity), alpha=F32(params.alpha), gamma=F32(params.gamma),
                            margin=F32(params.contact_margin), beta_lin=F32(params.beta_lin), beta_ang=F32(params.beta_ang),
                            lanes=tuple(params.primal_lanes))
    sl = slice(int(ws[0]), int(ws[1]))
    fz, rec, gi0, gm0 = w0["fz"], w0["rec"], w0["gi"], w0["gm"]
    st0 = oracle.state_from_records(rec)
    fz.run(st0)
    gr, gmur = fz.backward(d_final[sl])  # the replay's own adjoint (f64 trajectory)
    loss = lambda o: float((o * d_final[sl]).sum())
    order = np.argsort(-np.abs(gr[:, 8:15]).ravel())
    picks = [(int(k // 7), 8 + int(k % 7)) for k in order if (k % 7) != 3][:6]
    fd_rows = []
    for i, slot in picks + [(-1, -1)]:
        row = {"input": "friction" if i < 0 else f"body {i} slot {slot}"}
        for h in (1e-4, 1e-7):
            if i < 0:
                mu0 = rec[:, 3].copy()
                fd = (loss(fz.run(st0, mu0 + h)) - loss(fz.run(st0, mu0 - h))) / (2 * h)
            else:
                e = np.zeros_like(st0)
                e[i, slot] = h
                fd = (loss(fz.run(st0 + e)) - loss(fz.run(st0 - e))) / (2 * h)
            row[f"fd_h{h:.0e}"] = fd
        an = gmur.sum() if i < 0 else gr[i, slot]
        gp = gm0.sum() if i < 0 else gi0[i, slot]
        row.update({"oracle_replay": an, "gpu": gp,
                    "oracle_vs_fd1e-4": abs(an - row["fd_h1e-04"]) / abs(an),
                    "oracle_vs_fd1e-7": abs(an - row["fd_h1e-07"]) / abs(an),
                    "gpu_vs_fd1e-4": abs(gp - row["fd_h1e-04"]) / abs(row["fd_h1e-04"])})
        fd_rows.append(row)
    horizon = []
    for hs in (10, 20, 40, 60, steps):
        one = avbd_amd.GpuSolver3D(bodies[sl], params=params)
        one.forward(hs, every)
        gi, _ = one.backward(d_final[sl])
        one.destroy()
        m = oracle.Model(32)
        m.set_params(op)
        m.set_bodies(rec)
        m.record_tape(True)
        m.step(hs)
        fzh = oracle.Frozen(m, rec)
        fzh.run(st0)
        grh, _ = fzh.backward(d_final[sl])
        order = np.argsort(-np.abs(grh[:, 8:15]).ravel())
        pk = [(int(k // 7), 8 + int(k % 7)) for k in order if (k % 7) != 3][:6]
        fd = []
        for i, slot in pk:
            e = np.zeros_like(st0)
            e[i, slot] = 1e-6
            fd.append((loss(fzh.run(st0 + e)) - loss(fzh.run(st0 - e))) / 2e-6)
        fd = np.array(fd)
        g_gpu = np.array([gi[i, slot] for i, slot in pk])
        g_or = np.array([grh[i, slot] for i, slot in pk])
        row = {"steps": hs, "gpu_vs_fd": float(np.abs(g_gpu - fd).max() / np.abs(fd).max()),
               "oracle_f64_vs_fd": float(np.abs(g_or - fd).max() / np.abs(fd).max())}
        row["gpu_target_1e-3"] = "pass" if row["gpu_vs_fd"] <= 1e-3 else "fail"
        horizon.append(row)
        log(f"horizon {hs}: gpu vs fd {row['gpu_vs_fd']:.2e} ({row['gpu_target_1e-3']}), oracle f64 vs fd {row['oracle_f64_vs_fd']:.2e}")
    return {"grad_rel_err_vs_fd_by_horizon": horizon, "fd_world0": fd_rows}


def bench_step(args, workload):
    """The plain step on one scene: settle, autotune, time `--steps` steps, roofline, CPU baseline."""
    t = time.perf_counter()
    bodies, joints, info, hulls, substeps, iters = build_workload(workload, args.workers)
    log(f"scene {workload}: {len(bodies)} bodies, {len(joints)} joints, built in {time.perf_counter() - t:.1f} s")

    import torch
    import avbd_amd
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.copy_calibration_mb:
        src = torch.ones(args.copy_calibration_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 x {args.copy_calibration_mb} MiB read + {args.copy_calibration_mb} MiB written")
        del src, dst
    params = avbd_amd.default_params(dt=1.0 / (60.0 * substeps), iterations=iters, color_cap=args.color_cap, color_rounds=16,
                                     flags=avbd_amd.FLAG_DEFAULTS | avbd_amd.solver.FLAG_ADAPT_COLORS,
                                     primal_lanes=tuple(args.primal_lanes))
    n = len(bodies)
    solver = avbd_amd.GpuSolver3D(bodies, joints, info, params=params, hulls=hulls,
                                  capacity={"pairs": 12 * n, "manifolds": 12 * n, "contacts": 40 * n})
    t = time.perf_counter()
    tune_steps = 2
    total_lead = max(SETTLE_STEPS.get(workload, 0), args.warmup)
    tune_total = 14 * tune_steps if total_lead >= 14 * tune_steps else 0
    lead = total_lead - tune_total
    for done in range(0, lead, 50):  # in chunks, with a progress line each (profiled runs are slow)
        solver.steps(min(50, lead - done), substeps)
        c = solver.read_counters()
        log(f"lead-in {min(lead, done + 50)}/{total_lead} steps: {c['contacts']} contacts, {c['manifolds']} manifolds, "
            f"{c['colors']} colours")
    if args.sparse_tail is not None:
        solver.set_sparse_tail(args.sparse_tail)
    if tune_total and args.schedule == "auto":
        schedule, tuned_ms = solver.autotune(tune_steps, substeps)  # part of the lead-in: the simulation advances
    else:  # counter-collection runs distort the autotuner's timings: fix the schedule instead
        if args.schedule in ("flow", "tail", "lead"):
            coop = {"flow": 2, "tail": 3, "lead": 4}[args.schedule]
            solver.set_schedule(cooperative=coop, threads=64, blocks_per_cu=args.flow_blocks_per_cu)
        else:
            solver.set_schedule(cooperative=int(args.schedule == "coop"), threads=512, blocks_per_cu=1)
        schedule, tuned_ms = solver.schedule, solver.time_steps(tune_total, substeps) / max(1, tune_total)
    if args.graphs:
        solver.set_graphs(True)
        solver.steps(4, substeps)  # capture the substeps' launch sequences before the timed steps
    solver.sync()
    c = solver.read_counters()
    log(f"lead-in {total_lead} steps x {substeps} substeps (autotune: {schedule}, {tuned_ms:.2f} ms/step) "
        f"in {time.perf_counter() - t:.1f} s: {c['contacts']} contacts")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.steps(args.steps, substeps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    c = solver.read_counters()
    ev_ms = solver.time_steps(2, substeps)  # hipEvent-bracketed cross-check of the wall clock
    log(f"timed {args.steps} steps: {wall / args.steps * 1e3:.2f} ms/step")

    # roofline of the dominant kernel: HIP events around every launch of it, then bytes from counters
    ks = solver.profile_kernels(2, substeps)
    c2 = solver.read_counters()
    m2, k2, j2 = c2["manifolds"], c2["contacts"], len(joints)
    if ks["solve_launches"]:
        avg_ms = ks["solve_ms"] / ks["solve_launches"]
        if schedule.get("cooperative") == 2:
            kernel = "k_flow (cooperative dataflow: warm starts, iterations x (primal, owned duals), velocities)"
        else:
            kernel = "k_solve (cooperative: warm starts, iterations x (primal per colour, dual), velocities)"
        nbytes = solve_bytes(n, m2, k2, j2, iters)
    elif ks["primal_launches"] == 0 and ks["tail_launches"] == 0 and ks["dual_launches"]:
        # dataflow tail from the first colour: one k_flow_tail launch per iteration runs the whole
        # sweep and every dual (timed with the duals)
        avg_ms = ks["dual_ms"] / ks["dual_launches"]
        kernel = "k_flow_tail (dataflow sweep over every colour plus all duals, one launch per iteration)"
        nbytes = solve_bytes(n, m2, k2, j2, 1) - (368 * n + 304 * j2)
    elif ks["primal_launches"] == 0 and ks["tail_launches"]:
        # every colour small enough for the fused tail: one workgroup launch per sweep
        avg_ms = ks["tail_ms"] / ks["tail_launches"]
        kernel = "k_primal_tail (every colour of the sweep in one workgroup)"
        nbytes = primal_sweep_bytes(n, c2["manifolds"], c2["contacts"], len(joints))
    else:
        # one launch per colour per iteration (trailing small colours fused into k_primal_tail,
        # timed apart): a sweep's bytes over its k_primal launches
        avg_ms = ks["primal_ms"] / max(1, ks["primal_launches"])
        if schedule.get("cooperative") == 4:
            kernel = "k_flow_lead (the leading colours of one iteration as one dataflow launch; the trailing colours and the duals run in k_flow_tail)"
        elif schedule.get("cooperative") == 3:
            kernel = "k_primal (one leading colour of the sweep; the trailing colours and the duals run in k_flow_tail)"
        else:
            kernel = "k_primal (one colour of the sweep; trailing small colours run fused in k_primal_tail)"
        per_sweep = ks["primal_launches"] / (2 * substeps * iters)
        nbytes = primal_sweep_bytes(n, c2["manifolds"], c2["contacts"], len(joints)) / max(1.0, per_sweep)
    if avg_ms <= 0.0:
        raise SystemExit(f"bench: no timed launch of the solve kernels in the profiled steps ({ks})")
    achieved = nbytes / (avg_ms * 1e-3) / 1e9

    value = n * args.steps * substeps / wall
    out = {
        "metric": "body-substeps/s @100k bodies" if workload == "large-100k" else f"body-substeps/s ({workload})",
        "value": value,
        "unit": "body-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": total_lead,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": workload, "bodies": n, "joints": int(len(joints)), "substeps": substeps, "iterations": iters,
            "lead_in_steps": total_lead, "contacts": c["contacts"], "manifolds": c["manifolds"], "pairs": c["pairs"],
            "colors": c["colors"], "clashes": c["clashes"], "overflow": c["overflow"], "parallelism": "single-gpu",
            "event_ms_per_step": ev_ms / 2, "schedule": schedule, "color_cap": c["color_cap"],
            "primal_lanes": list(solver.params.primal_lanes),
        },
        "roofline": {
            "bound": "hbm", "limiter": "latency: a colour launch is one wavefront's dependent walk over its bodies' contacts",
            "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(workload),
            "traffic_source": f"profiles/{workload}_hbm_traffic.txt (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated)",
            "bytes_per_launch": nbytes, "avg_launch_ms": avg_ms,
            "launches": ks["solve_launches"] or ks["primal_launches"] or ks["tail_launches"] or ks["dual_launches"],
            "solve_blocks": ks["solve_blocks"],
        },
    }
    if not args.no_cpu_baseline:
        info_h = host_info()
        done, dt = cpu_baseline(solver, hulls, substeps, info_h["threads"], budget_s=10.0, max_substeps=40)
        done1, dt1 = cpu_baseline(solver, hulls, substeps, 1, budget_s=6.0, max_substeps=8)
        allcore = n * done / dt
        out["cpu_baseline"] = {
            "value": allcore, "unit": "body-substeps/s", "cores": info_h["threads"], "kind": "port",
            "sample": f"{done} substeps of the same {n}-body state (after the timed run) through the oracle's f32 "
                      f"restatement of the step, OpenMP over {info_h['threads']} host threads, {dt:.1f} s",
            "cpu_model": info_h["cpu_model"], "nproc": info_h["nproc"], "usable_cpus": info_h["usable_cpus"],
            "gpu_over_cpu": value / allcore,
            "one_thread": {"value": n * done1 / dt1, "cores": 1, "gpu_over_cpu": value * dt1 / (n * done1),
                           "sample": f"{done1} substeps on one host thread, {dt1:.1f} s"},
        }
        if workload == "pyramid-55":
            steps64, dt64 = f64_reference_baseline(bodies)
            out["cpu_baseline"]["f64_sequential"] = {
                "value": n * steps64 / dt64, "unit": "body-substeps/s", "cores": 1, "kind": "port",
                "sample": f"{steps64} steps of the pyramid from rest through the reference's f64 sequential solver "
                          f"(src/avbd3d/ref, restated in the oracle), one host thread, {dt64:.1f} s",
                "gpu_over_cpu": value * dt64 / (n * steps64)}
        else:
            out["cpu_baseline"]["f64_sequential"] = ("not run: the reference's f64 CPU solver (src/avbd3d/ref) has boxes "
                                                     "only; this scene has spheres and hulls")
        log(f"cpu baseline: {allcore:.0f} body-substeps/s on {info_h['threads']} threads "
            f"({info_h['cpu_model']}), {n * done1 / dt1:.0f} on one")
    solver.destroy()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=0,
                    help="lead-in steps before the timed window; at least the workload's own (300 for large-100k)")
    ap.add_argument("--workload", default="large-100k", choices=["large-100k", "pile-10k", "pyramid-55", "batched-4096", "adjoint-256"])
    ap.add_argument("--workers", type=int, default=8, help="host processes building the scene's hulls")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="large-100k: skip the batched and adjoint sub-objects")
    ap.add_argument("--color-cap", type=int, default=32)
    ap.add_argument("--color-cap-adjoint", type=int, default=16)
    ap.add_argument("--primal-lanes", type=int, nargs=4, default=[2 ** 31, 2 ** 15, 2 ** 12, 2 ** 11],
                    help="colour sizes from which 1, 2, 4, 8 threads per body are used in the primal (else 16)")
    ap.add_argument("--schedule", default="auto", choices=["auto", "colors", "coop", "flow", "tail", "lead"],
                    help="solve schedule of the large/pile workloads: autotuned, or fixed (for counter-collection runs)")
    ap.add_argument("--flow-blocks-per-cu", type=int, default=8, help="--schedule flow: workgroups per CU of the dataflow sweep")
    ap.add_argument("--graphs", action="store_true", help="replay each substep as a captured hipGraph (avbd_set_graphs)")
    ap.add_argument("--sparse-tail", type=int, default=None,
                    help="first colour of the fused tail (k_primal_tail, or the dataflow tail under --schedule tail); default automatic")
    ap.add_argument("--adjoint-worlds", type=int, default=256, help="adjoint-256 workload: worlds (smaller for profiling runs)")
    ap.add_argument("--adjoint-steps", type=int, default=120, help="adjoint-256 workload: steps")
    ap.add_argument("--adjoint-pair-only", action="store_true",
                    help="adjoint-256 workload: only two forward/backward pairs, no checks and no JSON line (for kernel traces)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="first run a plain device copy of this many MiB (the HBM counters' calibration kernel)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (worlds would shard across GPUs "
              "with no per-step exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batched-4096":
        print(json.dumps(bench_batched(args)), flush=True)
        return
    if args.workload == "adjoint-256":
        out = bench_adjoint(args, worlds=args.adjoint_worlds, steps=args.adjoint_steps)
        if out is not None:
            print(json.dumps(out), flush=True)
        return
    out = bench_step(args, args.workload)
    if args.workload == "large-100k" and not args.no_extras:
        b = bench_batched(args, steps=20)
        out["batched"] = {k: b[k] for k in ("metric", "value", "unit", "ms_per_step", "steps", "config", "looped_plain_step", "graphs",
                                            "plain_launches")}
        a = bench_adjoint(args, worlds=args.adjoint_worlds, steps=args.adjoint_steps, light=True)
        out["adjoint"] = {k: a[k] for k in ("metric", "value", "unit", "config", "adjoint", "gpu_worlds_steps_per_s")}
        out["adjoint"]["higher_is_better"] = False
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
