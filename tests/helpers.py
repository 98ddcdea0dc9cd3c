"""Shared helpers for the parity tests (test code; uses the oracle as the checker)."""
import math

import numpy as np

import oracle

PARITY_SCENES = {"Ground": 1, "Dynamic Friction": 2, "Static Friction": 3, "Rope": 5, "Heavy Rope": 6, "Spring": 7,
                 "Spring Ratio": 8, "Stack": 9, "Stack Ratio": 10, "Breakable": 13}


def lcg(seed):
    """The reference tests' generator: s = s*1103515245 + 12345 mod 2^32, /2^32."""
    s = [seed]

    def rand():
        s[0] = (s[0] * 1103515245 + 12345) & 0xFFFFFFFF
        return s[0] / 2 ** 32
    return rand


def qnormalize(q):
    q = np.asarray(q, np.float64)
    return q / np.sqrt((q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) + q[3] * q[3])


def sorted_contacts(man, con):
    """Manifolds sorted by (a, b) with their contact rows, for order-free comparison."""
    order = sorted(range(len(man)), key=lambda m: (int(man[m][0]), int(man[m][1])))
    out = []
    for m in order:
        first, cnt = int(man[m][2]), int(man[m][3]) & 15
        out.append((man[m], con[first:first + cnt]))
    return out


def contact_match(ref_man, ref_con, got_man, got_con, bodies, tol=1e-4, lam_tol=1e-3):
    """The reference test's contact match (tests-gpu/avbd3d-gpu.gpu.test.ts:57-86): every
    reference contact between bodies that can move found at the same anchors. Returns
    (expected, found, feature_mismatches, lambda_ok)."""
    index = {}
    for m in range(len(got_man)):
        index[(int(got_man[m][0]), int(got_man[m][1]))] = m
    expected = found = keymis = 0
    lam_ok = True
    for m in range(len(ref_man)):
        a, b = int(ref_man[m][0]), int(ref_man[m][1])
        if bodies[a, 19] <= 0 and bodies[b, 19] <= 0:
            continue
        for c in range(int(ref_man[m][2]), int(ref_man[m][2] + ref_man[m][3])):
            expected += 1
            gm = index.get((a, b))
            if gm is None:
                continue
            g0, gc = int(got_man[gm][2]), int(got_man[gm][3]) & 15
            for k in range(g0, g0 + gc):
                if np.abs(got_con[k, 0:3] - ref_con[c, 0:3]).max() < tol and np.abs(got_con[k, 4:7] - ref_con[c, 4:7]).max() < tol:
                    found += 1
                    if (int(got_con[k, 3]) & 0x7FFFFFFF) != (int(ref_con[c, 3]) & 0x7FFFFFFF):
                        keymis += 1
                    else:
                        scale = max(1.0, abs(ref_con[c, 12]))
                        if np.abs(got_con[k, 12:15] - ref_con[c, 12:15]).max() >= lam_tol * scale:
                            lam_ok = False
                    break
    return expected, found, keymis, lam_ok


def random_box_pairs(n_pairs=300, seed=7):
    """300 overlapping random box pairs far apart from each other (avbd3d-gpu.gpu.test.ts:101-118)."""
    rand = lcg(seed)
    ref = oracle.RefSolver()
    for p in range(n_pairs):
        base = [(p % 20) * 20, (p // 20) * 20, 0]
        for k in range(2):
            size = [0.4 + rand() * 1.6, 0.4 + rand() * 1.6, 0.4 + rand() * 1.6]
            off = [0, 0, 0] if k == 0 else [(rand() - 0.5) * 1.2, (rand() - 0.5) * 1.2, (rand() - 0.5) * 1.2]
            q = qnormalize([rand() - 0.5, rand() - 0.5, rand() - 0.5, rand() - 0.5])
            ref.add_body(size, 1, 0.5, [base[0] + off[0], base[1] + off[1], base[2] + off[2]], rot=q)
    return ref


def random_broadphase_scene(n=3000, seed=11):
    """3000 random boxes, a few large, one ignored pair (avbd3d-gpu.gpu.test.ts:144-157)."""
    rand = lcg(seed)
    ref = oracle.RefSolver()
    for i in range(n):
        big = i % 499 == 0
        s = [30, 30, 1] if big else [0.3 + rand(), 0.3 + rand(), 0.3 + rand()]
        dens = 1 if rand() < 0.9 else 0
        pos = [(rand() - 0.5) * 40, (rand() - 0.5) * 40, (rand() - 0.5) * 40]
        rot = None if big else qnormalize([rand() - 0.5, rand() - 0.5, rand() - 0.5, rand() - 0.5])
        ref.add_body(s, dens, 0.5, pos, rot=rot)
    ref.add_ignore(2, 1)
    return ref


def brute_force_pairs(bodies, ignore=()):
    """The broadphase predicate evaluated over every pair in f64 (avbd3d-gpu.gpu.test.ts:166-205)."""
    B = np.asarray(bodies, np.float64)
    n = len(B)

    def rot(q, v):
        u = q[:3]
        t = 2 * np.cross(u, v)
        return v + q[3] * t + np.cross(u, t)
    axes = np.array([[rot(B[i, 4:8], e) for e in np.eye(3)] for i in range(n)])
    half = np.array([np.abs(axes[i] * (B[i, 16:19, None] * 0.5)).sum(0) for i in range(n)])
    out = set()
    ign = {(max(a, b), min(a, b)) for a, b in ignore}
    for i in range(n):
        for j in range(i):
            if B[i, 19] <= 0 and B[j, 19] <= 0:
                continue
            if (i, j) in ign:
                continue
            d = B[i, 0:3] - B[j, 0:3]
            if np.dot(d, d) > (B[i, 23] + B[j, 23]) ** 2:
                continue
            if (np.abs(d) > half[i] + half[j]).any():
                continue
            ext = lambda k, nrm: sum(B[k, 16 + c] * 0.5 * abs(np.dot(nrm, axes[k][c])) for c in range(3))
            if any(abs(np.dot(d, nrm)) - ext(i, nrm) - ext(j, nrm) > 1e-3 for nrm in list(axes[i]) + list(axes[j])):
                continue
            out.add((i, j))
    return out


def oracle_params_of(solver):
    """The oracle's params equal to a GPU solver's (its f32 values, exactly)."""
    p = solver.params
    return oracle.make_params(dt=p.dt, gravity=p.gravity, iterations=p.iterations, alpha=p.alpha, beta_lin=p.beta_lin,
                              beta_ang=p.beta_ang, gamma=p.gamma, flags=p.flags & 0x7F, color_cap=p.color_cap,
                              color_rounds=p.color_rounds, margin=p.contact_margin, up=tuple(p.up),
                              lanes=tuple(p.primal_lanes))


def oracle_from_gpu(solver, hulls=None, threads=1):
    """An f32 oracle model holding a GPU solver's whole state: params (colour cap and lane rule as
    adapted), body and joint records, hulls, the last step's manifolds and contacts, the colour
    state, the step count and updateRefs' reference poses, so its next step is the GPU's next."""
    from avbd_amd import records
    m = oracle.Model(32)
    m.set_params(oracle_params_of(solver))
    m.set_bodies(records.bodies_to_oracle(solver.read_bodies()))
    if solver.joint_count:
        m.set_joints(solver.read_joints().astype(np.float64), solver._joint_info)
    if hulls is not None:
        m.set_hulls(hulls)
    man, con = solver.read_contacts_raw()
    m.set_prev(*records.contacts_to_oracle(man, con))
    m.set_color_state(solver.read_colors()[0])
    m.set_step_count(solver.read_counters()["step"])
    m.set_refs(solver.read_refs().astype(np.float64))
    m.set_threads(threads)
    return m


def canonical_contacts(man, con):
    """Manifolds sorted by (a, b), each one's `first` rewritten to its place in the sorted order,
    and the contact rows gathered in that order: two contact lists that differ only by emission
    order give equal arrays. Manifold word 3 keeps its count bits only (the step stamp aside)."""
    man = np.asarray(man)
    con = np.asarray(con)
    if len(man) == 0:
        return man.reshape(0, 8), con.reshape(0, 16)
    order = np.lexsort((man[:, 1], man[:, 0]))
    M = man[order].copy()
    cnt = (M[:, 3].astype(np.int64) & 15)
    offs = np.cumsum(cnt) - cnt
    idx = np.repeat(M[:, 2].astype(np.int64) - offs, cnt) + np.arange(int(cnt.sum()))
    M[:, 2] = offs
    M[:, 3] = cnt
    return M, con[idx]


# --- records as the reference's GpuSolver3D packs them (scene edits) ----------------------------
BIG = np.float32(3.0e38)


def fresh_joint():
    """freshJoint (solver.ts:132-138)."""
    o = np.zeros(40, np.float32)
    o[11] = BIG
    o[23] = BIG
    o[31] = BIG
    o[35] = 1.0
    o[39] = 1.0
    return o


def torque_arm(sa, sb):
    return np.float32(sum((float(x) + float(y)) ** 2 for x, y in zip(sa, sb)))


def spec_record(spec, sizes, fracture=math.inf, linear=False):
    """appendJoints' record of one spec (solver.ts:1661-1683)."""
    fin = lambda x: BIG if x == math.inf else (-BIG if x == -math.inf else np.float32(x))
    o = fresh_joint()
    o[3] = BIG
    o[7] = BIG if "angular" not in spec else fin(spec["angular"])
    limit = fin(spec.get("fracture", fracture))
    o[11] = -limit if spec.get("linear", linear) else limit
    if "yield" in spec:
        o[23] = fin(spec["yield"])
    if "breakBend" in spec:
        o[31] = fin(spec["breakBend"])
    if "rest" in spec:
        q = np.asarray(spec["rest"], np.float64)
        q = (q / math.sqrt(float((q * q).sum()))).astype(np.float32)
        o[32:36] = q
        o[36:40] = q
    o[24:27] = spec["rA"]
    o[28:31] = spec["rB"]
    o[15] = torque_arm(sizes[spec["a"]], sizes[spec["b"]])
    return o


def aux_factors(rate, dt):
    x = -(np.float32(rate) * np.float32(dt))
    return np.float32(math.exp(float(x)))


def new_box(pos, size, density, friction=0.5, rot=(0, 0, 0, 1)):
    from avbd_amd import records
    r = records.box_record(pos, size, density, friction, rot=rot).astype(np.float32)
    r[11] = 1.0  # a new occupant: no contact history (pack, solver.ts:1200)
    return r
