"""The oracle's colouring (src/avbd2d/gpu/wgsl-topology.ts:155-328 with the priorities of
avbd2d/gpu/layout.ts:107-115) pinned by a brute force, on the CPU.

Jones-Plassmann with first-fit colours a body once it beats every uncoloured neighbour, so all
its higher-priority neighbours are coloured by then and none of its lower ones: it takes the
smallest colour none of its higher-priority neighbours holds. From an uncoloured start with
enough rounds that is exactly the sequential greedy colouring in decreasing priority. The
priorities are lowbias32 (C. Wellons' published integer hash: xor-shift 16, multiply 0x7feb352d,
xor-shift 15, multiply 0x846ca68b, xor-shift 16) of index + 0x9e3779b9, ties by index, written
again here from the published steps. Graphs come from spheres whose overlaps are the edges."""
import itertools

import numpy as np

import oracle
from avbd_amd import records

M32 = 0xFFFFFFFF


def lowbias32(x):
    x &= M32
    x ^= x >> 16
    x = (x * 0x7FEB352D) & M32
    x ^= x >> 15
    x = (x * 0x846CA68B) & M32
    x ^= x >> 16
    return x


def greedy(n, edges):
    prio = [lowbias32(i + 0x9E3779B9) for i in range(n)]
    order = sorted(range(n), key=lambda i: (prio[i], i), reverse=True)
    nbr = [set() for _ in range(n)]
    for a, b in edges:
        nbr[a].add(b)
        nbr[b].add(a)
    col = [None] * n
    for i in order:
        used = {col[j] for j in nbr[i] if col[j] is not None}
        col[i] = next(c for c in itertools.count() if c not in used)
    return col


def sphere_graph(centres, r=0.5):
    """Sphere records (density 1, no gravity) and their overlap graph; no pair is within 3% of
    touching, so the narrowphase's answer is unambiguous."""
    c = np.asarray(centres, np.float64)
    edges = []
    for i in range(len(c)):
        for j in range(i):
            d = np.linalg.norm(c[i] - c[j])
            assert abs(d - 2 * r) > 0.03 * 2 * r
            if d < 2 * r:
                edges.append((i, j))
    recs = np.array([records.sphere_record(tuple(p), r, 1.0) for p in c])
    return recs, edges


def oracle_colours(recs):
    m = oracle.Model(32)
    m.set_params(oracle.make_params(gravity=0.0, color_cap=64, color_rounds=64, iterations=1))
    m.set_bodies(records.bodies_to_oracle(recs.astype(np.float32)))
    m.step(1)
    c = m.counts()
    assert c["clashes"] == 0
    return [int(x) & 0xFF for x in m.colors()], c


def test_lowbias32_known_values():
    """The published hash's fixed structure: 0 maps to 0, and it is a bijection on a sample."""
    assert lowbias32(0) == 0
    sample = [lowbias32(x) for x in range(1 << 16)]
    assert len(set(sample)) == len(sample)


def test_hand_built_graphs_colour_as_the_greedy_order():
    s = 0.9  # centre distance for touching spheres of radius 0.5
    path = [(k * s, 0, 0) for k in range(9)]
    star = [(0, 0, 0)] + [tuple(v * s) for v in np.vstack([np.eye(3), -np.eye(3)])]
    tet = [(0, 0, 0), (s, 0, 0), (s / 2, s * 3 ** 0.5 / 2, 0), (s / 2, s * 3 ** 0.5 / 6, s * (2 / 3) ** 0.5)]
    ring = [(2 * np.cos(a), 2 * np.sin(a), 0) for a in np.linspace(0, 2 * np.pi, 14, endpoint=False)]  # an even cycle
    for name, centres in (("path", path), ("star", star), ("K4", tet), ("cycle", ring)):
        recs, edges = sphere_graph(centres)
        got, c = oracle_colours(recs)
        want = greedy(len(recs), edges)
        assert got == want, (name, got, want)
        assert c["manifolds"] == len(edges), name
    assert sorted(oracle_colours(sphere_graph(tet)[0])[0]) == [0, 1, 2, 3]


def test_random_geometric_graphs_colour_as_the_greedy_order():
    rng = np.random.default_rng(5)
    for trial in range(6):
        n = int(rng.integers(60, 200))
        centres = []
        while len(centres) < n:
            p = rng.uniform(0, 6.0, 3)
            if all(abs(np.linalg.norm(p - q) - 1.0) > 0.03 for q in centres):
                centres.append(p)
        recs, edges = sphere_graph(centres)
        got, c = oracle_colours(recs)
        assert c["manifolds"] == len(edges), trial
        assert got == greedy(n, edges), trial
