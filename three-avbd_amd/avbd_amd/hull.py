"""Convex hull shapes: built once on the host from a point set, then packed into the hull buffer
the narrowphase reads (src/avbd3d/hull.ts convexHull / buildHull, solver.ts packHull).

The hull's triangles come from Qhull (scipy.spatial.ConvexHull); coplanar triangles are merged
into polygon faces (counter-clockwise seen from outside). Mass properties use signed tetrahedra
from the origin over fan-triangulated faces (hull.ts:380-398); vertices are stored in the
principal frame, centred on the centre of mass, since a body keeps its inertia as three
principal moments. Edges hold the two faces each separates (the Gauss-map test).
"""
import numpy as np

MAX_HULL_VERTICES = 32


def _faces(points, simplices, equations, tol=1e-9):
    """Qhull's triangles merged by plane into polygons: (outward normal, counter-clockwise loop)."""
    keys = np.round(np.asarray(equations) / tol).astype(np.int64)
    _, first, inverse = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    inverse = np.asarray(inverse).reshape(-1)
    tri = np.asarray(simplices).copy()
    n = np.asarray(equations)[:, :3]
    e1 = points[tri[:, 1]] - points[tri[:, 0]]
    e2 = points[tri[:, 2]] - points[tri[:, 0]]
    flip = np.einsum("ij,ij->i", np.cross(e1, e2), n) < 0
    tri[flip, 1], tri[flip, 2] = tri[flip, 2], tri[flip, 1].copy()
    polys = []
    for g in np.argsort(first):
        members = np.nonzero(inverse == g)[0]
        if len(members) == 1:
            polys.append((n[members[0]], [int(v) for v in tri[members[0]]]))
            continue
        directed = [(int(t[i]), int(t[(i + 1) % 3])) for t in tri[members] for i in range(3)]
        dset = set(directed)
        boundary = [(a, b) for a, b in directed if (b, a) not in dset]
        nxt = {a: b for a, b in boundary}
        start = min(nxt)
        loop = [start]
        while nxt[loop[-1]] != start:
            loop.append(nxt[loop[-1]])
            if len(loop) > len(boundary):
                raise ValueError("hull face boundary is not a single loop")
        polys.append((n[members[0]], loop))
    return polys


def convex_hull(points):
    """The convex hull of (N, 3) points and its mass properties (per unit density), or None when
    the points span no volume. Keys: vertices (V, 3) f32 principal frame, faces [(normal, d,
    verts)], edges (E, 4), volume, moments, size, radius, center, rotation (x, y, z, w)."""
    from scipy.spatial import ConvexHull
    pts = np.asarray(points, np.float64).reshape(-1, 3)
    if len(pts) < 4:
        return None
    try:
        h = ConvexHull(pts)
    except Exception:
        return None
    used = np.unique(h.simplices)
    if len(used) > MAX_HULL_VERTICES:
        raise ValueError(f"hull: {len(used)} vertices (at most {MAX_HULL_VERTICES})")
    remap = np.full(len(pts), -1, np.int64)
    remap[used] = np.arange(len(used))
    verts = pts[used]
    polys = _faces(verts, remap[h.simplices], h.equations)
    # signed tetrahedra from the origin over fan-triangulated faces (Blow & Binstock)
    fan = np.array([(loop[0], loop[k], loop[k + 1]) for _, loop in polys for k in range(1, len(loop) - 1)])
    a, b, c = verts[fan[:, 0]], verts[fan[:, 1]], verts[fan[:, 2]]
    det = np.einsum("ij,ij->i", a, np.cross(b, c))
    volume = float(det.sum() / 6)
    if not volume > 0:
        return None
    s = a + b + c
    com = (det[:, None] / 24 * s).sum(0) / volume
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    C = np.array([(det / 120 * (a[:, i] * a[:, j] + b[:, i] * b[:, j] + c[:, i] * c[:, j] + s[:, i] * s[:, j])).sum() for i, j in pairs])
    C -= volume * np.array([com[i] * com[j] for i, j in pairs])
    tr = C[0] + C[1] + C[2]
    I = np.array([[tr - C[0], -C[3], -C[4]], [-C[3], tr - C[1], -C[5]], [-C[4], -C[5], tr - C[2]]])
    values, R = np.linalg.eigh(I)
    if np.linalg.det(R) < 0:
        R[:, 2] = -R[:, 2]
    local = (verts - com) @ R
    faces = []
    for n, loop in polys:
        nl = n @ R
        faces.append((nl, float((local[loop] @ nl).max()), loop))
    face_of = {}
    for k, (_, _, loop) in enumerate(faces):
        for j, v in enumerate(loop):
            face_of[(v, loop[(j + 1) % len(loop)])] = k
    edges = [(v, w, k, face_of[(w, v)]) for (v, w), k in face_of.items() if v < w and (w, v) in face_of]
    edges.sort(key=lambda e: (e[2], e[0], e[1]))
    if len(faces) > 255 or len(edges) > 255:
        raise ValueError("hull: more than 255 faces or edges")
    size = 2 * np.abs(local).max(0)
    radius = float(np.linalg.norm(local, axis=1).max())
    return {"vertices": local.astype(np.float32), "faces": faces, "edges": np.array(edges, np.uint32).reshape(-1, 4),
            "volume": volume, "moments": values, "size": size, "radius": radius, "center": com,
            "rotation": quat_from_basis(R)}


def quat_from_basis(R):
    """Unit quaternion (x, y, z, w) of the rotation whose matrix columns are R's."""
    m = np.asarray(R, np.float64)
    t = np.trace(m)
    if t > 0:
        s = np.sqrt(t + 1.0) * 2
        q = [(m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s, 0.25 * s]
    elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
        s = np.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
        q = [0.25 * s, (m[0, 1] + m[1, 0]) / s, (m[0, 2] + m[2, 0]) / s, (m[2, 1] - m[1, 2]) / s]
    elif m[1, 1] > m[2, 2]:
        s = np.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
        q = [(m[0, 1] + m[1, 0]) / s, 0.25 * s, (m[1, 2] + m[2, 1]) / s, (m[0, 2] - m[2, 0]) / s]
    else:
        s = np.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
        q = [(m[0, 2] + m[2, 0]) / s, (m[1, 2] + m[2, 1]) / s, 0.25 * s, (m[1, 0] - m[0, 1]) / s]
    q = np.array(q)
    return q / np.linalg.norm(q)


def hull_words(shape, at):
    """The hull buffer layout (wgsl-hull.ts:14-20) as u32 words, header at vec4 offset `at`."""
    V = len(shape["vertices"])
    F = len(shape["faces"])
    index_count = sum(len(f[2]) for f in shape["faces"])
    I = (index_count + 3) // 4
    E = len(shape["edges"])
    out = np.zeros((2 + V + 2 * F + I + E) * 4, np.uint32)
    f32 = out.view(np.float32)
    vs, fs = at + 2, at + 2 + V
    is_, es = fs + 2 * F, fs + 2 * F + I
    out[0:8] = [vs, V, fs, F, es, E, is_, 0]
    for v in range(V):
        f32[(2 + v) * 4:(2 + v) * 4 + 3] = shape["vertices"][v]
    index = 0
    base_i = (2 + V + 2 * F) * 4
    for k, (n, d, loop) in enumerate(shape["faces"]):
        o = (2 + V + 2 * k) * 4
        f32[o:o + 3] = n
        f32[o + 3] = d
        out[o + 4] = index
        out[o + 5] = len(loop)
        for v in loop:
            out[base_i + index] = v
            index += 1
    for k, e in enumerate(shape["edges"]):
        o = (2 + V + 2 * F + I + k) * 4
        out[o:o + 4] = e
    return out


class HullBuffer:
    """Packs hulls one after another; a body references its hull by SHAPE_HULL + header offset."""

    def __init__(self):
        self.chunks = []
        self.top = 0

    def add(self, shape):
        at = self.top
        words = hull_words(shape, at)
        self.chunks.append(words)
        self.top += len(words) // 4
        return at

    def words(self):
        return np.concatenate(self.chunks) if self.chunks else np.zeros(4, np.uint32)
