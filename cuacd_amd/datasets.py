"""Seeded synthetic inputs for the benchmark workloads and the tests.

* ``ball_clouds`` - point clouds uniform in the unit ball (rejection sampling
  from [-1, 1]^3, rows kept in draw order);
* ``icosphere`` / ``noisy_icospheres`` - subdivided icospheres with per-vertex
  radial noise;
* ``cube_mesh`` / ``box_mesh`` / ``l_shape`` - closed-form test solids.
"""
import numpy as np


def ball_clouds(n_clouds=1024, n_pts=256, seed=0):
    """float32 clouds: for each cloud draw uniform(-1, 1, (512, 3)) chunks from one
    default_rng(seed) stream, keep rows with |p| <= 1 in order, truncate at n_pts."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_clouds):
        kept = []
        have = 0
        while have < n_pts:
            c = rng.uniform(-1.0, 1.0, (512, 3))
            c = c[np.einsum("ij,ij->i", c, c) <= 1.0]
            kept.append(c)
            have += len(c)
        out.append(np.concatenate(kept, 0)[:n_pts].astype(np.float32))
    return out


def icosphere(level=3):
    """Unit icosphere: (verts float64 [V,3], tris int32 [T,3]), outward winding.
    Level 3 has 642 vertices / 1280 triangles."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts), np.array(faces, np.int32)


def noisy_icospheres(n_meshes=512, level=3, sigma=0.15, seed=0):
    """Icospheres with per-vertex radial scale 1 + sigma*N(0,1), drawn mesh by
    mesh from one default_rng(seed) stream; float32 verts, int32 tris."""
    base_v, base_t = icosphere(level)
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_meshes):
        s = 1.0 + sigma * rng.standard_normal(len(base_v))
        out.append(((base_v * s[:, None]).astype(np.float32), base_t.copy()))
    return out


def box_mesh(lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """Closed box with 8 vertices and 12 outward triangles."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    t = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [1, 2, 6], [1, 6, 5], [2, 3, 7], [2, 7, 6], [3, 0, 4], [3, 4, 7]], np.int32)
    return v, t


def cube_mesh():
    return box_mesh()


def l_shape():
    """L-shaped prism (union of two boxes sharing a face), watertight, 12 verts."""
    v = np.array([[0, 0, 0], [2, 0, 0], [2, 1, 0], [1, 1, 0], [1, 2, 0], [0, 2, 0],
                  [0, 0, 1], [2, 0, 1], [2, 1, 1], [1, 1, 1], [1, 2, 1], [0, 2, 1]], np.float32)
    bottom = [(0, 2, 1), (0, 3, 2), (0, 5, 3), (3, 5, 4)]
    top = [(6, 7, 8), (6, 8, 9), (6, 9, 11), (9, 10, 11)]
    sides = []
    ring = [0, 1, 2, 3, 4, 5]
    for i in range(6):
        a, b = ring[i], ring[(i + 1) % 6]
        sides += [(a, b, b + 6), (a, b + 6, a + 6)]
    return v, np.array(bottom + top + sides, np.int32)


def voxel_solid(occ, origin=(0.0, 0.0, 0.0), h=1.0):
    """Closed surface of a union of voxels: occ is a bool [nx, ny, nz] grid of
    cells of edge h starting at origin. Every face between a filled and an
    empty cell becomes two outward triangles on the shared lattice vertices,
    so the result is watertight (2-manifold when no two filled cells touch
    only along an edge or a corner)."""
    occ = np.asarray(occ, bool)
    nx, ny, nz = occ.shape
    pad = np.zeros((nx + 2, ny + 2, nz + 2), bool)
    pad[1:-1, 1:-1, 1:-1] = occ
    vid = {}
    verts = []
    tris = []

    def v(i, j, k):
        key = (i, j, k)
        if key not in vid:
            vid[key] = len(verts)
            verts.append((origin[0] + h * i, origin[1] + h * j, origin[2] + h * k))
        return vid[key]

    # (axis, direction): the four corners of the face in CCW order seen from outside
    for i, j, k in zip(*np.nonzero(occ)):
        for ax in range(3):
            for sgn in (-1, 1):
                nb = [i + 1, j + 1, k + 1]
                nb[ax] += sgn
                if pad[tuple(nb)]:
                    continue
                u, w = (ax + 1) % 3, (ax + 2) % 3
                base = [i, j, k]
                if sgn > 0:
                    base[ax] += 1
                c = []
                for du, dw in ((0, 0), (1, 0), (1, 1), (0, 1)):
                    p = list(base)
                    p[u] += du
                    p[w] += dw
                    c.append(v(*p))
                if sgn < 0:
                    c = c[::-1]
                tris += [(c[0], c[1], c[2]), (c[0], c[2], c[3])]
    return np.array(verts, np.float32), np.array(tris, np.int32)


def hollow_tube():
    """2x2x2 box centred at the origin minus a 1x1 square through-hole along z
    (the ring-cap solid of the reference's tests/test_plane_cut.py:115-126).
    Volume 6."""
    occ = np.ones((4, 4, 4), bool)
    occ[1:3, 1:3, :] = False
    return voxel_solid(occ, (-1.0, -1.0, -1.0), 0.5)


def two_holes():
    """4x2x2 box centred at the origin minus two 0.5x0.5 through-holes along z
    at x = -1 and x = +1 (reference tests/test_plane_cut.py:128-141). Volume 15."""
    occ = np.ones((16, 8, 8), bool)
    occ[3:5, 3:5, :] = False
    occ[11:13, 3:5, :] = False
    return voxel_solid(occ, (-2.0, -1.0, -1.0), 0.25)


def l_shape_union():
    """Union of a 1x2x1 and a 1x1x1 box: the L of the reference's tests
    (tests/test_plane_cut.py:102-113), spanning [0,2]x[0,2]x[0,1]. Volume 3."""
    occ = np.zeros((2, 2, 1), bool)
    occ[0, :, 0] = True
    occ[1, 0, 0] = True
    return voxel_solid(occ, (0.0, 0.0, 0.0), 1.0)


def u_shape():
    """[0,3] x [0,2] x [0,1] minus the notch [1,2] x [1,2] x [0,1] (unit
    voxels): volume 5, two concave edges at (1,1,z) and (2,1,z)."""
    occ = np.ones((3, 2, 1), bool)
    occ[1, 1, 0] = False
    return voxel_solid(occ)


def comb_solid(teeth):
    """A comb of unit voxels: a row of 2 teeth - 1 cells along x with a tooth
    on every other cell. Each of the teeth - 1 gaps has two concave edges, the
    vertical lattice edges at (x, 1, z) on the gap's sides."""
    occ = np.zeros((2 * teeth - 1, 2, 1), bool)
    occ[:, 0, 0] = True
    occ[0::2, 1, 0] = True
    return voxel_solid(occ)


def chevron_prism():
    """The square [0,2]^2 with the V notch (0,2), (1,1), (2,2) removed,
    extruded over z in [0,1]: volume 3, hull gap 1. Its only concave edge is
    the vertical edge at (1,1); the plane x = 1 contains that edge but no
    face (the notch walls are sloped), so it splits the solid into two
    convex halves of volume 1.5."""
    poly = [(0, 0), (2, 0), (2, 2), (1, 1), (0, 2)]  # counter-clockwise
    v = np.array([(x, y, 0) for x, y in poly] + [(x, y, 1) for x, y in poly], np.float32)
    cap = [(0, 1, 3), (1, 2, 3), (0, 3, 4)]
    t = [(a, c, b) for a, b, c in cap] + [(a + 5, b + 5, c + 5) for a, b, c in cap]
    for i in range(5):
        a, b = i, (i + 1) % 5
        t += [(a, b, b + 5), (a, b + 5, a + 5)]
    return v, np.array(t, np.int32)


def cut_candidates(verts, tris, n_edges=16, n_total=30, threshold=3.49, seed=0):
    """Candidate planes (a, b, c, d) for one closed mesh, float32 [n_total, 4]:
    up to n_edges planes through sampled concave edges - edges whose interior
    dihedral angle exceeds `threshold` radians (pi + the angle between the two
    face normals when the second face's far vertex lies above the first face;
    reference rule cuda/la_concave.cu:213-233) - each the plane of one of the
    edge's two faces (alternately) through the edge midpoint, then axis cuts
    spaced evenly inside the bounding box (x, y, z in turn) up to n_total."""
    v = np.asarray(verts, np.float64).reshape(-1, 3)
    t = np.asarray(tris, np.int64).reshape(-1, 3)
    rng = np.random.default_rng(seed)
    planes = []
    if len(t):
        p0, p1, p2 = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
        fn = np.cross(p1 - p0, p2 - p0)
        ln = np.linalg.norm(fn, axis=1)
        ok = ln > 0
        fn[ok] /= ln[ok, None]
        # directed edges a->b of face f; the twin b->a names the neighbour
        a = t.reshape(-1)
        b = t[:, [1, 2, 0]].reshape(-1)
        f = np.repeat(np.arange(len(t)), 3)
        third = t[:, [2, 0, 1]].reshape(-1)
        key = a * (len(v) + 1) + b
        twin = b * (len(v) + 1) + a
        order = np.argsort(key)
        pos = np.searchsorted(key[order], twin)
        pos = np.minimum(pos, len(order) - 1)
        has = key[order][pos] == twin
        e = np.nonzero(has & (a < b))[0]
        g = order[pos[e]]
        f0, f1 = f[e], f[g]
        n0, n1 = fn[f0], fn[f1]
        up = np.einsum("ij,ij->i", v[third[g]] - v[a[e]], n0) > 0
        ang = np.arccos(np.clip(np.einsum("ij,ij->i", n0, n1), -1.0, 1.0))
        dih = np.where(up, np.pi + ang, np.pi - ang)
        cand = np.nonzero((dih > threshold) & ok[f0] & ok[f1])[0]
        if len(cand) > n_edges:
            cand = rng.choice(cand, n_edges, replace=False)
        for j, c in enumerate(cand):
            k = e[c]
            n = n0[c] if j % 2 == 0 else n1[c]
            mid = 0.5 * (v[a[k]] + v[b[k]])
            planes.append([n[0], n[1], n[2], -float(n @ mid)])
    lo, hi = v.min(0), v.max(0)
    n_axis = n_total - len(planes)
    for k in range(n_axis):
        ax = k % 3
        slot, per = k // 3, (n_axis - ax + 2) // 3
        pos = lo[ax] + (hi[ax] - lo[ax]) * (slot + 1) / (per + 1)
        pl = [0.0, 0.0, 0.0, -pos]
        pl[ax] = 1.0
        planes.append(pl)
    return np.asarray(planes, np.float32).reshape(-1, 4)
