"""Triangle meshes for the mesh-masker tests (triangle soups, (T, 3, 3) float arrays)."""

import numpy as np

from xlb_amd.workloads import icosphere  # noqa: F401  (re-exported for the tests)


def box_mesh(lo, hi):
    """The 12 triangles of the surface of the axis-aligned box [lo, hi]."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    c = np.array([[lo[0] if (k >> 0) & 1 == 0 else hi[0], lo[1] if (k >> 1) & 1 == 0 else hi[1],
                   lo[2] if (k >> 2) & 1 == 0 else hi[2]] for k in range(8)])
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for a, b, cc, d in quads:
        tris.append([c[a], c[b], c[cc]])
        tris.append([c[a], c[cc], c[d]])
    return np.array(tris)


def random_soup(shape, n, seed, size=3.0):
    rng = np.random.default_rng(seed)
    base = rng.uniform(1.0, np.asarray(shape, float) - 1.0 - size, size=(n, 1, 3))
    return base + rng.uniform(0.0, size, size=(n, 3, 3))


def sat_overlap_margin(tri, p):
    """Separating-axis test of a triangle against the unit box [p, p+1] in float64 (Akenine-Moller
    2001: 3 box axes, the triangle normal, 9 edge cross products).  Returns the largest separation
    over all axes, normalised by the axis length: < 0 overlap, > 0 separated."""
    tri = np.asarray(tri, float)
    c = np.asarray(p, float) + 0.5
    v = tri - c
    e = [v[1] - v[0], v[2] - v[1], v[0] - v[2]]
    axes = [np.eye(3)[k] for k in range(3)] + [np.cross(e[0], e[1])]
    axes += [np.cross(np.eye(3)[k], ej) for k in range(3) for ej in e]
    best = -np.inf
    for a in axes:
        n = np.linalg.norm(a)
        if n < 1e-12:
            continue
        a = a / n
        pv = v @ a
        r = 0.5 * np.abs(a).sum()
        sep = max(pv.min() - r, -r - pv.max())
        best = max(best, sep)
    return best
