"""CPU restatement of the source-mesh mapper (test oracle only).

The reference's likelihood takes the mesh and mapping as inputs; its upstream builds them with
autolens / PyAutoArray (external/get_dataset.py:38-60: `al.mesh.Delaunay()`, the mapper's
`pix_indexes_for_sub_slim_index`, `pix_weights_for_sub_slim_index` and the mesh's `neighbors`),
which are not part of the reference repository.  Their published algorithm is the Delaunay
triangulation of the source vertices by scipy.spatial.Delaunay (Qhull), the containing simplex of
each traced image pixel with its barycentric weights, and the mesh edges as neighbour lists; that is
what this module restates, with scipy as the triangulator.  Conventions fixed here (and by the device
mapper): vertices and positions are (y, x); a pixel outside the hull maps to its nearest vertex (first
on ties) with weight 1 and the other slots -1 / 0; a vertex coinciding with an earlier one is left out
of the mesh (Qhull leaves out one of each coincident pair; which one depends on its insertion order).

Pins (tests/test_oracle.py): every triangle's circumcircle is empty of vertices (brute force), the
triangle count is 2n - h - 2 (Euler, h hull vertices), the weights sum to 1 and reproduce the pixel,
and neighbour lists built from triangles equal those built from the edge list.
"""

from __future__ import annotations

import numpy as np


def delaunay_triangles(vertices):
    """Finite triangles i8[T, 3] of the Delaunay triangulation of vertices f8[n, 2] (y, x)."""
    from scipy.spatial import Delaunay

    return np.ascontiguousarray(Delaunay(np.asarray(vertices, dtype=np.float64)).simplices.astype(np.int64))


def neighbors_from_triangles(n, triangles, max_neighbors=None):
    """Sorted, -1 padded neighbour lists (numba.py:672-677 format) of the triangles' edges."""
    nb = [set() for _ in range(n)]
    for t in triangles:
        for e in range(3):
            a, b = int(t[e]), int(t[(e + 1) % 3])
            nb[a].add(b)
            nb[b].add(a)
    P = max_neighbors if max_neighbors is not None else max(1, max(len(x) for x in nb))
    out = np.full((n, P), -1, dtype=np.int64)
    sizes = np.zeros(n, dtype=np.int64)
    for i, x in enumerate(nb):
        v = sorted(x)[:P]
        out[i, : len(v)] = v
        sizes[i] = len(v)
    return out, sizes


def _area2(a, b, c):
    # twice the signed area in the (x, y) plane of points given as (y, x)
    return (b[..., 1] - a[..., 1]) * (c[..., 0] - a[..., 0]) - (c[..., 1] - a[..., 1]) * (b[..., 0] - a[..., 0])


def barycentric_mapping(vertices, triangles, positions):
    """pix_indexes i8[M, 3] / pix_weights f8[M, 3] of positions f8[M, 2] (y, x) in the triangles: the
    first triangle containing the pixel (all three sub-areas >= 0 after orienting it counter-clockwise),
    else the nearest vertex with weight 1."""
    V = np.asarray(vertices, dtype=np.float64)
    T = np.asarray(triangles, dtype=np.int64).copy()
    # counter-clockwise in (x, y)
    neg = _area2(V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]) < 0
    T[neg, 0], T[neg, 1] = T[neg, 1].copy(), T[neg, 0].copy()
    A, B, C = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    area = _area2(A, B, C)
    pos = np.asarray(positions, dtype=np.float64)
    M = pos.shape[0]
    idx = np.full((M, 3), -1, dtype=np.int64)
    w = np.zeros((M, 3))
    for m in range(M):
        p = pos[m]
        a0 = _area2(p, B, C)
        a1 = _area2(A, p, C)
        a2 = _area2(A, B, p)
        inside = np.nonzero((a0 >= 0) & (a1 >= 0) & (a2 >= 0))[0]
        if inside.size:
            t = inside[0]
            idx[m] = T[t]
            w[m] = (a0[t] / area[t], a1[t] / area[t], a2[t] / area[t])
        else:
            d = np.sum((V - p) ** 2, axis=1)
            idx[m, 0] = int(np.argmin(d))
            w[m, 0] = 1.0
    return idx, w


def mapping_dense(pix_indexes, pix_weights, S):
    """T f8[M, S] with T[m, idx] += w over the live (>= 0) entries."""
    M = pix_indexes.shape[0]
    T = np.zeros((M, S))
    for b in range(pix_indexes.shape[1]):
        live = pix_indexes[:, b] >= 0
        np.add.at(T, (np.nonzero(live)[0], pix_indexes[live, b]), pix_weights[live, b])
    return T
