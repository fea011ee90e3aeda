"""Source-mesh mapper on the device: what the reference's upstream builds before calling the
likelihood (external/get_dataset.py:38-60 -- a Delaunay mesh of each lens-model sample's source
vertices, its neighbour lists, and the image pixels' barycentric mapping into it).

`build_mappers` takes, per sample, the source vertices (y, x) and the image pixels' source-plane
positions, and returns the likelihood's inputs in the reference's formats: `neighbors` /
`neighbors_sizes` (numba.py:672-677, sorted, -1 padded) and `pix_indexes_for_sub_slim_index` /
`pix_weights_for_sub_slim_index` (numba.py:355-360; a pixel outside the mesh's hull maps to its
nearest vertex with weight 1).  All samples go through two launches (`aj_delaunay_mesh_batched`,
`aj_delaunay_locate_batched`); the results stay on the device unless asked for on the host.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from . import device as dv

MAX_VERTICES = 3072
STATUS = {1: "more than 3072 vertices", 2: "all vertices on one line", 3: "an insertion exceeded the cavity capacity",
          4: "an insertion met an inconsistent cavity (degenerate geometry)", 5: "a vertex has more than max_neighbors neighbours"}


@dataclass
class DeviceMappers:
    """Concatenated device results for n samples: sample s owns vertex rows offsets[s]:offsets[s+1]."""

    offsets: np.ndarray            # i8[n+1] (host)
    triangles: torch.Tensor        # i8[2 * total, 3]; sample s from row 2 * offsets[s], local vertex ids
    n_triangles: torch.Tensor      # i8[n]
    neighbors: torch.Tensor        # i8[total, P]
    neighbors_sizes: torch.Tensor  # i8[total]
    pix_indexes: torch.Tensor      # i8[n, M, 3]
    pix_weights: torch.Tensor      # f8[n, M, 3]

    def sample(self, s: int) -> dict:
        """Host arrays of sample s in the reference's formats (neighbour lists trimmed to the sample's
        largest degree, at least one column)."""
        a, b = int(self.offsets[s]), int(self.offsets[s + 1])
        nt = int(self.n_triangles[s].item())
        sizes = dv.to_host(self.neighbors_sizes[a:b])
        P = max(1, int(sizes.max()) if sizes.size else 1)
        return dict(triangles=dv.to_host(self.triangles[2 * a: 2 * a + nt]),
                    neighbors=np.ascontiguousarray(dv.to_host(self.neighbors[a:b, :P])),
                    neighbors_sizes=sizes,
                    pix_indexes_for_sub_slim_index=dv.to_host(self.pix_indexes[s]),
                    pix_weights_for_sub_slim_index=dv.to_host(self.pix_weights[s]))


def build_mappers(vertices, positions, max_neighbors: int = 32) -> DeviceMappers:
    """vertices: list of f8[S_s, 2] (3 <= S_s <= 3072); positions: f8[M, 2] shared by all samples, or a
    list of f8[M, 2] (one per sample), or f8[n, M, 2].  Raises ValueError with the failing sample when
    a triangulation cannot be built."""
    n = len(vertices)
    if n == 0:
        raise ValueError("no samples")
    sizes = np.array([np.asarray(v).shape[0] for v in vertices], dtype=np.int64)
    if sizes.min() < 3 or sizes.max() > MAX_VERTICES:
        raise ValueError(f"each sample needs 3..{MAX_VERTICES} vertices, got {sizes.min()}..{sizes.max()}")
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    for v in vertices:
        if np.asarray(v).ndim != 2 or np.asarray(v).shape[1] != 2:
            raise ValueError("vertices must be [S, 2] (y, x)")
    V = dv.to_device(np.concatenate([np.asarray(v, dtype=np.float64) for v in vertices]))
    if isinstance(positions, (list, tuple)):
        if len(positions) != n:
            raise ValueError("one position set per sample")
        pos = np.stack([np.asarray(p, dtype=np.float64) for p in positions])
        stride = pos.shape[1]
    else:
        pos = np.asarray(positions, dtype=np.float64)
        stride = 0 if pos.ndim == 2 else pos.shape[1]
        if pos.ndim == 3 and pos.shape[0] != n:
            raise ValueError("positions [n, M, 2] must have one set per sample")
    M = pos.shape[-2]
    if pos.shape[-1] != 2 or M == 0:
        raise ValueError("positions must be [..., M, 2] (y, x) with M > 0")
    P_d = dv.to_device(pos)
    off_d = dv.to_device(offsets)
    total = int(offsets[-1])
    P = int(max_neighbors)
    tri = dv.empty((2 * total, 3), torch.int64)
    ntri = dv.empty((n,), torch.int64)
    nbr = dv.empty((total, P), torch.int64)
    nsz = dv.empty((total,), torch.int64)
    status = dv.empty((n,), torch.int32)
    _lib.call("aj_delaunay_mesh_batched", V.data_ptr(), off_d.data_ptr(), offsets.ctypes.data, n, P, tri.data_ptr(),
              ntri.data_ptr(), nbr.data_ptr(), nsz.data_ptr(), status.data_ptr(), dv.stream_ptr())
    idx = dv.empty((n, M, 3), torch.int64)
    wts = dv.empty((n, M, 3))
    _lib.call("aj_delaunay_locate_batched", V.data_ptr(), off_d.data_ptr(), n, tri.data_ptr(), ntri.data_ptr(),
              P_d.data_ptr(), stride, M, idx.data_ptr(), wts.data_ptr(), dv.stream_ptr())
    st = dv.to_host(status)
    bad = np.nonzero(st)[0]
    if bad.size:
        s = int(bad[0])
        raise ValueError(f"sample {s}: {STATUS.get(int(st[s]), f'status {int(st[s])}')}")
    return DeviceMappers(offsets, tri, ntri, nbr, nsz, idx, wts)
