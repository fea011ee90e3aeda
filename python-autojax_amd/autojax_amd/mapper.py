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


class MapperBatch:
    """Device buffers for one set of samples; `launch` enqueues the two kernels on the current stream
    (no synchronisation), `result` checks every sample's status and returns the outputs."""

    def __init__(self, vertices, positions, max_neighbors: int = 32):
        n = len(vertices)
        if n == 0:
            raise ValueError("no samples")
        for v in vertices:
            if np.asarray(v).ndim != 2 or np.asarray(v).shape[1] != 2:
                raise ValueError("vertices must be [S, 2] (y, x)")
        sizes = np.array([np.asarray(v).shape[0] for v in vertices], dtype=np.int64)
        if sizes.min() < 3 or sizes.max() > MAX_VERTICES:
            raise ValueError(f"each sample needs 3..{MAX_VERTICES} vertices, got {sizes.min()}..{sizes.max()}")
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.V = dv.to_device(np.concatenate([np.asarray(v, dtype=np.float64) for v in vertices]))
        if isinstance(positions, (list, tuple)):
            if len(positions) != n:
                raise ValueError("one position set per sample")
            pos = np.stack([np.asarray(p, dtype=np.float64) for p in positions])
            self.stride = pos.shape[1]
        else:
            pos = np.asarray(positions, dtype=np.float64)
            self.stride = 0 if pos.ndim == 2 else pos.shape[1]
            if pos.ndim == 3 and pos.shape[0] != n:
                raise ValueError("positions [n, M, 2] must have one set per sample")
        self.M = pos.shape[-2]
        if pos.shape[-1] != 2 or self.M == 0:
            raise ValueError("positions must be [..., M, 2] (y, x) with M > 0")
        self.n = n
        self.pos = dv.to_device(pos)
        self.off_d = dv.to_device(self.offsets)
        total = int(self.offsets[-1])
        self.P = int(max_neighbors)
        self.tri = dv.empty((2 * total, 3), torch.int64)
        self.ntri = dv.empty((n,), torch.int64)
        self.nbr = dv.empty((total, self.P), torch.int64)
        self.nsz = dv.empty((total,), torch.int64)
        self.status = dv.empty((n,), torch.int32)
        self.idx = dv.empty((n, self.M, 3), torch.int64)
        self.wts = dv.empty((n, self.M, 3))

    def launch(self):
        _lib.call("aj_delaunay_mesh_batched", self.V.data_ptr(), self.off_d.data_ptr(), self.offsets.ctypes.data,
                  self.n, self.P, self.tri.data_ptr(), self.ntri.data_ptr(), self.nbr.data_ptr(), self.nsz.data_ptr(),
                  self.status.data_ptr(), dv.stream_ptr())
        _lib.call("aj_delaunay_locate_batched", self.V.data_ptr(), self.off_d.data_ptr(), self.n, self.tri.data_ptr(),
                  self.ntri.data_ptr(), self.pos.data_ptr(), self.stride, self.M, self.idx.data_ptr(),
                  self.wts.data_ptr(), dv.stream_ptr())

    def result(self) -> DeviceMappers:
        st = dv.to_host(self.status)
        bad = np.nonzero(st)[0]
        if bad.size:
            s = int(bad[0])
            raise ValueError(f"sample {s}: {STATUS.get(int(st[s]), f'status {int(st[s])}')}")
        return DeviceMappers(self.offsets, self.tri, self.ntri, self.nbr, self.nsz, self.idx, self.wts)


def build_mappers(vertices, positions, max_neighbors: int = 32) -> DeviceMappers:
    """vertices: list of f8[S_s, 2] (3 <= S_s <= 3072); positions: f8[M, 2] shared by all samples, or a
    list of f8[M, 2] (one per sample), or f8[n, M, 2].  Raises ValueError with the failing sample when
    a triangulation cannot be built."""
    mb = MapperBatch(vertices, positions, max_neighbors)
    mb.launch()
    return mb.result()
