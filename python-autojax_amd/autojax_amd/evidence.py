"""Batched log-evidence on one shared W-tilde (device API).

A *sample* is one lens model: a mapping (pix_indexes / pix_weights, [M, B]),
a source neighbour graph and a regularisation coefficient.  Many samples
share the dataset (grid, visibilities, W-tilde, dirty image), as in a
sampler that proposes lens models against fixed data.  ``SampleBatch``
uploads a list of samples once; ``log_evidence`` evaluates all of them with
one pass over W-tilde (numba.py:755-857 per sample, explicit coefficient).

log|H| uses a rank-one deflation when the neighbour graph is a connected,
symmetric graph with sizes equal to the listed neighbour counts: then
H 1 = 1e-8 * 1 exactly and
    log|H| = log|H + (beta/S) 1 1^T| - log(1e-8 + beta) + log(1e-8),
with beta = c^2 * mean degree, which removes the 1e-8 pivot that makes a
plain factorisation of H lose digits at large coefficients.
"""

from __future__ import annotations

from collections import deque

import numpy as np
import torch

from . import _lib
from . import device as dv
from .api import assignment_weights, validate_mapping, validate_neighbors


def deflation_beta(neighbors, neighbors_sizes, coefficient) -> float:
    """beta for the rank-one deflation of log|H|, or 0 if the graph does not allow it."""
    nb = np.asarray(neighbors)
    sz = np.asarray(neighbors_sizes)
    S = sz.size
    live = np.ones(nb.shape, bool)
    if nb.shape[1] > 1:
        live[:, 1:] = np.cumprod(nb[:, :-1] != -1, axis=1).astype(bool)
    live &= nb >= 0
    if np.any(live.sum(axis=1) != sz):
        return 0.0
    if np.any((nb == np.arange(S)[:, None]) & live):
        return 0.0
    adj = [nb[i][live[i]] for i in range(S)]
    seen = np.zeros(S, bool)
    seen[0] = True
    q = deque([0])
    while q:
        i = q.popleft()
        for k in adj[i]:
            if not seen[k]:
                seen[k] = True
                q.append(k)
    if not seen.all():
        return 0.0
    return float(coefficient) ** 2 * float(sz.mean())


class SampleBatch:
    """Samples uploaded to the device, ready for ``log_evidence``."""

    def __init__(self, samples, M, deflate=True, validate=True):
        if len(samples) == 0:
            raise ValueError("empty batch")
        B = samples[0].pix_indexes_for_sub_slim_index.shape[1]
        P = max(s.neighbors.shape[1] for s in samples)
        self.n = len(samples)
        self.M = int(M)
        self.B = int(B)
        self.P = int(P)
        self.S_host = np.array([s.pixels for s in samples], dtype=np.int64)
        idx = np.empty((self.n, M, B), np.int64)
        wts = np.empty((self.n, M, B), np.float64)
        nbr = np.full((int(self.S_host.sum()), P), -1, np.int64)
        nsz = np.empty(int(self.S_host.sum()), np.int64)
        off = 0
        for b, s in enumerate(samples):
            if s.pix_indexes_for_sub_slim_index.shape != (M, B):
                raise ValueError(f"sample {b}: mapping shape {s.pix_indexes_for_sub_slim_index.shape} != {(M, B)}")
            if validate:
                # the batched Cholesky factors the lower triangles: the regularisation matrix must be symmetric
                validate_mapping(s.pix_indexes_for_sub_slim_index, s.pixels)
                validate_neighbors(s.neighbors, s.neighbors_sizes, symmetric=True)
            idx[b] = s.pix_indexes_for_sub_slim_index
            # a row naming a source pixel twice: the dense mapping's assignment (numba.py:375, the last wins)
            wts[b] = assignment_weights(s.pix_indexes_for_sub_slim_index, s.pix_weights_for_sub_slim_index)
            S = s.pixels
            nbr[off : off + S, : s.neighbors.shape[1]] = s.neighbors
            nsz[off : off + S] = s.neighbors_sizes
            off += S
        coef = np.array([float(s.coefficient) for s in samples], np.float64)
        # mean degree of deflatable graphs (0 = plain factorisation of H)
        deg = np.array(
            [deflation_beta(s.neighbors, s.neighbors_sizes, 1.0) if deflate else 0.0 for s in samples], np.float64
        )
        self._setup(dv.to_device(idx), dv.to_device(wts), dv.to_device(nbr), dv.to_device(nsz), coef, deg)

    @classmethod
    def from_mappers(cls, mappers, coefficients, deflate=True):
        """A batch over the device mapper's output (`autojax_amd.mapper.build_mappers`): the mapping and
        neighbour lists stay on the device; only the neighbour lists' graph check for the deflated
        log|H| reads them back."""
        self = cls.__new__(cls)
        self.n = len(mappers.offsets) - 1
        self.M = int(mappers.pix_indexes.shape[1])
        self.B = int(mappers.pix_indexes.shape[2])
        self.P = int(mappers.neighbors.shape[1])
        self.S_host = np.diff(mappers.offsets).astype(np.int64)
        coef = np.asarray(coefficients, np.float64).reshape(self.n).copy()
        deg = np.zeros(self.n)
        if deflate:
            nb, sz = dv.to_host(mappers.neighbors), dv.to_host(mappers.neighbors_sizes)
            for b in range(self.n):
                a, e = int(mappers.offsets[b]), int(mappers.offsets[b + 1])
                deg[b] = deflation_beta(nb[a:e], sz[a:e], 1.0)
        self._setup(mappers.pix_indexes, mappers.pix_weights, mappers.neighbors, mappers.neighbors_sizes, coef, deg)
        return self

    def _setup(self, idx, wts, nbr, nsz, coef, deg):
        # incremented by every evaluation; a backward checks it to know the workspace's factors are its own
        self.evaluations = 0
        self.coef_host = coef
        self.deg_host = deg
        self.beta_host = self.coef_host**2 * self.deg_host
        self.idx = idx.contiguous()
        self.wts = wts.contiguous()
        self.nbr = nbr.contiguous()
        self.nsz = nsz.contiguous()
        self.ws = dv.workspace(
            _lib.query("aj_log_evidence_batched_workspace_bytes", self.M, self.B, self.n,
                       self.S_host.ctypes.data_as(_lib.C.c_void_p))
        )
        dev = self.idx.device
        self.out = torch.empty(self.n, dtype=torch.float64, device=dev)
        self.logdets = torch.empty((self.n, 2), dtype=torch.float64, device=dev)
        self.info = torch.zeros((self.n, 2), dtype=torch.int64, device=dev)
        self.recon = torch.empty(int(self.S_host.sum()), dtype=torch.float64, device=dev)

    def set_coefficients(self, coefficients):
        c = np.asarray(coefficients, np.float64).reshape(self.n).copy()
        self.coef_host = c
        self.beta_host = c**2 * self.deg_host

    def log_evidence(self, w_tilde, dirty_image, chi2_noise_norm, check=False):
        """Evaluate all samples; returns the device tensor [n] (valid after the stream runs).

        ``check`` synchronises and raises LinAlgError for a non-positive-definite matrix.
        """
        if w_tilde.shape != (self.M, self.M) or w_tilde.dtype != torch.float64:
            raise ValueError("w_tilde must be float64 [M, M] on the device")
        self.evaluations += 1
        _lib.call(
            "aj_log_evidence_batched",
            w_tilde.data_ptr(), self.M, dirty_image.data_ptr(), chi2_noise_norm.data_ptr(), self.n, self.B, self.P,
            self.S_host.ctypes.data_as(_lib.C.c_void_p), self.idx.data_ptr(), self.wts.data_ptr(),
            self.nbr.data_ptr(), self.nsz.data_ptr(), self.coef_host.ctypes.data_as(_lib.C.c_void_p),
            self.beta_host.ctypes.data_as(_lib.C.c_void_p), self.out.data_ptr(), self.recon.data_ptr(),
            self.logdets.data_ptr(), self.info.data_ptr(), self.ws.data_ptr(), self.ws.numel(), dv.stream_ptr(),
        )
        if check:
            self._check_info()
        return self.out

    def log_evidence_compact(self, w_compact, native_index_for_slim_index, dirty_image, chi2_noise_norm, check=False):
        """As ``log_evidence`` with W-tilde given by its lattice table w_compact [N, 2N-1] and the
        device int64 [M, 2] grid positions of the image pixels (never expanded to M x M)."""
        N = w_compact.shape[0]
        if w_compact.shape != (N, 2 * N - 1) or w_compact.dtype != torch.float64:
            raise ValueError("w_compact must be float64 [N, 2N-1] on the device")
        if native_index_for_slim_index.shape != (self.M, 2) or native_index_for_slim_index.dtype != torch.int64:
            raise ValueError("native_index_for_slim_index must be int64 [M, 2] on the device")
        self.evaluations += 1
        _lib.call(
            "aj_log_evidence_batched_compact",
            w_compact.data_ptr(), N, native_index_for_slim_index.data_ptr(), self.M, dirty_image.data_ptr(),
            chi2_noise_norm.data_ptr(), self.n, self.B, self.P, self.S_host.ctypes.data_as(_lib.C.c_void_p),
            self.idx.data_ptr(), self.wts.data_ptr(), self.nbr.data_ptr(), self.nsz.data_ptr(),
            self.coef_host.ctypes.data_as(_lib.C.c_void_p), self.beta_host.ctypes.data_as(_lib.C.c_void_p),
            self.out.data_ptr(), self.recon.data_ptr(), self.logdets.data_ptr(), self.info.data_ptr(),
            self.ws.data_ptr(), self.ws.numel(), dv.stream_ptr(),
        )
        if check:
            self._check_info()
        return self.out

    def _check_info(self):
        """Raise for a failed evaluation: RuntimeError when the persistent factorisation schedule gave up a
        dependency wait (info -1), LinAlgError for a matrix that is not positive definite (its first failing
        column, 1-based)."""
        if bool((self.info < 0).any()):
            raise RuntimeError("batched Cholesky (persistent schedule): a dependency wait exceeded its limit; "
                               "the evaluation is invalid")
        bad = torch.nonzero(self.info.any(dim=1)).flatten().tolist()
        if bad:
            raise np.linalg.LinAlgError(f"samples {bad}: matrix not positive definite, info={self.info[bad].tolist()}")

    def reconstructions(self):
        off = np.concatenate([[0], np.cumsum(self.S_host)])
        return [self.recon[off[b] : off[b + 1]] for b in range(self.n)]


def chi2_noise_norm(data, noise_map):
    """Device pair (chi.chi, noise normalisation) of numba.py:846-847 and :752."""
    d = torch.view_as_real(dv.to_device(data)).reshape(-1)
    n = torch.view_as_real(dv.to_device(noise_map)).reshape(-1)
    out = dv.empty((2,))
    _lib.call("aj_chi_squared_noise_normalization", d.data_ptr(), n.data_ptr(), d.numel() // 2, out.data_ptr(),
              dv.stream_ptr())
    return out


def w_tilde(noise_map_real, uv_wavelengths, grid_radians_slim, precision="fp64", out=None, workspace=None):
    """Device W-tilde [M, M] float64 at the requested arithmetic precision."""
    n = dv.to_device(noise_map_real)
    uv = dv.to_device(uv_wavelengths)
    g = dv.to_device(grid_radians_slim)
    K, M = n.shape[0], g.shape[0]
    prec = _lib.PRECISIONS[precision]
    if out is None:
        out = dv.empty((M, M))
    if workspace is None:
        workspace = dv.workspace(_lib.query("aj_w_tilde_workspace_bytes", K, M, prec))
    _lib.call("aj_w_tilde_curvature_interferometer_from", n.data_ptr(), uv.data_ptr(), g.data_ptr(), K, M, prec,
              out.data_ptr(), workspace.data_ptr(), workspace.numel(), dv.stream_ptr())
    return out


def dirty_image(data, noise_map, uv_wavelengths, grid_radians_slim):
    from .api import dirty_image_from

    return dirty_image_from(dv.to_device(data), dv.to_device(noise_map), dv.to_device(uv_wavelengths),
                            dv.to_device(grid_radians_slim))
