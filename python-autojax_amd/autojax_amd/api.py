"""Drop-in for the likelihood path of ``autojax.numba`` on MI355X.

Every public function here has the name, parameter names and parameter order
of the function of the same name in the reference's ``src/autojax/numba.py``
(cited per function), so that the reference's name-binding harness
(``tests.py:564-575``) can call it unchanged.  Arguments may be numpy arrays
(results come back as numpy, like numba) or torch device tensors (results
stay on the device, on the current stream).  Extra options are keyword-only
with defaults, so positional calls are the reference's.

All arithmetic runs in libautojax_amd (HIP, gfx950); this module validates
arguments, allocates outputs and workspaces, and maps errors to exceptions.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import device as dv

TWO_PI = 2.0 * np.pi
LOG_TWO_PI = np.log(TWO_PI)


# ---------------------------------------------------------------------------
# argument handling
# ---------------------------------------------------------------------------

def _is_t(x) -> bool:
    return isinstance(x, torch.Tensor)


def _dev(x, dtype, name, ndim=None):
    """Validate dtype/rank and return a contiguous device tensor."""
    if _is_t(x):
        want = {np.float64: torch.float64, np.int64: torch.int64, np.complex128: torch.complex128}[dtype]
        if x.dtype != want:
            raise TypeError(f"{name}: expected {want}, got {x.dtype}")
        if ndim is not None and x.dim() != ndim:
            raise ValueError(f"{name}: expected {ndim}-d, got shape {tuple(x.shape)}")
        return dv.to_device(x)
    a = np.asarray(x)
    if a.dtype != np.dtype(dtype):
        raise TypeError(f"{name}: expected {np.dtype(dtype)}, got {a.dtype}")
    if ndim is not None and a.ndim != ndim:
        raise ValueError(f"{name}: expected {ndim}-d, got shape {a.shape}")
    return dv.to_device(a)


def _host(x):
    return x.detach().cpu().numpy() if _is_t(x) else np.asarray(x)


def _out(t, like_torch: bool):
    return t if like_torch else dv.to_host(t)


def _any_t(*xs) -> bool:
    return any(_is_t(x) for x in xs)


def _view_f8(c: torch.Tensor) -> torch.Tensor:
    """complex128 [K] -> interleaved float64 [2K] view."""
    return torch.view_as_real(c).reshape(-1)


def _live(idx):
    """Entries before a row's first -1 (numba.py:372: a -1 ends the row)."""
    live = np.ones(idx.shape, bool)
    if idx.shape[1] > 1:
        live[:, 1:] = np.cumprod(idx[:, :-1] != -1, axis=1).astype(bool)
    return live & (idx != -1)


def validate_mapping(pix_indexes_for_sub_slim_index, pixels, unique=False):
    """Indices must lie in [-1, pixels) (the reference writes out of bounds for other values,
    numba.py:371-375).  Repeated indices within a row are accepted: the dense mapping assigns them (the
    last one wins, numba.py:375) and the sparse forms sum them (numba.py:585-586); ``unique=True`` rejects
    them instead."""
    idx = _host(pix_indexes_for_sub_slim_index)
    if idx.size and (idx.min() < -1 or idx.max() >= pixels):
        raise ValueError(f"pix_indexes_for_sub_slim_index out of range [-1, {pixels})")
    if unique and has_repeated_indices(idx):
        raise ValueError("pix_indexes_for_sub_slim_index repeats a source pixel within a row")


def has_repeated_indices(pix_indexes_for_sub_slim_index) -> bool:
    """True when a row names one source pixel twice before its first -1."""
    idx = _host(pix_indexes_for_sub_slim_index)
    if idx.ndim != 2 or idx.shape[1] < 2:
        return False
    s = np.where(_live(idx), idx, -2 - np.arange(idx.shape[1])[None, :])
    ss = np.sort(s, axis=1)
    return bool(np.any((ss[:, 1:] == ss[:, :-1]) & (ss[:, 1:] >= 0)))


def assignment_weights(pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index):
    """Weights for which summing the mapping's entries gives the dense mapping matrix of numba.py:368-375,
    which assigns: where a row names a source pixel more than once, every entry but the last gets weight 0.
    Returns the input unchanged when no row repeats an index."""
    idx = _host(pix_indexes_for_sub_slim_index)
    if not has_repeated_indices(idx):
        return pix_weights_for_sub_slim_index
    w = np.array(_host(pix_weights_for_sub_slim_index), dtype=np.float64, copy=True)
    live = _live(idx)
    B = idx.shape[1]
    for b in range(B - 1):
        later = (idx[:, b + 1:] == idx[:, b:b + 1]) & live[:, b + 1:]
        w[later.any(axis=1) & live[:, b], b] = 0.0
    if _is_t(pix_weights_for_sub_slim_index):
        return torch.from_numpy(w).to(pix_weights_for_sub_slim_index.device)
    return w


def neighbors_symmetric(neighbors, neighbors_sizes) -> bool:
    """True when k lists i whenever i lists k (before the rows' first -1): then the regularisation matrix of
    numba.py:686-694 is symmetric."""
    nb = _host(neighbors)
    S = _host(neighbors_sizes).size
    live = _live(nb) & (nb >= 0)
    i = np.repeat(np.arange(S), nb.shape[1])[live.ravel()]
    k = nb.ravel()[live.ravel()]
    a = set(zip(i.tolist(), k.tolist()))
    return all((kk, ii) in a for ii, kk in a)


def validate_neighbors(neighbors, neighbors_sizes, symmetric=False):
    """Shapes and index range ([-1, S)); with ``symmetric`` also that the lists form a symmetric relation."""
    nb = _host(neighbors)
    sz = _host(neighbors_sizes)
    S = sz.size
    if nb.shape[0] != S:
        raise ValueError(f"neighbors has {nb.shape[0]} rows, neighbors_sizes {S}")
    if nb.size and (nb.min() < -1 or nb.max() >= S):
        raise ValueError(f"neighbors out of range [-1, {S})")
    if symmetric and not neighbors_symmetric(nb, sz):
        raise ValueError("neighbors is not a symmetric relation: the regularisation matrix would not be symmetric")


# ---------------------------------------------------------------------------
# numba.py geometry helpers (host-side, no GPU work)
# ---------------------------------------------------------------------------

def mask_2d_centres_from(shape_native, pixel_scales, centre):
    """numba.py:13-44."""
    return (
        0.5 * (shape_native[0] - 1) - (centre[0] / pixel_scales[0]),
        0.5 * (shape_native[1] - 1) + (centre[1] / pixel_scales[1]),
    )


def mask_2d_circular_from(shape_native, pixel_scales, radius, centre):
    """numba.py:50-86 (True = masked)."""
    cy, cx = mask_2d_centres_from(shape_native, pixel_scales, centre)
    ys, xs = np.indices(shape_native)
    return (radius * radius) < (np.square((ys - cy) * pixel_scales[0]) + np.square((xs - cx) * pixel_scales[1]))


# ---------------------------------------------------------------------------
# W-tilde and data-side sums
# ---------------------------------------------------------------------------

def w_tilde_curvature_interferometer_from(noise_map_real, uv_wavelengths, grid_radians_slim, *, precision="fp64"):
    """numba.py:155-207 -> f8[M, M].

    ``precision``: "fp64" (default, the reference's arithmetic); "int8x3",
    "int8x4", "int8x5" (exact integer MFMA on 3, 4 or 5 signed 8-bit digits of FP64
    cos/sin features, all digit products that reach FP64); "int8crt" (the
    five-digit fixed-point features, their exact integer Gram matrix from its
    residues modulo 12 moduli, one int8 MFMA product each); "int8crt11" (36-bit
    features, 11 moduli); "fp32" or
    "bf16x3" (FP32 cos/sin, FP32-accumulated MFMA).
    """
    torch_in = _any_t(noise_map_real, uv_wavelengths, grid_radians_slim)
    n = _dev(noise_map_real, np.float64, "noise_map_real", 1)
    uv = _dev(uv_wavelengths, np.float64, "uv_wavelengths", 2)
    g = _dev(grid_radians_slim, np.float64, "grid_radians_slim", 2)
    K, M = n.shape[0], g.shape[0]
    if uv.shape != (K, 2) or g.shape != (M, 2):
        raise ValueError("shape mismatch: noise_map_real [K], uv_wavelengths [K,2], grid_radians_slim [M,2]")
    prec = _lib.PRECISIONS[precision]
    out = dv.empty((M, M))
    ws = dv.workspace(_lib.query("aj_w_tilde_workspace_bytes", K, M, prec))
    _lib.call("aj_w_tilde_curvature_interferometer_from", n.data_ptr(), uv.data_ptr(), g.data_ptr(), K, M, prec,
              out.data_ptr(), ws.data_ptr(), ws.numel(), dv.stream_ptr())
    return _out(out, torch_in)


def validate_native_index(native_index_for_slim_index, grid_size):
    """Host check of the lattice positions: shape [M, 2], every (row, column) in [0, grid_size)."""
    ni = _host(native_index_for_slim_index)
    if ni.ndim != 2 or ni.shape[1] != 2:
        raise ValueError("native_index_for_slim_index must be [M, 2]")
    if ni.size and (ni.min() < 0 or ni.max() >= int(grid_size)):
        raise ValueError(f"native_index_for_slim_index outside the {int(grid_size)} x {int(grid_size)} grid")


def w_compact_curvature_interferometer_from(grid_size, noise_map_real, uv_wavelengths, pixel_scale):
    """numba.py:210-234 -> f8[N, 2N-1]: the distinct W-tilde values of a regular N x N grid
    (pixel_scale in arcsec), exact FP64 on MFMA."""
    torch_in = _any_t(noise_map_real, uv_wavelengths)
    n = _dev(noise_map_real, np.float64, "noise_map_real", 1)
    uv = _dev(uv_wavelengths, np.float64, "uv_wavelengths", 2)
    K, N = n.shape[0], int(grid_size)
    if uv.shape != (K, 2) or N <= 0:
        raise ValueError("shape mismatch: noise_map_real [K], uv_wavelengths [K,2], grid_size > 0")
    out = dv.empty((N, 2 * N - 1))
    ws = dv.workspace(_lib.query("aj_w_compact_workspace_bytes", K, N))
    _lib.call("aj_w_compact_curvature_interferometer_from", n.data_ptr(), uv.data_ptr(), K, N, float(pixel_scale),
              out.data_ptr(), ws.data_ptr(), ws.numel(), dv.stream_ptr())
    return _out(out, torch_in)


def w_tilde_via_compact_from(w_compact, native_index_for_slim_index):
    """numba.py:237-255 -> f8[M, M]."""
    torch_in = _any_t(w_compact, native_index_for_slim_index)
    C = _dev(w_compact, np.float64, "w_compact", 2)
    N = C.shape[0]
    if C.shape[1] != 2 * N - 1:
        raise ValueError("w_compact must be [N, 2N-1]")
    validate_native_index(native_index_for_slim_index, N)
    ni = _dev(native_index_for_slim_index, np.int64, "native_index_for_slim_index", 2)
    M = ni.shape[0]
    out = dv.empty((M, M))
    _lib.call("aj_w_tilde_via_compact_from", C.data_ptr(), N, ni.data_ptr(), M, out.data_ptr(), dv.stream_ptr())
    return _out(out, torch_in)


def w_tilde_via_preload_from(w_tilde_preload, native_index_for_slim_index):
    """numba.py:259-291 -> f8[M, M]: W[i, j] = w_tilde_preload[y_j - y_i, x_j - x_i] for every pair (i, j)
    (numba.py:282-291), negative offsets wrapping as numpy indexes do."""
    torch_in = _any_t(w_tilde_preload, native_index_for_slim_index)
    Pm = _dev(w_tilde_preload, np.float64, "w_tilde_preload", 2)
    Py, Px = Pm.shape
    ni_h = _host(native_index_for_slim_index)
    if ni_h.ndim != 2 or ni_h.shape[1] != 2:
        raise ValueError("native_index_for_slim_index must be [M, 2]")
    if ni_h.size and (np.ptp(ni_h[:, 0]) >= Py or np.ptp(ni_h[:, 1]) >= Px):
        raise IndexError("pixel offsets exceed the preload table")
    ni = _dev(native_index_for_slim_index, np.int64, "native_index_for_slim_index", 2)
    M = ni.shape[0]
    out = dv.empty((M, M))
    _lib.call("aj_w_tilde_via_preload_from", Pm.data_ptr(), Py, Px, ni.data_ptr(), M, out.data_ptr(), dv.stream_ptr())
    return _out(out, torch_in)


def w_tilde_curvature_preload_interferometer_from(noise_map_real, uv_wavelengths, shape_masked_pixels_2d,
                                                  grid_radians_2d):
    """original.py:230-403 -> f8[2Y, 2X]: the preload table of W-tilde values by pixel offset over the
    (Y, X) = shape_masked_pixels_2d corner of grid_radians_2d f8[Gy, Gx, 2] ((y, x) radians), one FP64
    sum over the visibilities per entry; read back by w_tilde_via_preload_from and
    curvature_matrix_via_w_tilde_curvature_preload_interferometer_from."""
    torch_in = _any_t(noise_map_real, uv_wavelengths, grid_radians_2d)
    n = _dev(noise_map_real, np.float64, "noise_map_real", 1)
    uv = _dev(uv_wavelengths, np.float64, "uv_wavelengths", 2)
    g = _dev(grid_radians_2d, np.float64, "grid_radians_2d", 3)
    K = n.shape[0]
    Y, X = (int(v) for v in shape_masked_pixels_2d)
    if uv.shape != (K, 2) or g.shape[2] != 2:
        raise ValueError("shape mismatch: noise_map_real [K], uv_wavelengths [K,2], grid_radians_2d [Gy,Gx,2]")
    if not (0 < Y <= g.shape[0] and 0 < X <= g.shape[1]):
        raise ValueError(f"shape_masked_pixels_2d {(Y, X)} outside the {tuple(g.shape[:2])} grid")
    out = dv.empty((2 * Y, 2 * X))
    ws = dv.workspace(_lib.query("aj_w_tilde_preload_workspace_bytes", K, Y, X))
    _lib.call("aj_w_tilde_curvature_preload_interferometer_from", n.data_ptr(), uv.data_ptr(), K, Y, X, g.data_ptr(),
              g.shape[0], g.shape[1], out.data_ptr(), ws.data_ptr(), ws.numel(), dv.stream_ptr())
    return _out(out, torch_in)


def w_tilde_data_interferometer_from(visibilities_real, noise_map_real, uv_wavelengths, grid_radians_slim):
    """numba.py:95-152 -> f8[M]."""
    torch_in = _any_t(visibilities_real, noise_map_real, uv_wavelengths, grid_radians_slim)
    v = _dev(visibilities_real, np.float64, "visibilities_real", 1)
    n = _dev(noise_map_real, np.float64, "noise_map_real", 1)
    uv = _dev(uv_wavelengths, np.float64, "uv_wavelengths", 2)
    g = _dev(grid_radians_slim, np.float64, "grid_radians_slim", 2)
    K, M = v.shape[0], g.shape[0]
    out = dv.empty((M,))
    ws = dv.workspace(_lib.query("aj_visibility_sum_workspace_bytes", K, M))
    _lib.call("aj_w_tilde_data_interferometer_from", v.data_ptr(), n.data_ptr(), uv.data_ptr(), g.data_ptr(), K, M,
              out.data_ptr(), ws.data_ptr(), ws.numel(), dv.stream_ptr())
    return _out(out, torch_in)


def w_tilde_with_data_interferometer_from(noise_map_real, uv_wavelengths, grid_radians_slim, *, data=None,
                                          visibilities_real=None, precision="fp64"):
    """W-tilde with the data-side sums of the same visibilities, in one pass.

    Fuses numba.py:155-207 (W-tilde) with numba.py:95-152 (``w_tilde_data``)
    and the dirty image, whose cos/sin of each (pixel, visibility) phase the
    W-tilde features already hold.  Returns ``(w_tilde, dirty_image,
    w_tilde_data)``; ``dirty_image`` is computed from the complex ``data`` with
    n_k = ``noise_map_real[k]`` (as :func:`dirty_image_from`), ``w_tilde_data``
    from ``visibilities_real`` (as :func:`w_tilde_data_interferometer_from`);
    either is None when its input is.  ``precision`` as in
    :func:`w_tilde_curvature_interferometer_from`; the data sums are FP64 at
    every precision.
    """
    torch_in = _any_t(noise_map_real, uv_wavelengths, grid_radians_slim, data, visibilities_real)
    n = _dev(noise_map_real, np.float64, "noise_map_real", 1)
    uv = _dev(uv_wavelengths, np.float64, "uv_wavelengths", 2)
    g = _dev(grid_radians_slim, np.float64, "grid_radians_slim", 2)
    K, M = n.shape[0], g.shape[0]
    if uv.shape != (K, 2) or g.shape != (M, 2):
        raise ValueError("shape mismatch: noise_map_real [K], uv_wavelengths [K,2], grid_radians_slim [M,2]")
    d = v = None
    if data is not None:
        d = _view_f8(_dev(data, np.complex128, "data", 1))
        if d.shape[0] != 2 * K:
            raise ValueError("data length must equal noise_map_real length")
    if visibilities_real is not None:
        v = _dev(visibilities_real, np.float64, "visibilities_real", 1)
        if v.shape[0] != K:
            raise ValueError("visibilities_real length must equal noise_map_real length")
    prec = _lib.PRECISIONS[precision]
    W = dv.empty((M, M))
    y = dv.empty((M,)) if d is not None else None
    z = dv.empty((M,)) if v is not None else None
    ws = dv.workspace(_lib.query("aj_w_tilde_with_data_workspace_bytes", K, M, prec))
    ptr = lambda t: None if t is None else t.data_ptr()  # noqa: E731
    _lib.call("aj_w_tilde_with_data_interferometer_from", n.data_ptr(), uv.data_ptr(), g.data_ptr(), K, M, prec,
              ptr(d), ptr(v), W.data_ptr(), ptr(y), ptr(z), ws.data_ptr(), ws.numel(), dv.stream_ptr())
    return (_out(W, torch_in), None if y is None else _out(y, torch_in), None if z is None else _out(z, torch_in))


def dirty_image_from(data, noise_map, uv_wavelengths, grid_radians_slim):
    """Dirty image consistent with W-tilde (no reference counterpart; data preparation)."""
    torch_in = _any_t(data, noise_map, uv_wavelengths, grid_radians_slim)
    d = _view_f8(_dev(data, np.complex128, "data", 1))
    n = _view_f8(_dev(noise_map, np.complex128, "noise_map", 1))
    uv = _dev(uv_wavelengths, np.float64, "uv_wavelengths", 2)
    g = _dev(grid_radians_slim, np.float64, "grid_radians_slim", 2)
    K, M = uv.shape[0], g.shape[0]
    out = dv.empty((M,))
    ws = dv.workspace(_lib.query("aj_visibility_sum_workspace_bytes", K, M))
    _lib.call("aj_dirty_image_from", d.data_ptr(), n.data_ptr(), uv.data_ptr(), g.data_ptr(), K, M, out.data_ptr(),
              ws.data_ptr(), ws.numel(), dv.stream_ptr())
    return _out(out, torch_in)


# ---------------------------------------------------------------------------
# mapping, data vector, curvature
# ---------------------------------------------------------------------------

def mapping_matrix_from(pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index, pixels):
    """numba.py:293-376 -> f8[M, pixels] (assignment; rows stop at the first -1)."""
    torch_in = _any_t(pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index)
    validate_mapping(pix_indexes_for_sub_slim_index, pixels, unique=False)
    idx = _dev(pix_indexes_for_sub_slim_index, np.int64, "pix_indexes_for_sub_slim_index", 2)
    w = _dev(pix_weights_for_sub_slim_index, np.float64, "pix_weights_for_sub_slim_index", 2)
    M, B = idx.shape
    out = dv.empty((M, int(pixels)))
    _lib.call("aj_mapping_matrix_from", idx.data_ptr(), w.data_ptr(), M, B, int(pixels), out.data_ptr(),
              dv.stream_ptr())
    return _out(out, torch_in)


def data_vector_from(mapping_matrix, dirty_image):
    """numba.py:379-406 -> f8[S] = dirty_image @ mapping_matrix."""
    torch_in = _any_t(mapping_matrix, dirty_image)
    T = _dev(mapping_matrix, np.float64, "mapping_matrix", 2)
    y = _dev(dirty_image, np.float64, "dirty_image", 1)
    M, S = T.shape
    if y.shape[0] != M:
        raise ValueError("dirty_image length must equal mapping_matrix rows")
    out = dv.empty((S,))
    _lib.call("aj_data_vector_from", T.data_ptr(), y.data_ptr(), M, S, out.data_ptr(), dv.stream_ptr())
    return _out(out, torch_in)


def curvature_matrix_via_w_tilde_from(w_tilde, mapping_matrix):
    """numba.py:409-435 -> f8[S, S] = T^T W T (dense, FP64 MFMA)."""
    torch_in = _any_t(w_tilde, mapping_matrix)
    W = _dev(w_tilde, np.float64, "w_tilde", 2)
    T = _dev(mapping_matrix, np.float64, "mapping_matrix", 2)
    M, S = T.shape
    if W.shape != (M, M):
        raise ValueError("w_tilde must be [M, M] with M = mapping_matrix rows")
    out = dv.empty((S, S))
    ws = dv.workspace(_lib.query("aj_curvature_dense_workspace_bytes", M, S))
    _lib.call("aj_curvature_matrix_via_w_tilde_from", W.data_ptr(), T.data_ptr(), M, S, out.data_ptr(), ws.data_ptr(),
              ws.numel(), dv.stream_ptr())
    return _out(out, torch_in)


def sparse_mapping_matrix_transpose_matmul(matrix, pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index,
                                           pixels):
    """numba.py:552-587 -> f8[pixels, S2] = T^T @ matrix."""
    torch_in = _any_t(matrix, pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index)
    validate_mapping(pix_indexes_for_sub_slim_index, pixels)
    X = _dev(matrix, np.float64, "matrix", 2)
    idx = _dev(pix_indexes_for_sub_slim_index, np.int64, "pix_indexes_for_sub_slim_index", 2)
    w = _dev(pix_weights_for_sub_slim_index, np.float64, "pix_weights_for_sub_slim_index", 2)
    M, S2 = X.shape
    B = idx.shape[1]
    if idx.shape[0] != M:
        raise ValueError("matrix rows must equal mapping rows")
    out = dv.empty((int(pixels), S2))
    ws = dv.workspace(_lib.query("aj_sparse_transpose_matmul_workspace_bytes", M, B, int(pixels)))
    _lib.call("aj_sparse_mapping_matrix_transpose_matmul", X.data_ptr(), idx.data_ptr(), w.data_ptr(), M, S2, B,
              int(pixels), out.data_ptr(), ws.data_ptr(), ws.numel(), dv.stream_ptr())
    return _out(out, torch_in)


def curvature_matrix_via_w_wilde_sparse_mapping_matrix_from(w_tilde, pix_indexes_for_sub_slim_index,
                                                            pix_weights_for_sub_slim_index, pixels):
    """numba.py:590-614 -> f8[pixels, pixels] (sparse mapping; w_tilde symmetric)."""
    torch_in = _any_t(w_tilde, pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index)
    validate_mapping(pix_indexes_for_sub_slim_index, pixels)
    W = _dev(w_tilde, np.float64, "w_tilde", 2)
    idx = _dev(pix_indexes_for_sub_slim_index, np.int64, "pix_indexes_for_sub_slim_index", 2)
    w = _dev(pix_weights_for_sub_slim_index, np.float64, "pix_weights_for_sub_slim_index", 2)
    M, B = idx.shape
    if W.shape != (M, M):
        raise ValueError("w_tilde must be [M, M]")
    out = dv.empty((int(pixels), int(pixels)))
    ws = dv.workspace(_lib.query("aj_curvature_sparse_workspace_bytes", M, B, int(pixels)))
    _lib.call("aj_curvature_matrix_via_w_wilde_sparse_mapping_matrix_from", W.data_ptr(), idx.data_ptr(),
              w.data_ptr(), M, B, int(pixels), out.data_ptr(), ws.data_ptr(), ws.numel(), dv.stream_ptr())
    return _out(out, torch_in)


def curvature_matrix_via_w_compact_sparse_mapping_matrix_from(w_compact, native_index_for_slim_index,
                                                              pix_indexes_for_sub_slim_index,
                                                              pix_weights_for_sub_slim_index, pixels):
    """numba.py:617-647 -> f8[pixels, pixels]: the sparse two-stage product with W-tilde read from
    the lattice table (never expanded to M x M)."""
    torch_in = _any_t(w_compact, native_index_for_slim_index, pix_indexes_for_sub_slim_index,
                      pix_weights_for_sub_slim_index)
    validate_mapping(pix_indexes_for_sub_slim_index, pixels)
    C = _dev(w_compact, np.float64, "w_compact", 2)
    N = C.shape[0]
    if C.shape[1] != 2 * N - 1:
        raise ValueError("w_compact must be [N, 2N-1]")
    validate_native_index(native_index_for_slim_index, N)
    ni = _dev(native_index_for_slim_index, np.int64, "native_index_for_slim_index", 2)
    idx = _dev(pix_indexes_for_sub_slim_index, np.int64, "pix_indexes_for_sub_slim_index", 2)
    w = _dev(pix_weights_for_sub_slim_index, np.float64, "pix_weights_for_sub_slim_index", 2)
    M, B = idx.shape
    if ni.shape[0] != M:
        raise ValueError("native_index_for_slim_index and the mapping disagree on M")
    out = dv.empty((int(pixels), int(pixels)))
    ws = dv.workspace(_lib.query("aj_curvature_sparse_workspace_bytes", M, B, int(pixels)))
    _lib.call("aj_curvature_matrix_via_w_compact_sparse_mapping_matrix_from", C.data_ptr(), N, ni.data_ptr(),
              idx.data_ptr(), w.data_ptr(), M, B, int(pixels), out.data_ptr(), ws.data_ptr(), ws.numel(),
              dv.stream_ptr())
    return _out(out, torch_in)


def curvature_matrix_via_w_tilde_curvature_preload_interferometer_from(curvature_preload,
                                                                       pix_indexes_for_sub_slim_index,
                                                                       pix_size_for_sub_slim_index,
                                                                       pix_weights_for_sub_slim_index,
                                                                       native_index_for_slim_index, pix_pixels):
    """original.py:600-700 -> f8[pix_pixels, pix_pixels] = T^T W T with W read from the preload table
    as w_tilde_via_preload_from reads it (never expanded to M x M).  Mapping entries at or past
    pix_size_for_sub_slim_index[m] are not used, as in the reference's loops."""
    torch_in = _any_t(curvature_preload, pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index,
                      native_index_for_slim_index)
    idx_h = _host(pix_indexes_for_sub_slim_index)
    size_h = _host(pix_size_for_sub_slim_index)
    for a, name in ((idx_h, "pix_indexes_for_sub_slim_index"), (size_h, "pix_size_for_sub_slim_index")):
        if a.dtype != np.int64:
            raise TypeError(f"{name}: expected int64, got {a.dtype}")
    idx_h = np.array(idx_h, copy=True)
    if idx_h.ndim != 2 or size_h.shape != (idx_h.shape[0],):
        raise ValueError("pix_indexes_for_sub_slim_index [M, B] and pix_size_for_sub_slim_index [M] disagree")
    if size_h.size and (size_h.min() < 0 or size_h.max() > idx_h.shape[1]):
        raise ValueError("pix_size_for_sub_slim_index outside [0, B]")
    idx_h[np.arange(idx_h.shape[1])[None, :] >= size_h[:, None]] = -1
    validate_mapping(idx_h, pix_pixels)
    Pm = _dev(curvature_preload, np.float64, "curvature_preload", 2)
    Py, Px = Pm.shape
    ni_h = _host(native_index_for_slim_index)
    if ni_h.ndim != 2 or ni_h.shape != (idx_h.shape[0], 2):
        raise ValueError("native_index_for_slim_index must be [M, 2]")
    if ni_h.size and (np.ptp(ni_h[:, 0]) >= Py or np.ptp(ni_h[:, 1]) >= Px):
        raise IndexError("pixel offsets exceed the preload table")
    ni = _dev(native_index_for_slim_index, np.int64, "native_index_for_slim_index", 2)
    idx = _dev(idx_h, np.int64, "pix_indexes_for_sub_slim_index", 2)
    w = _dev(pix_weights_for_sub_slim_index, np.float64, "pix_weights_for_sub_slim_index", 2)
    M, B = idx.shape
    if w.shape != (M, B):
        raise ValueError("pix_weights_for_sub_slim_index must match pix_indexes_for_sub_slim_index")
    S = int(pix_pixels)
    out = dv.empty((S, S))
    ws = dv.workspace(_lib.query("aj_curvature_sparse_workspace_bytes", M, B, S))
    _lib.call("aj_curvature_matrix_via_w_tilde_curvature_preload_interferometer_from", Pm.data_ptr(), Py, Px,
              ni.data_ptr(), idx.data_ptr(), w.data_ptr(), M, B, S, out.data_ptr(), ws.data_ptr(), ws.numel(),
              dv.stream_ptr())
    return _out(out, torch_in)


def curvature_matrix_via_w_compact_sparse_mapping_matrix_direct_from(w_compact, native_index_for_slim_index,
                                                                     pix_indexes_for_sub_slim_index,
                                                                     pix_weights_for_sub_slim_index, pixels):
    """numba.py:439-497 -> f8[pixels, pixels]: the same F as
    curvature_matrix_via_w_compact_sparse_mapping_matrix_from (the reference's four-loop form); the
    device computes it by the same two sparse products from the lattice table."""
    return curvature_matrix_via_w_compact_sparse_mapping_matrix_from(
        w_compact, native_index_for_slim_index, pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index,
        pixels)


# ---------------------------------------------------------------------------
# regularisation, solve, normalisation, likelihood
# ---------------------------------------------------------------------------

def constant_regularization_matrix_from(coefficient, neighbors, neighbors_sizes):
    """numba.py:656-695 -> f8[S, S]."""
    torch_in = _any_t(neighbors, neighbors_sizes)
    validate_neighbors(neighbors, neighbors_sizes, symmetric=False)
    nb = _dev(neighbors, np.int64, "neighbors", 2)
    sz = _dev(neighbors_sizes, np.int64, "neighbors_sizes", 1)
    S, P = nb.shape
    out = dv.empty((S, S))
    _lib.call("aj_constant_regularization_matrix_from", float(coefficient), nb.data_ptr(), sz.data_ptr(), S, P,
              out.data_ptr(), dv.stream_ptr())
    return _out(out, torch_in)


def _lu_solve(A, D=None):
    """(x, log|det A|, info) on the device: LU with partial pivoting of the general matrix A (np.linalg.solve /
    slogdet); x is None when D is None (log-determinant only); info 0 or the 1-based zero pivot."""
    S = A.shape[0]
    x = dv.empty((S,)) if D is not None else None
    ld = dv.empty((1,))
    info = torch.zeros(1, dtype=torch.int64, device=A.device)
    ws = dv.workspace(_lib.query("aj_solve_workspace_bytes", S))
    _lib.call("aj_reconstruction_positive_negative_from", None if D is None else D.data_ptr(), A.data_ptr(), S,
              None if x is None else x.data_ptr(), ld.data_ptr(), info.data_ptr(), ws.data_ptr(), ws.numel(),
              dv.stream_ptr())
    return x, ld, info


def reconstruction_positive_negative_from(data_vector, curvature_reg_matrix):
    """numba.py:698-731 -> f8[S] solving A s = D as np.linalg.solve does: LU with partial pivoting of the
    whole matrix (symmetric or not, definite or not).

    Raises numpy.linalg.LinAlgError when A is singular (an exact zero pivot), as numpy does.
    """
    torch_in = _any_t(data_vector, curvature_reg_matrix)
    D = _dev(data_vector, np.float64, "data_vector", 1)
    A = _dev(curvature_reg_matrix, np.float64, "curvature_reg_matrix", 2)
    S = D.shape[0]
    if A.shape != (S, S):
        raise ValueError("curvature_reg_matrix must be [S, S]")
    out, _, info = _lu_solve(A, D)
    bad = int(info.item())
    if bad:
        raise np.linalg.LinAlgError(f"Singular matrix (zero pivot in column {bad})")
    return _out(out, torch_in)


def noise_normalization_complex_from(noise_map):
    """numba.py:734-752 -> float."""
    n = _view_f8(_dev(noise_map, np.complex128, "noise_map", 1))
    out = dv.empty((1,))
    _lib.call("aj_noise_normalization_complex_from", n.data_ptr(), n.numel() // 2, out.data_ptr(), dv.stream_ptr())
    return float(out.item())


def _loglike_composed(y, d, n, F, idx, w_assign, nb, sz, S):
    """numba.py:818-857 stage by stage on the device, for the inputs the Cholesky pipeline does not take
    (neighbour lists that are not symmetric, or curvature_reg_matrix / regularization_matrix not positive
    definite): H (numba.py:686-694), A = F + H, D = T^T y with the dense mapping's assignment, s and both
    log-determinants by LU with partial pivoting (np.linalg.solve / slogdet), chi^2 and the noise
    normalisation.  Raises LinAlgError when A or H is singular."""
    H = constant_regularization_matrix_from(1.0, nb, sz)
    A = F + H
    D = sparse_mapping_matrix_transpose_matmul(y.reshape(-1, 1), idx, w_assign, S).reshape(-1)
    s, ld_a, inf_a = _lu_solve(A, D)
    _, ld_h, inf_h = _lu_solve(H)
    c2nn = dv.empty((2,))
    _lib.call("aj_chi_squared_noise_normalization", d.data_ptr(), n.data_ptr(), d.numel() // 2, c2nn.data_ptr(),
              dv.stream_ptr())
    if int(inf_a.item()) or int(inf_h.item()):
        which = "curvature_reg_matrix" if int(inf_a.item()) else "regularization_matrix"
        raise np.linalg.LinAlgError(f"Singular matrix ({which})")
    return float((-0.5 * (c2nn[0] - s @ D + ld_a[0] - ld_h[0] + c2nn[1])).item())


def log_likelihood_function_via_w_tilde_from(dirty_image, data, noise_map, w_tilde, pix_indexes_for_sub_slim_index,
                                             pix_weights_for_sub_slim_index, neighbors, neighbors_sizes):
    """numba.py:755-857 -> float (coefficient 1, numba.py:819).

    Symmetric neighbour lists (every mesh's): F = T^T W T by source-pixel gathers, FP64 Cholesky of
    A = F + H and of H on the device (the reference's LU solve and two slogdets, in the form that holds for
    symmetric positive definite matrices).  Otherwise, or when A or H is not positive definite, the
    reference's own steps (LU solve and LU log-determinants) on the device.  Rows that name a source pixel
    twice follow the dense mapping's assignment (numba.py:375: the last entry wins).
    """
    S = int(_host(neighbors_sizes).size)
    validate_mapping(pix_indexes_for_sub_slim_index, S)
    validate_neighbors(neighbors, neighbors_sizes)
    w_assign = assignment_weights(pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index)
    y = _dev(dirty_image, np.float64, "dirty_image", 1)
    d = _view_f8(_dev(data, np.complex128, "data", 1))
    n = _view_f8(_dev(noise_map, np.complex128, "noise_map", 1))
    W = _dev(w_tilde, np.float64, "w_tilde", 2)
    idx = _dev(pix_indexes_for_sub_slim_index, np.int64, "pix_indexes_for_sub_slim_index", 2)
    w = _dev(w_assign, np.float64, "pix_weights_for_sub_slim_index", 2)
    nb = _dev(neighbors, np.int64, "neighbors", 2)
    sz = _dev(neighbors_sizes, np.int64, "neighbors_sizes", 1)
    M, B = idx.shape
    K = d.numel() // 2
    P = nb.shape[1]
    if W.shape != (M, M) or y.shape[0] != M or n.numel() != 2 * K or w.shape != (M, B):
        raise ValueError("shape mismatch between dirty_image, w_tilde, mapping, data and noise_map")
    if neighbors_symmetric(neighbors, neighbors_sizes):
        out = dv.empty((1,))
        info = torch.zeros(2, dtype=torch.int64, device=out.device)
        ws = dv.workspace(_lib.query("aj_log_likelihood_workspace_bytes", M, K, B, S, P))
        _lib.call("aj_log_likelihood_function_via_w_tilde_from", y.data_ptr(), d.data_ptr(), n.data_ptr(),
                  W.data_ptr(), idx.data_ptr(), w.data_ptr(), nb.data_ptr(), sz.data_ptr(), M, K, B, S, P,
                  out.data_ptr(), info.data_ptr(), ws.data_ptr(), ws.numel(), dv.stream_ptr())
        if not any(info.tolist()):
            return float(out.item())
    F = curvature_matrix_via_w_wilde_sparse_mapping_matrix_from(W, idx, w, S)
    return _loglike_composed(y, d, n, F, idx, w, nb, sz, S)


def log_likelihood_function_via_w_compact_from(dirty_image, data, noise_map, w_compact, native_index_for_slim_index,
                                               pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index,
                                               neighbors, neighbors_sizes):
    """numba.py:860-967 -> float (coefficient 1, numba.py:923): as
    log_likelihood_function_via_w_tilde_from with W-tilde read from the lattice table.  The reference forms
    F by the sparse products (repeated indices within a row summed, numba.py:930-936) and D from the dense
    mapping (assigned, numba.py:947-953); inputs with such rows, neighbour lists that are not symmetric, or
    a matrix that is not positive definite take the stage-by-stage route with exactly those semantics."""
    S = int(_host(neighbors_sizes).size)
    validate_mapping(pix_indexes_for_sub_slim_index, S)
    validate_neighbors(neighbors, neighbors_sizes)
    C = _dev(w_compact, np.float64, "w_compact", 2)
    N = C.shape[0]
    if C.shape[1] != 2 * N - 1:
        raise ValueError("w_compact must be [N, 2N-1]")
    validate_native_index(native_index_for_slim_index, N)
    ni = _dev(native_index_for_slim_index, np.int64, "native_index_for_slim_index", 2)
    y = _dev(dirty_image, np.float64, "dirty_image", 1)
    d = _view_f8(_dev(data, np.complex128, "data", 1))
    n = _view_f8(_dev(noise_map, np.complex128, "noise_map", 1))
    idx = _dev(pix_indexes_for_sub_slim_index, np.int64, "pix_indexes_for_sub_slim_index", 2)
    w = _dev(pix_weights_for_sub_slim_index, np.float64, "pix_weights_for_sub_slim_index", 2)
    nb = _dev(neighbors, np.int64, "neighbors", 2)
    sz = _dev(neighbors_sizes, np.int64, "neighbors_sizes", 1)
    M, B = idx.shape
    K = d.numel() // 2
    P = nb.shape[1]
    if ni.shape[0] != M or y.shape[0] != M or n.numel() != 2 * K or w.shape != (M, B):
        raise ValueError("shape mismatch between dirty_image, native indexes, mapping, data and noise_map")
    repeated = has_repeated_indices(pix_indexes_for_sub_slim_index)
    if not repeated and neighbors_symmetric(neighbors, neighbors_sizes):
        out = dv.empty((1,))
        info = torch.zeros(2, dtype=torch.int64, device=out.device)
        ws = dv.workspace(_lib.query("aj_log_likelihood_workspace_bytes", M, K, B, S, P))
        _lib.call("aj_log_likelihood_function_via_w_compact_from", y.data_ptr(), d.data_ptr(), n.data_ptr(),
                  C.data_ptr(), N, ni.data_ptr(), idx.data_ptr(), w.data_ptr(), nb.data_ptr(), sz.data_ptr(), M, K,
                  B, S, P, out.data_ptr(), info.data_ptr(), ws.data_ptr(), ws.numel(), dv.stream_ptr())
        if not any(info.tolist()):
            return float(out.item())
    F = curvature_matrix_via_w_compact_sparse_mapping_matrix_from(C, ni, idx, w, S)
    w_assign = _dev(assignment_weights(pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index), np.float64,
                    "pix_weights_for_sub_slim_index", 2)
    return _loglike_composed(y, d, n, F, idx, w_assign, nb, sz, S)
