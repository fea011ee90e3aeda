"""FP64 numpy restatement of the reference's likelihood path (test oracle only).

Each function keeps the name, parameter names and parameter order of the
function in ``src/autojax/numba.py`` (for the preload form, ``original.py``) it
restates, and cites the lines.  The
arithmetic is plain FP64 numpy/LAPACK; loops that the reference writes as
explicit loops are vectorised where that cannot change the value beyond FP64
rounding order.
"""

from __future__ import annotations

import numpy as np

# numba.py:6-7
TWO_PI = 2.0 * np.pi
LOG_TWO_PI = np.log(TWO_PI)


def mask_2d_centres_from(shape_native, pixel_scales, centre):
    """numba.py:13-44: (y, x) scaled centre of a mask."""
    return (
        0.5 * (shape_native[0] - 1) - (centre[0] / pixel_scales[0]),
        0.5 * (shape_native[1] - 1) + (centre[1] / pixel_scales[1]),
    )


def mask_2d_circular_from(shape_native, pixel_scales, radius, centre):
    """numba.py:50-86: True = masked; unmasked where r <= radius."""
    cy, cx = mask_2d_centres_from(shape_native, pixel_scales, centre)
    ys, xs = np.indices(shape_native)
    return (radius * radius) < (
        np.square((ys - cy) * pixel_scales[0]) + np.square((xs - cx) * pixel_scales[1])
    )


def w_tilde_data_interferometer_from(visibilities_real, noise_map_real, uv_wavelengths, grid_radians_slim):
    """numba.py:95-152.

    res_i = sum_k (n_k^2 / v_k)^2 cos(2 pi (g_i1 u_k0 + g_i0 u_k1)).  The weight
    is the reference's own expression, kept as written (numba.py:149).
    """
    g_2pi = TWO_PI * grid_radians_slim
    w_k = np.square(np.square(noise_map_real) / visibilities_real)
    res = np.zeros(grid_radians_slim.shape[0])
    for k0 in range(0, uv_wavelengths.shape[0], 4096):
        uv = uv_wavelengths[k0 : k0 + 4096]
        ph = g_2pi[:, 1:2] * uv[None, :, 0] + g_2pi[:, 0:1] * uv[None, :, 1]
        res += np.cos(ph) @ w_k[k0 : k0 + 4096]
    return res


def w_tilde_curvature_interferometer_from(noise_map_real, uv_wavelengths, grid_radians_slim, k_block=256):
    """numba.py:155-207.

    W_ij = sum_k n_k^-2 cos(dg_ij1 u_k0 + dg_ij0 u_k1), dg = 2 pi (g_i - g_j).
    Same pairing as numba.py:204: grid column 1 with uv column 0, grid column
    0 with uv column 1.  Streams K in blocks; memory is O(M^2).
    """
    M = grid_radians_slim.shape[0]
    g_2pi = TWO_PI * grid_radians_slim
    d0 = g_2pi[:, None, 0] - g_2pi[None, :, 0]
    d1 = g_2pi[:, None, 1] - g_2pi[None, :, 1]
    w = np.zeros((M, M))
    inv_n2 = np.reciprocal(np.square(noise_map_real))
    for k in range(uv_wavelengths.shape[0]):
        w += np.cos(d1 * uv_wavelengths[k, 0] + d0 * uv_wavelengths[k, 1]) * inv_n2[k]
    return w


def w_tilde_rows_from(noise_map_real, uv_wavelengths, grid_radians_slim, rows, k_block=4096, threads=None):
    """Rows `rows` of numba.py:155-207's W (the same element sum, W_ij = sum_k n_k^-2 cos(dg_ij1 u_k0 +
    dg_ij0 u_k1), dg = 2 pi (g_i - g_j)) for sizes whose full M x M matrix the oracle cannot build in
    seconds.  K is cut into blocks of k_block evaluated on a thread pool (numpy releases the GIL in its
    ufuncs); the blocks' partial sums are added in block order, so the result does not depend on the
    thread count.  Memory O(threads len(rows) M k_block)."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    if noise_map_real.shape[0] != uv_wavelengths.shape[0]:
        raise ValueError("noise_map_real and uv_wavelengths must have the same number of visibilities")
    g_2pi = TWO_PI * grid_radians_slim
    rows = np.asarray(rows)
    d0 = g_2pi[rows, None, 0] - g_2pi[None, :, 0]
    d1 = g_2pi[rows, None, 1] - g_2pi[None, :, 1]
    inv_n2 = np.reciprocal(np.square(noise_map_real))

    def part(k0):
        u = uv_wavelengths[k0:k0 + k_block]
        return np.cos(d1[:, :, None] * u[None, None, :, 0] + d0[:, :, None] * u[None, None, :, 1]) @ inv_n2[k0:k0 + k_block]

    starts = range(0, uv_wavelengths.shape[0], k_block)
    nt = threads or min(16, len(os.sched_getaffinity(0)))
    out = np.zeros((rows.size, grid_radians_slim.shape[0]))
    with ThreadPoolExecutor(max_workers=nt) as pool:
        for p in pool.map(part, starts):
            out += p
    return out


def w_compact_curvature_interferometer_from(grid_size, noise_map_real, uv_wavelengths, pixel_scale):
    """numba.py:210-234: lattice form, N x (2N-1) unique W values."""
    N = grid_size
    OFFSET = N - 1
    N_DIFF = 2 * N - 1
    TWOPI_D = (np.pi * np.pi * pixel_scale) / 324000.0
    d0 = (TWOPI_D * np.arange(N, dtype=np.float64)).reshape(-1, 1)
    d1 = TWOPI_D * (np.arange(N_DIFF, dtype=np.float64) - OFFSET)
    out = np.zeros((N, N_DIFF))
    inv_n2 = np.square(np.reciprocal(noise_map_real))
    for k in range(uv_wavelengths.shape[0]):
        out += np.cos(d1 * uv_wavelengths[k, 0] - d0 * uv_wavelengths[k, 1]) * inv_n2[k]
    return out


def w_tilde_via_compact_from(w_compact, native_index_for_slim_index):
    """numba.py:237-255: expand the lattice form to M x M."""
    OFFSET = w_compact.shape[0] - 1
    i0 = native_index_for_slim_index[:, 0]
    i1 = native_index_for_slim_index[:, 1]
    m = i0[:, None] - i0[None, :]
    n = np.where(m >= 0, i1[:, None] - i1[None, :], i1[None, :] - i1[:, None]) + OFFSET
    return w_compact[np.abs(m), n]


def w_compact_matmul_sparse_mapping_matrix_from(w_compact, native_index_for_slim_index, pix_indexes_for_sub_slim_index,
                                                pix_weights_for_sub_slim_index, pixels):
    """numba.py:500-549: Omega = W T with W read from the lattice table (loops; small sizes only)."""
    M, B = pix_indexes_for_sub_slim_index.shape
    OFFSET = w_compact.shape[0] - 1
    omega = np.zeros((M, pixels))
    r = native_index_for_slim_index[:, 0]
    c = native_index_for_slim_index[:, 1]
    for m1 in range(M):
        n1 = r[m1] - r
        n2 = np.where(n1 >= 0, c[m1] - c, c - c[m1]) + OFFSET
        w_row = w_compact[np.abs(n1), n2]
        for m2 in range(M):
            for b2 in range(B):
                s2 = pix_indexes_for_sub_slim_index[m2, b2]
                if s2 == -1:
                    break
                omega[m1, s2] += w_row[m2] * pix_weights_for_sub_slim_index[m2, b2]
    return omega


def curvature_matrix_via_w_compact_sparse_mapping_matrix_from(w_compact, native_index_for_slim_index,
                                                              pix_indexes_for_sub_slim_index,
                                                              pix_weights_for_sub_slim_index, pixels):
    """numba.py:617-647: F = T^T Omega with Omega from the lattice table."""
    omega = w_compact_matmul_sparse_mapping_matrix_from(w_compact, native_index_for_slim_index,
                                                        pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index,
                                                        pixels)
    return sparse_mapping_matrix_transpose_matmul(omega, pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index,
                                                  pixels)


def w_tilde_via_preload_from(w_tilde_preload, native_index_for_slim_index):
    """numba.py:259-291: W[i, j] = preload[y_j - y_i, x_j - x_i] for every pair (i, j) (numba.py:282-291;
    negative offsets wrap as numpy indexes).  original.py:431-447 reads the table for j >= i only and
    mirrors; the two agree when the table is point-symmetric."""
    y = native_index_for_slim_index[:, 0]
    x = native_index_for_slim_index[:, 1]
    return w_tilde_preload[y[None, :] - y[:, None], x[None, :] - x[:, None]]


def _preload_quadrant(noise_map_real, uv_wavelengths, dy, dx):
    """sum_k n_k^-2 cos(2 pi (dx u_k0 + dy u_k1)) for offset arrays dy, dx (original.py:356-361)."""
    w = noise_map_real**-2.0
    ph = TWO_PI * (dx.reshape(-1, 1) * uv_wavelengths[None, :, 0] + dy.reshape(-1, 1) * uv_wavelengths[None, :, 1])
    return (np.cos(ph) @ w).reshape(dy.shape)


def w_tilde_curvature_preload_interferometer_from(noise_map_real, uv_wavelengths, shape_masked_pixels_2d,
                                                  grid_radians_2d):
    """original.py:230-403: the 2Y x 2X preload table over the Y x X corner of the grid; each of the four
    loops of the reference is one quadrant, with offsets from its own corner (row Y and column X stay 0)."""
    Y, X = shape_masked_pixels_2d
    g = grid_radians_2d[0:Y, 0:X]
    out = np.zeros((2 * Y, 2 * X))
    # original.py:355-361: [i, j], offsets from g[0, 0]
    out[:Y, :X] = _preload_quadrant(noise_map_real, uv_wavelengths, g[0, 0, 0] - g[:, :, 0], g[0, 0, 1] - g[:, :, 1])
    j = np.arange(1, X)
    i = np.arange(1, Y)
    # original.py:363-374: [i, -j], j >= 1, offsets from g[0, -1] to g[i, X - j - 1]
    if X > 1:
        gp = g[:, X - j - 1]
        out[:Y, 2 * X - j] = _preload_quadrant(noise_map_real, uv_wavelengths, g[0, -1, 0] - gp[:, :, 0],
                                               g[0, -1, 1] - gp[:, :, 1])
    # original.py:376-387: [-i, j], i >= 1, offsets from g[-1, 0] to g[Y - i - 1, j]
    if Y > 1:
        gp = g[Y - i - 1, :]
        out[2 * Y - i, :X] = _preload_quadrant(noise_map_real, uv_wavelengths, g[-1, 0, 0] - gp[:, :, 0],
                                               g[-1, 0, 1] - gp[:, :, 1])
    # original.py:389-400: [-i, -j], i, j >= 1, offsets from g[-1, -1] to g[Y - i - 1, X - j - 1]
    if X > 1 and Y > 1:
        gp = g[np.ix_(Y - i - 1, X - j - 1)]
        out[np.ix_(2 * Y - i, 2 * X - j)] = _preload_quadrant(noise_map_real, uv_wavelengths,
                                                              g[-1, -1, 0] - gp[:, :, 0], g[-1, -1, 1] - gp[:, :, 1])
    return out


def curvature_matrix_via_w_tilde_curvature_preload_interferometer_from(curvature_preload,
                                                                       pix_indexes_for_sub_slim_index,
                                                                       pix_size_for_sub_slim_index,
                                                                       pix_weights_for_sub_slim_index,
                                                                       native_index_for_slim_index, pix_pixels):
    """original.py:600-700, loop for loop: pixel pairs ip0 < ip1 read the table at the offset of ip1
    from ip0, the result is added to its transpose, then each pixel's own pairs add preload[0, 0]."""
    idx, size, wts, nat = (pix_indexes_for_sub_slim_index, pix_size_for_sub_slim_index,
                           pix_weights_for_sub_slim_index, native_index_for_slim_index)
    M = nat.shape[0]
    F = np.zeros((pix_pixels, pix_pixels))
    for ip0 in range(M):
        y0, x0 = nat[ip0]
        # all ip1 > ip0 at once: the table value of each pair, then each (ip0_pix, ip1_pix) mapping pair
        ip1 = np.arange(ip0 + 1, M)
        if ip1.size == 0:
            continue
        val = curvature_preload[nat[ip1, 0] - y0, nat[ip1, 1] - x0]
        for a in range(size[ip0]):
            for b in range(idx.shape[1]):
                live = b < size[ip1]
                np.add.at(F[idx[ip0, a]], idx[ip1[live], b], val[live] * wts[ip0, a] * wts[ip1[live], b])
    F = F + F.T
    pre = curvature_preload[0, 0]
    for ip0 in range(M):
        for a in range(size[ip0]):
            for b in range(size[ip0]):
                s0, s1 = idx[ip0, a], idx[ip0, b]
                v = pre * wts[ip0, a] * wts[ip0, b]
                if s0 > s1:
                    F[s0, s1] += v
                    F[s1, s0] += v
                elif s0 == s1:
                    F[s0, s1] += v
    return F


def curvature_matrix_via_w_compact_sparse_mapping_matrix_direct_from(w_compact, native_index_for_slim_index,
                                                                     pix_indexes_for_sub_slim_index,
                                                                     pix_weights_for_sub_slim_index, pixels):
    """numba.py:439-497: F[s1, s2] += t1 W[m1, m2] t2 over all pixel pairs (four loops; tiny sizes only)."""
    M, B = pix_indexes_for_sub_slim_index.shape
    W = w_tilde_via_compact_from(w_compact, native_index_for_slim_index)
    F = np.zeros((pixels, pixels))
    for m1 in range(M):
        for m2 in range(M):
            for b1 in range(B):
                s1 = pix_indexes_for_sub_slim_index[m1, b1]
                if s1 == -1:
                    break
                for b2 in range(B):
                    s2 = pix_indexes_for_sub_slim_index[m2, b2]
                    if s2 == -1:
                        break
                    F[s1, s2] += pix_weights_for_sub_slim_index[m1, b1] * W[m1, m2] * pix_weights_for_sub_slim_index[m2, b2]
    return F


def mapping_matrix_from(pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index, pixels):
    """numba.py:293-376: dense T[m, s] = t (assignment), rows stop at the first -1."""
    M, B = pix_indexes_for_sub_slim_index.shape
    T = np.zeros((M, pixels))
    for m in range(M):
        for b in range(B):
            s = pix_indexes_for_sub_slim_index[m, b]
            if s == -1:
                break
            T[m, s] = pix_weights_for_sub_slim_index[m, b]
    return T


def data_vector_from(mapping_matrix, dirty_image):
    """numba.py:379-406: D = dirty @ T."""
    return dirty_image @ mapping_matrix


def curvature_matrix_via_w_tilde_from(w_tilde, mapping_matrix):
    """numba.py:409-435: F = T^T W T (dense)."""
    return mapping_matrix.T @ w_tilde @ mapping_matrix


def sparse_mapping_matrix_transpose_matmul(matrix, pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index, pixels):
    """numba.py:552-587: T^T @ matrix from the sparse mapping (rows stop at -1)."""
    M, S2 = matrix.shape
    B = pix_indexes_for_sub_slim_index.shape[1]
    F = np.zeros((pixels, S2))
    for m in range(M):
        for b in range(B):
            s1 = pix_indexes_for_sub_slim_index[m, b]
            if s1 == -1:
                break
            F[s1, :] += pix_weights_for_sub_slim_index[m, b] * matrix[m, :]
    return F


def curvature_matrix_via_w_wilde_sparse_mapping_matrix_from(w_tilde, pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index, pixels):
    """numba.py:590-614: F = T^T (T^T W)^T with two sparse products."""
    TW = sparse_mapping_matrix_transpose_matmul(w_tilde, pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index, pixels)
    return sparse_mapping_matrix_transpose_matmul(
        np.ascontiguousarray(TW.T), pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index, pixels
    )


def constant_regularization_matrix_from(coefficient, neighbors, neighbors_sizes):
    """numba.py:656-695: H = diag(1e-8 + c^2 sizes); H[i, k] -= c^2 per neighbour until -1."""
    S, P = neighbors.shape
    c2 = coefficient * coefficient
    H = np.diag(1e-8 + c2 * neighbors_sizes.astype(np.float64))
    for i in range(S):
        for j in range(P):
            k = neighbors[i, j]
            if k == -1:
                break
            H[i, k] -= c2
    return H


def reconstruction_positive_negative_from(data_vector, curvature_reg_matrix):
    """numba.py:698-731: s = A^-1 D via LAPACK LU."""
    return np.linalg.solve(curvature_reg_matrix, data_vector)


def noise_normalization_complex_from(noise_map):
    """numba.py:734-752: 2 (K log 2 pi + sum over 2K entries of log|n|)."""
    return 2.0 * (noise_map.size * LOG_TWO_PI + np.log(np.abs(noise_map.view(np.float64))).sum())


def log_likelihood_function_via_w_tilde_from(
    dirty_image,
    data,
    noise_map,
    w_tilde,
    pix_indexes_for_sub_slim_index,
    pix_weights_for_sub_slim_index,
    neighbors,
    neighbors_sizes,
):
    """numba.py:755-857 (lambda = 1 hard-coded at numba.py:819).

    logL = -1/2 (chi.chi - s.D + log|A| - log|H| + noise_norm), A = F + H.
    """
    S = neighbors_sizes.size
    coefficient = 1.0
    noise_normalization = noise_normalization_complex_from(noise_map)
    T = mapping_matrix_from(pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index, S)
    F = curvature_matrix_via_w_tilde_from(w_tilde, T)
    H = constant_regularization_matrix_from(coefficient, neighbors, neighbors_sizes)
    A = F + H
    D = data_vector_from(T, dirty_image)
    s = reconstruction_positive_negative_from(D, A)
    chi = data.view(np.float64) / noise_map.view(np.float64)
    reg_chi2 = chi @ chi - s @ D
    return -0.5 * (reg_chi2 + np.linalg.slogdet(A)[1] - np.linalg.slogdet(H)[1] + noise_normalization)


def log_likelihood_original_form(dirty_image, data, noise_map, w_tilde, pix_indexes_for_sub_slim_index,
                                 pix_weights_for_sub_slim_index, neighbors, neighbors_sizes, coefficient=1.0):
    """original.py:890-923: chi^2 = s^T F s - 2 s^T D + sum chi^2 ; + s^T H s."""
    S = neighbors_sizes.size
    T = mapping_matrix_from(pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index, S)
    D = data_vector_from(T, dirty_image)
    F = curvature_matrix_via_w_tilde_from(w_tilde, T)
    H = constant_regularization_matrix_from(coefficient, neighbors, neighbors_sizes)
    A = F + H
    s = np.linalg.solve(A, D)
    chi2 = s @ F @ s - 2.0 * s @ D + np.sum(data.real**2 / noise_map.real**2) + np.sum(data.imag**2 / noise_map.imag**2)
    reg = s @ H @ s
    nn = np.sum(np.log(2 * np.pi * noise_map.real**2)) + np.sum(np.log(2 * np.pi * noise_map.imag**2))
    return float(-0.5 * (chi2 + reg + np.linalg.slogdet(A)[1] - np.linalg.slogdet(H)[1] + nn))


def log_likelihood_cholesky_form(dirty_image, data, noise_map, w_tilde, pix_indexes_for_sub_slim_index,
                                 pix_weights_for_sub_slim_index, neighbors, neighbors_sizes, coefficient=1.0):
    """jax.py:1032-1058: Cholesky half-log-determinants of H and A."""
    import scipy.linalg as sla

    S = neighbors_sizes.size
    T = mapping_matrix_from(pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index, S)
    F = curvature_matrix_via_w_tilde_from(w_tilde, T)
    H = constant_regularization_matrix_from(coefficient, neighbors, neighbors_sizes)
    A = F + H
    D = data_vector_from(T, dirty_image)
    half_log_h = np.log(np.diag(np.linalg.cholesky(H))).sum()
    L = np.linalg.cholesky(A)
    s = sla.cho_solve((L, True), D)
    half_log_a = np.log(np.diag(L)).sum()
    chi_r = data.real / noise_map.real
    chi_i = data.imag / noise_map.imag
    r = chi_r @ chi_r + chi_i @ chi_i - s @ D
    nn = noise_normalization_complex_from(noise_map)
    return -0.5 * (r + nn) + half_log_h - half_log_a


# ---------------------------------------------------------------------------
# Explicit-lambda forms used by the batched path (the reference fixes lambda
# = 1; these are the same formulas with the coefficient as a parameter).
# ---------------------------------------------------------------------------


def log_evidence_from_parts(F, D, chi_squared, noise_normalization, neighbors, neighbors_sizes, coefficient,
                            logdet_h=None):
    """numba.py:839-857 with an explicit coefficient.  Returns (logL, s, logdet A, logdet H)."""
    H = constant_regularization_matrix_from(coefficient, neighbors, neighbors_sizes)
    A = F + H
    L = np.linalg.cholesky(A)
    y = np.linalg.solve(L, D)
    s = np.linalg.solve(L.T, y)
    ld_a = 2.0 * np.log(np.diag(L)).sum()
    if logdet_h is None:
        logdet_h = np.linalg.slogdet(H)[1]
    return -0.5 * (chi_squared - s @ D + ld_a - logdet_h + noise_normalization), s, ld_a, logdet_h


def log_likelihood_function_via_w_compact_from(dirty_image, data, noise_map, w_compact, native_index_for_slim_index,
                                               pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index,
                                               neighbors, neighbors_sizes):
    """numba.py:860-967: the log-likelihood of numba.py:755-857 with F from the lattice table."""
    S = neighbors_sizes.size
    F = curvature_matrix_via_w_compact_sparse_mapping_matrix_from(w_compact, native_index_for_slim_index,
                                                                  pix_indexes_for_sub_slim_index,
                                                                  pix_weights_for_sub_slim_index, S)
    H = constant_regularization_matrix_from(1.0, neighbors, neighbors_sizes)
    A = F + H
    T = mapping_matrix_from(pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index, S)
    D = data_vector_from(T, dirty_image)
    s = reconstruction_positive_negative_from(D, A)
    chi = data.view(np.float64) / noise_map.view(np.float64)
    return -0.5 * (chi @ chi - s @ D + np.linalg.slogdet(A)[1] - np.linalg.slogdet(H)[1]
                   + noise_normalization_complex_from(noise_map))
