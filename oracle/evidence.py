"""Independent checks that pin the oracle (test infrastructure only).

None of these restate the reference's code; each computes the same quantity
by a different route, so that a dropped term, a wrong sign or a transposed
operand in ``restated`` fails a comparison.
"""

from __future__ import annotations

import math

import numpy as np


def w_tilde_bruteforce(noise_map_real, uv_wavelengths, grid_radians_slim):
    """Triple loop with math.fsum (exactly rounded sums) -- tiny sizes only.

    W_ij = sum_k n_k^-2 cos(2 pi [(g_i1 - g_j1) u_k0 + (g_i0 - g_j0) u_k1])
    (the formula in the docstring of numba.py:175-176).
    """
    M = len(grid_radians_slim)
    K = len(noise_map_real)
    W = [[0.0] * M for _ in range(M)]
    for i in range(M):
        for j in range(M):
            dy = grid_radians_slim[i][0] - grid_radians_slim[j][0]
            dx = grid_radians_slim[i][1] - grid_radians_slim[j][1]
            W[i][j] = math.fsum(
                math.cos(2.0 * math.pi * (dx * uv_wavelengths[k][0] + dy * uv_wavelengths[k][1]))
                / (noise_map_real[k] * noise_map_real[k])
                for k in range(K)
            )
    return np.array(W)


def design_matrix(uv_wavelengths, grid_radians_slim):
    """Real 2K x M response of a pixel image: rows cos(phi_k . x), then sin."""
    ph = 2.0 * np.pi * (uv_wavelengths[:, 0:1] * grid_radians_slim[None, :, 1] + uv_wavelengths[:, 1:2] * grid_radians_slim[None, :, 0])
    return np.concatenate([np.cos(ph), np.sin(ph)], axis=0)


def gaussian_evidence(data, noise_map, uv_wavelengths, grid_radians_slim, T, H):
    """log N(d; 0, Phi H^-1 Phi^T + N) for the 2K real data d = [Re d, Im d].

    Phi = R T with R the design matrix above and N = diag(sigma^2) with
    sigma = Re noise for both components.  This is the Bayesian evidence of
    the linear model with prior s ~ N(0, H^-1); the reference's
    log-likelihood (numba.py:846-857) equals it when the dirty image is
    R^T N^-1 d (datasets.dirty_image_numpy).  Direct dense computation in
    dimension 2K -- tiny sizes only.
    """
    R = design_matrix(uv_wavelengths, grid_radians_slim)
    Phi = R @ T
    sig = np.concatenate([noise_map.real, noise_map.real])
    d = np.concatenate([data.real, data.imag])
    C = Phi @ np.linalg.solve(H, Phi.T) + np.diag(sig**2)
    sign, logdet = np.linalg.slogdet(2.0 * np.pi * C)
    assert sign > 0
    return -0.5 * (d @ np.linalg.solve(C, d) + logdet)


def grid_laplacian_logdet(n, coefficient, eps=1e-8):
    """log|H| for the 4-connected n x n grid: eigenvalues c^2 (4 - 2cos(pi p/n) - 2cos(pi q/n)) + eps.

    The path-graph Laplacian has eigenvalues 2 - 2 cos(pi p / n), p = 0..n-1,
    and the grid Laplacian is their Kronecker sum.
    """
    mu = 2.0 - 2.0 * np.cos(np.pi * np.arange(n) / n)
    lam = coefficient**2 * (mu[:, None] + mu[None, :]) + eps
    return float(np.sum(np.log(lam)))


def logdet_eigh(H):
    """log|H| from the eigenvalues of the symmetric matrix H."""
    ev = np.linalg.eigvalsh(H)
    assert ev.min() > 0
    return float(np.sum(np.log(ev)))


def logdet_deflated(H, beta, eps=1e-8):
    """log|H| = log|H + (beta/S) 1 1^T| - log(eps + beta) + log eps, valid when H 1 = eps 1."""
    S = H.shape[0]
    L = np.linalg.cholesky(H + beta / S)
    return 2.0 * np.log(np.diag(L)).sum() - np.log(eps + beta) + np.log(eps)


def longdouble_cholesky_logdet(A):
    """log|A| by a Cholesky in numpy longdouble (x87 80-bit, ~19 digits) -- small S only."""
    A = np.array(A, dtype=np.longdouble)
    n = A.shape[0]
    L = np.zeros_like(A)
    for j in range(n):
        d = A[j, j] - np.dot(L[j, :j], L[j, :j])
        L[j, j] = np.sqrt(d)
        L[j + 1 :, j] = (A[j + 1 :, j] - L[j + 1 :, :j] @ L[j, :j]) / L[j, j]
    return float(2.0 * np.sum(np.log(np.diag(L))))


def torch_log_evidence(coefficient, weights, idx, W, dirty, chi2, nn, neighbors, neighbors_sizes):
    """Dense FP64 torch restatement of numba.py:818-857 with an explicit coefficient, for autograd
    cross-checks (coefficient and weights may require grad).  Runs on the device of `weights` (W and
    dirty may be numpy arrays or tensors); dense T, so sizes up to a few thousand source pixels."""
    import torch

    dev = weights.device
    M, B = idx.shape
    S = neighbors_sizes.shape[0]
    live = np.ones(idx.shape, bool)
    if B > 1:
        live[:, 1:] = np.cumprod(idx[:, :-1] != -1, axis=1).astype(bool)
    live &= idx >= 0
    rows, cols = np.nonzero(live)
    T = torch.zeros((M, S), dtype=torch.float64, device=dev)
    r_t, c_t = torch.as_tensor(rows, device=dev), torch.as_tensor(cols, device=dev)
    T = T.index_put((r_t, torch.as_tensor(idx[rows, cols], device=dev)), weights[r_t, c_t])
    Wt = torch.as_tensor(W, device=dev)
    F = T.T @ (Wt @ T)
    D = T.T @ torch.as_tensor(dirty, device=dev)
    Lap = np.diag(np.asarray(neighbors_sizes, dtype=np.float64))
    for i in range(S):
        for k in neighbors[i]:
            if k == -1:
                break
            Lap[i, k] -= 1.0
    Lap = torch.as_tensor(Lap, device=dev)
    H = coefficient**2 * Lap + 1e-8 * torch.eye(S, dtype=torch.float64, device=dev)
    A = F + H
    s = torch.linalg.solve(A, D)
    return -0.5 * (chi2 - s @ D + torch.linalg.slogdet(A)[1] - torch.linalg.slogdet(H)[1] + nn)


def torch_barycentric(vertices, positions, triangles):
    """Barycentric weights by areas, torch (autograd w.r.t. vertices)."""
    import torch

    V = vertices[torch.as_tensor(triangles)]  # [M,3,2] (y, x)
    P = torch.as_tensor(positions)

    def a2(p0, p1, p2):
        return (p1[:, 1] - p0[:, 1]) * (p2[:, 0] - p0[:, 0]) - (p2[:, 1] - p0[:, 1]) * (p1[:, 0] - p0[:, 0])

    A = a2(V[:, 0], V[:, 1], V[:, 2])
    return torch.stack([a2(P, V[:, 1], V[:, 2]) / A, a2(V[:, 0], P, V[:, 2]) / A, a2(V[:, 0], V[:, 1], P) / A], 1)
