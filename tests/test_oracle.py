"""Pins the CPU oracle (oracle/restated.py) against independent routes.

Each check is chosen so that a plausible mistake in the restatement (a
swapped grid/uv pairing, a dropped term, a sign, a transposed operand, the
-1 break rule) fails it.  CPU only.
"""

from pathlib import Path

import numpy as np
import pytest

from autojax_amd import datasets as ds
from oracle import evidence as ev
from oracle import restated as R

GOLD = Path(__file__).resolve().parent / "golden"


def small_problem(seed=3, n_pix=60, K=50, S=12, B=3, uv_max=3e5):
    rng = np.random.default_rng(seed)
    grid, native = ds.grid_for_pixel_count(n_pix, 0.1)
    uv = ds.uv_disc_uniform(rng, K, uv_max)
    data, noise = ds.visibilities(rng, K, 0.5, 1.5)
    mesh = ds.delaunay_source(rng, S)
    idx, w = ds.barycentric_mapping(rng, grid.shape[0], mesh.triangles)
    return grid, native, uv, data, noise, idx[:, :B], w[:, :B], mesh


# --------------------------------------------------------------------- W-tilde

def test_w_tilde_matches_fsum_golden():
    g = np.load(GOLD / "tiny_wtilde.npz")
    W = R.w_tilde_curvature_interferometer_from(g["noise_map_real"], g["uv_wavelengths"], g["grid_radians_slim"])
    np.testing.assert_allclose(W, g["w_tilde"], rtol=0, atol=1e-12 * np.abs(g["w_tilde"]).max())


def test_w_tilde_rows_match_fsum_golden_and_full_oracle():
    """The row-subset oracle against the brute-force fsum golden (every row, in blocks smaller than K)
    and against the full-matrix oracle on the 716-pixel workload."""
    g = np.load(GOLD / "tiny_wtilde.npz")
    M = g["grid_radians_slim"].shape[0]
    rows = R.w_tilde_rows_from(g["noise_map_real"], g["uv_wavelengths"], g["grid_radians_slim"], np.arange(M),
                               k_block=5)
    np.testing.assert_allclose(rows, g["w_tilde"], rtol=0, atol=1e-12 * np.abs(g["w_tilde"]).max())
    d = ds.config1(K=300)
    W = R.w_tilde_curvature_interferometer_from(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim)
    sel = np.array([0, 5, 357, 715])
    Wr = R.w_tilde_rows_from(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim, sel, k_block=64)
    np.testing.assert_allclose(Wr, W[sel], rtol=0, atol=1e-12 * np.abs(W).max())


@pytest.fixture(scope="module")
def c_oracle():
    """The C restatement (oracle/wtilde_cpu.c), built with `make -C oracle` when it is not there yet."""
    import subprocess

    from oracle import wtilde_c

    if not wtilde_c.LIB_PATH.exists():
        subprocess.run(["make", "-C", str(wtilde_c.LIB_PATH.parent), "-s"], check=True)
    return wtilde_c


def test_c_oracle_matches_fsum_golden_and_numpy_oracle(c_oracle):
    """The C/OpenMP loop (full matrix, any thread count) and its row-subset form against the brute-force
    fsum golden and the numpy restatement on the 716-pixel workload; a transposed grid pairing fails."""
    g = np.load(GOLD / "tiny_wtilde.npz")
    args = (g["noise_map_real"], g["uv_wavelengths"], g["grid_radians_slim"])
    tol = 1e-12 * np.abs(g["w_tilde"]).max()
    for threads in (1, 3):
        np.testing.assert_allclose(c_oracle.w_tilde(*args, threads=threads), g["w_tilde"], rtol=0, atol=tol)
    M = g["grid_radians_slim"].shape[0]
    np.testing.assert_allclose(c_oracle.w_tilde_rows(*args, np.arange(M), threads=2), g["w_tilde"], rtol=0, atol=tol)
    swapped = c_oracle.w_tilde(args[0], args[1], args[2][:, ::-1].copy(), threads=2)
    assert np.abs(swapped - g["w_tilde"]).max() > 1e-3 * np.abs(g["w_tilde"]).max()
    d = ds.config1(K=300)
    W = R.w_tilde_curvature_interferometer_from(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim)
    Wc = c_oracle.w_tilde(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim, threads=4)
    np.testing.assert_allclose(Wc, W, rtol=0, atol=1e-12 * np.abs(W).max())
    sel = np.array([0, 5, 300, 357, 715])
    Wr = c_oracle.w_tilde_rows(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim, sel, threads=4)
    np.testing.assert_allclose(Wr, W[sel], rtol=0, atol=1e-12 * np.abs(W).max())
    with pytest.raises(ValueError):
        c_oracle.w_tilde_rows(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim, [716])


def test_w_tilde_pairing_is_not_transposed():
    # swapping the (y, x) grid columns must change W for a generic uv set
    g = np.load(GOLD / "tiny_wtilde.npz")
    W = R.w_tilde_curvature_interferometer_from(g["noise_map_real"], g["uv_wavelengths"], g["grid_radians_slim"][:, ::-1].copy())
    assert np.abs(W - g["w_tilde"]).max() > 1e-3 * np.abs(g["w_tilde"]).max()


def test_w_tilde_closed_forms():
    grid, native, uv, data, noise, idx, w, mesh = small_problem()
    n = noise.real
    W = R.w_tilde_curvature_interferometer_from(n, uv, grid)
    # diagonal: cos(0) = 1 for every k
    np.testing.assert_allclose(np.diag(W), np.sum(1 / n**2), rtol=1e-13)
    np.testing.assert_allclose(W, W.T, rtol=0, atol=1e-12 * W.max())
    assert np.linalg.eigvalsh(W).min() > -1e-9 * W.max()
    # K = 1: rank-2 matrix w (c c^T + s s^T)
    W1 = R.w_tilde_curvature_interferometer_from(n[:1], uv[:1], grid)
    ph = 2 * np.pi * (grid[:, 1] * uv[0, 0] + grid[:, 0] * uv[0, 1])
    c, s = np.cos(ph), np.sin(ph)
    np.testing.assert_allclose(W1, (np.outer(c, c) + np.outer(s, s)) / n[0] ** 2, rtol=0, atol=1e-12)
    # translation invariance
    W2 = R.w_tilde_curvature_interferometer_from(n, uv, grid + np.array([1e-5, -2e-5]))
    np.testing.assert_allclose(W2, W, rtol=0, atol=1e-9 * W.max())


def test_w_tilde_matches_compact_lattice_form():
    # numba.py:210-255: independent lattice formulation on the same grid
    rng = np.random.default_rng(5)
    N, ps = 9, 0.1
    grid, native = ds.circular_grid(N, ps, 0.45)
    K = 30
    uv = ds.uv_disc_uniform(rng, K, 4e5)
    n = rng.uniform(0.5, 1.5, K)
    W = R.w_tilde_curvature_interferometer_from(n, uv, grid)
    wc = R.w_compact_curvature_interferometer_from(N, n, uv, ps)
    Wc = R.w_tilde_via_compact_from(wc, native)
    np.testing.assert_allclose(Wc, W, rtol=0, atol=1e-10 * W.max())


def test_compact_curvature_and_likelihood_match_dense_route():
    # numba.py:500-549 / :617-647 / :860-967 against the dense route with the brute-force W
    rng = np.random.default_rng(15)
    N, ps = 8, 0.1
    grid, native = ds.circular_grid(N, ps, 0.4)
    M, K, S = grid.shape[0], 25, 9
    uv = ds.uv_disc_uniform(rng, K, 4e5)
    n = rng.uniform(0.5, 1.5, K)
    W = ev.w_tilde_bruteforce(n, uv, grid)
    wc = R.w_compact_curvature_interferometer_from(N, n, uv, ps)
    idx = np.array([rng.choice(S, 3, replace=False) for _ in range(M)])  # distinct per row, as in the reference
    idx[::5, 2] = -1
    wts = rng.uniform(0.1, 1.0, (M, 3))
    F = R.curvature_matrix_via_w_compact_sparse_mapping_matrix_from(wc, native, idx, wts, S)
    T = R.mapping_matrix_from(idx, wts, S)
    np.testing.assert_allclose(F, T.T @ W @ T, rtol=0, atol=1e-11 * np.abs(F).max())
    Fd = R.curvature_matrix_via_w_compact_sparse_mapping_matrix_direct_from(wc, native, idx, wts, S)
    np.testing.assert_allclose(Fd, F, rtol=0, atol=1e-11 * np.abs(F).max())
    # a preload table holding the same lattice values gives the same W (numba.py:259-291)
    O = N - 1
    pre = np.zeros((2 * N, 2 * N))
    for a in range(-O, N):
        for b in range(-O, N):
            pre[-a, -b] = wc[a, b + O] if a >= 0 else wc[-a, -b + O]
    np.testing.assert_allclose(R.w_tilde_via_preload_from(pre, native), W, rtol=0, atol=1e-10 * W.max())
    mesh = ds.rectangular_source(3)
    data = rng.standard_normal(K) + 1j * rng.standard_normal(K)
    noise = n * (1 + 1j)
    y = ds.dirty_image_numpy(data, noise, uv, grid)
    got = R.log_likelihood_function_via_w_compact_from(y, data, noise, wc, native, idx, wts, *mesh)
    want = R.log_likelihood_function_via_w_tilde_from(y, data, noise, W, idx, wts, *mesh)
    assert abs(got - want) < 1e-10 * abs(want)


def test_w_tilde_data_bruteforce():
    rng = np.random.default_rng(6)
    grid, _ = ds.circular_grid(5, 0.2, 0.45)
    K = 20
    uv = ds.uv_disc_uniform(rng, K, 2e5)
    v = rng.uniform(0.5, 2.0, K)
    n = rng.uniform(0.5, 1.5, K)
    got = R.w_tilde_data_interferometer_from(v, n, uv, grid)
    import math

    for i in range(grid.shape[0]):
        want = math.fsum((n[k] ** 2 / v[k]) ** 2 * math.cos(2 * math.pi * (grid[i, 1] * uv[k, 0] + grid[i, 0] * uv[k, 1])) for k in range(K))
        assert abs(got[i] - want) < 1e-12 * max(1.0, abs(want))


# ------------------------------------------------------------- mapping, D, F

def test_mapping_matrix_break_rule_and_assignment():
    idx = np.array([[0, 2, -1], [1, -1, 2], [2, 2, 0]], np.int64)
    w = np.array([[0.5, 0.25, 9.0], [1.0, 7.0, 7.0], [0.1, 0.2, 0.3]])
    T = R.mapping_matrix_from(idx, w, 3)
    want = np.array([[0.5, 0, 0.25], [0, 1.0, 0], [0.3, 0, 0.2]])  # entries after -1 ignored; last assignment wins
    np.testing.assert_array_equal(T, want)


def test_data_vector_and_curvature_routes():
    grid, native, uv, data, noise, idx, w, mesh = small_problem()
    S = mesh.vertices.shape[0]
    T = R.mapping_matrix_from(idx, w, S)
    y = np.random.default_rng(1).standard_normal(grid.shape[0])
    D = R.data_vector_from(T, y)
    Dl = np.array([sum(T[m, s] * y[m] for m in range(T.shape[0])) for s in range(S)])
    np.testing.assert_allclose(D, Dl, rtol=1e-12, atol=1e-12)
    W = R.w_tilde_curvature_interferometer_from(noise.real, uv, grid)
    F = R.curvature_matrix_via_w_tilde_from(W, T)
    Fs = R.curvature_matrix_via_w_wilde_sparse_mapping_matrix_from(W, idx, w, S)
    np.testing.assert_allclose(Fs, F, rtol=0, atol=1e-11 * np.abs(F).max())
    # relabelling source pixels permutes F
    perm = np.random.default_rng(2).permutation(S)
    inv = np.argsort(perm)
    Fp = R.curvature_matrix_via_w_tilde_from(W, R.mapping_matrix_from(inv[idx], w, S))
    np.testing.assert_allclose(Fp[np.ix_(inv, inv)], F, rtol=0, atol=1e-11 * np.abs(F).max())


def test_curvature_one_hot_block_sums():
    rng = np.random.default_rng(4)
    M, S = 40, 6
    W = rng.standard_normal((M, M))
    W = W + W.T
    idx = rng.integers(0, S, (M, 1)).astype(np.int64)
    F = R.curvature_matrix_via_w_tilde_from(W, R.mapping_matrix_from(idx, np.ones((M, 1)), S))
    for a in range(S):
        for b in range(S):
            want = W[np.ix_(idx[:, 0] == a, idx[:, 0] == b)].sum()
            assert abs(F[a, b] - want) < 1e-10


# -------------------------------------------------------------- regulariser

def test_regularization_grid_closed_form_golden():
    g = np.load(GOLD / "grid20_logdet.npz")
    nb, sz = ds.rectangular_source(20)
    for lam, want in zip(g["lambdas"], g["logdet"]):
        H = R.constant_regularization_matrix_from(lam, nb, sz)
        np.testing.assert_allclose(H, H.T, rtol=0, atol=0)
        np.testing.assert_allclose(H.sum(axis=1), 1e-8, rtol=1e-6, atol=1e-12 * lam**2)
        err = abs(ev.logdet_eigh(H) - want)
        if lam <= 1.0:
            assert err < 1e-9 * abs(want) + 1e-9
        else:
            # the 1e-8 eigenvalue drowns in round-off of an O(lam^2) spectrum: why log|H| is deflated
            assert err > 1e-9 * abs(want)


def test_regularization_matches_original_loop_form():
    # original.py:703-743 loops to neighbors_sizes[i] instead of breaking at -1
    nb, sz = ds.rectangular_source(5)
    H = R.constant_regularization_matrix_from(0.7, nb, sz)
    S = sz.size
    Ho = np.zeros((S, S))
    for i in range(S):
        Ho[i, i] += 1e-8
        for j in range(sz[i]):
            Ho[i, i] += 0.49
            Ho[i, nb[i, j]] -= 0.49
    np.testing.assert_allclose(H, Ho, rtol=1e-15, atol=1e-15)


def test_deflated_logdet_beats_plain_at_large_lambda():
    g = np.load(GOLD / "grid20_logdet.npz")
    nb, sz = ds.rectangular_source(20)
    for lam, want in zip(g["lambdas"], g["logdet"]):
        H = R.constant_regularization_matrix_from(lam, nb, sz)
        got = ev.logdet_deflated(H, lam**2 * sz.mean())
        assert abs(got - want) < 1e-8 * abs(want) + 1e-8, (lam, got, want)


# ---------------------------------------------------- solve, norm, likelihood

def test_solve_lu_vs_cholesky():
    rng = np.random.default_rng(8)
    X = rng.standard_normal((30, 30))
    A = X @ X.T + 30 * np.eye(30)
    D = rng.standard_normal(30)
    s = R.reconstruction_positive_negative_from(D, A)
    np.testing.assert_allclose(A @ s, D, rtol=1e-12, atol=1e-12)
    L = np.linalg.cholesky(A)
    np.testing.assert_allclose(np.linalg.solve(L.T, np.linalg.solve(L, D)), s, rtol=1e-12)


def test_noise_normalization_forms():
    rng = np.random.default_rng(9)
    n = rng.uniform(0.5, 2, 50) + 1j * rng.uniform(0.5, 2, 50)
    a = R.noise_normalization_complex_from(n)
    b = np.sum(np.log(2 * np.pi * n.real**2)) + np.sum(np.log(2 * np.pi * n.imag**2))  # original.py:781-798
    assert abs(a - b) < 1e-11 * abs(b)


def test_log_likelihood_equals_gaussian_evidence_golden():
    g = np.load(GOLD / "tiny_evidence.npz")
    W = R.w_tilde_curvature_interferometer_from(g["noise_map"].real.copy(), g["uv_wavelengths"], g["grid_radians_slim"])
    S = g["neighbors_sizes"].size
    T = R.mapping_matrix_from(g["pix_indexes_for_sub_slim_index"], g["pix_weights_for_sub_slim_index"], S)
    chi = g["data"].view(np.float64) / g["noise_map"].view(np.float64)
    nn = R.noise_normalization_complex_from(g["noise_map"])
    F = R.curvature_matrix_via_w_tilde_from(W, T)
    D = R.data_vector_from(T, g["dirty_image"])
    for lam, want in zip(g["lambdas"], g["log_evidence"]):
        got, *_ = R.log_evidence_from_parts(F, D, chi @ chi, nn, g["neighbors"], g["neighbors_sizes"], lam)
        # the dense 2K x 2K evidence carries H's 1e-8 eigenvalue as a 1e8 prior variance, so it
        # holds ~8 digits; a dropped or mis-signed term is O(1) or larger
        assert abs(got - want) < 3e-8 * abs(want), (lam, got, want)
    # the lambda = 1 drop-in form of numba.py:755-857
    got1 = R.log_likelihood_function_via_w_tilde_from(g["dirty_image"], g["data"], g["noise_map"], W,
                                                      g["pix_indexes_for_sub_slim_index"],
                                                      g["pix_weights_for_sub_slim_index"], g["neighbors"],
                                                      g["neighbors_sizes"])
    assert abs(got1 - g["log_evidence"][1]) < 3e-8 * abs(got1)


def test_three_reference_forms_agree():
    grid, native, uv, data, noise, idx, w, mesh = small_problem(seed=21)
    dirty = ds.dirty_image_numpy(data, noise, uv, grid)
    W = R.w_tilde_curvature_interferometer_from(noise.real, uv, grid)
    args = (dirty, data, noise, W, idx, w, mesh.neighbors, mesh.neighbors_sizes)
    a = R.log_likelihood_function_via_w_tilde_from(*args)
    b = R.log_likelihood_original_form(*args)
    c = R.log_likelihood_cholesky_form(*args)
    assert abs(a - b) < 1e-10 * abs(a) and abs(a - c) < 1e-10 * abs(a)


def test_wrong_sign_dirty_image_breaks_evidence():
    # a planted mistake in the oracle inputs: flipping the sin term must break the evidence identity
    g = np.load(GOLD / "tiny_evidence.npz")
    data = g["data"].real - 1j * g["data"].imag
    dirty_bad = ds.dirty_image_numpy(data, g["noise_map"], g["uv_wavelengths"], g["grid_radians_slim"])
    W = R.w_tilde_curvature_interferometer_from(g["noise_map"].real.copy(), g["uv_wavelengths"], g["grid_radians_slim"])
    got = R.log_likelihood_function_via_w_tilde_from(dirty_bad, g["data"], g["noise_map"], W,
                                                     g["pix_indexes_for_sub_slim_index"],
                                                     g["pix_weights_for_sub_slim_index"], g["neighbors"],
                                                     g["neighbors_sizes"])
    assert abs(got - g["log_evidence"][1]) > 1e-6


# ------------------------------------------------------- preload form (original.py)

def test_preload_table_reproduces_w_tilde():
    # original.py:230-403 then numba.py:259-291 against the brute-force W on the reference harness's
    # grid (tests.py:291-306; the reference's own equivalence test, tests.py:720-731).  A table whose
    # second quadrant takes its offsets from the wrong corner must fail.
    d = ds.reference_convention_dataset(seed=21, K=60, S=20)
    n, uv, N = d.noise_map_real, d.uv_wavelengths, d.n_side
    W = ev.w_tilde_bruteforce(n, uv, d.grid_radians_slim)
    pre = R.w_tilde_curvature_preload_interferometer_from(n, uv, (N, N), d.grid_radians_2d)
    assert pre.shape == (2 * N, 2 * N)
    assert not pre[N].any() and not pre[:, N].any()
    Wp = R.w_tilde_via_preload_from(pre, d.native_index_for_slim_index)
    np.testing.assert_allclose(Wp, W, rtol=0, atol=1e-10 * W.max())
    bad = pre.copy()
    g, j = d.grid_radians_2d, np.arange(1, N)
    gp = g[:, N - j - 1]
    bad[:N, 2 * N - j] = R._preload_quadrant(n, uv, g[0, 0, 0] - gp[:, :, 0], g[0, 0, 1] - gp[:, :, 1])
    assert not np.allclose(R.w_tilde_via_preload_from(bad, d.native_index_for_slim_index), W, rtol=0,
                           atol=1e-6 * W.max())


def test_preload_curvature_matches_dense_route():
    # original.py:600-700 (pixel pairs, transpose, own-pixel terms; rows cut by pix_size) against
    # T^T W T with the brute-force W (the reference's TestCurvatureMatrix equivalence, tests.py:797-809)
    d = ds.reference_convention_dataset(seed=22, K=40, S=16)
    n, uv, N = d.noise_map_real, d.uv_wavelengths, d.n_side
    W = ev.w_tilde_bruteforce(n, uv, d.grid_radians_slim)
    pre = R.w_tilde_curvature_preload_interferometer_from(n, uv, (N, N), d.grid_radians_2d)
    s = d.samples[0]
    M, S = d.M, s.pixels
    size = np.full(M, 3, dtype=np.int64)
    size[::7] = 2
    size[::11] = 1
    F = R.curvature_matrix_via_w_tilde_curvature_preload_interferometer_from(
        pre, s.pix_indexes_for_sub_slim_index, size, s.pix_weights_for_sub_slim_index, d.native_index_for_slim_index, S)
    idx = s.pix_indexes_for_sub_slim_index.copy()
    idx[np.arange(3)[None, :] >= size[:, None]] = -1
    T = R.mapping_matrix_from(idx, s.pix_weights_for_sub_slim_index, S)
    np.testing.assert_allclose(F, T.T @ W @ T, rtol=0, atol=1e-10 * np.abs(F).max())


def test_preload_gather_reads_every_pair():
    # numba.py:282-291 reads the table for every (i, j), with numpy's wrapping of negative offsets: against
    # the reference's double loop written out with Python indexing on a table that is not point-symmetric
    # (so the lower triangle is not the mirror of the upper one); a mirrored gather (original.py:431-447)
    # would fail this
    d = ds.reference_convention_dataset(seed=23, K=8, S=10)
    N = d.n_side
    nat = d.native_index_for_slim_index[:50]
    pre = np.random.default_rng(0).standard_normal((2 * N, 2 * N))
    W = R.w_tilde_via_preload_from(pre, nat)
    loop = np.empty_like(W)
    for i in range(nat.shape[0]):
        for j in range(nat.shape[0]):
            loop[i, j] = pre[int(nat[j, 0] - nat[i, 0]), int(nat[j, 1] - nat[i, 1])]
    np.testing.assert_array_equal(W, loop)
    assert not np.array_equal(W, W.T)
    mirrored = np.triu(W) + np.triu(W, 1).T
    assert not np.array_equal(mirrored, loop)


# ------------------------------------------------------- source-mesh mapper (upstream of the reference)

def test_mapper_oracle_delaunay_property_and_euler():
    from scipy.spatial import ConvexHull

    from oracle import mapper as MP

    rng = np.random.default_rng(40)
    V = rng.standard_normal((150, 2))
    T = MP.delaunay_triangles(V)
    h = len(ConvexHull(V).vertices)
    assert T.shape[0] == 2 * len(V) - h - 2
    # no vertex strictly inside any triangle's circumcircle
    for t in T:
        a, b, c = V[t]
        ax, ay, bx, by, cx, cy = a[1], a[0], b[1], b[0], c[1], c[0]
        d = 2 * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by))
        ux = ((ax**2 + ay**2) * (by - cy) + (bx**2 + by**2) * (cy - ay) + (cx**2 + cy**2) * (ay - by)) / d
        uy = ((ax**2 + ay**2) * (cx - bx) + (bx**2 + by**2) * (ax - cx) + (cx**2 + cy**2) * (bx - ax)) / d
        r2 = (ax - ux) ** 2 + (ay - uy) ** 2
        d2 = (V[:, 1] - ux) ** 2 + (V[:, 0] - uy) ** 2
        assert np.all(d2 >= r2 * (1 - 1e-9))
    # neighbour lists from triangles equal those from the edge list
    nb, sz = MP.neighbors_from_triangles(len(V), T)
    edges = np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]])
    nb2, sz2 = ds.neighbors_from_edges(len(V), edges)
    np.testing.assert_array_equal(sz, sz2)
    np.testing.assert_array_equal(nb, nb2)


def test_mapper_oracle_barycentric_reproduces_points_and_nearest_outside():
    from oracle import mapper as MP

    rng = np.random.default_rng(41)
    V = rng.random((80, 2))
    T = MP.delaunay_triangles(V)
    P = rng.uniform(-0.2, 1.2, (400, 2))
    idx, w = MP.barycentric_mapping(V, T, P)
    live = idx[:, 1] >= 0
    assert 0 < live.sum() < len(P)
    np.testing.assert_allclose(w.sum(axis=1), 1.0, rtol=0, atol=1e-12)
    rec = np.einsum("mb,mbk->mk", w[live], V[idx[live]])
    np.testing.assert_allclose(rec, P[live], rtol=0, atol=1e-12)
    assert np.all(w[live] >= -1e-15)
    far = ~live
    d = ((P[far, None, :] - V[None, :, :]) ** 2).sum(-1)
    np.testing.assert_array_equal(idx[far, 0], d.argmin(axis=1))
    assert np.all(w[far, 0] == 1.0) and np.all(idx[far, 1:] == -1)
    # a plant: weights in the wrong vertex order no longer reproduce the points
    rec_bad = np.einsum("mb,mbk->mk", w[live][:, [1, 0, 2]], V[idx[live]])
    assert not np.allclose(rec_bad, P[live], atol=1e-6)
