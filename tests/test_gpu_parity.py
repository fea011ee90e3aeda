"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle and the golden fixtures."""

from pathlib import Path

import numpy as np
import pytest

from autojax_amd import datasets as ds
from oracle import evidence as ev
from oracle import restated as R

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"

# normwise tolerances of W-tilde against the FP64 oracle, relative to max|W| (= the diagonal)
WT_TOL = {"fp64": 1e-12, "int8x4": 5e-8, "int8x5": 5e-12, "int8crt": 5e-12, "int8crt11": 3e-11, "int8x3": 3e-7,
          "fp32": 3e-6, "bf16x3": 3e-4}


@pytest.fixture(scope="module")
def cfg1_small():
    """The 716-pixel workload with K cut to 600 visibilities (oracle W-tilde in seconds)."""
    d = ds.config1(K=600)
    W = R.w_tilde_curvature_interferometer_from(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim)
    d.ensure_dirty_image()
    return d, W


@pytest.fixture(scope="module")
def delaunay_small():
    d = ds.small_delaunay(31, 700, 1500, (40, 120), n_samples=6, lam_range=(0.01, 100.0))
    W = R.w_tilde_curvature_interferometer_from(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim)
    d.ensure_dirty_image()
    return d, W


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_w_tilde_fp64_matches_fsum_golden(gpu):
    from autojax_amd import api

    g = np.load(GOLD / "tiny_wtilde.npz")
    W = api.w_tilde_curvature_interferometer_from(g["noise_map_real"], g["uv_wavelengths"], g["grid_radians_slim"])
    print("tiny golden rel err", _rel(W, g["w_tilde"]), flush=True)
    assert _rel(W, g["w_tilde"]) < 1e-13


@pytest.mark.parametrize("precision", ["fp64", "int8x4", "int8x5", "int8crt", "int8crt11", "int8x3", "fp32", "bf16x3"])
def test_w_tilde_modes_vs_oracle(gpu, cfg1_small, precision):
    from autojax_amd import api

    d, Wo = cfg1_small
    W = api.w_tilde_curvature_interferometer_from(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim,
                                                  precision=precision)
    err = _rel(W, Wo)
    print(f"{precision}: normwise rel err {err:.3e}", flush=True)
    assert err < WT_TOL[precision]
    np.testing.assert_array_equal(W, W.T)


@pytest.mark.parametrize("precision", ["fp64", "int8x5", "int8crt", "int8crt11"])
def test_w_tilde_edge_sizes(gpu, precision):
    """Image-pixel counts around the 256-row tile edge and tiny ones (1, 2, 255, 256, 257) with 0, 1 and 65
    visibilities: zeros without visibilities, the oracle's W otherwise (a single visibility gives the
    rank-2 matrix w cos(phi_i - phi_j))."""
    from autojax_amd import api

    rng = np.random.default_rng(60)
    for M in (1, 2, 255, 256, 257):
        grid, _ = ds.grid_for_pixel_count(max(M, 3), 0.05)
        grid = np.ascontiguousarray(grid[:M])
        for K in (0, 1, 65):
            uv = ds.uv_loguniform(rng, K, 1e3, 5e5) if K else np.zeros((0, 2))
            n = rng.uniform(0.5, 2.0, K)
            W = api.w_tilde_curvature_interferometer_from(n, uv, grid, precision=precision)
            assert W.shape == (M, M)
            if K == 0:
                assert not W.any()
                continue
            Wo = R.w_tilde_curvature_interferometer_from(n, uv, grid)
            err = _rel(W, Wo)
            assert err < WT_TOL[precision], (M, K, err)
            np.testing.assert_array_equal(W, W.T)


def test_w_tilde_full_size_properties(gpu):
    """The batching workload's size (M 5,000, K 1e6): the mixed path's W (int8crt11) against the FP64
    feature SYRK (norm-wise 2e-13), exactly symmetric, its diagonal equal to sum_k n_k^-2 (1e-12
    relative), and two builds bitwise equal."""
    import torch

    from autojax_amd import device as dv
    from autojax_amd import evidence as E

    rng = np.random.default_rng(5)
    grid, _ = ds.grid_for_pixel_count(5000, 0.05)
    K = 1_000_000
    uv = ds.uv_loguniform(rng, K, 1e3, 5e5)
    _, noise = ds.visibilities(rng, K, 0.5, 2.0)
    nr = np.ascontiguousarray(noise.real)
    n_d, uv_d, g_d = dv.to_device(nr), dv.to_device(uv), dv.to_device(grid)
    W = E.w_tilde(n_d, uv_d, g_d, "int8crt11")
    W2 = E.w_tilde(n_d, uv_d, g_d, "int8crt11")
    W64 = E.w_tilde(n_d, uv_d, g_d, "fp64")
    torch.cuda.synchronize()
    err = float(((W - W64).abs().max() / W64.abs().max()).item())
    diag = float(np.sum(nr**-2.0))
    derr = float(((torch.diagonal(W) - diag).abs().max() / diag).item())
    print(f"int8crt11 vs fp64 {err:.3e}, diagonal {derr:.3e}", flush=True)
    assert err < 2e-13 and derr < 1e-12
    assert torch.equal(W, W.T) and torch.equal(W, W2)


def test_w_tilde_cfg2_rows_vs_oracle(gpu):
    """The plain-method workload at full size (M 7,860, K 1e5, the generator of config 2): rows 0, M/2 and
    M-1 of the FP64 W against the oracle's row sum, within 1e-10 of max|W| (the tolerance that workload
    states), and of the mixed path (int8crt11) within its mode tolerance; both W exactly symmetric."""
    import torch

    from autojax_amd import device as dv
    from autojax_amd import evidence as E

    d = ds.config2()
    rows = np.array([0, d.M // 2, d.M - 1])
    want = R.w_tilde_rows_from(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim, rows, k_block=1024)
    n_d, uv_d, g_d = dv.to_device(d.noise_map_real), dv.to_device(d.uv_wavelengths), dv.to_device(d.grid_radians_slim)
    scale = np.abs(want).max()
    for precision, tol in (("fp64", 1e-10), ("int8crt11", WT_TOL["int8crt11"])):
        W = E.w_tilde(n_d, uv_d, g_d, precision)
        torch.cuda.synchronize()
        assert torch.equal(W, W.T)
        err = np.abs(W[torch.from_numpy(rows).cuda()].cpu().numpy() - want).max() / scale
        print(f"cfg2 {precision}: rows vs oracle {err:.3e} (M {d.M}, K {d.K})", flush=True)
        assert err < tol


def test_w_tilde_multichunk_rows(gpu):
    """K spanning several 4096-visibility chunks; oracle on a row subset at full K."""
    from autojax_amd import api

    d = ds.config1(K=9000)
    W = api.w_tilde_curvature_interferometer_from(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim)
    rows = np.array([0, 1, 57, 300, 715])
    g = d.grid_radians_slim
    n = d.noise_map_real
    for i in rows:
        dg = 2 * np.pi * (g[i] - g)
        ph = dg[:, 1:2] * d.uv_wavelengths[None, :, 0] + dg[:, 0:1] * d.uv_wavelengths[None, :, 1]
        want = np.cos(ph) @ (1 / n**2)
        assert np.abs(W[i] - want).max() < 1e-12 * want.max()


def test_w_tilde_crt_multichunk(gpu):
    """The residue (CRT) mode over three 16384-visibility chunks, the last one ragged: against the
    oracle on a row subset, against the five-digit mode (same fixed-point grid) and the FP64 mode on
    the whole matrix, and bitwise run to run."""
    from autojax_amd import api

    d = ds.config1(K=40000)
    args = (d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim)
    W = api.w_tilde_curvature_interferometer_from(*args, precision="int8crt")
    np.testing.assert_array_equal(W, api.w_tilde_curvature_interferometer_from(*args, precision="int8crt"))
    np.testing.assert_array_equal(W, W.T)
    W5 = api.w_tilde_curvature_interferometer_from(*args, precision="int8x5")
    W64 = api.w_tilde_curvature_interferometer_from(*args, precision="fp64")
    W11 = api.w_tilde_curvature_interferometer_from(*args, precision="int8crt11")
    print(f"int8crt vs int8x5 {_rel(W, W5):.3e}, vs fp64 {_rel(W, W64):.3e}; int8crt11 vs fp64 {_rel(W11, W64):.3e}",
          flush=True)
    assert _rel(W, W5) < 5e-12 and _rel(W, W64) < 5e-12 and _rel(W11, W64) < 3e-11
    np.testing.assert_array_equal(W11, W11.T)
    g, n = d.grid_radians_slim, d.noise_map_real
    for i in (0, 1, 255, 256, 511, 715):
        dg = 2 * np.pi * (g[i] - g)
        ph = dg[:, 1:2] * d.uv_wavelengths[None, :, 0] + dg[:, 0:1] * d.uv_wavelengths[None, :, 1]
        want = np.cos(ph) @ (1 / n**2)
        assert np.abs(W[i] - want).max() < 5e-12 * want.max()


def test_w_tilde_crt_variants_agree(gpu):
    """All residue SYRK kernels give bitwise-identical W-tilde for one chunk size (two chunks, the last
    one ragged, M not a multiple of 256); other chunk sizes agree to FP64 rounding (the per-chunk scale
    and the FP64 accumulation order change)."""
    from autojax_amd import _lib, api

    d = ds.config1(K=20000)
    args = (d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim)
    out = []
    try:
        for v in (0, 1, 2, 3, 3):
            _lib.call("aj_w_tilde_crt_configure", v, 16384)
            out.append(api.w_tilde_curvature_interferometer_from(*args, precision="int8crt"))
        for kc in (4096, 32768):
            _lib.call("aj_w_tilde_crt_configure", -1, kc)
            W = api.w_tilde_curvature_interferometer_from(*args, precision="int8crt")
            print(f"chunk {kc}: {_rel(W, out[0]):.3e}", flush=True)
            assert _rel(W, out[0]) < 1e-13
    finally:
        _lib.call("aj_w_tilde_crt_configure", -1, 0)
    for W in out[1:]:
        np.testing.assert_array_equal(W, out[0])


def test_w_tilde_crt_persistent_schedule(gpu):
    """The persistent residue SYRK (variant 3: every CU runs whole units in rounds, then pieces of the
    units left over, summed by a fix-up) equals the one-workgroup-per-unit kernel bitwise, at a size
    with full rounds and split units (M = 2500: 55 tiles, 605 units of 11 moduli) and a ragged last
    chunk, in both residue modes, on two calls in a row."""
    from autojax_amd import _lib, api

    rng = np.random.default_rng(11)
    grid, _ = ds.grid_for_pixel_count(2500, 0.05)
    K = 20000
    uv = ds.uv_loguniform(rng, K, 1e3, 5e5)
    _, noise = ds.visibilities(rng, K, 0.5, 2.0)
    args = (np.ascontiguousarray(noise.real), uv, grid)
    try:
        for precision in ("int8crt11", "int8crt"):
            _lib.call("aj_w_tilde_crt_configure", 0, 16384)
            ref = api.w_tilde_curvature_interferometer_from(*args, precision=precision)
            _lib.call("aj_w_tilde_crt_configure", 3, 16384)
            for _ in range(2):
                W = api.w_tilde_curvature_interferometer_from(*args, precision=precision)
                print(f"{precision}: max |dW| {np.abs(W - ref).max():.3e}", flush=True)
                np.testing.assert_array_equal(W, ref)
    finally:
        _lib.call("aj_w_tilde_crt_configure", -1, 0)


def test_w_tilde_concurrent_host_threads(gpu):
    """Four host threads building W-tilde at the same time (int8crt11, int8crt, fp64, int8x4), each on
    its own stream with its own workspace, two builds each: every build equals the one made alone, bit
    for bit (the persistent SYRK's schedule tables and the other host-side scratch are per thread)."""
    import threading

    import torch

    from autojax_amd import _lib
    from autojax_amd import device as dv
    from autojax_amd import evidence as E

    rng = np.random.default_rng(12)
    grid, _ = ds.grid_for_pixel_count(1500, 0.05)
    K = 20000
    uv = ds.uv_loguniform(rng, K, 1e3, 5e5)
    _, noise = ds.visibilities(rng, K, 0.5, 2.0)
    n_d, uv_d, g_d = dv.to_device(np.ascontiguousarray(noise.real)), dv.to_device(uv), dv.to_device(grid)
    M = grid.shape[0]
    precs = ("int8crt11", "int8crt", "fp64", "int8x4")
    alone = {p: E.w_tilde(n_d, uv_d, g_d, p).cpu().numpy() for p in precs}
    torch.cuda.synchronize()
    out, errs = {}, []

    def work(p):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                ws = dv.workspace(_lib.query("aj_w_tilde_workspace_bytes", K, M, _lib.PRECISIONS[p]))
                Ws = [dv.empty((M, M)) for _ in range(2)]
                for W in Ws:
                    E.w_tilde(n_d, uv_d, g_d, p, out=W, workspace=ws)
                s.synchronize()
                out[p] = [W.cpu().numpy() for W in Ws]
        except Exception as e:  # reported below, from the main thread
            errs.append((p, repr(e)))

    threads = [threading.Thread(target=work, args=(p,)) for p in precs]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errs, errs
    for p in precs:
        for W in out[p]:
            print(f"{p}: max |dW| vs alone {np.abs(W - alone[p]).max():.3e}", flush=True)
            np.testing.assert_array_equal(W, alone[p])


def test_w_tilde_run_to_run_bitwise(gpu, cfg1_small):
    from autojax_amd import api

    d, _ = cfg1_small
    a = api.w_tilde_curvature_interferometer_from(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim, precision="int8x3")
    b = api.w_tilde_curvature_interferometer_from(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim, precision="int8x3")
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", ["int8x3", "int8x4", "int8x5"])
def test_w_tilde_integer_variants_agree(gpu, cfg1_small, precision):
    """Every integer SYRK kernel variant computes the same sums in the same order: bitwise equal W-tilde,
    also when called twice (no stale LDS or register state between stages or calls)."""
    from autojax_amd import _lib, api

    d, Wo = cfg1_small
    args = (d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim)
    ref = None
    try:
        for v in range(_lib.query("aj_w_tilde_variant_count")):
            _lib.call("aj_w_tilde_set_variant", v)
            try:
                W = api.w_tilde_curvature_interferometer_from(*args, precision=precision)
            except ValueError as e:  # a layout that does not fit in LDS for this digit count
                print(f"variant {v}: {e}", flush=True)
                continue
            W2 = api.w_tilde_curvature_interferometer_from(*args, precision=precision)
            print(f"{precision} variant {v}: rel err {_rel(W, Wo):.3e}", flush=True)
            np.testing.assert_array_equal(W, W2)
            assert _rel(W, Wo) < WT_TOL[precision]
            if ref is None:
                ref = W
            np.testing.assert_array_equal(W, ref)
    finally:
        _lib.call("aj_w_tilde_set_variant", -1)


def test_data_side_sums(gpu, cfg1_small):
    from autojax_amd import api

    d, _ = cfg1_small
    y = api.dirty_image_from(d.data, d.noise_map, d.uv_wavelengths, d.grid_radians_slim)
    np.testing.assert_allclose(y, d.dirty_image, rtol=0, atol=1e-11 * np.abs(d.dirty_image).max())
    v = np.ascontiguousarray(d.data.real) + 3.0
    got = api.w_tilde_data_interferometer_from(v, d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim)
    want = R.w_tilde_data_interferometer_from(v, d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-11 * np.abs(want).max())


def test_mapping_data_vector_curvature(gpu, delaunay_small):
    from autojax_amd import api

    d, W = delaunay_small
    smp = d.samples[0]
    S = smp.pixels
    T = api.mapping_matrix_from(smp.pix_indexes_for_sub_slim_index, smp.pix_weights_for_sub_slim_index, S)
    To = R.mapping_matrix_from(smp.pix_indexes_for_sub_slim_index, smp.pix_weights_for_sub_slim_index, S)
    np.testing.assert_array_equal(T, To)
    D = api.data_vector_from(T, d.dirty_image)
    np.testing.assert_allclose(D, R.data_vector_from(To, d.dirty_image), rtol=1e-12, atol=1e-12 * np.abs(D).max())
    Fo = R.curvature_matrix_via_w_tilde_from(W, To)
    F1 = api.curvature_matrix_via_w_tilde_from(W, T)
    F2 = api.curvature_matrix_via_w_wilde_sparse_mapping_matrix_from(W, smp.pix_indexes_for_sub_slim_index,
                                                                      smp.pix_weights_for_sub_slim_index, S)
    print("F dense", _rel(F1, Fo), "F sparse", _rel(F2, Fo), flush=True)
    assert _rel(F1, Fo) < 1e-12 and _rel(F2, Fo) < 1e-12
    X = np.random.default_rng(0).standard_normal((W.shape[0], 17))
    G = api.sparse_mapping_matrix_transpose_matmul(X, smp.pix_indexes_for_sub_slim_index,
                                                   smp.pix_weights_for_sub_slim_index, S)
    Go = R.sparse_mapping_matrix_transpose_matmul(X, smp.pix_indexes_for_sub_slim_index,
                                                  smp.pix_weights_for_sub_slim_index, S)
    assert _rel(G, Go) < 1e-13


def test_regularization_solve_noise(gpu, delaunay_small):
    from autojax_amd import api

    d, W = delaunay_small
    smp = d.samples[1]
    H = api.constant_regularization_matrix_from(1.7, smp.neighbors, smp.neighbors_sizes)
    # equal up to the rounding of a fused multiply-add in 1e-8 + c^2 * size
    np.testing.assert_allclose(H, R.constant_regularization_matrix_from(1.7, smp.neighbors, smp.neighbors_sizes),
                               rtol=2e-16, atol=0)
    To = R.mapping_matrix_from(smp.pix_indexes_for_sub_slim_index, smp.pix_weights_for_sub_slim_index, smp.pixels)
    A = R.curvature_matrix_via_w_tilde_from(W, To) + H
    Dv = R.data_vector_from(To, d.dirty_image)
    s = api.reconstruction_positive_negative_from(Dv, A)
    so = R.reconstruction_positive_negative_from(Dv, A)
    print("solve rel", _rel(s, so), flush=True)
    assert _rel(s, so) < 1e-9
    nn = api.noise_normalization_complex_from(d.noise_map)
    assert abs(nn - R.noise_normalization_complex_from(d.noise_map)) < 1e-12 * abs(nn)
    # negative definite: solved, as np.linalg.solve does (numba.py:731), giving -s
    np.testing.assert_allclose(api.reconstruction_positive_negative_from(Dv, -A), -so, rtol=0,
                               atol=1e-9 * np.abs(so).max())
    singular = A.copy()
    singular[:, 3] = 0.0
    with pytest.raises(np.linalg.LinAlgError):
        api.reconstruction_positive_negative_from(Dv, singular)


def test_log_likelihood_golden_evidence(gpu):
    from autojax_amd import api

    g = np.load(GOLD / "tiny_evidence.npz")
    W = api.w_tilde_curvature_interferometer_from(g["noise_map"].real.copy(), g["uv_wavelengths"], g["grid_radians_slim"])
    got = api.log_likelihood_function_via_w_tilde_from(g["dirty_image"], g["data"], g["noise_map"], W,
                                                       g["pix_indexes_for_sub_slim_index"],
                                                       g["pix_weights_for_sub_slim_index"], g["neighbors"],
                                                       g["neighbors_sizes"])
    want = g["log_evidence"][1]
    print("golden evidence", got, want, flush=True)
    assert abs(got - want) < 3e-8 * abs(want)


def test_log_likelihood_cfg1_vs_oracle(gpu, cfg1_small):
    from autojax_amd import api

    d, W = cfg1_small
    smp = d.samples[0]
    args = (d.dirty_image, d.data, d.noise_map, W, smp.pix_indexes_for_sub_slim_index,
            smp.pix_weights_for_sub_slim_index, smp.neighbors, smp.neighbors_sizes)
    got = api.log_likelihood_function_via_w_tilde_from(*args)
    want = R.log_likelihood_function_via_w_tilde_from(*args)
    print("cfg1 logL", got, want, got - want, flush=True)
    assert abs(got - want) < 1e-10 * abs(want)


def test_log_likelihood_cfg2_vs_oracle(gpu):
    """The plain-method workload at full size (M 7,860, K 1e5, a 1,000-vertex Delaunay source, B = 3):
    the single-sample log-likelihood drop-in (numba.py:755-857, lambda 1) on the device against the
    oracle's (dense T, LU solve, two slogdets) with the same FP64 W and dirty image, within 1e-10
    relative (the workload's bound)."""
    import torch

    from autojax_amd import api
    from autojax_amd import device as dv
    from autojax_amd import evidence as E

    d = ds.config2()
    smp = d.samples[0]
    n_d, uv_d, g_d = dv.to_device(d.noise_map_real), dv.to_device(d.uv_wavelengths), dv.to_device(d.grid_radians_slim)
    W = E.w_tilde(n_d, uv_d, g_d, "fp64")
    y = E.dirty_image(d.data, d.noise_map, d.uv_wavelengths, d.grid_radians_slim)
    torch.cuda.synchronize()
    args = (y.cpu().numpy(), d.data, d.noise_map, W.cpu().numpy(), smp.pix_indexes_for_sub_slim_index,
            smp.pix_weights_for_sub_slim_index, smp.neighbors, smp.neighbors_sizes)
    got = api.log_likelihood_function_via_w_tilde_from(*args)
    want = R.log_likelihood_function_via_w_tilde_from(*args)
    print(f"cfg2 logL {got:.12g} vs oracle {want:.12g} (rel {abs(got - want) / abs(want):.3e}; M {d.M}, "
          f"S {smp.pixels})", flush=True)
    assert abs(got - want) < 1e-10 * abs(want)


def test_batched_log_evidence_full_size_vs_oracle(gpu):
    """The batching workload at full size (M 5,000, K 1e6, 64 samples, S ~ U{300..1500}, lambda in
    [0.01, 100]): the batched log-evidence from the FP64 W within 1e-9 relative, and from the mixed
    path's W (int8crt11) within 1e-3 absolute, of the oracle's on six samples (the smallest lambda, the
    largest and the smallest S, three more), the oracle given the FP64 W and the same dirty image; and the
    default factorisation schedule bitwise equal to schedules 0 and 3 at this size."""
    import torch

    from autojax_amd import device as dv
    from autojax_amd import evidence as E

    d = ds.config5()
    n_d, uv_d, g_d = dv.to_device(d.noise_map_real), dv.to_device(d.uv_wavelengths), dv.to_device(d.grid_radians_slim)
    batch = E.SampleBatch(d.samples, d.M)
    y = E.dirty_image(d.data, d.noise_map, d.uv_wavelengths, d.grid_radians_slim)
    c2nn = E.chi2_noise_norm(d.data, d.noise_map)
    W64 = E.w_tilde(n_d, uv_d, g_d, "fp64")
    out = {"fp64": batch.log_evidence(W64, y, c2nn, check=True).cpu().numpy()}
    # the factorisation schedules at this size (the left-looking one switches from 128-row to 32-row tasks
    # in its last steps) give the same bits as the three-launches-per-step one
    from autojax_amd import _lib

    ref = (out["fp64"].copy(), batch.logdets.cpu().numpy().copy(), batch.recon.cpu().numpy().copy())
    try:
        for mode in (0, 3, 5):
            _lib.call("aj_log_evidence_configure", mode)
            le = batch.log_evidence(W64, y, c2nn, check=True).cpu().numpy()
            for a, b in zip((le, batch.logdets.cpu().numpy(), batch.recon.cpu().numpy()), ref):
                np.testing.assert_array_equal(a, b)
    finally:
        _lib.call("aj_log_evidence_configure", -1)
    Wm = E.w_tilde(n_d, uv_d, g_d, "int8crt11")
    out["int8crt11"] = batch.log_evidence(Wm, y, c2nn, check=True).cpu().numpy()
    torch.cuda.synchronize()
    Wh, yh = W64.cpu().numpy(), y.cpu().numpy()
    del Wm
    lam = np.array([s.coefficient for s in d.samples])
    S = np.array([s.pixels for s in d.samples])
    sel = list(dict.fromkeys([int(lam.argmin()), int(S.argmax()), int(S.argmin()), 7, 30, 55]))
    chi = d.data.view(np.float64) / d.noise_map.view(np.float64)
    nn = R.noise_normalization_complex_from(d.noise_map)
    for b in sel:
        smp = d.samples[b]
        To = R.mapping_matrix_from(smp.pix_indexes_for_sub_slim_index, smp.pix_weights_for_sub_slim_index, smp.pixels)
        F = R.curvature_matrix_via_w_tilde_from(Wh, To)
        D = R.data_vector_from(To, yh)
        H = R.constant_regularization_matrix_from(smp.coefficient, smp.neighbors, smp.neighbors_sizes)
        ldh = ev.logdet_deflated(H, smp.coefficient**2 * smp.neighbors_sizes.mean())
        want = R.log_evidence_from_parts(F, D, chi @ chi, nn, smp.neighbors, smp.neighbors_sizes, smp.coefficient,
                                         logdet_h=ldh)[0]
        print(f"sample {b} S={smp.pixels} lam={smp.coefficient:.3g}: oracle {want:.12g}, fp64 "
              f"{out['fp64'][b] - want:.2e}, int8crt11 {out['int8crt11'][b] - want:.2e}", flush=True)
        assert abs(out["fp64"][b] - want) < 1e-9 * max(1.0, abs(want))
        assert abs(out["int8crt11"][b] - want) < 1e-3


def test_batched_rejects_samples_beyond_the_solve_limit(gpu):
    """A sample of 14,273 source pixels, one above what the solve kernel holds in LDS, is rejected with a
    ValueError naming the limit (the entry point checks it before its first launch)."""
    import torch

    from autojax_amd import evidence as E

    S, M = 14273, 64
    rng = np.random.default_rng(13)
    idx = rng.integers(0, S, (M, 1))
    w = np.ones((M, 1))
    nbr = np.full((S, 2), -1, np.int64)
    nbr[1:, 0] = np.arange(S - 1)
    nbr[:-1, 1] = np.arange(1, S)
    nbr[0] = [1, -1]
    sizes = np.full(S, 2, np.int64)
    sizes[[0, -1]] = 1
    batch = E.SampleBatch([ds.Sample(idx, w, nbr, sizes, 1.0)], M)
    W = torch.eye(M, dtype=torch.float64, device="cuda")
    y = torch.zeros(M, dtype=torch.float64, device="cuda")
    c2nn = torch.zeros(2, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="14272"):
        batch.log_evidence(W, y, c2nn)


def test_batched_log_evidence_vs_oracle(gpu, delaunay_small):
    import torch

    from autojax_amd import evidence as E

    d, W = delaunay_small
    Wd = torch.from_numpy(W).cuda()
    batch = E.SampleBatch(d.samples, d.M)
    c2nn = E.chi2_noise_norm(d.data, d.noise_map)
    y = torch.from_numpy(d.dirty_image).cuda()
    out = batch.log_evidence(Wd, y, c2nn, check=True).cpu().numpy()
    out2 = batch.log_evidence(Wd, y, c2nn, check=True).cpu().numpy()
    np.testing.assert_array_equal(out, out2)
    chi = d.data.view(np.float64) / d.noise_map.view(np.float64)
    nn = R.noise_normalization_complex_from(d.noise_map)
    logdets = batch.logdets.cpu().numpy()
    for b, smp in enumerate(d.samples):
        To = R.mapping_matrix_from(smp.pix_indexes_for_sub_slim_index, smp.pix_weights_for_sub_slim_index, smp.pixels)
        F = R.curvature_matrix_via_w_tilde_from(W, To)
        D = R.data_vector_from(To, d.dirty_image)
        H = R.constant_regularization_matrix_from(smp.coefficient, smp.neighbors, smp.neighbors_sizes)
        ldh = ev.logdet_deflated(H, smp.coefficient**2 * smp.neighbors_sizes.mean())
        want, s, lda, _ = R.log_evidence_from_parts(F, D, chi @ chi, nn, smp.neighbors, smp.neighbors_sizes,
                                                     smp.coefficient, logdet_h=ldh)
        print(f"sample {b} S={smp.pixels} lam={smp.coefficient:.3g}: {out[b]:.12g} vs {want:.12g} "
              f"d={out[b]-want:.2e} ldA {logdets[b,0]-lda:.2e} ldH {logdets[b,1]-ldh:.2e}", flush=True)
        assert abs(out[b] - want) < 1e-9 * max(1.0, abs(want))


@pytest.mark.parametrize("n_pix", [12000, 21000])
def test_batched_log_evidence_large_images_vs_oracle(gpu, n_pix):
    """The batched evaluation where one row of W-tilde fills the LDS (12,000 pixels: one row per Omega
    workgroup) and where it does not fit (21,000: Omega gathered from global memory), three small Delaunay
    samples against the oracle's dense products (numba.py:552-614 and :755-857), within 1e-9 relative."""
    import torch

    from autojax_amd import device as dv
    from autojax_amd import evidence as E

    d = ds.small_delaunay(41, n_pix, 200, (60, 120), n_samples=3, lam_range=(0.1, 10.0))
    d.ensure_dirty_image()
    n_d, uv_d, g_d = dv.to_device(d.noise_map_real), dv.to_device(d.uv_wavelengths), dv.to_device(d.grid_radians_slim)
    Wd = E.w_tilde(n_d, uv_d, g_d, "fp64")
    batch = E.SampleBatch(d.samples, d.M)
    c2nn = E.chi2_noise_norm(d.data, d.noise_map)
    y = torch.from_numpy(d.dirty_image).cuda()
    out = batch.log_evidence(Wd, y, c2nn, check=True).cpu().numpy()
    W = Wd.cpu().numpy()
    del Wd
    chi = d.data.view(np.float64) / d.noise_map.view(np.float64)
    nn = R.noise_normalization_complex_from(d.noise_map)
    for b, smp in enumerate(d.samples):
        To = R.mapping_matrix_from(smp.pix_indexes_for_sub_slim_index, smp.pix_weights_for_sub_slim_index, smp.pixels)
        F = R.curvature_matrix_via_w_tilde_from(W, To)
        D = R.data_vector_from(To, d.dirty_image)
        H = R.constant_regularization_matrix_from(smp.coefficient, smp.neighbors, smp.neighbors_sizes)
        ldh = ev.logdet_deflated(H, smp.coefficient**2 * smp.neighbors_sizes.mean())
        want, _, _, _ = R.log_evidence_from_parts(F, D, chi @ chi, nn, smp.neighbors, smp.neighbors_sizes,
                                                   smp.coefficient, logdet_h=ldh)
        print(f"M {d.M} sample {b} S={smp.pixels}: {out[b]:.12g} vs {want:.12g} d={out[b] - want:.2e}", flush=True)
        assert abs(out[b] - want) < 1e-9 * max(1.0, abs(want))


def test_batched_cholesky_modes_bitwise_equal(gpu, delaunay_small):
    """The six ways the batch factors its matrices (three launches per step; one cooperative launch,
    without and with lookahead; one launch per step with lookahead; left-looking, one launch per step; the
    left-looking tasks in one persistent launch with lookahead, with its default switch to 32-row tasks, with
    256-row tasks in every block column and with 32-row tasks in every block column, twice each) give bitwise
    the same log-evidence,
    log-determinants and reconstructions, also with samples ending inside a step's block (S 97, 31, 300) and
    one whose block columns take several strips (S 700)."""
    import torch

    from autojax_amd import _lib
    from autojax_amd import evidence as E

    d, W = delaunay_small
    Wd = torch.from_numpy(W).cuda()
    y = torch.from_numpy(d.dirty_image).cuda()
    c2nn = E.chi2_noise_norm(d.data, d.noise_map)
    extra = [ds.small_delaunay(80 + i, d.M, 10, S, lam_range=(0.1, 10.0)).samples[0]
             for i, S in enumerate((97, 31, 300, 700))]
    batch = E.SampleBatch(d.samples + extra, d.M)
    out = {}
    runs = [(m, -1) for m in range(6)] + [(5, 0), (5, 1 << 20)]
    try:
        for mode, small in runs:
            _lib.call("aj_log_evidence_configure", mode)
            _lib.call("aj_log_evidence_flow_configure", small, -1)
            for rep in range(2 if mode == 5 else 1):
                le = batch.log_evidence(Wd, y, c2nn, check=True).clone()
                out[(mode, small, rep)] = (le.cpu().numpy(), batch.logdets.cpu().numpy().copy(),
                                           batch.recon.cpu().numpy().copy())
    finally:
        _lib.call("aj_log_evidence_configure", -1)
        _lib.call("aj_log_evidence_flow_configure", -1, -1)
    ref = out[(0, -1, 0)]
    for key, got in out.items():
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b, err_msg=f"schedule {key}")
    print("cholesky modes bitwise equal:", ref[0][:4], flush=True)


def test_persistent_cholesky_wait_limit_reports_failure(gpu, delaunay_small):
    """The persistent schedule's dependency wait is bounded: with a spin limit of zero polls a task that
    finds its dependencies incomplete gives up at once, sets the error flag and the batch raises
    RuntimeError instead of returning numbers; if no task had to wait, the results are those of the
    default schedule.  Every kernel still exits normally (the limit is set from the host)."""
    import torch

    from autojax_amd import _lib
    from autojax_amd import evidence as E

    d, W = delaunay_small
    Wd = torch.from_numpy(W).cuda()
    y = torch.from_numpy(d.dirty_image).cuda()
    c2nn = E.chi2_noise_norm(d.data, d.noise_map)
    extra = [ds.small_delaunay(90 + i, d.M, 10, S, lam_range=(0.1, 10.0)).samples[0] for i, S in enumerate((700, 650))]
    batch = E.SampleBatch(d.samples + extra, d.M)
    ref = batch.log_evidence(Wd, y, c2nn, check=True).cpu().numpy()
    try:
        _lib.call("aj_log_evidence_configure", 5)
        _lib.call("aj_log_evidence_flow_configure", -1, 0)
        try:
            got = batch.log_evidence(Wd, y, c2nn, check=True).cpu().numpy()
        except RuntimeError as e:
            print("wait limit reported:", e, flush=True)
            assert bool((batch.info == -1).all())
        else:
            print("no task had to wait", flush=True)
            np.testing.assert_array_equal(got, ref)
        _lib.call("aj_log_evidence_flow_configure", -1, -1)
        np.testing.assert_array_equal(batch.log_evidence(Wd, y, c2nn, check=True).cpu().numpy(), ref)
    finally:
        _lib.call("aj_log_evidence_configure", -1)
        _lib.call("aj_log_evidence_flow_configure", -1, -1)
