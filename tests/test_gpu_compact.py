"""GPU parity of the lattice (w_compact) path, numba.py:210-255, :617-647, :860-967, and of the preload
form, original.py:230-403, :600-700."""

import numpy as np
import pytest
import torch

from autojax_amd import datasets as ds
from oracle import restated as R

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def lattice():
    """30 x 30 grid at 0.1", radius 1.5" (716 pixels), 2,000 visibilities, one 300-vertex Delaunay source."""
    rng = np.random.default_rng(21)
    N, ps = 30, 0.1
    grid, native = ds.circular_grid(N, ps, 1.5)
    K = 2000
    uv = ds.uv_loguniform(rng, K, 1e3, 5e5)
    data, noise = ds.visibilities(rng, K, 0.5, 2.0)
    mesh = ds.delaunay_source(rng, 300)
    idx, w = ds.barycentric_mapping(rng, grid.shape[0], mesh.triangles)
    wc = R.w_compact_curvature_interferometer_from(N, noise.real.copy(), uv, ps)
    y = ds.dirty_image_numpy(data, noise, uv, grid)
    return dict(N=N, ps=ps, grid=grid, native=native, uv=uv, data=data, noise=noise, mesh=mesh, idx=idx, w=w, wc=wc,
                y=y)


def test_w_compact_vs_oracle(gpu, lattice):
    from autojax_amd import api

    L = lattice
    got = api.w_compact_curvature_interferometer_from(L["N"], L["noise"].real.copy(), L["uv"], L["ps"])
    err = _rel(got, L["wc"])
    print(f"w_compact rel err {err:.3e}", flush=True)
    assert got.shape == (L["N"], 2 * L["N"] - 1)
    assert err < 1e-13


def test_w_compact_split_visibilities_and_tail(gpu):
    """K spread over many visibility slices with a ragged last chunk; larger grid (several tiles)."""
    from autojax_amd import api

    rng = np.random.default_rng(22)
    N, ps, K = 100, 0.05, 5003
    uv = ds.uv_loguniform(rng, K, 1e3, 5e5)
    n = rng.uniform(0.5, 2.0, K)
    got = api.w_compact_curvature_interferometer_from(N, n, uv, ps)
    want = R.w_compact_curvature_interferometer_from(N, n, uv, ps)
    err = _rel(got, want)
    print(f"w_compact N=100 K=5003 rel err {err:.3e}", flush=True)
    assert err < 1e-13
    again = api.w_compact_curvature_interferometer_from(N, n, uv, ps)
    np.testing.assert_array_equal(got, again)


def test_w_tilde_via_compact_is_the_lattice_gather(gpu, lattice):
    from autojax_amd import api

    L = lattice
    got = api.w_tilde_via_compact_from(L["wc"], L["native"])
    np.testing.assert_array_equal(got, R.w_tilde_via_compact_from(L["wc"], L["native"]))


def test_w_tilde_via_preload_is_the_wrapped_gather(gpu, lattice):
    from autojax_amd import api

    L = lattice
    N = L["N"]
    pre = np.random.default_rng(3).standard_normal((2 * N, 2 * N))
    got = api.w_tilde_via_preload_from(pre, L["native"])
    np.testing.assert_array_equal(got, R.w_tilde_via_preload_from(pre, L["native"]))
    # every pair read from the table (numba.py:282-291): a table that is not point-symmetric gives a W that is
    # not symmetric
    assert not np.array_equal(got, got.T)


@pytest.fixture(scope="module")
def refgrid():
    """The reference harness's grid (30 x 30 at 0.2", corner 9.9", radius 3": 716 pixels), 3,000
    visibilities, a 120-vertex Delaunay source with rows cut short by pix_size."""
    d = ds.reference_convention_dataset(seed=31, K=3000, S=120)
    N = d.n_side
    pre = R.w_tilde_curvature_preload_interferometer_from(d.noise_map_real, d.uv_wavelengths, (N, N),
                                                          d.grid_radians_2d)
    size = np.full(d.M, 3, dtype=np.int64)
    size[::5] = 2
    size[::13] = 1
    return d, pre, size


def test_preload_table_vs_oracle(gpu, refgrid):
    """original.py:230-403 on the device: the 2N x 2N table against the oracle (FP64, norm-wise 1e-10),
    a sub-corner of the grid (Y, X) = (17, 23), and no visibilities (all zero)."""
    from autojax_amd import api

    d, pre, _ = refgrid
    N = d.n_side
    got = api.w_tilde_curvature_preload_interferometer_from(d.noise_map_real, d.uv_wavelengths, (N, N),
                                                            d.grid_radians_2d)
    err = _rel(got, pre)
    print(f"preload table rel err {err:.3e}", flush=True)
    assert got.shape == (2 * N, 2 * N) and err < 1e-10
    assert not got[N].any() and not got[:, N].any()
    sub = api.w_tilde_curvature_preload_interferometer_from(d.noise_map_real, d.uv_wavelengths, (17, 23),
                                                            d.grid_radians_2d)
    want = R.w_tilde_curvature_preload_interferometer_from(d.noise_map_real, d.uv_wavelengths, (17, 23),
                                                           d.grid_radians_2d)
    assert _rel(sub, want) < 1e-10
    empty = api.w_tilde_curvature_preload_interferometer_from(np.zeros(0), np.zeros((0, 2)), (N, N),
                                                              d.grid_radians_2d)
    assert empty.shape == (2 * N, 2 * N) and not empty.any()


def test_preload_curvature_vs_oracle_and_expanded(gpu, refgrid):
    """original.py:600-700 on the device (W read from the table, rows cut by pix_size): against the
    oracle's loop form to 1e-10 norm-wise, and within 1e-13 of the sparse product over the W-tilde that
    w_tilde_via_preload_from expands (numba.py:282-291 reads each pair's own offset, original.py:600-700
    the pairs ip0 < ip1 only: the reference-built table is point-symmetric up to rounding)."""
    from autojax_amd import api

    d, pre, size = refgrid
    s = d.samples[0]
    args = (pre, s.pix_indexes_for_sub_slim_index, size, s.pix_weights_for_sub_slim_index,
            d.native_index_for_slim_index, s.pixels)
    got = api.curvature_matrix_via_w_tilde_curvature_preload_interferometer_from(*args)
    want = R.curvature_matrix_via_w_tilde_curvature_preload_interferometer_from(*args)
    err = _rel(got, want)
    print(f"F via preload rel err {err:.3e}", flush=True)
    assert err < 1e-10
    idx = s.pix_indexes_for_sub_slim_index.copy()
    idx[np.arange(3)[None, :] >= size[:, None]] = -1
    W = api.w_tilde_via_preload_from(pre, d.native_index_for_slim_index)
    F2 = api.curvature_matrix_via_w_wilde_sparse_mapping_matrix_from(W, idx, s.pix_weights_for_sub_slim_index, s.pixels)
    np.testing.assert_allclose(got, F2, rtol=0, atol=1e-13 * np.abs(F2).max())


def test_compact_w_tilde_equals_feature_syrk(gpu, lattice):
    """Two independent device routes to the same W: lattice table + gather vs the FP64 feature SYRK."""
    from autojax_amd import api

    L = lattice
    wc = api.w_compact_curvature_interferometer_from(L["N"], L["noise"].real.copy(), L["uv"], L["ps"])
    Wc = api.w_tilde_via_compact_from(wc, L["native"])
    Ws = api.w_tilde_curvature_interferometer_from(L["noise"].real.copy(), L["uv"], L["grid"])
    err = _rel(Wc, Ws)
    print(f"compact vs SYRK rel err {err:.3e}", flush=True)
    assert err < 1e-12


def test_curvature_and_likelihood_via_compact(gpu, lattice):
    from autojax_amd import api

    L = lattice
    mesh = L["mesh"]
    S = mesh.neighbors_sizes.size
    F = api.curvature_matrix_via_w_compact_sparse_mapping_matrix_from(L["wc"], L["native"], L["idx"], L["w"], S)
    W = R.w_tilde_via_compact_from(L["wc"], L["native"])
    T = R.mapping_matrix_from(L["idx"], L["w"], S)
    Fo = T.T @ W @ T
    print(f"F via compact rel err {_rel(F, Fo):.3e}", flush=True)
    assert _rel(F, Fo) < 1e-12
    args = (L["y"], L["data"], L["noise"], L["wc"], L["native"], L["idx"], L["w"], mesh.neighbors,
            mesh.neighbors_sizes)
    got = api.log_likelihood_function_via_w_compact_from(*args)
    want = R.log_likelihood_function_via_w_tilde_from(L["y"], L["data"], L["noise"], W, L["idx"], L["w"],
                                                       mesh.neighbors, mesh.neighbors_sizes)
    print(f"logL via compact {got:.12f} vs oracle {want:.12f}", flush=True)
    assert abs(got - want) < 1e-10 * abs(want)


def test_batched_compact_equals_batched_dense(gpu):
    from autojax_amd import api
    from autojax_amd import device as dv
    from autojax_amd import evidence as E

    d = ds.small_delaunay(23, 900, 3000, (60, 200), n_samples=5, lam_range=(0.1, 10.0))
    N = d.n_side
    wc = api.w_compact_curvature_interferometer_from(N, dv.to_device(d.noise_map_real), dv.to_device(d.uv_wavelengths),
                                                     d.pixel_scale)
    native = dv.to_device(np.ascontiguousarray(d.native_index_for_slim_index.astype(np.int64)))
    W = api.w_tilde_via_compact_from(wc, native)
    y = E.dirty_image(d.data, d.noise_map, d.uv_wavelengths, d.grid_radians_slim)
    c2nn = E.chi2_noise_norm(d.data, d.noise_map)
    batch = E.SampleBatch(d.samples, d.M)
    dense = batch.log_evidence(W, y, c2nn, check=True).cpu().numpy()
    comp = batch.log_evidence_compact(wc, native, y, c2nn, check=True).cpu().numpy()
    print("dense", dense, "compact", comp, flush=True)
    np.testing.assert_array_equal(comp, dense)
