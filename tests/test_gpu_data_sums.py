"""GPU parity of W-tilde built together with the data-side sums (aj_w_tilde_with_data_interferometer_from):
the W-tilde bits equal to the plain build's at the same precision, and the dirty image and w_tilde_data
against the CPU sums, at visibility counts that end inside a 512-visibility block and past the chunk
boundaries of both feature paths; optional outputs, K = 0, and the full-size workload against the
stand-alone GPU sums (whose phases are evaluated on their own, absolute rather than centre-relative)."""

import numpy as np
import pytest

from autojax_amd import datasets as ds
from oracle import restated as R

pytestmark = pytest.mark.gpu

# normwise tolerance of the sums, relative to their largest entry (as the stand-alone sums' test)
TOL = 1e-11


def _case(K, seed=0):
    d = ds.config1(seed=seed, K=K)
    np.testing.assert_array_equal(d.noise_map_real, d.noise_map.real)
    v = np.ascontiguousarray(d.data.real) + 3.0  # real visibilities away from zero: the weight divides by them
    return d, v


def _close(got, want, what):
    err = float(np.abs(got - want).max() / np.abs(want).max())
    print(f"{what}: normwise rel err {err:.3e}", flush=True)
    assert err < TOL


# K = 300: one partial 512-visibility block; 4,097: one past the FP64 path's 4,096-visibility chunk;
# 16,900: past the residue path's 16,384-visibility chunk, ending mid-block
@pytest.mark.parametrize("K", [300, 4097, 16900])
@pytest.mark.parametrize("precision", ["fp64", "int8crt11", "int8crt", "fp32"])
def test_fused_data_sums_vs_cpu(gpu, K, precision):
    from autojax_amd import api

    d, v = _case(K)
    args = (d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim)
    print(f"K={K} M={d.M} precision={precision}", flush=True)
    W, y, z = api.w_tilde_with_data_interferometer_from(*args, data=d.data, visibilities_real=v, precision=precision)
    np.testing.assert_array_equal(W, api.w_tilde_curvature_interferometer_from(*args, precision=precision))
    _close(y, ds.dirty_image_numpy(d.data, d.noise_map, d.uv_wavelengths, d.grid_radians_slim), "dirty image")
    _close(z, R.w_tilde_data_interferometer_from(v, *args), "w_tilde_data")


@pytest.mark.parametrize("precision", ["fp64", "int8crt11"])
def test_fused_optional_outputs(gpu, precision):
    """Each output alone has the bits it has when both are requested (the other's coefficients are zero
    but the sums are separate); with neither, W-tilde alone."""
    from autojax_amd import api

    d, v = _case(5000, seed=3)
    args = (d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim)
    W, y, z = api.w_tilde_with_data_interferometer_from(*args, data=d.data, visibilities_real=v, precision=precision)
    W1, y1, z1 = api.w_tilde_with_data_interferometer_from(*args, data=d.data, precision=precision)
    assert z1 is None
    np.testing.assert_array_equal(y1, y)
    np.testing.assert_array_equal(W1, W)
    W2, y2, z2 = api.w_tilde_with_data_interferometer_from(*args, visibilities_real=v, precision=precision)
    assert y2 is None
    np.testing.assert_array_equal(z2, z)
    W3, y3, z3 = api.w_tilde_with_data_interferometer_from(*args, precision=precision)
    assert y3 is None and z3 is None
    np.testing.assert_array_equal(W3, W)


@pytest.mark.parametrize("precision", ["fp64", "int8crt11"])
def test_fused_no_visibilities(gpu, precision):
    from autojax_amd import api

    g = ds.config1(K=10).grid_radians_slim
    e = np.zeros(0)
    W, y, z = api.w_tilde_with_data_interferometer_from(e, np.zeros((0, 2)), g, data=np.zeros(0, np.complex128),
                                                        visibilities_real=e, precision=precision)
    assert not W.any() and not y.any() and not z.any()
    assert W.shape == (g.shape[0], g.shape[0]) and y.shape == z.shape == (g.shape[0],)


def test_fused_rejects_mismatched_data(gpu):
    from autojax_amd import api

    d, v = _case(300)
    with pytest.raises(ValueError):
        api.w_tilde_with_data_interferometer_from(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim,
                                                  data=d.data[:-1])
    with pytest.raises(ValueError):
        api.w_tilde_with_data_interferometer_from(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim,
                                                  visibilities_real=v[:-1])


def test_fused_full_size(gpu):
    """The batching workload's sizes (M 5,000, K 1e6): the fused sums against the stand-alone GPU sums,
    and W-tilde bitwise equal to the plain build, for the FP64 and the mixed (int8crt11) path."""
    import torch

    from autojax_amd import api
    from autojax_amd import device as dv

    d = ds.config5(n_samples=1)
    v = np.ascontiguousarray(d.data.real) + 3.0
    n_d, uv_d, g_d = dv.to_device(d.noise_map_real), dv.to_device(d.uv_wavelengths), dv.to_device(d.grid_radians_slim)
    data_d, v_d = dv.to_device(d.data), dv.to_device(v)
    y0 = api.dirty_image_from(data_d, dv.to_device(d.noise_map), uv_d, g_d).cpu().numpy()
    z0 = api.w_tilde_data_interferometer_from(v_d, n_d, uv_d, g_d).cpu().numpy()
    for precision in ("int8crt11", "fp64"):
        print(f"full size {precision}", flush=True)
        W, y, z = api.w_tilde_with_data_interferometer_from(n_d, uv_d, g_d, data=data_d, visibilities_real=v_d,
                                                            precision=precision)
        W0 = api.w_tilde_curvature_interferometer_from(n_d, uv_d, g_d, precision=precision)
        assert torch.equal(W, W0)
        del W, W0
        _close(y.cpu().numpy(), y0, f"{precision} dirty image")
        _close(z.cpu().numpy(), z0, f"{precision} w_tilde_data")
