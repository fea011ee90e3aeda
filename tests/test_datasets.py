"""The reference's dataset file format (tests.py:396-452): write, read back, reject a wrong grid.  CPU."""

import numpy as np
import pytest

from autojax_amd import datasets as ds


def test_reference_grid_and_mask_sizes():
    # tests.py:26, :158, :210-212: 30 x 30 at 0.2", radius 3.0" -> the DataLoaded pixel count 716
    # (numba.py:811-816)
    mask = ds.reference_mask(30, 0.2, 3.0)
    assert int((~mask).sum()) == 716
    g = ds.reference_grid_radians_2d(30, 0.2)
    d = 0.2 * ds.ARCSEC
    assert g[0, 0, 0] == pytest.approx(9.9 * ds.ARCSEC) and g[0, 0, 1] == pytest.approx(-9.9 * ds.ARCSEC)
    np.testing.assert_allclose(g[1, 0, 0] - g[0, 0, 0], -d, rtol=1e-12)
    np.testing.assert_allclose(g[0, 1, 1] - g[0, 0, 1], d, rtol=1e-12)


def test_npz_round_trip(tmp_path):
    d = ds.reference_convention_dataset(seed=4, K=64, S=24)
    f = tmp_path / "data.npz"
    ds.save_npz(d, f)
    with np.load(f, allow_pickle=False) as z:
        assert sorted(z.files) == sorted(ds.NPZ_KEYS)
    e = ds.load_npz(f)
    for name in ("grid_radians_slim", "native_index_for_slim_index", "uv_wavelengths", "data", "noise_map",
                 "dirty_image", "grid_radians_2d"):
        np.testing.assert_array_equal(getattr(e, name), getattr(d, name))
    for name in ("pix_indexes_for_sub_slim_index", "pix_weights_for_sub_slim_index", "neighbors", "neighbors_sizes"):
        np.testing.assert_array_equal(getattr(e.samples[0], name), getattr(d.samples[0], name))


def test_npz_rejects_missing_key_and_wrong_grid(tmp_path):
    d = ds.reference_convention_dataset(seed=5, K=16, S=12)
    f = tmp_path / "data.npz"
    ds.save_npz(d, f)
    with pytest.raises(ValueError, match="image pixels"):
        ds.load_npz(f, grid_size=20)
    g = tmp_path / "partial.npz"
    with np.load(f, allow_pickle=False) as z:
        np.savez(g, **{k: z[k] for k in z.files if k != "neighbors"})
    with pytest.raises(ValueError, match="missing"):
        ds.load_npz(g)
