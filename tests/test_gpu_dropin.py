"""Drop-in fidelity on the GPU: the inputs the reference accepts besides symmetric positive definite ones.

* reconstruction_positive_negative_from solves as np.linalg.solve (numba.py:731): LU with partial pivoting
  of the whole matrix, so symmetric indefinite and non-symmetric nonsingular matrices are solved; an
  exactly singular one raises LinAlgError.  Sizes around the 32-column panel edge.
* log_likelihood_function_via_w_tilde_from (numba.py:755-857) with one-sided neighbour lists (H not
  symmetric), with rows that name a source pixel twice (the dense mapping assigns, numba.py:375), and with
  a curvature matrix that makes A indefinite (slogdet's log|det A|); the compact form (numba.py:860-967)
  with repeated indices (its F sums them, its D assigns them).  Each against the oracle, which follows the
  reference's steps (dense mapping, np.linalg.solve, np.linalg.slogdet).
"""

import numpy as np
import pytest

from autojax_amd import datasets as ds
from oracle import restated as R

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("n", [1, 2, 31, 32, 33, 65, 257, 1000])
def test_solve_general_matrices_as_numpy(gpu, n):
    from autojax_amd import api
    from autojax_amd.api import _dev, _lu_solve

    rng = np.random.default_rng(n)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = rng.uniform(0.5, 2.0, n) * rng.choice([-1.0, 1.0], n)
    cases = {
        "spd": (Q * np.abs(ev)) @ Q.T,
        "symmetric indefinite": (Q * ev) @ Q.T,
        "general": rng.standard_normal((n, n)) + 0.1 * n * np.eye(n) * (n % 2),
    }
    b = rng.standard_normal(n)
    for name, A in cases.items():
        x = api.reconstruction_positive_negative_from(b, A)
        want = np.linalg.solve(A, b)
        err = _rel(x, want)
        _, ld, info = _lu_solve(_dev(A, np.float64, "A", 2))
        sign, want_ld = np.linalg.slogdet(A)
        print(f"n={n} {name}: solve rel {err:.2e}, log|det| {ld.item():.12g} vs {want_ld:.12g}", flush=True)
        assert int(info.item()) == 0
        assert err < 1e-10 * max(1.0, np.linalg.cond(A) / 1e3)
        assert abs(ld.item() - want_ld) < 1e-10 * max(1.0, abs(want_ld))


def test_solve_singular_raises(gpu):
    from autojax_amd import api

    rng = np.random.default_rng(5)
    A = rng.standard_normal((40, 40))
    A[17] = 0.0  # a zero row: an exact zero pivot however the rows are swapped
    with pytest.raises(np.linalg.LinAlgError, match="Singular"):
        api.reconstruction_positive_negative_from(rng.standard_normal(40), A)


@pytest.fixture(scope="module")
def problem():
    d = ds.small_delaunay(71, 300, 400, 80, n_samples=1)
    W = R.w_tilde_curvature_interferometer_from(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim)
    d.ensure_dirty_image()
    return d, W


def _args(d, W, idx, w, nb, sz):
    return (d.dirty_image, d.data, d.noise_map, W, idx, w, nb, sz)


def test_log_likelihood_one_sided_neighbours(gpu, problem):
    from autojax_amd import api

    d, W = problem
    s = d.samples[0]
    nb = s.neighbors.copy()
    sz = s.neighbors_sizes.copy()
    # drop the second entry of every third row: those pixels no longer list a neighbour that lists them
    for i in range(0, s.pixels, 3):
        if sz[i] >= 2:
            nb[i, 1:-1] = nb[i, 2:]
            nb[i, -1] = -1
            sz[i] -= 1
    assert not api.neighbors_symmetric(nb, sz)
    args = _args(d, W, s.pix_indexes_for_sub_slim_index, s.pix_weights_for_sub_slim_index, nb, sz)
    got = api.log_likelihood_function_via_w_tilde_from(*args)
    want = R.log_likelihood_function_via_w_tilde_from(*args)
    print(f"one-sided neighbours: {got:.12f} vs {want:.12f}", flush=True)
    assert abs(got - want) < 1e-10 * abs(want)


def test_log_likelihood_repeated_indices(gpu, problem):
    from autojax_amd import api

    d, W = problem
    s = d.samples[0]
    idx = s.pix_indexes_for_sub_slim_index.copy()
    idx[::4, 2] = idx[::4, 0]  # every fourth row names its first source pixel again
    args = _args(d, W, idx, s.pix_weights_for_sub_slim_index, s.neighbors, s.neighbors_sizes)
    got = api.log_likelihood_function_via_w_tilde_from(*args)
    want = R.log_likelihood_function_via_w_tilde_from(*args)
    plain = R.log_likelihood_function_via_w_tilde_from(*_args(d, W, s.pix_indexes_for_sub_slim_index,
                                                              s.pix_weights_for_sub_slim_index, s.neighbors,
                                                              s.neighbors_sizes))
    print(f"repeated indices: {got:.12f} vs {want:.12f} (without repeats {plain:.6f})", flush=True)
    assert abs(got - want) < 1e-10 * abs(want)
    assert abs(want - plain) > 1e-6 * abs(plain)


def test_log_likelihood_indefinite_curvature(gpu, problem):
    """A W-tilde of the wrong sign makes A = F + H indefinite but nonsingular: the reference still returns
    -1/2 (chi^2 - s.D + log|det A| - log|det H| + NN) from its LU solve and slogdet."""
    from autojax_amd import api

    d, W = problem
    s = d.samples[0]
    args = _args(d, -W, s.pix_indexes_for_sub_slim_index, s.pix_weights_for_sub_slim_index, s.neighbors,
                 s.neighbors_sizes)
    got = api.log_likelihood_function_via_w_tilde_from(*args)
    want = R.log_likelihood_function_via_w_tilde_from(*args)
    print(f"indefinite A: {got:.12f} vs {want:.12f}", flush=True)
    assert abs(got - want) < 1e-9 * abs(want)


def test_compact_log_likelihood_repeated_indices(gpu):
    from autojax_amd import api

    d = ds.reference_convention_dataset(seed=72, K=500, S=60)
    N = d.n_side
    wc = R.w_compact_curvature_interferometer_from(N, d.noise_map_real, d.uv_wavelengths, d.pixel_scale)
    d.ensure_dirty_image()
    s = d.samples[0]
    idx = s.pix_indexes_for_sub_slim_index.copy()
    idx[1::3, 1] = idx[1::3, 0]
    args = (d.dirty_image, d.data, d.noise_map, wc, d.native_index_for_slim_index, idx,
            s.pix_weights_for_sub_slim_index, s.neighbors, s.neighbors_sizes)
    got = api.log_likelihood_function_via_w_compact_from(*args)
    want = R.log_likelihood_function_via_w_compact_from(*args)
    print(f"compact, repeated indices: {got:.12f} vs {want:.12f}", flush=True)
    assert abs(got - want) < 1e-10 * abs(want)
