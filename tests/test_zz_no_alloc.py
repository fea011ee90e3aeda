"""The allocation-free solve, aj_reconstruction_positive_negative_from_noalloc, through the C-ABI: on two calls
in a row (the second smaller, so it runs on scratch left by the first) it gives bitwise the results of
aj_reconstruction_positive_negative_from with a fresh caller workspace (the same kernels), which match
np.linalg.solve / slogdet; the largest size it takes (S = 2048) works and S = 2049 is rejected before any
GPU work."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

S_MAX = 2048


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + np.sqrt(n) * np.eye(n)  # general, nonsymmetric, well conditioned
    return A, rng.standard_normal(n)


def _solve(A, b, noalloc):
    from autojax_amd import _lib
    from autojax_amd import device as dv

    n = b.size
    Ad, bd = dv.to_device(A), dv.to_device(b)
    x = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    ld = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    info = torch.full((1,), -7, dtype=torch.int64, device="cuda")
    if noalloc:
        _lib.call("aj_reconstruction_positive_negative_from_noalloc", bd.data_ptr(), Ad.data_ptr(), n, x.data_ptr(),
                  ld.data_ptr(), info.data_ptr(), dv.stream_ptr())
    else:
        ws = torch.zeros(_lib.query("aj_solve_workspace_bytes", n), dtype=torch.uint8, device="cuda")
        _lib.call("aj_reconstruction_positive_negative_from", bd.data_ptr(), Ad.data_ptr(), n, x.data_ptr(),
                  ld.data_ptr(), info.data_ptr(), ws.data_ptr(), ws.numel(), dv.stream_ptr())
    torch.cuda.synchronize()
    return x.cpu().numpy(), float(ld.item()), int(info.item())


def test_noalloc_solve_matches_workspace_entry(gpu):
    for n, seed in ((1500, 1), (700, 2), (S_MAX, 3)):
        A, b = _inputs(n, seed)
        x0, ld0, i0 = _solve(A, b, noalloc=False)
        x1, ld1, i1 = _solve(A, b, noalloc=True)
        want = np.linalg.solve(A, b)
        err = float(np.abs(x1 - want).max() / np.abs(want).max())
        print(f"S={n}: noalloc vs workspace entry bitwise {np.array_equal(x0, x1) and ld0 == ld1}, vs numpy {err:.2e}",
              flush=True)
        assert i0 == 0 and i1 == 0
        np.testing.assert_array_equal(x1, x0)
        assert ld1 == ld0
        assert err < 1e-12
        assert abs(ld1 - np.linalg.slogdet(A)[1]) < 1e-10 * abs(ld1)


def test_noalloc_solve_rejects_sizes_beyond_its_scratch(gpu):
    from autojax_amd import _lib

    with pytest.raises(ValueError, match="2048"):
        _lib.call("aj_reconstruction_positive_negative_from_noalloc", 1, 1, S_MAX + 1, 1, None, 1, None)
