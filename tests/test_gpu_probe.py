"""The register-only MFMA probe (aj_mfma_peak_probe) is exact on the device: with small integer
operands every product and sum is exact, so n iterations give exactly n times one iteration."""

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", [0, 1])
def test_mfma_probe_linear_in_iterations(gpu, kind):
    """int8 (16x16x64, 32 accumulators per wave) and FP64 (16x16x4, 16 accumulators per wave): 37 and
    1,000 iterations against one, on launches of 3 and 8 workgroups; the one-iteration values are not
    all zero."""
    import torch

    from autojax_amd import _lib
    from autojax_amd import device as dv

    dtype = torch.int64 if kind == 0 else torch.float64
    for wg, iters in ((3, 37), (8, 1000)):
        one = torch.full((wg * 512,), -1, dtype=dtype, device="cuda")
        out = torch.full((wg * 512,), -1, dtype=dtype, device="cuda")
        _lib.call("aj_mfma_peak_probe", kind, wg, 1, one.data_ptr(), dv.stream_ptr())
        _lib.call("aj_mfma_peak_probe", kind, wg, iters, out.data_ptr(), dv.stream_ptr())
        torch.cuda.synchronize()
        print(f"kind {kind}: workgroups {wg} iters {iters}: one[0] {one[0].item()} out[0] {out[0].item()}", flush=True)
        assert bool((one != 0).any().item())
        assert bool((out == one * iters).all().item())
