"""GPU gradients of the log-evidence against torch FP64 autograd (CPU, dense restatement) and
central finite differences of the GPU forward."""

import numpy as np
import pytest
import torch

from autojax_amd import datasets as ds
from oracle import evidence as ev
from oracle import restated as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prob():
    d = ds.small_delaunay(41, 400, 600, (30, 60), n_samples=3, lam_range=(0.3, 3.0))
    W = R.w_tilde_curvature_interferometer_from(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim)
    d.ensure_dirty_image()
    chi = d.data.view(np.float64) / d.noise_map.view(np.float64)
    nn = R.noise_normalization_complex_from(d.noise_map)
    return d, W, float(chi @ chi), nn


def _gpu_eval(d, W, coefs, weights=None):
    from autojax_amd import evidence as E
    from autojax_amd import torch_api as TA

    batch = E.SampleBatch(d.samples, d.M)
    Wd = torch.from_numpy(W).cuda()
    y = torch.from_numpy(d.dirty_image).cuda()
    c2nn = E.chi2_noise_norm(d.data, d.noise_map)
    c = torch.tensor(coefs, dtype=torch.float64, requires_grad=True)
    w = torch.from_numpy(np.stack([s.pix_weights_for_sub_slim_index for s in d.samples]) if weights is None else weights)
    w = w.cuda().requires_grad_(True)
    out = TA.log_evidence(c, w, Wd, y, c2nn, batch)
    return out, c, w


def test_grad_matches_torch_autograd(gpu, prob):
    d, W, chi2, nn = prob
    coefs = [s.coefficient for s in d.samples]
    out, c, w = _gpu_eval(d, W, coefs)
    out.sum().backward()
    for b, smp in enumerate(d.samples):
        ct = torch.tensor(smp.coefficient, dtype=torch.float64, requires_grad=True)
        wt = torch.tensor(smp.pix_weights_for_sub_slim_index, requires_grad=True)
        ref = ev.torch_log_evidence(ct, wt, smp.pix_indexes_for_sub_slim_index, W, d.dirty_image, chi2, nn,
                                    smp.neighbors, smp.neighbors_sizes)
        ref.backward()
        print(f"sample {b}: logL {out[b].item():.10f} vs {ref.item():.10f}; dc {c.grad[b].item():.10e} vs "
              f"{ct.grad.item():.10e}", flush=True)
        assert abs(out[b].item() - ref.item()) < 1e-8 * abs(ref.item())
        assert abs(c.grad[b].item() - ct.grad.item()) < 1e-7 * abs(ct.grad.item()) + 1e-9
        gw = w.grad[b].cpu().numpy()
        err = np.abs(gw - wt.grad.numpy()).max() / np.abs(wt.grad.numpy()).max()
        print(f"  dweights rel err {err:.2e}", flush=True)
        assert err < 1e-7


def test_grad_matches_central_fd(gpu, prob):
    d, W, chi2, nn = prob
    coefs = np.array([s.coefficient for s in d.samples])
    out, c, w = _gpu_eval(d, W, coefs)
    out.sum().backward()
    h = 1e-6
    for b in range(len(d.samples)):
        cp, cm = coefs.copy(), coefs.copy()
        cp[b] += h * coefs[b]
        cm[b] -= h * coefs[b]
        fd = (_gpu_eval(d, W, cp)[0][b].item() - _gpu_eval(d, W, cm)[0][b].item()) / (2 * h * coefs[b])
        print(f"sample {b}: dc {c.grad[b].item():.10e} fd {fd:.10e}", flush=True)
        assert abs(c.grad[b].item() - fd) < 1e-5 * abs(fd)
    # weights use h = 1e-5: at 1e-6 the FP64 rounding of logL (~1e-13 relative, |logL| ~ 2e3) would
    # put ~1e-4 relative noise into the difference quotient
    h = 1e-5
    w0 = np.stack([s.pix_weights_for_sub_slim_index for s in d.samples])
    rng = np.random.default_rng(0)
    for _ in range(4):
        b, m, k = rng.integers(len(d.samples)), rng.integers(d.M), rng.integers(3)
        wp, wm = w0.copy(), w0.copy()
        wp[b, m, k] += h
        wm[b, m, k] -= h
        fd = (_gpu_eval(d, W, coefs, wp)[0][b].item() - _gpu_eval(d, W, coefs, wm)[0][b].item()) / (2 * h)
        g = w.grad[b, m, k].item()
        print(f"dT[{b},{m},{k}] {g:.8e} fd {fd:.8e}", flush=True)
        assert abs(g - fd) < 1e-5 * max(abs(fd), 1e-3 * np.abs(w.grad.cpu().numpy()).max())


def test_vertex_gradient(gpu, prob):
    from autojax_amd import torch_api as TA

    d, W, chi2, nn = prob
    smp = d.samples[0]
    tri = smp.pix_indexes_for_sub_slim_index
    pos = TA.traced_positions(smp.mesh.vertices, tri, smp.pix_weights_for_sub_slim_index)
    V = torch.tensor(smp.mesh.vertices, device="cuda", requires_grad=True)
    t = TA.barycentric_weights(V, torch.from_numpy(pos).cuda(), torch.from_numpy(tri).cuda())
    np.testing.assert_allclose(t.detach().cpu().numpy(), smp.pix_weights_for_sub_slim_index, rtol=0, atol=1e-9)
    g = torch.randn(t.shape, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(1))
    (t * g).sum().backward()
    Vt = torch.tensor(smp.mesh.vertices, requires_grad=True)
    tt = ev.torch_barycentric(Vt, pos, tri)
    (tt * g.cpu()).sum().backward()
    err = np.abs(V.grad.cpu().numpy() - Vt.grad.numpy()).max() / np.abs(Vt.grad.numpy()).max()
    print("vertex vjp rel err", err, flush=True)
    assert err < 1e-9


def test_grad_sample_larger_than_image(gpu):
    """A sample with more source pixels than image pixels (M 64, S 200: F has rank <= 64, A = F + H is
    positive definite through H): value and gradients against torch FP64 autograd of the dense form.  The
    backward's scratch holds per-sample vectors of both lengths, so S > M must not write past it."""
    d = ds.small_delaunay(43, 64, 300, 200, n_samples=1, lam_range=(0.5, 2.0))
    W = R.w_tilde_curvature_interferometer_from(d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim)
    d.ensure_dirty_image()
    chi = d.data.view(np.float64) / d.noise_map.view(np.float64)
    nn = R.noise_normalization_complex_from(d.noise_map)
    smp = d.samples[0]
    assert smp.pixels > d.M
    out, c, w = _gpu_eval(d, W, [smp.coefficient])
    out.sum().backward()
    ct = torch.tensor(smp.coefficient, dtype=torch.float64, requires_grad=True)
    wt = torch.tensor(smp.pix_weights_for_sub_slim_index, requires_grad=True)
    ref = ev.torch_log_evidence(ct, wt, smp.pix_indexes_for_sub_slim_index, W, d.dirty_image, float(chi @ chi), nn,
                                smp.neighbors, smp.neighbors_sizes)
    ref.backward()
    err_w = np.abs(w.grad[0].cpu().numpy() - wt.grad.numpy()).max() / np.abs(wt.grad.numpy()).max()
    print(f"M {d.M} S {smp.pixels}: logL {out[0].item():.10f} vs {ref.item():.10f}; dc {c.grad[0].item():.8e} vs "
          f"{ct.grad.item():.8e}; dweights {err_w:.2e}", flush=True)
    assert abs(out[0].item() - ref.item()) < 1e-8 * abs(ref.item())
    assert abs(c.grad[0].item() - ct.grad.item()) < 1e-7 * abs(ct.grad.item()) + 1e-9
    assert err_w < 1e-7


def test_backward_after_a_second_evaluation_raises(gpu, prob):
    """Two forward passes on one SampleBatch, then backward of the first: its factors were overwritten by the
    second, so backward raises instead of returning the second evaluation's gradients."""
    from autojax_amd import evidence as E
    from autojax_amd import torch_api as TA

    d, W, chi2, nn = prob
    batch = E.SampleBatch(d.samples, d.M)
    Wd = torch.from_numpy(W).cuda()
    y = torch.from_numpy(d.dirty_image).cuda()
    c2nn = E.chi2_noise_norm(d.data, d.noise_map)
    w = torch.from_numpy(np.stack([s.pix_weights_for_sub_slim_index for s in d.samples])).cuda().requires_grad_(True)
    c1 = torch.tensor([s.coefficient for s in d.samples], dtype=torch.float64, requires_grad=True)
    c2 = (2.0 * c1).detach().requires_grad_(True)
    first = TA.log_evidence(c1, w, Wd, y, c2nn, batch)
    second = TA.log_evidence(c2, w, Wd, y, c2nn, batch)
    with pytest.raises(RuntimeError, match="evaluated again"):
        first.sum().backward(retain_graph=True)
    second.sum().backward()
    assert torch.isfinite(c2.grad).all()
