"""Gradients of the log-evidence at the full differentiation workload (M 7,860 image pixels,
K 1e5 visibilities, one 1,000-vertex Delaunay source, lambda log-uniform in [0.01, 100]) against
central finite differences of the GPU forward.

Criterion, per parameter group: max_i |fd_i - g_i| / max_i |g_i| <= 1e-5 over the checked entries
(the entries of largest |g| plus random ones).  Steps: lambda h = 1e-6 lambda; vertices
h = 1e-7 x the vertex coordinate scale; mapping weights h = 1e-5, because |logL| ~ 3e5 carries
~1e-14 relative rounding noise, which at h = 1e-6 alone is ~5e-3 absolute in the quotient.
"""

import numpy as np
import pytest
import torch

from autojax_amd import datasets as dsm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full():
    from autojax_amd import device as dv
    from autojax_amd import evidence as E
    from autojax_amd import torch_api as TA

    d = dsm.config4()
    smp = d.samples[0]
    noise, uv, grid = (dv.to_device(x) for x in (d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim))
    W = E.w_tilde(noise, uv, grid, "fp64")
    y = E.dirty_image(d.data, d.noise_map, d.uv_wavelengths, d.grid_radians_slim)
    c2nn = E.chi2_noise_norm(d.data, d.noise_map)
    batch = E.SampleBatch([smp], d.M)
    tri = torch.from_numpy(smp.pix_indexes_for_sub_slim_index).cuda()
    pos = torch.from_numpy(TA.traced_positions(smp.mesh.vertices, smp.pix_indexes_for_sub_slim_index,
                                               smp.pix_weights_for_sub_slim_index)).cuda()
    print(f"cfg4: M={d.M} K={d.K} S={smp.pixels} lambda={smp.coefficient:.4g}", flush=True)

    def f(lam, V=None, w=None):
        if w is None:
            w = TA.barycentric_weights(torch.from_numpy(smp.mesh.vertices).cuda() if V is None else V, pos, tri)
        c = torch.tensor([lam], dtype=torch.float64)
        return TA.log_evidence(c, w.reshape(1, d.M, 3), W, y, c2nn, batch)[0].item()

    V = torch.tensor(smp.mesh.vertices, device="cuda", requires_grad=True)
    c = torch.tensor([smp.coefficient], dtype=torch.float64, requires_grad=True)
    w = TA.barycentric_weights(V, pos, tri)
    w.retain_grad()
    out = TA.log_evidence(c, w.reshape(1, d.M, 3), W, y, c2nn, batch)
    out.sum().backward()
    return d, smp, f, c.grad.item(), w.grad.cpu().numpy(), V.grad.cpu().numpy()


def _entries(g, n_top, n_rand, seed):
    top = [np.unravel_index(i, g.shape) for i in np.argsort(-np.abs(g).ravel())[:n_top]]
    rng = np.random.default_rng(seed)
    return top + [tuple(int(rng.integers(n)) for n in g.shape) for _ in range(n_rand)]


def test_dlambda_full(gpu, full):
    d, smp, f, g, _, _ = full
    lam, h = smp.coefficient, 1e-6
    fd = (f(lam * (1 + h)) - f(lam * (1 - h))) / (2 * h * lam)
    print(f"dlogL/dlambda {g:.10e} fd {fd:.10e} rel {abs(g - fd) / abs(fd):.2e}", flush=True)
    assert abs(g - fd) <= 1e-5 * abs(fd)


def test_dweights_full(gpu, full):
    d, smp, f, _, gw, _ = full
    w0 = torch.from_numpy(smp.pix_weights_for_sub_slim_index).cuda()
    h = 1e-5
    ent = _entries(gw, 4, 4, 0)
    errs = []
    for m, k in ent:
        wp, wm = w0.clone(), w0.clone()
        wp[m, k] += h
        wm[m, k] -= h
        fd = (f(smp.coefficient, w=wp) - f(smp.coefficient, w=wm)) / (2 * h)
        errs.append(abs(fd - gw[m, k]))
        print(f"dT[{m},{k}] {gw[m, k]:.8e} fd {fd:.8e}", flush=True)
    err = max(errs) / np.abs([gw[e] for e in ent]).max()
    print(f"weights: norm-wise rel err {err:.2e}", flush=True)
    assert err <= 1e-5


def test_dvertices_full(gpu, full):
    d, smp, f, _, _, gV = full
    V0 = torch.from_numpy(smp.mesh.vertices).cuda()
    h = 1e-7 * np.abs(smp.mesh.vertices).max()
    ent = _entries(gV, 4, 4, 1)
    errs = []
    for s, j in ent:
        Vp, Vm = V0.clone(), V0.clone()
        Vp[s, j] += h
        Vm[s, j] -= h
        fd = (f(smp.coefficient, V=Vp) - f(smp.coefficient, V=Vm)) / (2 * h)
        errs.append(abs(fd - gV[s, j]))
        print(f"dV[{s},{j}] {gV[s, j]:.8e} fd {fd:.8e}", flush=True)
    err = max(errs) / np.abs([gV[e] for e in ent]).max()
    print(f"vertices: norm-wise rel err {err:.2e}", flush=True)
    assert err <= 1e-5
