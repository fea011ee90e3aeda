"""Differentiable log-evidence (torch.autograd.Function over hand-written HIP kernels).

``log_evidence(coefficients, weights, w_tilde, dirty_image, chi2_noise_norm, batch)``
returns the log-evidence of every sample of ``batch`` and differentiates with
respect to the per-sample regularisation coefficients and mapping weights.
``barycentric_weights(vertices, positions, triangles)`` builds the mapping
weights from source-mesh vertices with the traced image-pixel positions held
fixed, and differentiates with respect to the vertices; composing the two
gives dlogL/dvertices.  Both backwards run in libautojax_amd: the evidence
backward reuses the Cholesky factors of the forward pass for the adjoint
solve and the trace terms.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import device as dv


def check_factors_current(batch, evaluation):
    """The backward reads the Cholesky factors, Omega and reconstruction that the batch's workspace holds;
    they belong to the batch's last evaluation only.  Raises if the batch was evaluated after the forward
    pass being differentiated (its numbers would be silently wrong)."""
    if batch.evaluations != evaluation:
        raise RuntimeError(
            f"the batch was evaluated again (evaluation {batch.evaluations}) after the forward pass being "
            f"differentiated (evaluation {evaluation}); its factors are gone: evaluate again, or use one "
            "SampleBatch per graph")


class _LogEvidence(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coefficients, weights, w_tilde, dirty_image, chi2_noise_norm, batch):
        if weights.shape != (batch.n, batch.M, batch.B) or weights.dtype != torch.float64 or not weights.is_cuda:
            raise ValueError("weights must be a float64 device tensor [n, M, B]")
        coef = coefficients.detach().to("cpu", torch.float64).reshape(batch.n)
        batch.set_coefficients(coef.numpy())
        batch.wts = weights.detach().contiguous()
        out = batch.log_evidence(w_tilde, dirty_image, chi2_noise_norm).clone()
        ctx.batch = batch
        ctx.evaluation = batch.evaluations
        ctx.coef_device = coefficients.device
        ctx.save_for_backward(coef.to(weights.device), batch.wts, w_tilde, dirty_image)
        return out

    @staticmethod
    def backward(ctx, grad):
        coef_dev, wts, w_tilde, dirty = ctx.saved_tensors
        b = ctx.batch
        check_factors_current(b, ctx.evaluation)
        dcoef = torch.empty(b.n, dtype=torch.float64, device=wts.device)
        dw = torch.empty_like(wts)
        ws = dv.workspace(_lib.query("aj_log_evidence_backward_workspace_bytes", b.M, b.n,
                                     b.S_host.ctypes.data_as(_lib.C.c_void_p)))
        _lib.call("aj_log_evidence_batched_backward", w_tilde.data_ptr(), b.M, dirty.data_ptr(), b.n, b.B, b.P,
                  b.S_host.ctypes.data_as(_lib.C.c_void_p), b.idx.data_ptr(), wts.data_ptr(), b.nbr.data_ptr(),
                  b.nsz.data_ptr(), coef_dev.data_ptr(), b.recon.data_ptr(), b.ws.data_ptr(), b.ws.numel(),
                  dcoef.data_ptr(), dw.data_ptr(), ws.data_ptr(), ws.numel(), dv.stream_ptr())
        g = grad.to(wts.device, torch.float64)
        return (dcoef * g).to(ctx.coef_device), dw * g[:, None, None], None, None, None, None


def log_evidence(coefficients, weights, w_tilde, dirty_image, chi2_noise_norm, batch):
    """Differentiable batched log-evidence: coefficients [n] (any device), weights [n, M, B] on the GPU."""
    return _LogEvidence.apply(coefficients, weights, w_tilde, dirty_image, chi2_noise_norm, batch)


class _Barycentric(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vertices, positions, triangles):
        M = positions.shape[0]
        V = vertices.detach().contiguous()
        Pp = positions.detach().contiguous()
        tri = triangles.contiguous()
        out = torch.empty((M, 3), dtype=torch.float64, device=V.device)
        _lib.call("aj_barycentric_weights", V.data_ptr(), Pp.data_ptr(), tri.data_ptr(), M, out.data_ptr(),
                  dv.stream_ptr())
        ctx.save_for_backward(V, Pp, tri)
        return out

    @staticmethod
    def backward(ctx, grad):
        V, Pp, tri = ctx.saved_tensors
        M, S = Pp.shape[0], V.shape[0]
        g = grad.contiguous().to(torch.float64)
        dV = torch.empty_like(V)
        ws = dv.workspace(_lib.query("aj_barycentric_vjp_workspace_bytes", M, S))
        _lib.call("aj_barycentric_vjp", V.data_ptr(), Pp.data_ptr(), tri.data_ptr(), M, S, g.data_ptr(), dV.data_ptr(),
                  ws.data_ptr(), ws.numel(), dv.stream_ptr())
        return dV, None, None


def barycentric_weights(vertices, positions, triangles):
    """Weights [M, 3] of fixed points ``positions`` [M, 2] in triangles [M, 3] of ``vertices`` [S, 2]."""
    return _Barycentric.apply(vertices, positions, triangles)


def traced_positions(vertices, triangles, weights):
    """p_m = sum_b t_mb v_tri[m,b] (host helper used to fix the points of a generated mapping)."""
    V = np.asarray(vertices)
    return np.einsum("mb,mbk->mk", np.asarray(weights), V[np.asarray(triangles)])
