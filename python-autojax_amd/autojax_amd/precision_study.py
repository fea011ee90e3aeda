"""Precision study: per-stage error of each W-tilde arithmetic mode against the FP64 GPU path.

    python -m autojax_amd.precision_study [--config cfg3] [--modes int8crt,int8crt11,int8x5,int8x4,int8x3,fp32,bf16x3]
                                          [--lams 0.01,1,100]

Mode ``lattice_fp64`` builds the FP64 W-tilde by an independent route (the exact lattice table of
numba.py:210-255, expanded): its differences from the FP64 feature SYRK are the FP64 path's own
rounding spread, the floor against which the other modes' errors are read.

For every mode, W-tilde is rebuilt at that precision and everything after it runs in FP64 (as in
the product path).  Reported, against the same quantities from the FP64 W-tilde:

  W   max |dW| / max |W|                         (norm-wise; the maximum is the diagonal)
  F   max |dF| / max |F|,   F = T^T W T          (sparse two-stage product, numba.py:590-614)
  s   max |ds| / max |s|,   s = (F + H)^-1 D     per regularization coefficient
  logL  |d log-evidence|                          per regularization coefficient (numba.py:755-857)

Workloads: cfg3 = M 10,000 image pixels, K 1e6 visibilities, one 1,000-vertex Delaunay source
(datasets.config3); cfg5 = M 5,000, the first sample of the batching workload.  Prints one JSON
line per mode.
"""

from __future__ import annotations

import argparse
import json
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="cfg3", choices=["cfg3", "cfg5"])
    ap.add_argument("--K", type=int, default=1_000_000)
    ap.add_argument("--modes", default="int8crt,int8crt11,int8x5,int8x4,int8x3,fp32,bf16x3")
    ap.add_argument("--lams", default="0.01,1,100")
    a = ap.parse_args(argv)

    import torch

    from . import _lib, api
    from . import datasets as dsm
    from . import device as dv
    from . import evidence as E

    _lib.load()
    t0 = time.perf_counter()
    ds = dsm.config3(K=a.K) if a.config == "cfg3" else dsm.config5(K=a.K, n_samples=1)
    smp = ds.samples[0]
    print(f"# {a.config}: M={ds.M} K={ds.K} S={smp.pixels} ({time.perf_counter() - t0:.1f}s)", flush=True)
    lams = [float(x) for x in a.lams.split(",")]
    noise, uv, grid = (dv.to_device(x) for x in (ds.noise_map_real, ds.uv_wavelengths, ds.grid_radians_slim))
    y = E.dirty_image(ds.data, ds.noise_map, ds.uv_wavelengths, ds.grid_radians_slim)
    c2nn = E.chi2_noise_norm(ds.data, ds.noise_map)
    batch = E.SampleBatch([smp] * len(lams), ds.M)
    batch.set_coefficients(lams)
    idx, wts, S = smp.pix_indexes_for_sub_slim_index, smp.pix_weights_for_sub_slim_index, smp.pixels
    D = api.sparse_mapping_matrix_transpose_matmul(y.reshape(-1, 1), idx, wts, S).reshape(-1)
    H = [api.constant_regularization_matrix_from(c, smp.neighbors, smp.neighbors_sizes) for c in lams]

    native = dv.to_device(ds.native_index_for_slim_index)

    def stages(prec):
        t = time.perf_counter()
        if prec == "lattice_fp64":
            # the same FP64 W-tilde by another summation order: the lattice table, expanded
            wc = api.w_compact_curvature_interferometer_from(ds.n_side, noise, uv, ds.pixel_scale)
            W = api.w_tilde_via_compact_from(wc, native)
        else:
            W = E.w_tilde(noise, uv, grid, prec)
        torch.cuda.synchronize()
        t_w = time.perf_counter() - t
        F = api.curvature_matrix_via_w_wilde_sparse_mapping_matrix_from(W, idx, wts, S)
        s = []
        for h in H:
            try:
                s.append(api.reconstruction_positive_negative_from(D, F + torch.as_tensor(h, device=F.device)))
            except np.linalg.LinAlgError:  # F + H not positive definite at this precision
                s.append(None)
        logl = batch.log_evidence(W, y, c2nn).cpu().numpy()
        bad = batch.info.any(dim=1).cpu().numpy()
        return W, F, s, logl, bad, t_w

    W0, F0, s0, L0, bad0, t0w = stages("fp64")
    print(json.dumps({"mode": "fp64", "w_tilde_s": round(t0w, 3), "logL": L0.tolist(), "nonpd": bad0.tolist()}), flush=True)
    for prec in a.modes.split(","):
        W, F, s, L, bad, tw = stages(prec)
        rec = {
            "mode": prec,
            "w_tilde_s": round(tw, 3),
            "W": float((W - W0).abs().max() / W0.abs().max()),
            "F": float((F - F0).abs().max() / F0.abs().max()),
            "s": {str(c): (None if si is None else float((si - s0i).abs().max() / s0i.abs().max()))
                  for c, si, s0i in zip(lams, s, s0)},
            "logL": {str(c): float(abs(l - l0)) for c, l, l0 in zip(lams, L, L0)},
            "nonpd": {str(c): bool(b) for c, b in zip(lams, bad)},
        }
        print(json.dumps(rec), flush=True)
        del W, F, s


if __name__ == "__main__":
    main()
