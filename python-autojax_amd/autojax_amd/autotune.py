"""Autotuners: the W-tilde SYRK's integer kernel variant for an image size and precision
(``tune_w_tilde``), and the batched log-evidence's Cholesky schedule for a batch's sizes
(``tune_batch``).

The variants (LDS layout of the digit slices, register or direct-to-LDS staging, 16x16x64 or
32x32x32 MFMA, 4 or 8 waves per 64x128 tile) compute bitwise-identical sums; only their speed
differs with the digit count and M.  ``tune_w_tilde`` times each one on a short synthetic build
(a few 4096-visibility chunks at the caller's M, inputs resident on the device), sets the
fastest process-wide through ``aj_w_tilde_set_variant`` and returns the timings.  Results are
cached per (M, precision) for the life of the process.
"""

from __future__ import annotations

import numpy as np

from . import _lib
from . import device as dv

_CACHE: dict = {}
INTEGER_MODES = ("int8x3", "int8x4", "int8x5")


def tune_w_tilde(grid_radians_slim, precision: str, chunks: int = 3, reps: int = 2, verbose: bool = False):
    """Time every integer SYRK variant for this grid and precision; select the fastest.

    Returns {variant: ms per build of ``chunks`` x 4096 visibilities} (variants that do not fit
    on the device for this digit count are left out).  Non-integer precisions have one kernel and
    return {} without timing anything.
    """
    import torch

    if precision not in INTEGER_MODES:
        return {}
    g = dv.to_device(np.asarray(grid_radians_slim, dtype=np.float64))
    M = g.shape[0]
    key = (M, precision)
    if key in _CACHE:
        _lib.call("aj_w_tilde_set_variant", _CACHE[key][0])
        return _CACHE[key][1]
    K = 4096 * chunks
    rng = np.random.default_rng(1)
    r = 1e3 * (5e2 ** rng.random(K))
    th = 2 * np.pi * rng.random(K)
    uv = dv.to_device(np.stack([r * np.cos(th), r * np.sin(th)], 1))
    noise = dv.to_device(rng.uniform(0.5, 2.0, K))
    prec = _lib.PRECISIONS[precision]
    out = dv.empty((M, M))
    ws = dv.workspace(_lib.query("aj_w_tilde_workspace_bytes", K, M, prec))

    def build():
        _lib.call("aj_w_tilde_curvature_interferometer_from", noise.data_ptr(), uv.data_ptr(), g.data_ptr(), K, M,
                  prec, out.data_ptr(), ws.data_ptr(), ws.numel(), dv.stream_ptr())

    times = {}
    for v in range(_lib.query("aj_w_tilde_variant_count")):
        _lib.call("aj_w_tilde_set_variant", v)
        try:
            build()
        except ValueError:  # this variant's LDS footprint does not fit for this digit count
            continue
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            build()
        b.record()
        torch.cuda.synchronize()
        times[v] = a.elapsed_time(b) / reps
        if verbose:
            print(f"autotune {precision} M={M}: variant {v} {times[v]:.3f} ms", flush=True)
    best = min(times, key=times.get)
    _lib.call("aj_w_tilde_set_variant", best)
    _CACHE[key] = (best, times)
    return times


CHOL_SCHEDULES = {0: "three launches per step", 1: "cooperative", 2: "cooperative, lookahead",
                  3: "one launch per step, lookahead", 4: "left-looking, one launch per step"}
_BATCH_CACHE: dict = {}


def tune_batch(batch, w_tilde, dirty_image, chi2_noise_norm, reps: int = 3, verbose: bool = False):
    """Time every Cholesky schedule of the batched log-evidence (``aj_log_evidence_configure``: the same
    factors bit for bit, in 3 launches per 32-column step, one cooperative launch without or with lookahead,
    or one launch per step with lookahead) on this batch, whose matrix sizes decide which is fastest (how
    many steps, how much trailing work per step); select the fastest process-wide and return
    {schedule: ms per evaluation}.  Cached per (M, sorted source-pixel counts) for the life of the process.
    """
    import torch

    key = (batch.M, tuple(sorted(int(s) for s in batch.S_host)))
    if key in _BATCH_CACHE:
        _lib.call("aj_log_evidence_configure", _BATCH_CACHE[key][0])
        return _BATCH_CACHE[key][1]
    times = {}
    for mode in CHOL_SCHEDULES:
        _lib.call("aj_log_evidence_configure", mode)
        batch.log_evidence(w_tilde, dirty_image, chi2_noise_norm)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            batch.log_evidence(w_tilde, dirty_image, chi2_noise_norm)
        b.record()
        torch.cuda.synchronize()
        times[mode] = a.elapsed_time(b) / reps
        if verbose:
            print(f"autotune batch n={batch.n}: schedule {mode} ({CHOL_SCHEDULES[mode]}) {times[mode]:.3f} ms",
                  flush=True)
    best = min(times, key=times.get)
    _lib.call("aj_log_evidence_configure", best)
    _BATCH_CACHE[key] = (best, times)
    return times
