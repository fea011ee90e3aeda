"""Autotuners: the W-tilde SYRK's kernel for an image size and precision (``tune_w_tilde``), and the
batched log-evidence's Cholesky schedule for a batch's sizes (``tune_batch``).

Digit modes (int8x3/4/5): the variants (LDS layout of the digit slices, register or direct-to-LDS
staging, 16x16x64 or 32x32x32 MFMA, 4 or 8 waves per 64x128 tile) compute bitwise-identical sums.
Residue modes (int8crt, int8crt11): the residue SYRK's variant (0 one workgroup per unit, 1 the same
without the MFMA split around the barrier, 2 two stages per barrier, 3 persistent) and the visibilities
per exact chunk (8,192 / 12,288 / 16,384), which give the same W-tilde per chunk size (the chunk size
changes the fixed-point scale of each chunk's features).  ``tune_w_tilde`` times each choice on a short
synthetic build at the caller's M (inputs resident on the device), sets the fastest process-wide
(``aj_w_tilde_set_variant`` / ``aj_w_tilde_crt_configure``) and returns the timings.  Results are cached
per (M, precision) for the life of the process.
"""

from __future__ import annotations

import numpy as np

from . import _lib
from . import device as dv

_CACHE: dict = {}
INTEGER_MODES = ("int8x3", "int8x4", "int8x5")
RESIDUE_MODES = ("int8crt", "int8crt11")
CRT_VARIANTS = (0, 1, 2, 3)
CRT_CHUNKS = (8192, 12288, 16384)


def tune_w_tilde(grid_radians_slim, precision: str, chunks: int = 3, reps: int = 2, verbose: bool = False):
    """Time every integer SYRK variant for this grid and precision; select the fastest.

    Returns {variant: ms per build of ``chunks`` x 4096 visibilities} (variants that do not fit
    on the device for this digit count are left out); for the residue modes {"v<variant>/c<chunk>": ms
    per build of 49,152 visibilities}.  FP64, FP32 and split-BF16 have one kernel and return {} without
    timing anything.
    """
    import torch

    if precision in RESIDUE_MODES:
        return _tune_crt(grid_radians_slim, precision, reps=reps, verbose=verbose)
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
                  3: "one launch per step, lookahead", 4: "left-looking, one launch per step",
                  5: "left-looking, one persistent launch with lookahead"}
# schedule 5's switch to 32-row tasks (aj_log_evidence_flow_configure small_per_cu) tried by the tuner
FLOW_SMALL = (1, 2)
_BATCH_CACHE: dict = {}


def _set_schedule(choice):
    mode, small = choice
    _lib.call("aj_log_evidence_configure", mode)
    _lib.call("aj_log_evidence_flow_configure", small, -1)


def set_schedule(name) -> None:
    """Select a schedule by its ``tune_batch`` key ("4", "5/s2", or an int mode) process-wide."""
    text = str(name)
    if "/s" in text:
        mode, small = text.split("/s")
        _set_schedule((int(mode), int(small)))
    else:
        _set_schedule((int(text), -1))


def tune_batch(batch, w_tilde, dirty_image, chi2_noise_norm, reps: int = 3, verbose: bool = False):
    """Time every Cholesky schedule of the batched log-evidence (``aj_log_evidence_configure``: the same
    factors bit for bit, in 3 launches per 32-column step, one cooperative launch without or with lookahead,
    one launch per step with lookahead, one left-looking launch per step, or the left-looking tasks in one
    persistent launch, the last with each switch value of ``FLOW_SMALL``) on this batch, whose matrix sizes
    decide which is fastest; select the fastest process-wide and return {schedule: ms per evaluation} (keys
    "4" or "5/s1").  Cached per (M, sorted source-pixel counts) for the life of the process.
    """
    import torch

    key = (batch.M, tuple(sorted(int(s) for s in batch.S_host)))
    if key in _BATCH_CACHE:
        _set_schedule(_BATCH_CACHE[key][0])
        return _BATCH_CACHE[key][1]
    choices = [(m, -1) for m in CHOL_SCHEDULES if m != 5] + [(5, s) for s in FLOW_SMALL]
    times = {}
    best, best_ms = None, None
    try:
        for choice in choices:
            _set_schedule(choice)
            batch.log_evidence(w_tilde, dirty_image, chi2_noise_norm, check=True)
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(reps):
                batch.log_evidence(w_tilde, dirty_image, chi2_noise_norm)
            b.record()
            torch.cuda.synchronize()
            ms = a.elapsed_time(b) / reps
            name = str(choice[0]) if choice[0] != 5 else f"5/s{choice[1]}"
            times[name] = ms
            if verbose:
                print(f"autotune batch n={batch.n}: schedule {name} ({CHOL_SCHEDULES[choice[0]]}) {ms:.3f} ms",
                      flush=True)
            if best_ms is None or ms < best_ms:
                best, best_ms = choice, ms
    finally:
        _set_schedule(best if best is not None else (-1, -1))
    _BATCH_CACHE[key] = (best, times)
    return times


def _tune_crt(grid_radians_slim, precision: str, reps: int = 2, verbose: bool = False):
    """Residue modes: time each (SYRK variant, chunk size) on 49,152 visibilities (a whole number of chunks
    of every size) at the caller's M; select the fastest through ``aj_w_tilde_crt_configure``."""
    import torch

    g = dv.to_device(np.asarray(grid_radians_slim, dtype=np.float64))
    M = g.shape[0]
    key = (M, precision)
    if key in _CACHE:
        v, c = _CACHE[key][0]
        _lib.call("aj_w_tilde_crt_configure", v, c)
        return _CACHE[key][1]
    K = 49152
    rng = np.random.default_rng(1)
    r = 1e3 * (5e2 ** rng.random(K))
    th = 2 * np.pi * rng.random(K)
    uv = dv.to_device(np.stack([r * np.cos(th), r * np.sin(th)], 1))
    noise = dv.to_device(rng.uniform(0.5, 2.0, K))
    prec = _lib.PRECISIONS[precision]
    out = dv.empty((M, M))
    max_chunk = 16384 if precision == "int8crt11" else 32768
    times = {}
    best, best_ms = None, None
    try:
        for c in CRT_CHUNKS:
            if c > max_chunk:
                continue
            for v in CRT_VARIANTS:
                _lib.call("aj_w_tilde_crt_configure", v, c)
                ws = dv.workspace(_lib.query("aj_w_tilde_workspace_bytes", K, M, prec))

                def build():
                    _lib.call("aj_w_tilde_curvature_interferometer_from", noise.data_ptr(), uv.data_ptr(), g.data_ptr(),
                              K, M, prec, out.data_ptr(), ws.data_ptr(), ws.numel(), dv.stream_ptr())

                build()
                torch.cuda.synchronize()
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record()
                for _ in range(reps):
                    build()
                b.record()
                torch.cuda.synchronize()
                ms = a.elapsed_time(b) / reps
                times[f"v{v}/c{c}"] = ms
                if verbose:
                    print(f"autotune {precision} M={M}: variant {v}, chunk {c}: {ms:.3f} ms", flush=True)
                if best_ms is None or ms < best_ms:
                    best, best_ms = (v, c), ms
                del ws
    finally:
        _lib.call("aj_w_tilde_crt_configure", *(best if best is not None else (-1, 0)))
    _CACHE[key] = (best, times)
    return times
