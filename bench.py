"""Log-evidence throughput on one MI355X: mixed-precision batched path vs plain FP64 GPU path.

Workload (the batching configuration): one dataset with M = 5,000 image
pixels and K = 1e6 visibilities, 64 lens-model samples with S ~ U{300..1500}
Delaunay source pixels each (own mesh, own barycentric mapping, B = 3) and
lambda log-uniform in [0.01, 100].  Synthetic data (datasets.config5).

One step = build W-tilde for the dataset + log-evidence of all 64 samples:
  mixed  : W-tilde at --precision (default int8crt11: FP64 cos/sin features
           rounded to 36-bit integers, their exact integer Gram matrix per
           16384-visibility chunk from its residues modulo 11 moduli, one int8
           MFMA product each, recombined in FP64), then one batched evaluation
           of all samples.
  plain  : W-tilde in FP64 (FP64 MFMA), then 64 sequential single-sample
           evaluations.
Both use FP64 for F, H, Cholesky, solves and log-determinants.
value = log-evidence evaluations / s of the mixed path (64 per step).

Also measured and reported beside it ("lattice_fp64"): the lattice path of
the reference's own API for regular pixel grids (numba.py:210-255, :860-967):
the N x (2N-1) table of distinct W-tilde values in exact FP64 (FP64 MFMA over
per-visibility phasors), then the same batched evaluation reading W-tilde from
that table.

roofline: the dominant kernel is the W-tilde SYRK; its algorithmic work is
2 M^2 K flops per W-tilde (lower triangle, cos and sin terms), timed with HIP
events on the library's stream around every SYRK launch.  "traffic" (HBM
bytes per launch) comes from the FETCH_SIZE / WRITE_SIZE counter summary in
profiles/ when it is present for this kernel and workload.

python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "python-autojax_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

import numpy as np  # noqa: E402

PEAKS_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0, "int8x3": 5000.0, "int8x4": 5000.0, "int8x5": 5000.0,
                "int8crt": 5000.0, "int8crt11": 5000.0}
RAW_FACTOR = {"fp64": 1, "fp32": 1, "bf16x3": 3, "int8x3": 9, "int8x4": 15, "int8x5": 22, "int8crt": 12,
              "int8crt11": 11}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


I8_KERNELS = {0: "syrk_i8_kernel<{ns},{ng},0>", 1: "syrk_i8_kernel<{ns},{ng},1>", 2: "syrk_i8_kernel<{ns},{ng},2>",
              3: "syrk_i8_kernel<{ns},{ng},3>", 4: "syrk_i8_kernel<{ns},{ng},4>", 5: "syrk_i8_dma_kernel<{ns},{ng}>",
              6: "syrk_i8_k32_kernel<{ns},{ng}>", 7: "syrk_i8_w8_kernel<{ns},{ng}>"}
I8_DIGITS = {"int8x3": (3, 5), "int8x4": (4, 6), "int8x5": (5, 7)}


CRT_KERNELS = {0: "crt::syrk_pipe8_kernel<1>", 1: "crt::syrk_pipe8_kernel<0>", 2: "crt::syrk_ring8_kernel",
               3: "crt::syrk_sk_kernel<1>"}


def syrk_kernel_name(precision, variant):
    if precision.startswith("int8crt"):
        # variant: the tuner's "v<variant>/c<chunk>" key, or None (the default, persistent form)
        v = int(variant.split("/")[0][1:]) if variant else 3
        return CRT_KERNELS[v]
    if precision not in I8_DIGITS:
        return f"syrk_kernel<{precision}>"
    ns, ng = I8_DIGITS[precision]
    if variant is None:
        variant = 7 if ns == 5 else 0
    return I8_KERNELS[variant].format(ns=ns, ng=ng)


def counter_traffic(precision, M, kernel, chunk=None):
    """HBM bytes per SYRK launch from the committed counter summary for this precision and size
    (profiles/counters_w_tilde_syrk_<precision>_M<M>.md, lines 'kernel: <name>',
    'traffic_bytes_per_launch: <float>' and, for the residue modes, 'chunk_visibilities: <int>'), if it
    was measured on the same kernel (and chunk size); else None."""
    p = ROOT / "profiles" / f"counters_w_tilde_syrk_{precision}_M{M}.md"
    if not p.exists():
        return None
    fields = {}
    for line in p.read_text().splitlines():
        key, sep, val = line.strip().strip("`").partition(":")
        if sep and key in ("kernel", "traffic_bytes_per_launch", "chunk_visibilities"):
            fields[key] = val.strip().strip("`")
    if fields.get("kernel") != kernel or "traffic_bytes_per_launch" not in fields:
        return None
    if chunk is not None and fields.get("chunk_visibilities") != str(chunk):
        return None
    return float(fields["traffic_bytes_per_launch"])


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for row in f:
                if row.startswith("model name"):
                    return row.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ds, k_sub, n_like, threads):
    """Oracle timed on the host: C/OpenMP W-tilde on k_sub visibilities (scaled to K) and the
    numpy likelihood (dense T, two DGEMMs, LU solve, two slogdets) on n_like samples (scaled), with
    `threads` OpenMP threads and the BLAS limited to the same count."""
    import ctypes

    from threadpoolctl import threadpool_limits

    from oracle import restated as R

    lib = ctypes.CDLL(str(ROOT / "oracle" / "liboracle_wtilde.so"))
    lib.oracle_w_tilde.restype = ctypes.c_int
    lib.oracle_w_tilde.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]
    M, K = ds.M, ds.K
    n = np.ascontiguousarray(ds.noise_map_real[:k_sub])
    uv = np.ascontiguousarray(ds.uv_wavelengths[:k_sub])
    g = np.ascontiguousarray(ds.grid_radians_slim)
    W = np.empty((M, M))
    t0 = time.perf_counter()
    rc = lib.oracle_w_tilde(n.ctypes.data, uv.ctypes.data, g.ctypes.data, k_sub, M, threads, W.ctypes.data)
    t_w = (time.perf_counter() - t0) * (K / k_sub)
    assert rc == 0
    y = ds.dirty_image
    with threadpool_limits(limits=threads):
        t0 = time.perf_counter()
        for smp in ds.samples[:n_like]:
            R.log_likelihood_function_via_w_tilde_from(y, ds.data, ds.noise_map, W, smp.pix_indexes_for_sub_slim_index,
                                                       smp.pix_weights_for_sub_slim_index, smp.neighbors,
                                                       smp.neighbors_sizes)
        t_like = (time.perf_counter() - t0) / n_like * len(ds.samples)
    return len(ds.samples) / (t_w + t_like), t_w, t_like


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="int8crt11", choices=list(PEAKS_TFLOPS))
    ap.add_argument("--no-autotune", action="store_true", help="keep the default SYRK kernel variant")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--samples", type=int, default=64)
    ap.add_argument("--K", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-k-sub", type=int, default=160)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch

    from autojax_amd import _lib
    from autojax_amd import datasets as dsm
    from autojax_amd import device as dv
    from autojax_amd import evidence as E

    _lib.load()
    t0 = time.perf_counter()
    ds = dsm.config5(K=args.K, n_samples=args.samples)
    log(f"dataset: M={ds.M} K={ds.K} samples={len(ds.samples)} S={[s.pixels for s in ds.samples[:8]]}... "
        f"({time.perf_counter() - t0:.1f}s)")
    noise = dv.to_device(ds.noise_map_real)
    uv = dv.to_device(ds.uv_wavelengths)
    grid = dv.to_device(ds.grid_radians_slim)
    y = E.dirty_image(ds.data, ds.noise_map, ds.uv_wavelengths, ds.grid_radians_slim)
    c2nn = E.chi2_noise_norm(ds.data, ds.noise_map)
    batch = E.SampleBatch(ds.samples, ds.M)
    singles = [E.SampleBatch([s], ds.M) for s in ds.samples]
    M, K = ds.M, ds.K
    W_mixed = dv.empty((M, M))
    W_plain = dv.empty((M, M))
    torch.cuda.synchronize()
    log("inputs resident")
    tuned = {}
    from autojax_amd import autotune

    # the mixed path's SYRK kernel (and, for the residue modes, its chunk size), tuned at this M
    if not args.no_autotune:
        tuned = autotune.tune_w_tilde(ds.grid_radians_slim, args.precision)
        if tuned:
            log("autotune W-tilde: " + ", ".join(f"{v}={t:.2f}ms" for v, t in sorted(tuned.items(), key=str)))
    # (the residue modes' workspace depends on the tuned chunk size)
    ws_mixed = dv.workspace(_lib.query("aj_w_tilde_workspace_bytes", K, M, _lib.PRECISIONS[args.precision]))
    ws_plain = dv.workspace(_lib.query("aj_w_tilde_workspace_bytes", K, M, 0))

    # Cholesky schedule of the batched evaluation, tuned on the 64-sample batch and, for the plain path's
    # one-sample evaluations, on its median-size sample (each path then runs with its own)
    sched = {"batch": "5", "single": "5"}
    tuned_batch, tuned_single = {}, {}
    if not args.no_autotune:
        E.w_tilde(noise, uv, grid, "fp64", out=W_plain, workspace=ws_plain)
        tuned_batch = autotune.tune_batch(batch, W_plain, y, c2nn)
        mid = singles[int(np.argsort(batch.S_host)[len(singles) // 2])]
        tuned_single = autotune.tune_batch(mid, W_plain, y, c2nn)
        sched = {"batch": min(tuned_batch, key=tuned_batch.get), "single": min(tuned_single, key=tuned_single.get)}
        log("autotune batch schedule: " + ", ".join(f"{m}={t:.2f}ms" for m, t in sorted(tuned_batch.items())) +
            "; single: " + ", ".join(f"{m}={t:.2f}ms" for m, t in sorted(tuned_single.items())))

    def mixed_step():
        autotune.set_schedule(sched["batch"])
        E.w_tilde(noise, uv, grid, args.precision, out=W_mixed, workspace=ws_mixed)
        return batch.log_evidence(W_mixed, y, c2nn)

    def plain_step():
        autotune.set_schedule(sched["single"])
        E.w_tilde(noise, uv, grid, "fp64", out=W_plain, workspace=ws_plain)
        return torch.cat([s.log_evidence(W_plain, y, c2nn).clone() for s in singles])

    def timed(fn, steps, warmup, label):
        for i in range(warmup):
            fn()
            torch.cuda.synchronize()
            log(f"{label} warmup {i + 1}/{warmup} done")
        _lib.timing_enable(True)
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for i in range(steps):
            out = fn()
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / steps
        timers = {k: _lib.timing_read(k) for k in _lib.TIMERS}
        _lib.timing_enable(False)
        log(f"{label}: {ms:.2f} ms/step; timers " + ", ".join(f"{k}={v[0] / steps:.2f}ms/{v[1]}" for k, v in timers.items()))
        return ms, out, timers

    N = ds.n_side
    native = dv.to_device(np.ascontiguousarray(ds.native_index_for_slim_index.astype(np.int64)))
    wc = dv.empty((N, 2 * N - 1))
    ws_wc = dv.workspace(_lib.query("aj_w_compact_workspace_bytes", K, N))

    def lattice_step():
        autotune.set_schedule(sched["batch"])
        _lib.call("aj_w_compact_curvature_interferometer_from", noise.data_ptr(), uv.data_ptr(), K, N,
                  float(ds.pixel_scale), wc.data_ptr(), ws_wc.data_ptr(), ws_wc.numel(), dv.stream_ptr())
        return batch.log_evidence_compact(wc, native, y, c2nn)

    def build_compact():
        _lib.call("aj_w_compact_curvature_interferometer_from", noise.data_ptr(), uv.data_ptr(), K, N,
                  float(ds.pixel_scale), wc.data_ptr(), ws_wc.data_ptr(), ws_wc.numel(), dv.stream_ptr())

    ms_mixed, out_mixed, tm = timed(mixed_step, args.steps, args.warmup, "mixed")
    mixed_ev = out_mixed.cpu().numpy()  # the batch's output buffer is reused by the lattice step
    info_bad = int(batch.info.any(dim=1).sum().item())
    ms_lat, out_lat, tl = timed(lattice_step, args.steps, args.warmup, "lattice")
    lat_ev = out_lat.cpu().numpy()
    lat_bad = int(batch.info.any(dim=1).sum().item())
    ms_wc, _, _ = timed(build_compact, args.steps, 1, "w_compact")
    ms_plain, out_plain, tp = timed(plain_step, args.plain_steps, 1, "plain")
    plain_ev = out_plain.cpu().numpy()
    dlogl = float(np.abs(mixed_ev - plain_ev).max())

    # the data-side sums (dirty image and w_tilde_data) fused into the mixed path's W-tilde build, against
    # the build alone and the two stand-alone sums (HIP events on the library's stream)
    from autojax_amd import api

    def event_ms(fn, reps=2):
        fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    data_d, nmap_d = dv.to_device(ds.data), dv.to_device(ds.noise_map)
    vis_d = dv.to_device(np.ascontiguousarray(ds.data.real) + 3.0)
    sums = {
        "w_tilde_with_sums": event_ms(lambda: api.w_tilde_with_data_interferometer_from(
            noise, uv, grid, data=data_d, visibilities_real=vis_d, precision=args.precision)),
        "w_tilde_alone": event_ms(lambda: E.w_tilde(noise, uv, grid, args.precision, out=W_mixed, workspace=ws_mixed)),
        "dirty_image_standalone": event_ms(lambda: api.dirty_image_from(data_d, nmap_d, uv, grid)),
        "w_tilde_data_standalone": event_ms(lambda: api.w_tilde_data_interferometer_from(vis_d, noise, uv, grid)),
    }
    sums = {k: round(v, 3) for k, v in sums.items()}
    sums["fused_marginal"] = round(sums["w_tilde_with_sums"] - sums["w_tilde_alone"], 3)
    log("data-side sums (ms): " + ", ".join(f"{k}={v}" for k, v in sums.items()))
    info_bad = max(info_bad, lat_bad)
    n = len(ds.samples)
    value = n / (ms_mixed / 1e3)
    plain_value = n / (ms_plain / 1e3)

    # roofline of the dominant kernel: the W-tilde SYRK launches of the mixed path
    syrk_ms, syrk_launches = tm["wt_syrk"]
    algo_flops_step = 2.0 * M * M * K  # lower-triangle SYRK count, cos and sin terms
    achieved = algo_flops_step * args.steps / (syrk_ms / 1e3) / 1e12
    peak = PEAKS_TFLOPS[args.precision]
    best_variant = min(tuned, key=tuned.get) if tuned else None
    kernel = syrk_kernel_name(args.precision, best_variant)
    crt_chunk = (int(best_variant.split("/c")[1]) if best_variant else 16384) if args.precision.startswith("int8crt") else None
    # the batched evaluation's own span (HIP events around its launches); the rest of the step is W-tilde
    batch_ms = tm["batch"][0] / args.steps
    wt_ms = ms_mixed - batch_ms

    chol_flops = float(sum(2.0 * float(S) ** 3 / 3.0 for S in batch.S_host))
    line = {
        "metric": "log-evidence evals/sec (mixed+batched vs plain FP64 GPU) at <=1e-3 |dlogL|",
        "value": round(value, 3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_mixed, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("i8+f64" if args.precision.startswith("int8") else "f64" if args.precision == "fp64" else
                  args.precision + "+f64") + f" ({args.precision} W-tilde on MFMA, f64 likelihood)",
        "data": "synthetic",
        "config": {
            "workload": "batched log-evidence: M=5000 px, K=1e6 vis, 64 Delaunay samples S~U{300..1500}, B=3, "
                        "lambda~logU[0.01,100] (datasets.config5)",
            "M": M, "K": K, "samples": n, "sum_S": int(batch.S_host.sum()), "precision": args.precision,
        },
        "plain_fp64": {"value": round(plain_value, 3), "ms_per_step": round(ms_plain, 3),
                       "w_tilde_ms": round((tp["wt_syrk"][0] + tp["wt_features"][0]) / args.plain_steps, 3)},
        "speedup_vs_plain": round(value / plain_value, 3),
        "max_abs_dlogL_vs_plain": dlogl,
        "within_1e-3": bool(dlogl <= 1e-3),
        "nonpd_samples": info_bad,
        "breakdown_ms": {"w_tilde_syrk": round(tm["wt_syrk"][0] / args.steps, 3),
                         "w_tilde_features": round(tm["wt_features"][0] / args.steps, 3),
                         "w_tilde_combine": round(tm["wt_combine"][0] / args.steps, 3),
                         "omega": round(tm["omega"][0] / args.steps, 3),
                         "fgather": round(tm["fgather"][0] / args.steps, 3),
                         "cholesky": round(tm["chol"][0] / args.steps, 3),
                         "solve": round(tm["solve"][0] / args.steps, 3),
                         "w_tilde_total": round(wt_ms, 3),
                         "batch_total": round(batch_ms, 3)},
        "evals_per_s_excluding_w_tilde": round(n / (batch_ms / 1e3), 3),
        "data_side_sums_ms": sums,
        # blocked Cholesky of every A = F + H and every H of the batch: S^3 / 3 flops each
        "factorisation": {"flops": chol_flops, "tflops": round(chol_flops / tm["chol"][0] * args.steps / 1e9, 3),
                          "frac_fp64_peak": round(chol_flops / tm["chol"][0] * args.steps / 1e9 / 78.6, 4)},
        "lattice_fp64": {"value": round(n / (ms_lat / 1e3), 3), "ms_per_step": round(ms_lat, 3),
                         "w_compact_ms": round(ms_wc, 3),
                         "speedup_vs_plain": round(ms_plain / ms_lat, 3),
                         "max_abs_dlogL_vs_plain": float(np.abs(lat_ev - plain_ev).max()),
                         "grid": f"{N} x {N} lattice, table {N} x {2 * N - 1}"},
        "roofline": {"bound": "mfma", "kernel": kernel, "achieved": round(achieved, 3),
                     "peak": peak, "unit": "TFLOP/s", "frac": round(achieved / peak, 5),
                     "raw_mfma_frac": round(achieved * RAW_FACTOR[args.precision] / peak, 5),
                     "avg_launch_ms": round(syrk_ms / max(syrk_launches, 1), 4), "launches_per_step": syrk_launches // args.steps,
                     "traffic": counter_traffic(args.precision, M, kernel, crt_chunk)},
        "autotune_ms": {"w_tilde_syrk_variant": {str(v): round(t, 3) for v, t in sorted(tuned.items())},
                        "w_tilde_selected": str(best_variant),
                        "cholesky_schedule_batch": {str(m): round(t, 3) for m, t in sorted(tuned_batch.items())},
                        "cholesky_schedule_single": {str(m): round(t, 3) for m, t in sorted(tuned_single.items())},
                        "selected": sched},
    }
    if not args.no_cpu_baseline:
        ds.dirty_image = y.cpu().numpy()
        affinity = len(os.sched_getaffinity(0))
        # a GPU box allots 16 host threads to its one GPU (OMP_NUM_THREADS there); each W-tilde thread also
        # holds an M x M accumulator (200 MB at M = 5,000)
        threads = min(affinity, 16)
        cv, t_w, t_like = cpu_baseline(ds, args.cpu_k_sub, 2, threads)
        k1 = max(1, args.cpu_k_sub // 10)
        cv1, t_w1, t_like1 = cpu_baseline(ds, k1, 1, 1)
        line["cpu_baseline"] = {"value": round(cv, 6), "unit": "evals/s", "cores": threads, "kind": "port",
                                "sample": f"C/OpenMP W-tilde on {args.cpu_k_sub} of {K} visibilities scaled x{K // args.cpu_k_sub} "
                                          f"({t_w:.1f}s), numpy likelihood on 2 of {n} samples scaled ({t_like:.1f}s); "
                                          f"per-thread W-tilde accumulator {M * M * 8 / 1e6:.0f} MB",
                                "value_1thread": round(cv1, 7),
                                "sample_1thread": f"W-tilde on {k1} visibilities scaled x{K // k1} ({t_w1:.1f}s), "
                                                  f"likelihood on 1 of {n} samples scaled ({t_like1:.1f}s)",
                                "cpu": cpu_model(), "affinity_cores": affinity,
                                "cores_note": "16 host threads per GPU on the box (its CPU share for one GPU)"}
        log(f"cpu baseline: {cv:.4g} evals/s ({threads} threads), {cv1:.4g} evals/s (1 thread)")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
