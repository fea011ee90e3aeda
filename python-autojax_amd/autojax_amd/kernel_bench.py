"""Single-kernel drivers for timing and hardware-counter runs (rocprofv3).

    python -m autojax_amd.kernel_bench wtilde --M 5000 --K 131072 --precision int8x4 --reps 3
    python -m autojax_amd.kernel_bench batch --M 5000 --samples 64 --K 8192
    python -m autojax_amd.kernel_bench copy --mib 1024
    python -m autojax_amd.kernel_bench intmm --M 8192
    python -m autojax_amd.kernel_bench mapper --M 5000 --samples 64
    python -m autojax_amd.kernel_bench mfma
    python -m autojax_amd.kernel_bench fill --mib 1024
    python -m autojax_amd.kernel_bench vendor --M 5120 --K 32768
    python -m autojax_amd.kernel_bench grad --samples 64

``wtilde`` builds W-tilde on synthetic data of the batching workload's
generator; ``batch`` evaluates the batched log-evidence of a sample batch on
one W-tilde; ``copy`` runs a plain device copy of known size, the reference
for scaling the FETCH_SIZE / WRITE_SIZE counters; ``intmm`` times the vendor
library's int8 GEMM (torch._int_mm, M x M x M) as the device's achievable
int8 matrix rate; ``mapper`` builds the source meshes and mappings of a sample batch
(S ~ U{300..1500} vertices each, the M image pixels as positions) on the device; ``mfma`` runs the
library's register-only MFMA loops (int8 16x16x64 and FP64 16x16x4, aj_mfma_peak_probe) as the device's
achievable matrix rates; ``fill`` writes a buffer of known size (achievable write bandwidth); ``vendor``
times the vendor libraries on the Gram shape C = X X^T, X of M rows and K columns: int8 (torch._int_mm),
fp8 (torch._scaled_mm), FP64 (torch matmul) full products and rocBLAS dsyrk (the lower triangle, FP64);
``grad`` times the batched log-evidence forward and its hand-written backward (gradients with respect to
the coefficients and the mapping weights) on the differentiation workload (--samples 1: M 7,860, K 1e5,
S 1,000) or the batching workload's samples (--samples 64).
Each prints one JSON line per measurement with the per-call time measured with HIP events.
"""

from __future__ import annotations

import argparse
import json
import sys

import numpy as np


def flow_summary(rec):
    """Summary of a persistent-schedule task trace (rows: taken, ready, done, matrix, kind, block column, task,
    hardware id; wall clock at 100 MHz): span, busy fraction of the workgroups, time spent waiting, and per
    kind the task count and mean run time, in microseconds."""
    t0 = rec[:, 0].min()
    span = (rec[:, 2].max() - t0) / 100.0
    run = (rec[:, 2] - rec[:, 1]) / 100.0
    wait = (rec[:, 1] - rec[:, 0]) / 100.0
    units = len(np.unique(rec[:, 7]))
    out = {"tasks": int(len(rec)), "span_us": round(float(span), 1), "units": units,
           "busy_frac": round(float(run.sum() / (span * max(units, 1))), 3),
           "wait_frac": round(float(wait.sum() / (span * max(units, 1))), 3)}
    for kind, name in ((0, "diag0"), (1, "strip256"), (2, "strip32")):
        m = rec[:, 4] == kind
        if m.any():
            out[name] = {"n": int(m.sum()), "mean_us": round(float(run[m].mean()), 2),
                         "max_us": round(float(run[m].max()), 2), "wait_mean_us": round(float(wait[m].mean()), 2)}
    # the last matrix to finish: its block columns' finish times
    last = rec[np.argmax(rec[:, 2])]
    mm = rec[rec[:, 3] == last[3]]
    ks = np.unique(mm[:, 5])
    # 256-row tasks' phases (slots 8-12): setup, blocks 0 .. k-2, second wait, last block, substitution, rest
    m = (rec[:, 4] == 1) & (rec[:, 8] > 0) & (rec[:, 12] > 0)
    if m.any():
        r = rec[m]
        k = r[:, 5]
        two = k >= 1
        ph = {"setup": (r[:, 8] - r[:, 1]), "solve": (r[:, 12] - r[:, 11]), "store_rest": (r[:, 2] - r[:, 12])}
        ph["update"] = np.where(two, r[:, 9] - r[:, 8], r[:, 11] - r[:, 8])
        ph["wait2"] = np.where(two, r[:, 10] - r[:, 9], 0)
        ph["last_block"] = np.where(two, r[:, 11] - r[:, 10], 0)
        first = r[:, 6] == 0
        out["strip256_phases_us"] = {
            name: {"others": round(float(v[~first].mean() / 100.0), 2) if (~first).any() else None,
                   "first": round(float(v[first].mean() / 100.0), 2) if first.any() else None}
            for name, v in ph.items()}
        out["strip256_update_us_per_block"] = round(float((ph["update"][two & ~first] / np.maximum(k[two & ~first] - 1, 1)).mean() / 100.0), 3) if (two & ~first).any() else None
        # the first task's next diagonal block (slots 13-15): its update by block k, factorisation, stage store
        nx = first & (r[:, 13] > 0) & (r[:, 14] > 0) & (r[:, 15] > 0)
        if nx.any():
            q = r[nx]
            out["next_block_us"] = {
                "stores_to_update": round(float((q[:, 13] - q[:, 12]).mean() / 100.0), 2),
                "factor": round(float((q[:, 14] - q[:, 13]).mean() / 100.0), 2),
                "stage_store": round(float((q[:, 15] - q[:, 14]).mean() / 100.0), 2),
                "to_complete": round(float((q[:, 2] - q[:, 15]).mean() / 100.0), 2)}
    out["last_matrix"] = {"matrix": int(last[3]), "columns": int(len(ks)),
                          "first_start_us": round(float((mm[:, 1].min() - t0) / 100.0), 1),
                          "column_end_us": [round(float((mm[mm[:, 5] == k][:, 2].max() - t0) / 100.0), 1)
                                            for k in ks[::4]]}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["wtilde", "batch", "copy", "intmm", "mapper", "mfma", "fill", "vendor", "grad"])
    ap.add_argument("--M", type=int, default=5000)
    ap.add_argument("--K", type=int, default=131072)
    ap.add_argument("--precision", default="int8x4")
    ap.add_argument("--samples", type=int, default=64)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--workgroups", type=int, default=1024, help="mfma: workgroups of 512 threads")
    ap.add_argument("--iters", type=int, default=0, help="mfma: loop iterations (0 = 5000 int8, 2500 fp64)")
    ap.add_argument("--kinds", default="int8,fp64", help="mfma: probes to run (int8, fp64)")
    ap.add_argument("--variant", type=int, default=None, help="integer SYRK kernel variant (LDS layout)")
    ap.add_argument("--chunk", type=int, default=0, help="int8crt: visibilities per exact chunk (0 = default)")
    ap.add_argument("--check", action="store_true", help="compare W-tilde with kernel variant 0 (bitwise)")
    ap.add_argument("--group-mib", type=int, default=0, help="cap on the features of one SYRK launch (0 = default)")
    ap.add_argument("--modes", default="0,1,2,3,4,5", help="batch: Cholesky schedules to time (aj_log_evidence_configure)")
    ap.add_argument("--flow-small", default="-1", help="batch: schedule 5's 32-row-task switch values "
                                                       "(aj_log_evidence_flow_configure small_per_cu)")
    ap.add_argument("--trace", default=None, help="batch: trace schedule 5's tasks once per switch value and save "
                                                  "the records to this .npz path (a summary is printed)")
    ap.add_argument("--against", default=None, help="also build W-tilde at this precision and report the "
                                                    "normwise difference max|dW| / max|W|")
    a = ap.parse_args(argv)

    import torch

    from . import _lib
    from . import datasets as dsm
    from . import device as dv
    from . import evidence as E

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            fn()
        e.record()
        torch.cuda.synchronize()
        return s.elapsed_time(e) / reps

    if a.precision.startswith("int8crt"):
        _lib.call("aj_w_tilde_crt_configure", -1 if a.variant is None else a.variant, a.chunk)
    elif a.variant is not None:
        _lib.call("aj_w_tilde_set_variant", a.variant)
    _lib.call("aj_w_tilde_set_group_bytes", a.group_mib << 20)
    if a.what == "copy":
        n = a.mib * (1 << 20) // 8
        x = torch.empty(n, dtype=torch.float64, device="cuda").uniform_()
        y = torch.empty_like(x)
        ms = timed(lambda: y.copy_(x), a.reps)
        print(json.dumps({"kernel": "copy", "bytes_read": n * 8, "bytes_written": n * 8, "ms": ms,
                          "GBps": 2 * n * 8 / ms / 1e6}), flush=True)
        return
    if a.what == "fill":
        n = a.mib * (1 << 20) // 8
        x = torch.empty(n, dtype=torch.float64, device="cuda")
        ms = timed(lambda: x.fill_(1.0), a.reps)
        print(json.dumps({"kernel": "fill", "bytes_written": n * 8, "ms": ms, "GBps": n * 8 / ms / 1e6}), flush=True)
        return
    if a.what == "mfma":
        wg = a.workgroups
        for kind, name, ops_per, dtype, chains in (
                (0, "int8 v_mfma_i32_16x16x64_i8", 16 * 16 * 64 * 2, torch.int64, 32),
                (1, "fp64 v_mfma_f64_16x16x4f64", 16 * 16 * 4 * 2, torch.float64, 16)):
            if ("int8", "fp64")[kind] not in a.kinds.split(","):
                continue
            iters = a.iters or (5000 if kind == 0 else 2500)
            one = torch.empty(wg * 512, dtype=dtype, device="cuda")
            _lib.call("aj_mfma_peak_probe", kind, wg, 1, one.data_ptr(), dv.stream_ptr())
            out = torch.empty(wg * 512, dtype=dtype, device="cuda")
            ms = timed(lambda: _lib.call("aj_mfma_peak_probe", kind, wg, iters, out.data_ptr(), dv.stream_ptr()), a.reps)
            ops = wg * 8 * iters * chains * ops_per
            exact = bool((out == one * iters).all().item())
            print(json.dumps({"kernel": "mfma_probe", "instruction": name, "workgroups": wg, "iters": iters, "ms": ms,
                              "T_ops_per_s": ops / ms / 1e9, "exact": exact}), flush=True)
        return
    if a.what == "batch":
        d = dsm.config5(K=a.K, n_samples=a.samples)
        noise, uv, grid = (dv.to_device(x) for x in (d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim))
        W = E.w_tilde(noise, uv, grid, "fp64")
        y = E.dirty_image(d.data, d.noise_map, d.uv_wavelengths, d.grid_radians_slim)
        c2nn = E.chi2_noise_norm(d.data, d.noise_map)
        batch = E.SampleBatch(d.samples, d.M)
        traces = {}
        runs = [(m, int(f)) for m in (int(x) for x in a.modes.split(","))
                for f in (a.flow_small.split(",") if m == 5 else ["-1"])]
        for mode, small in runs:
            _lib.call("aj_log_evidence_configure", mode)
            _lib.call("aj_log_evidence_flow_configure", small, -1)
            ms = timed(lambda: batch.log_evidence(W, y, c2nn), a.reps)
            _lib.timing_enable(True)
            for _ in range(a.reps):
                batch.log_evidence(W, y, c2nn)
            stages = {k: round(_lib.timing_read(k)[0] / a.reps, 4) for k in ("omega", "fgather", "chol", "solve")}
            _lib.timing_enable(False)
            batch._check_info()
            print(json.dumps({"kernel": "batched_log_evidence", "chol_mode": mode, "flow_small_per_cu": small,
                              "M": d.M, "samples": a.samples, "sum_S": int(batch.S_host.sum()), "ms": ms,
                              "stages_ms": stages}), flush=True)
            if mode == 5 and a.trace:
                cap = 1 << 18
                buf = torch.full((cap, 16), -1, dtype=torch.int64, device="cuda")
                _lib.call("aj_log_evidence_flow_trace", buf.data_ptr(), cap)
                try:
                    batch.log_evidence(W, y, c2nn)
                    torch.cuda.synchronize()
                finally:
                    _lib.call("aj_log_evidence_flow_trace", None, 0)
                rec = buf.cpu().numpy()
                rec = rec[rec[:, 2] >= 0]
                traces[f"small{small}"] = rec
                print(json.dumps({"kernel": "flow_trace", "flow_small_per_cu": small, **flow_summary(rec)}),
                      flush=True)
        _lib.call("aj_log_evidence_configure", -1)
        _lib.call("aj_log_evidence_flow_configure", -1, -1)
        if traces:
            np.savez(a.trace, S=batch.S_host, **traces)
        return
    if a.what == "grad":
        from . import torch_api as TA

        d = dsm.config4() if a.samples == 1 else dsm.config5(K=a.K, n_samples=a.samples)
        noise, uv, grid = (dv.to_device(x) for x in (d.noise_map_real, d.uv_wavelengths, d.grid_radians_slim))
        W = E.w_tilde(noise, uv, grid, "fp64")
        y = E.dirty_image(d.data, d.noise_map, d.uv_wavelengths, d.grid_radians_slim)
        c2nn = E.chi2_noise_norm(d.data, d.noise_map)
        batch = E.SampleBatch(d.samples, d.M)
        n = len(d.samples)
        w0 = torch.from_numpy(np.stack([s.pix_weights_for_sub_slim_index for s in d.samples])).cuda()
        c0 = torch.tensor([s.coefficient for s in d.samples], dtype=torch.float64)
        state = {}

        def fwd():
            c = c0.clone().requires_grad_(True)
            w = w0.clone().requires_grad_(True)
            state["out"] = TA.log_evidence(c, w, W, y, c2nn, batch)
            state["cw"] = (c, w)

        def bwd():
            state["out"].sum().backward()

        fwd()
        bwd()
        torch.cuda.synchronize()
        s0, s1, s2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        tf = tb = 0.0
        for _ in range(a.reps):
            s0.record()
            fwd()
            s1.record()
            bwd()
            s2.record()
            torch.cuda.synchronize()
            tf += s0.elapsed_time(s1)
            tb += s1.elapsed_time(s2)
        print(json.dumps({"kernel": "log_evidence_grad", "M": d.M, "K": d.K, "samples": n,
                          "sum_S": int(batch.S_host.sum()), "forward_ms": tf / a.reps, "backward_ms": tb / a.reps}),
              flush=True)
        return
    if a.what == "vendor":
        import ctypes as C

        n, k = a.M, a.K
        xi = torch.randint(-128, 128, (n, k), dtype=torch.int8, device="cuda")
        ms = timed(lambda: torch._int_mm(xi, xi.t()), a.reps)
        print(json.dumps({"kernel": "torch._int_mm", "M": n, "K": k, "ms": ms, "T_ops_per_s_full": 2.0 * n * n * k / ms / 1e9}),
              flush=True)
        x8 = xi.to(torch.float16).to(torch.float8_e4m3fn)
        one = torch.ones((), device="cuda")
        ms = timed(lambda: torch._scaled_mm(x8, x8.t(), scale_a=one, scale_b=one, out_dtype=torch.bfloat16), a.reps)
        print(json.dumps({"kernel": "torch._scaled_mm fp8", "M": n, "K": k, "ms": ms,
                          "T_flops_per_s_full": 2.0 * n * n * k / ms / 1e9}), flush=True)
        del xi, x8
        xd = torch.rand((n, k), dtype=torch.float64, device="cuda")
        ms = timed(lambda: xd @ xd.t(), a.reps)
        print(json.dumps({"kernel": "torch matmul fp64", "M": n, "K": k, "ms": ms,
                          "T_flops_per_s_full": 2.0 * n * n * k / ms / 1e9}), flush=True)
        # rocBLAS dsyrk: column-major C (n x n, lower) = A^T A with A = the row-major X seen as K x n
        rb = C.CDLL("librocblas.so")
        handle = C.c_void_p()
        assert rb.rocblas_create_handle(C.byref(handle)) == 0
        assert rb.rocblas_set_stream(handle, C.c_void_p(dv.stream_ptr())) == 0
        cd = torch.empty((n, n), dtype=torch.float64, device="cuda")
        alpha, beta = C.c_double(1.0), C.c_double(0.0)
        fill_lower, op_transpose = 122, 112

        def syrk():
            st = rb.rocblas_dsyrk(handle, fill_lower, op_transpose, n, k, C.byref(alpha), C.c_void_p(xd.data_ptr()), k,
                                  C.byref(beta), C.c_void_p(cd.data_ptr()), n)
            assert st == 0, st

        ms = timed(syrk, a.reps)
        rb.rocblas_destroy_handle(handle)
        print(json.dumps({"kernel": "rocblas_dsyrk lower", "M": n, "K": k, "ms": ms,
                          "T_flops_per_s_triangle": 1.0 * n * n * k / ms / 1e9}), flush=True)
        return
    if a.what == "intmm":
        n = a.M
        x = torch.randint(-128, 128, (n, n), dtype=torch.int8, device="cuda")
        y = torch.randint(-128, 128, (n, n), dtype=torch.int8, device="cuda")
        ms = timed(lambda: torch._int_mm(x, y), a.reps)
        print(json.dumps({"kernel": "torch._int_mm", "M": n, "ms": ms, "TOPS": 2.0 * n ** 3 / ms / 1e9}), flush=True)
        return
    rng = np.random.default_rng(5)
    grid, _ = dsm.grid_for_pixel_count(a.M, 0.05)
    uv = dsm.uv_loguniform(rng, a.K, 1e3, 5e5)
    data, noise = dsm.visibilities(rng, a.K, 0.5, 2.0)
    n_d = dv.to_device(np.ascontiguousarray(noise.real))
    uv_d = dv.to_device(uv)
    g_d = dv.to_device(grid)
    M, K = grid.shape[0], a.K
    W = dv.empty((M, M))
    if a.what == "wtilde":
        ws = dv.workspace(_lib.query("aj_w_tilde_workspace_bytes", K, M, _lib.PRECISIONS[a.precision]))
        ms = timed(lambda: E.w_tilde(n_d, uv_d, g_d, a.precision, out=W, workspace=ws), a.reps)
        rec = {"kernel": "w_tilde", "precision": a.precision, "variant": a.variant, "chunk": a.chunk,
               "group_mib": a.group_mib,
               "M": M, "K": K, "ms": ms,
               "algorithmic_tflops": 2.0 * M * M * K / ms / 1e9}
        if a.against:
            Wb = E.w_tilde(n_d, uv_d, g_d, a.against)
            rec["against"] = a.against
            rec["normwise_diff"] = float(((W - Wb).abs().max() / Wb.abs().max()).item())
        if a.check:
            if a.precision.startswith("int8crt"):
                _lib.call("aj_w_tilde_crt_configure", 0, a.chunk)
            else:
                _lib.call("aj_w_tilde_set_variant", 0)
            W0 = E.w_tilde(n_d, uv_d, g_d, a.precision, workspace=ws)
            rec["max_abs_diff_vs_variant0"] = float((W - W0).abs().max().item())
        print(json.dumps(rec), flush=True)
        return
    if a.what == "mapper":
        from . import mapper as mp

        r = float(np.abs(grid).max())
        verts = []
        for _ in range(a.samples):
            S = int(rng.integers(300, 1501))
            rad = r * np.sqrt(rng.random(S))
            th = 2.0 * np.pi * rng.random(S)
            verts.append(np.stack([rad * np.sin(th), rad * np.cos(th)], 1))
        mb = mp.MapperBatch(verts, grid)
        ms = timed(mb.launch, a.reps)
        mb.result()
        print(json.dumps({"kernel": "mapper", "M": M, "samples": a.samples, "sum_S": int(mb.offsets[-1]), "ms": ms}),
              flush=True)
        return
    E.w_tilde(n_d, uv_d, g_d, "fp64", out=W)
    samples = []
    for _ in range(a.samples):
        S = int(rng.integers(300, 1501))
        mesh = dsm.delaunay_source(rng, S)
        idx, w = dsm.barycentric_mapping(rng, M, mesh.triangles)
        lam = float(np.exp(rng.uniform(np.log(0.01), np.log(100.0))))
        samples.append(dsm.Sample(idx, w, mesh.neighbors, mesh.neighbors_sizes, lam, mesh))
    batch = E.SampleBatch(samples, M)
    y = E.dirty_image(data, noise, uv, grid)
    c2 = E.chi2_noise_norm(data, noise)
    ms = timed(lambda: batch.log_evidence(W, y, c2), a.reps)
    print(json.dumps({"kernel": "log_evidence_batched", "M": M, "samples": a.samples,
                      "sum_S": int(batch.S_host.sum()), "ms": ms}), flush=True)


if __name__ == "__main__":
    sys.exit(main())
