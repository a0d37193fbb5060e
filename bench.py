"""Benchmark: randomized SVD seconds-to-solution on the 100,000 x 20,000 Zipf
workload at a top-50 relative residual <= 1e-5, with the SpMM roofline.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision f32|bf16] [--no-cpu]

A step is one whole solve from the call (matrix already resident in HBM) to
U, S and V^T in host memory: rank 50, oversample 10, QR normaliser, power
iterations until the Rayleigh-Ritz residual estimate of the top 50 triplets
is <= 1e-5. value is the median of the timed steps after the warm-up; the
device-resident time (U and V^T left in HBM), the upload and the cold first
solve are reported beside it. The returned triplets are re-checked in FP64 on
the CPU outside the timed region:
max_i sqrt(|A v_i - s_i u_i|^2 + |A^T u_i - s_i v_i|^2) / s_i.
The other configurations (LAS2, batched masks, adjoint) follow in the same
line as extra keys unless --no-extras.

Prints one JSON line (rank 0).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "single-svdlib_amd"))

import numpy as np  # noqa: E402
import scipy.sparse as sp  # noqa: E402

METRIC = "SVD seconds-to-solution at rel residual ≤1e-5; SpMV/SpMM HBM GB/s vs peak"
HBM_PEAK_GBS = 8000.0
RANK, OVERSAMPLE, TOL = 50, 10, 1e-5


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def residual(a, u, s, vt):
    a = sp.csr_matrix(a, dtype=np.float64)
    v = vt.T.astype(np.float64)
    u = u.astype(np.float64)
    r1 = a @ v - u * s
    r2 = a.T @ u - v * s
    per = np.sqrt((r1 * r1).sum(0) + (r2 * r2).sum(0))
    return float(np.max(per / s)), float(np.max(per) / float(s[0]))


def traffic_from_profiles(name="spmm_cfg2_pmc.md"):
    """HBM bytes per SpMM (or SpMV) launch measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), recorded in profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def cpu_info():
    """CPU model and logical CPU count of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def cpu_baseline(a, iters_needed, threads):
    """The C restatement of the reference's Rayon path (oracle/, OpenMP: row-
    parallel A X, per-chunk partials of A^T Y with determine_chunk_size and a
    serial reduction) in the reference's f32 instantiation, run as one whole
    randomized_svd (rank 50, oversample 10, QR normaliser, f32 products, QR
    and SVD) with iters_needed plain power iterations: the count the
    reference's own algorithm needs for a top-50 residual of 1e-5 on this
    matrix (measured on the GPU with the shift and momentum switched off).
    The returned triplets' residual is recomputed in FP64."""
    import oracle
    op = oracle.OracleOp(a, nthreads=threads)
    om = np.random.default_rng(0).normal(size=(a.shape[1], RANK + OVERSAMPLE))
    t0 = time.perf_counter()
    u, s, vt = oracle.randomized_svd(op, RANK, OVERSAMPLE, iters_needed, "QR", omega=om, dtype=np.float32)
    wall = time.perf_counter() - t0
    r_rel, r_norm = residual(a, u, s.astype(np.float64), vt)
    model, ncpu = cpu_info()
    return {"value": wall, "unit": "s", "cores": threads, "kind": "port",
            "sample": (f"the whole f32 randomized_svd (C restatement of the reference's Rayon decomposition, "
                       f"{threads} OpenMP threads) with {iters_needed} plain power iterations on the same matrix; "
                       f"FP64 residual of its top-50 triplets {r_rel:.2e} (|A|-relative {r_norm:.2e})"),
            "cpu_model": model, "logical_cpus": ncpu, "threads": threads, "power_iterations": iters_needed,
            "residual_fp64": r_rel}


# ---------------------------------------------------------------------------
# Other measured paths (python bench.py --workload ...). Each returns one JSON
# object with its own metric, the roofline of its dominant kernel class and a
# CPU baseline where one is affordable.
# ---------------------------------------------------------------------------

def _kernel_stats(lib, ctx, classes, steps):
    ms, nl, by = C.c_double(), C.c_int64(), C.c_double()
    out = {}
    from single_svdlib import _ffi
    for name, k in classes:
        _ffi.check(lib.svdh_ctx_kernel_time(ctx, k, C.byref(ms), C.byref(nl), C.byref(by)))
        out[name] = {"ms_per_step": ms.value / steps, "launches_per_step": nl.value / steps,
                     "avg_launch_us": ms.value * 1e3 / max(nl.value, 1), "bytes": by.value, "_ms": ms.value,
                     "_n": nl.value}
    return out


def _roofline(total_bytes, total_ms, launches):
    avg_ms = total_ms / max(launches, 1)
    per = total_bytes / max(launches, 1)
    ach = per / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": None, "algorithmic_bytes_per_launch": per, "avg_launch_ms": avg_ms}


def _triplet_residual(a, u_cols, s, vt_rows):
    a = sp.csr_matrix(a, dtype=np.float64)
    v = np.asarray(vt_rows, dtype=np.float64).T
    u = np.asarray(u_cols, dtype=np.float64)
    r1 = a @ v - u * s
    r2 = a.T @ u - v * s
    per = np.sqrt((r1 * r1).sum(0) + (r2 * r2).sum(0))
    return float(np.max(per / s))


def bench_las2(args):
    """lanczos::svd_dim_seed on the 2,000 x 1,500 (k = 10) or 500,000 x 30,000
    (k = 100) workload, FP64 and FP32-SpMV/FP64-recurrence, U and V^T left in
    HBM (svdh_las2_device; the host hand-over is timed beside it); the
    dominant kernel is the SpMV (A x and A^T y)."""
    import single_svdlib as sl
    from single_svdlib import _ffi, synth
    import oracle
    cfg1 = args.workload == "las2-cfg1"
    a = synth.config1() if cfg1 else synth.config3(np.float64)
    k = 10 if cfg1 else 100
    lib = _ffi.lib()
    res = {}
    m, n = a.shape
    u_dev = _ffi.DeviceArray((k, m))
    vt_dev = _ffi.DeviceArray((k, n))
    s_host = np.zeros(k)
    precisions = (("f32", np.float32),) if args.profile_run else (("f64", np.float64), ("f32", np.float32))
    for prec, dt in precisions:
        op = sl.CsrMatrix.from_scipy(a.astype(dt))
        ctx = op._ctx
        pcode = _ffi.PREC_F64 if prec == "f64" else _ffi.PREC_F32

        def solve_device():
            meta = _ffi.Result_()
            _ffi.check(lib.svdh_las2_device(ctx, op._handle, k, 0, -1e-30, 1e-30, 1e-6, 42, None, pcode, u_dev.ptr,
                                            m, _ffi.ptr(s_host), vt_dev.ptr, n, C.byref(meta)))
            return meta

        for _ in range(args.warmup):
            solve_device()
        _ffi.check(lib.svdh_ctx_synchronize(ctx))
        times = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            solve_device()
            times.append(time.perf_counter() - t0)
        if args.profile_run:
            _ffi.check(lib.svdh_ctx_synchronize(ctx))
            nb = 256 << 20  # the calibration kernel: a plain copy of 256 MiB
            src, dst = _ffi.DeviceArray((nb // 8,)), _ffi.DeviceArray((nb // 8,))
            for _ in range(2):
                if _ffi.hip().hipMemcpy(dst.ptr, src.ptr, nb, 3) != 0:
                    raise RuntimeError("device-to-device copy failed")
            log(f"las2 {prec}: {np.median(times) * 1e3:.1f} ms; calibration copy done (2 x 268435456 bytes)")
            return {"metric": "LAS2 seconds-to-solution", "value": float(np.median(times)), "unit": "s",
                    "profile_run": True}
        # the same solve handing U, s, V^T over in host memory (PCIe-inclusive),
        # through svdh_las2 as the reference's SvdRec is returned: one warm-up
        # (the first call fills the pinned result pool), then the median of three
        def solve_host():
            res = _ffi.Result_()
            _ffi.check(lib.svdh_las2(ctx, op._handle, k, 0, -1e-30, 1e-30, 1e-6, 42, None, pcode, C.byref(res)))
            return res

        lib.svdh_result_free(C.byref(solve_host()))
        t_hosts = []
        for _ in range(3):
            t0 = time.perf_counter()
            res_h = solve_host()
            t_hosts.append(time.perf_counter() - t0)
            lib.svdh_result_free(C.byref(res_h))
        t_host = float(np.median(t_hosts))
        _ffi.check(lib.svdh_ctx_profile(ctx, 1))
        rec = sl.lanczos.svd_dim_seed(op, k, 42)
        st = _kernel_stats(lib, ctx, (("spmv", _ffi.KC_SPMV), ("blas1", _ffi.KC_BLAS1), ("gemv", _ffi.KC_GEMV),
                                      ("spmm", _ffi.KC_SPMM), ("spmm_t", _ffi.KC_SPMM_T), ("dense", _ffi.KC_DENSE)),
                           1)
        _ffi.check(lib.svdh_ctx_profile(ctx, 0))
        d = rec.diagnostics
        r = _triplet_residual(a, rec.u.T, rec.s.astype(np.float64), rec.vt)
        res[prec] = {"s_per_solve": float(np.median(times)), "host_output_s": t_host, "lanczos_steps": d.lanczos_steps,
                     "ritz_values_stabilized": d.ritz_values_stabilized, "significant_values": d.significant_values,
                     "residual_fp64": r, "sigma_max": float(rec.s[0]), "sigma_min": float(rec.s[-1]),
                     "roofline_spmv": _roofline(st["spmv"]["bytes"], st["spmv"]["_ms"], st["spmv"]["_n"]),
                     "kernels_ms": {kname: v["ms_per_step"] for kname, v in st.items()}}
        log(f"las2 {prec}: {res[prec]['s_per_solve'] * 1e3:.1f} ms (host outputs {t_host * 1e3:.1f} ms), "
            f"steps {d.lanczos_steps}, residual {r:.2e}")
        del op
    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        o = oracle.OracleOp(a.astype(np.float64), nthreads=threads)
        if cfg1:
            t0 = time.perf_counter()
            oracle.las2(o, k, 0, (-1e-30, 1e-30), 1e-6, 42)
            cpu = {"value": time.perf_counter() - t0, "unit": "s", "cores": threads, "kind": "port",
                   "sample": "the FP64 C oracle's full LAS2 solve of the same matrix"}
        else:
            x = np.random.default_rng(0).normal(size=a.shape[1])
            reps = 10
            t0 = time.perf_counter()
            for _ in range(reps):
                o.opa(o.opa(x, False), True)
            per = (time.perf_counter() - t0) / reps
            ops = res["f64"]["lanczos_steps"] + 2 * k  # one A^T A x per step, A v and A^T A v per Ritz vector
            cpu = {"value": per * ops, "unit": "s", "cores": threads, "kind": "port",
                   "sample": (f"FP64 oracle A^T(A x) timed over {reps} calls ({per * 1e3:.1f} ms each), projected "
                              f"to {ops} products (Lanczos steps + 2 per Ritz vector); reorthogonalisation and "
                              "eigensolves not included")}
    f32 = res["f32"]
    if not cfg1:
        f32["roofline_spmv"]["traffic"] = traffic_from_profiles("spmv_cfg3_pmc.md")
    return {"metric": f"LAS2 seconds-to-solution, top-{k} singular triplets", "value": f32["s_per_solve"],
            "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": f32["s_per_solve"] * 1e3, "higher_is_better": False, "scaling": "strong",
            "vs_baseline": None, "dtype": "f32 values, f64 recurrence", "data": "synthetic",
            "config": {"workload": ("lanczos::svd_dim_seed(k=10, seed=42) on the 2,000 x 1,500 uniform CSR"
                                    if cfg1 else "lanczos::svd_dim_seed(k=100, seed=42) on the 500,000 x 30,000 "
                                    "Pareto-row CSR (nnz 19,698,610)"), "k": k, "nnz": int(a.nnz)},
            "roofline": f32["roofline_spmv"], "cpu_baseline": cpu, "precisions": res}


def _tuned_configs(nops):
    """The autotuner's choices for the batched SpMM rounds of nops operators
    (lanes-per-short-row shift, gathers in flight, workgroups per CU with 0 =
    the static launch, items per claim), read back from its tuning file."""
    path = os.environ.get("SVDH_TUNE_FILE") or os.path.join(os.path.expanduser("~"), ".cache", "svdh", "tune.txt")
    out = {}
    try:
        with open(path) as f:
            for line in f:
                key, cfg, ms = line.rstrip("\n").split("\t")
                if f" ops={nops} " in key:
                    d = key.split(" dir=")[1].split()[0]
                    dl, uu, wpc, claim = (int(x) for x in cfg.split())
                    out[f"dir{d}"] = {"lanes_shift": dl, "gathers_in_flight": uu, "workgroups_per_cu": wpc,
                                      "claim": claim, "launch": "static" if wpc == 0 else "persistent",
                                      "tuning_ms": float(ms)}
    except (OSError, ValueError):
        pass
    return out


def bench_batched(args):
    """64 randomized SVDs (rank 20, oversample 10, q = 3) over column masks of
    the 200,000 x 25,000 matrix, every SpMM round in one launch, U and V^T left
    in HBM, against the same 64 solves called one by one (also device outputs)."""
    import single_svdlib as sl
    from single_svdlib import _ffi, synth
    from single_svdlib.lanczos.masked import MaskedCSRMatrix
    a, masks = synth.config4(np.float32)
    t0 = time.perf_counter()
    base = sl.CsrMatrix.from_scipy(a)
    views = [MaskedCSRMatrix.with_columns(base, c) for c in masks]
    log(f"cfg4 uploaded and 64 masks built in {time.perf_counter() - t0:.1f}s")
    lib = _ffi.lib()
    ctx = base._ctx
    rank, over, q = 20, 10, 3
    m = a.shape[0]
    nb = len(views)
    us = [_ffi.DeviceArray((m, rank)) for _ in views]
    vts = [_ffi.DeviceArray((rank, len(c))) for c in masks]
    handles = (C.c_void_p * nb)(*[v._handle for v in views])
    up = (C.c_void_p * nb)(*[u.ptr for u in us])
    vp = (C.c_void_p * nb)(*[v.ptr for v in vts])
    ldu = (C.c_int64 * nb)(*[rank] * nb)
    ldvt = (C.c_int64 * nb)(*[len(c) for c in masks])

    def batched():
        metas = (_ffi.Result_ * nb)()
        _ffi.check(lib.svdh_rsvd_batched_device(ctx, nb, handles, rank, over, q, _ffi.NORM_QR, 7, 1, None,
                                                _ffi.PREC_F32, up, ldu, vp, ldvt, metas))
        s0 = [np.ctypeslib.as_array(metas[b].s, (rank,)).copy() for b in range(nb)]
        for b in range(nb):
            lib.svdh_result_free(C.byref(metas[b]))
        return s0

    # the first call tunes the persistent SpMM launch (unless the tuning file
    # already holds this workload): timed on its own
    t0 = time.perf_counter()
    batched()
    t_tune = time.perf_counter() - t0
    for _ in range(args.warmup):
        batched()
    _ffi.check(lib.svdh_ctx_synchronize(ctx))

    def timed():
        ts = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            sv = batched()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), sv

    t_batch, svals = timed()
    variants = {}
    for name, env in (("persistent_default_config", {"SVDH_BATCH_TUNE": "0"}),
                      ("static_launch", {"SVDH_BATCH_PERSISTENT": "0"})):
        os.environ.update(env)
        try:
            batched()
            variants[name], _ = timed()
        finally:
            for key in env:
                del os.environ[key]
    _ffi.check(lib.svdh_ctx_profile(ctx, 1))
    batched()
    st = _kernel_stats(lib, ctx, (("spmm", _ffi.KC_SPMM), ("spmm_t", _ffi.KC_SPMM_T), ("dense", _ffi.KC_DENSE)), 1)
    _ffi.check(lib.svdh_ctx_profile(ctx, 0))
    s_host = np.zeros(rank)

    def one(b):
        meta = _ffi.Result_()
        _ffi.check(lib.svdh_rsvd_device(ctx, views[b]._handle, rank, over, q, _ffi.NORM_QR, 0, 7, 1, None,
                                        _ffi.PREC_F32, 0.0, q, us[b].ptr, rank, _ffi.ptr(s_host), vts[b].ptr,
                                        len(masks[b]), C.byref(meta)))
    for b in range(nb):
        one(b)
    t0 = time.perf_counter()
    for b in range(nb):
        one(b)
    t_loop = time.perf_counter() - t0
    # results in host memory through svdh_rsvd_batched (one SvdRec per mask,
    # pinned pool): one warm-up, then the median of three
    handles = (C.c_void_p * nb)(*[v._handle for v in views])

    def batched_host():
        outs = (_ffi.Result_ * nb)()
        _ffi.check(lib.svdh_rsvd_batched(ctx, nb, handles, rank, over, q, _ffi.NORM_QR, 7, 1, None, _ffi.PREC_F32,
                                         outs))
        return outs

    def free_all(outs):
        for i in range(nb):
            lib.svdh_result_free(C.byref(outs[i]))

    free_all(batched_host())
    t_hosts = []
    for _ in range(3):
        t0 = time.perf_counter()
        outs = batched_host()
        t_hosts.append(time.perf_counter() - t0)
        free_all(outs)
    t_host = float(np.median(t_hosts))
    recs = sl.randomized.randomized_svd_batched(views, rank, over, q, "QR", seed=7)
    same = all(np.array_equal(r.s, sv.astype(r.s.dtype)) for r, sv in zip(recs, svals))
    # algorithmic bytes of one batched round: the shared CSR once, the remap
    # tables, every in-panel and out-panel (f32, l = 30)
    n = a.shape[1]
    l = rank + over
    sizes = np.array([len(c) for c in masks])
    fwd = a.nnz * 8 + (m + 1) * 4 + nb * n * 4 + int(sizes.sum()) * l * 4 + nb * m * l * 4
    rounds = st["spmm"]["_n"] + st["spmm_t"]["_n"]
    total_ms = st["spmm"]["_ms"] + st["spmm_t"]["_ms"]
    roof = _roofline(fwd * rounds, total_ms, rounds)
    check = []
    for i in (0, 1):
        sub = a[:, masks[i]].tocsr()
        check.append(_triplet_residual(sub, recs[i].u, recs[i].s.astype(np.float64), recs[i].vt))
    log(f"batched {t_batch * 1e3:.1f} ms (tuned persistent; default config "
        f"{variants['persistent_default_config'] * 1e3:.1f} ms, static launch {variants['static_launch'] * 1e3:.1f} ms, "
        f"first call with tuning {t_tune * 1e3:.0f} ms) vs one by one {t_loop * 1e3:.1f} ms; host outputs "
        f"{t_host * 1e3:.1f} ms")
    return {"metric": "batched masked randomized SVD seconds (64 masks)", "value": t_batch, "unit": "s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_batch * 1e3,
            "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": ("randomized_svd (rank 20, oversample 10, 3 power iterations, QR) over 64 "
                                    "column masks (sizes 500..8,000) of the 200,000 x 25,000 Poisson(60) CSR, "
                                    "U and V^T left in HBM"), "masks": nb, "nnz": int(a.nnz)},
            "roofline": roof, "cpu_baseline": None, "one_by_one_s": t_loop, "speedup_vs_one_by_one": t_loop / t_batch,
            "untuned_s": variants, "first_call_with_tuning_s": t_tune, "tuned_configs": _tuned_configs(nb),
            "host_output_s": t_host, "host_output_equals_device_output": same,
            "residual_q3_first_masks": check,
            "kernels_ms": {kname: v["ms_per_step"] for kname, v in st.items()}}


def bench_adjoint(args):
    """backward() of sum of the top-32 singular values and of |U_32^T y|^2
    with respect to the CSR values of the 100,000 x 20,000 matrix. The
    factors come from the GPU randomized SVD with the BF16 sketch ladder
    (rank 50, oversample 10, tol 1e-5); the backward runs with BF16 factor
    panels in the SDDMM (FP32 accumulation) and in FP32. Accuracy on the
    1,000 nonzeros of the FP64 CPU fixture (tests/golden/cfg5_grad.json) when
    it is present."""
    import single_svdlib as sl
    from single_svdlib import _ffi, synth
    a = synth.config2(np.float32)
    lib = _ffi.lib()
    k = 32
    y = np.random.default_rng(20261205).standard_normal(a.shape[0])  # the fixture's y
    op = sl.CsrMatrix.from_scipy(a, keep_bf16=True)
    ctx = op._ctx
    t0 = time.perf_counter()
    rec = sl.randomized.randomized_svd(op, 50, 10, 3, "QR", False, 42, False, tol=1e-5, max_power_iters=400,
                                       precision="bf16")
    t_fwd = time.perf_counter() - t0
    fix = None
    path = os.path.join(ROOT, "tests", "golden", "cfg5_grad.json")
    if os.path.exists(path):
        with open(path) as f:
            fix = json.load(f)
    out = {}
    for prec in ("bf16", "f32"):
        for loss in ("sigma_sum", "projection"):
            for _ in range(args.warmup):
                rec.backward(op, loss=loss, k=k, y=y, precision=prec, grad_dtype=np.float32)
            _ffi.check(lib.svdh_ctx_synchronize(ctx))
            ts = []
            for _ in range(args.steps):
                t0 = time.perf_counter()
                g, info = rec.backward(op, loss=loss, k=k, y=y, precision=prec, grad_dtype=np.float32)
                ts.append(time.perf_counter() - t0)
            r = {"s": float(np.median(ts)), "cg_iterations": info["cg_iterations"]}
            if fix is not None:
                idx = np.asarray(fix["idx"])
                ref = np.asarray(fix["grad_" + loss])
                r["max_err_vs_fp64_fixture"] = float(np.max(np.abs(g[idx] - ref)) / np.max(np.abs(ref)))
                pos = np.asarray(fix["fd_positions"])
                fd = np.asarray(fix["fd_" + loss])
                r["max_err_vs_central_differences"] = float(np.max(np.abs(g[idx[pos]] - fd)) / np.max(np.abs(ref)))
            out[f"{loss}_{prec}"] = r
    _ffi.check(lib.svdh_ctx_profile(ctx, 1))
    rec.backward(op, loss="sigma_sum", k=k, y=y, precision="bf16", grad_dtype=np.float32)
    st = _kernel_stats(lib, ctx, (("sddmm", _ffi.KC_SDDMM),), 1)
    _ffi.check(lib.svdh_ctx_profile(ctx, 0))
    roof = _roofline(st["sddmm"]["bytes"], st["sddmm"]["_ms"], st["sddmm"]["_n"])
    log(f"adjoint: forward (BF16 ladder) {t_fwd * 1e3:.1f} ms; " +
        ", ".join(f"{key} {v['s'] * 1e3:.2f} ms" for key, v in out.items()))
    return {"metric": "adjoint seconds (gradient of the top-32 singular values w.r.t. the CSR values)",
            "value": out["sigma_sum_bf16"]["s"], "unit": "s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": out["sigma_sum_bf16"]["s"] * 1e3, "higher_is_better": False,
            "scaling": "strong", "vs_baseline": None, "dtype": "bf16 factors, f32 accumulation", "data": "synthetic",
            "config": {"workload": ("backward() of sum of the top-32 singular values and of |U_32^T y|^2 on the "
                                    "100,000 x 20,000 Zipf CSR, factors from randomized_svd with the BF16 sketch "
                                    "(rank 50 + 10, tol 1e-5)"), "k": k, "nnz": int(a.nnz)},
            "roofline": roof, "cpu_baseline": None, "forward_s": t_fwd, "losses": out}


def bench_ingest(args):
    """COO input (svdh_coo_create): the 100,000 x 20,000 Zipf matrix as FP32
    triplets in a scrambled order with 1 % of the entries repeated (split into
    two halves), converted to the device CSR operator. The device conversion
    (radix sort by (row, column), run sums, row pointers) is timed with HIP
    events; the whole call also includes the uploads, the CSR download and
    the operator construction. CPU baseline: the oracle's COO -> CSR (C, one
    thread) on the same triplets."""
    import oracle
    import single_svdlib as sl
    from single_svdlib import _ffi, synth
    a = synth.config2(np.float32).tocoo()
    rng = np.random.default_rng(20261208)
    rep = rng.choice(a.nnz, a.nnz // 100, replace=False)
    half = (a.data[rep] * np.float32(0.5)).astype(np.float32)
    rows = np.concatenate([a.row, a.row[rep]]).astype(np.int64)
    cols = np.concatenate([a.col, a.col[rep]]).astype(np.int64)
    vals = np.concatenate([a.data, half]).astype(np.float32)
    vals[rep] = (a.data[rep] - half).astype(np.float32)
    perm = rng.permutation(rows.size)
    rows, cols, vals = rows[perm], cols[perm], vals[perm]
    m, n = a.shape
    lib = _ffi.lib()
    ctx = _ffi.context(0)
    for _ in range(args.warmup):
        sl.CooMatrix(rows, cols, vals, (m, n))
    _ffi.check(lib.svdh_ctx_synchronize(ctx))
    _ffi.check(lib.svdh_ctx_profile(ctx, 1))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        op = sl.CooMatrix(rows, cols, vals, (m, n))
    _ffi.check(lib.svdh_ctx_synchronize(ctx))
    total = (time.perf_counter() - t0) / args.steps
    st = _kernel_stats(lib, ctx, (("ingest", _ffi.KC_INGEST),), args.steps)
    _ffi.check(lib.svdh_ctx_profile(ctx, 0))
    dev_s = st["ingest"]["ms_per_step"] * 1e-3
    u = op.nnz()
    algo = rows.size * (8 + 8 + 4) + (m + 1) * 4 + u * (4 + 4)
    roof = _roofline(algo * args.steps, st["ingest"]["_ms"], args.steps)
    t0 = time.perf_counter()
    rp, oc, ov = oracle.coo_to_csr(m, n, rows, cols, vals)
    cpu_s = time.perf_counter() - t0
    ok = ov.size == u
    log(f"COO ingest: device conversion {dev_s * 1e3:.2f} ms, whole call {total * 1e3:.1f} ms, "
        f"{rows.size} triplets -> {u} entries (oracle {ov.size}); CPU oracle {cpu_s * 1e3:.0f} ms")
    return {"metric": "COO ingestion seconds (device conversion to CSR)", "value": dev_s, "unit": "s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": dev_s * 1e3, "higher_is_better": False,
            "scaling": "strong", "vs_baseline": None, "dtype": "int64 indices, f32 values", "data": "synthetic",
            "config": {"workload": ("svdh_coo_create of the 100,000 x 20,000 Zipf matrix as scrambled FP32 triplets "
                                    "with 1 % repeated pairs"), "triplets": int(rows.size), "entries": int(u),
                       "entries_match_oracle": bool(ok)},
            "roofline": roof,
            "cpu_baseline": {"value": cpu_s, "unit": "s", "cores": 1, "kind": "port",
                             "sample": "the oracle's COO -> CSR (two counting sorts, run sums) on the same triplets"},
            "whole_call_s": total}


def _probe_lib():
    from single_svdlib._ffi import PKG_DIR
    lib = C.CDLL(os.path.join(PKG_DIR, "libsvdh_probe.so"))
    lib.svdh_probe_gather.restype = C.c_int32
    lib.svdh_probe_gather.argtypes = [C.c_int32, C.POINTER(C.c_int32), C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                      C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int64), C.c_int32,
                                      C.POINTER(C.c_double)]
    return lib


def staged_columns():
    """The library's choice of popular columns for the headline workload
    (FP32 panel rows of 60 values, 256 bytes apart): A X stages as many as
    SVDH_SPMM_HOT_LDS bytes of LDS hold (default 81,920 -> 320); A^T Y runs
    SVDH_SPMM_THOT_H of them on the matrix cores (128 unless 256)."""
    lds = int(os.environ.get("SVDH_SPMM_HOT_LDS", "81920"))
    h_ax = max(4096, min(155648, lds)) // 256
    h_aty = 256 if os.environ.get("SVDH_SPMM_THOT_H") == "256" else 128
    return h_ax, h_aty


def gather_ceilings(a, h_ax=None, h_aty=None):
    """Gather ceilings (ms) of the two SpMM directions of the headline
    workload, FP32 rows of 60 values: A X's unstaged column stream (the h_ax
    most popular columns come from LDS: as many 256-byte panel rows as
    SVDH_SPMM_HOT_LDS bytes hold, default 81,920 -> 320) and A^T Y's unstaged
    row stream (the h_aty most popular columns run on the matrix cores; the
    library's choice, SVDH_SPMM_THOT_H, default 128) grouped by 8 source
    blocks (as its segment schedule runs), best of U in (4, 8) at 16 waves per
    CU (the configurations the full sweep of --workload gather-ceiling
    selects)."""
    d_ax, d_aty = staged_columns()
    h_ax = d_ax if h_ax is None else h_ax
    h_aty = d_aty if h_aty is None else h_aty
    lib = _probe_lib()
    m, n = a.shape
    counts = np.bincount(a.indices, minlength=n)
    order = np.argsort(-counts, kind="stable")
    is_hot = np.zeros(n, dtype=bool)
    is_hot[order[:h_ax]] = True
    ax = np.ascontiguousarray(a.indices[~is_hot[a.indices]].astype(np.int32))
    is_hot[:] = False
    is_hot[order[:h_aty]] = True
    csc = a.tocsc()
    keep = np.repeat(~is_hot, np.diff(csc.indptr))
    aty = csc.indices[keep].astype(np.int32)
    bw = (m + 7) // 8
    blk = aty // bw
    aty = np.ascontiguousarray(aty[np.argsort(blk, kind="stable")])
    off = np.zeros(9, dtype=np.int64)
    np.cumsum(np.bincount(blk, minlength=8), out=off[1:])
    out = {}
    for name, st, nsrc, o in (("spmm_ax", ax, n, None), ("spmm_aty", aty, m, off)):
        best = None
        for u in (4, 8):
            ms = C.c_double()
            rc = lib.svdh_probe_gather(0, st.ctypes.data_as(C.POINTER(C.c_int32)), st.size, nsrc, 256, 240, u, 256 * 16,
                                       5, o.ctypes.data_as(C.POINTER(C.c_int64)) if o is not None else None,
                                       8 if o is not None else 1, C.byref(ms))
            if rc != 0:
                raise RuntimeError(f"probe failed with {rc}")
            best = ms.value if best is None else min(best, ms.value)
        out[name] = {"ceiling_ms": best, "entries": int(st.size), "bytes": int(st.size) * 240}
    return out


def bench_gather_ceiling(args):
    """The measured ceiling of the SpMM's gathers on the 100,000 x 20,000
    workload (libsvdh_probe.so: one panel row per stream entry, 16 bytes per
    lane, no arithmetic and no output). Streams: the column indices of A in
    CSR order (A X gathers rows of the 20,000-row panel) and the row indices
    of A in CSC order (A^T Y gathers rows of the 100,000-row panel), each with
    and without the 256 most popular columns (the SpMM serves those from LDS
    or the matrix cores). FP32 rows of 60 values (240 bytes, 256 apart) and
    BF16 rows (120 bytes, 128 apart); sweeps of loads in flight per row group
    and waves per CU; the best rate of each stream is its ceiling."""
    from single_svdlib import synth
    from single_svdlib._ffi import PKG_DIR
    lib = C.CDLL(os.path.join(PKG_DIR, "libsvdh_probe.so"))
    lib.svdh_probe_gather.restype = C.c_int32
    lib.svdh_probe_gather.argtypes = [C.c_int32, C.POINTER(C.c_int32), C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                      C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int64), C.c_int32,
                                      C.POINTER(C.c_double)]
    a = synth.config2(np.float32)
    m, n = a.shape
    counts = np.bincount(a.indices, minlength=n)
    order = np.argsort(-counts, kind="stable")
    h_ax, h_aty = staged_columns()
    hot_ax = np.zeros(n, dtype=bool)
    hot_ax[order[:h_ax]] = True
    hot_aty = np.zeros(n, dtype=bool)
    hot_aty[order[:h_aty]] = True
    csc = a.tocsc()
    streams = {
        "ax_all": (a.indices.astype(np.int32), n),
        "ax_unstaged": (a.indices[~hot_ax[a.indices]].astype(np.int32), n),
        "aty_all": (csc.indices.astype(np.int32), m),
        "aty_unstaged": (np.concatenate([csc.indices[csc.indptr[c]:csc.indptr[c + 1]] for c in range(n)
                                         if not hot_aty[c]]).astype(np.int32), m),
    }
    # the same unstaged streams grouped by eight equal source blocks, each
    # block gathered by the workgroups of one XCD (the SpMM segment schedule)
    for name in ("ax_unstaged", "aty_unstaged"):
        st, nsrc = streams[name]
        bw = (nsrc + 7) // 8
        blk = st // bw
        order = np.argsort(blk, kind="stable")
        off = np.zeros(9, dtype=np.int64)
        np.cumsum(np.bincount(blk, minlength=8), out=off[1:])
        streams[name + "_xcd8"] = (st[order].astype(np.int32), nsrc, off)
    out = {}
    for name, spec in streams.items():
        st, nsrc = spec[0], spec[1]
        off = spec[2] if len(spec) > 2 else None
        st = np.ascontiguousarray(st)
        offp = off.ctypes.data_as(C.POINTER(C.c_int64)) if off is not None else None
        for dt, row, ld in (("f32", 240, 256), ("bf16", 120, 128)):
            best = None
            for u in (2, 4, 8):
                for wpc in (8, 16, 32):
                    ms = C.c_double()
                    rc = lib.svdh_probe_gather(0, st.ctypes.data_as(C.POINTER(C.c_int32)), st.size, nsrc, ld, row, u,
                                               256 * wpc, 5, offp, 8 if off is not None else 1, C.byref(ms))
                    if rc != 0:
                        raise RuntimeError(f"probe failed with {rc}")
                    tbs = st.size * row / (ms.value * 1e-3) / 1e12
                    if best is None or tbs > best["tb_s"]:
                        best = {"tb_s": tbs, "ms": ms.value, "u": u, "waves_per_cu": wpc}
            best["entries"] = int(st.size)
            best["bytes"] = int(st.size) * row
            out[f"{name}_{dt}"] = best
            log(f"gather ceiling {name} {dt}: {best['tb_s']:.1f} TB/s ({best['ms'] * 1e3:.1f} us, U={best['u']}, "
                f"{best['waves_per_cu']} waves/CU, {st.size} entries)")
    return {"metric": "SpMM gather ceiling (TB/s of panel rows gathered, 100,000 x 20,000 Zipf CSR)",
            "value": out["ax_unstaged_f32"]["tb_s"], "unit": "TB/s", "n_gpus": 1, "higher_is_better": True,
            "data": "synthetic", "config": {"workload": "panel-row gathers of the SpMM column streams, no arithmetic"},
            "streams": out}


OTHER_WORKLOADS = {"las2-cfg1": bench_las2, "las2-cfg3": bench_las2, "batched": bench_batched,
                   "adjoint": bench_adjoint, "coo-ingest": bench_ingest, "gather-ceiling": bench_gather_ceiling}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="f32", choices=["f32", "bf16"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="only the headline solve (no other configurations)")
    ap.add_argument("--profile-run", action="store_true",
                    help="for rocprofv3 runs: the timed solves plus one 256 MiB device-to-device copy (the calibration "
                         "kernel for FETCH_SIZE / WRITE_SIZE), without the correctness and reference extras")
    ap.add_argument("--max-iters", type=int, default=400)
    ap.add_argument("--workload", default="rsvd",
                    choices=["rsvd", "las2-cfg1", "las2-cfg3", "batched", "adjoint", "coo-ingest", "gather-ceiling"],
                    help="rsvd (default, the headline metric) or one of the other measured paths")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the multi-GPU sharding of the "
              "path (row blocks of A, one all-reduce per product) is designed, not built or measured",
              file=sys.stderr)
        sys.exit(2)
    if args.workload != "rsvd":
        print(json.dumps(OTHER_WORKLOADS[args.workload](args)))
        return

    import single_svdlib as sl
    from single_svdlib import _ffi, synth

    t0 = time.perf_counter()
    a = synth.config2(np.float32)
    t_gen = time.perf_counter() - t0
    log(f"generated cfg2 {a.shape} nnz={a.nnz} in {t_gen:.1f}s")
    t0 = time.perf_counter()
    op = sl.CsrMatrix.from_scipy(a, keep_bf16=args.precision == "bf16")
    lib = _ffi.lib()
    ctx = op._ctx
    _ffi.check(lib.svdh_ctx_synchronize(ctx))
    t_upload = time.perf_counter() - t0
    log(f"uploaded in {t_upload:.2f}s")
    prec = _ffi.PREC_BF16 if args.precision == "bf16" else _ffi.PREC_F32

    m, n = a.shape
    k = min(RANK, m, n)
    u_dev = _ffi.DeviceArray((m, k))
    vt_dev = _ffi.DeviceArray((k, n))
    s_host = np.zeros(k)

    def solve(tol, q):
        """One solve with U and V^T left in HBM."""
        meta = _ffi.Result_()
        _ffi.check(lib.svdh_rsvd_device(ctx, op._handle, RANK, OVERSAMPLE, q, _ffi.NORM_QR, 0, 42, 1, None, prec,
                                        tol, args.max_iters, u_dev.ptr, k, _ffi.ptr(s_host), vt_dev.ptr, n,
                                        C.byref(meta)))
        return meta.power_iters_run, meta.residual_estimate

    def solve_host(tol, q):
        """One solve returning U, S, V^T in host memory (the measured step):
        randomized_svd's SvdRec, through svdh_rsvd."""
        res = _ffi.Result_()
        _ffi.check(lib.svdh_rsvd(ctx, op._handle, RANK, OVERSAMPLE, q, _ffi.NORM_QR, 0, 42, 1, None, prec,
                                 tol, args.max_iters, C.byref(res)))
        return res

    # the first solve after the upload builds the operator's schedules and
    # popular-column layouts: timed on its own
    t0 = time.perf_counter()
    first = solve_host(TOL, 3)
    t_cold = time.perf_counter() - t0
    lib.svdh_result_free(C.byref(first))
    log(f"cold first solve {t_cold * 1e3:.1f} ms")
    for i in range(args.warmup):
        lib.svdh_result_free(C.byref(solve_host(TOL, 3)))
    _ffi.check(lib.svdh_ctx_synchronize(ctx))
    # timed: K solves to results on the host, each timed on its own; value is
    # their median
    times, iters = [], []
    t_all = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        res = solve_host(TOL, 3)
        times.append(time.perf_counter() - t0)
        iters.append((res.power_iters_run, res.residual_estimate))
        if i < args.steps - 1:
            lib.svdh_result_free(C.byref(res))
    _ffi.check(lib.svdh_ctx_synchronize(ctx))
    wall = (time.perf_counter() - t_all) / args.steps
    per_step = float(np.median(times))
    rec_host = sl.utils.result_to_svdrec(res, np.float64)
    log(f"timed {args.steps} host-output solves: median {per_step * 1e3:.2f} ms (min {min(times) * 1e3:.2f}, "
        f"max {max(times) * 1e3:.2f}), power iterations {sorted(set(i[0] for i in iters))}")

    # the same solve with U and V^T left in HBM
    dev_times = []
    for i in range(args.steps):
        t0 = time.perf_counter()
        solve(TOL, 3)
        dev_times.append(time.perf_counter() - t0)
    t_dev = float(np.median(dev_times))
    log(f"device-resident solves: median {t_dev * 1e3:.2f} ms")

    # The same device-resident solves again with HIP events around every
    # kernel launch on the library's stream: per-kernel averages for the
    # roofline.
    _ffi.check(lib.svdh_ctx_profile(ctx, 1))
    t0 = time.perf_counter()
    for i in range(args.steps):
        solve(TOL, 3)
    _ffi.check(lib.svdh_ctx_synchronize(ctx))
    per_step_prof = (time.perf_counter() - t0) / args.steps
    log(f"profiled {args.steps} solves: {per_step_prof * 1e3:.1f} ms each")
    ms, nl, by = C.c_double(), C.c_int64(), C.c_double()
    kc = {}
    for name, kk in (("spmm_ax", _ffi.KC_SPMM), ("spmm_aty", _ffi.KC_SPMM_T), ("dense", _ffi.KC_DENSE),
                     ("spmv", _ffi.KC_SPMV)):
        _ffi.check(lib.svdh_ctx_kernel_time(ctx, kk, C.byref(ms), C.byref(nl), C.byref(by)))
        kc[name] = {"ms_per_solve": ms.value / args.steps, "launches_per_solve": nl.value / args.steps,
                    "avg_launch_us": ms.value * 1e3 / max(nl.value, 1),
                    "gbs": by.value / (ms.value * 1e-3) / 1e9 if ms.value else 0.0,
                    "_ms": ms.value, "_n": nl.value, "_bytes": by.value}
    # the dominant kernel: SpMM, both directions (A X and A^T Y) together
    sp_ms = kc["spmm_ax"]["_ms"] + kc["spmm_aty"]["_ms"]
    sp_n = kc["spmm_ax"]["_n"] + kc["spmm_aty"]["_n"]
    sp_bytes = kc["spmm_ax"]["_bytes"] + kc["spmm_aty"]["_bytes"]
    spmm_avg_ms = sp_ms / max(sp_n, 1)
    spmm_bytes = sp_bytes / max(sp_n, 1)
    achieved = spmm_bytes / (spmm_avg_ms * 1e-3) / 1e9 if sp_n else 0.0
    per_dir = {}
    for name in ("spmm_ax", "spmm_aty"):
        v = kc[name]
        per_dir[name] = {"achieved": v["gbs"], "frac": v["gbs"] / HBM_PEAK_GBS,
                         "algorithmic_bytes_per_launch": v["_bytes"] / max(v["_n"], 1),
                         "avg_launch_ms": v["_ms"] / max(v["_n"], 1)}
    for v in kc.values():
        for key in ("_ms", "_n", "_bytes"):
            v.pop(key)
    _ffi.check(lib.svdh_ctx_profile(ctx, 0))
    if not args.profile_run:
        # the measured ceiling of each direction's gathers (libsvdh_probe.so)
        try:
            ceil = gather_ceilings(a)
            for name, c in ceil.items():
                per_dir[name]["gather_ceiling_ms"] = c["ceiling_ms"]
                per_dir[name]["gather_ceiling_tb_s"] = c["bytes"] / (c["ceiling_ms"] * 1e-3) / 1e12
                per_dir[name]["frac_of_gather_ceiling"] = c["ceiling_ms"] / per_dir[name]["avg_launch_ms"]
            log("gather ceilings: " + ", ".join(f"{k} {v['gather_ceiling_ms'] * 1e3:.1f} us "
                                                f"({v['frac_of_gather_ceiling']:.2f} of it)" for k, v in per_dir.items()))
        except Exception as e:
            log(f"gather ceiling probe failed: {e!r}")

    if args.profile_run:
        # plain copy of known size: 256 MiB read + 256 MiB written
        nb = 256 << 20
        src, dst = _ffi.DeviceArray((nb // 8,)), _ffi.DeviceArray((nb // 8,))
        for _ in range(2):
            if _ffi.hip().hipMemcpy(dst.ptr, src.ptr, nb, 3) != 0:
                raise RuntimeError("device-to-device copy failed")
        log("calibration copy done (2 x 268435456 bytes)")
        print(json.dumps({"metric": METRIC, "value": per_step, "unit": "s", "n_gpus": 1, "steps": args.steps,
                          "warmup": args.warmup, "ms_per_step": per_step * 1e3, "profile_run": True,
                          "kernels": kc}))
        return

    # correctness of the timed configuration (outside the timed region): the
    # last timed solve's host triplets in FP64
    r_rel, r_norm = residual(a, rec_host.u, rec_host.s, rec_host.vt)
    log(f"FP64 residual of the returned top-{RANK}: per-sigma {r_rel:.2e}, |A|-relative {r_norm:.2e}")

    # The reference's algorithm (plain power iterations, no shift or momentum)
    # to the same tolerance on the GPU: its iteration count sizes the CPU
    # baseline, and its time shows what the acceleration buys.
    os.environ["SVDH_RSVD_SHIFT"] = "0"
    os.environ["SVDH_RSVD_MOMENTUM"] = "0"
    try:
        plain_t = []
        for _ in range(3):
            t0 = time.perf_counter()
            r = solve_host(TOL, 3)
            plain_t.append(time.perf_counter() - t0)
            plain = (r.power_iters_run, r.residual_estimate)
            lib.svdh_result_free(C.byref(r))
        t_plain = float(np.median(plain_t))
    finally:
        del os.environ["SVDH_RSVD_SHIFT"]
        del os.environ["SVDH_RSVD_MOMENTUM"]
    log(f"plain power iterations: {plain[0]} iterations, {t_plain * 1e3:.1f} ms")

    # the configuration as specified (3 power iterations) for reference
    q3_t = []
    for _ in range(5):
        t0 = time.perf_counter()
        res3 = solve_host(0.0, 3)
        q3_t.append(time.perf_counter() - t0)
        if len(q3_t) < 5:
            lib.svdh_result_free(C.byref(res3))
    t_q3 = float(np.median(q3_t))
    rec3 = sl.utils.result_to_svdrec(res3, np.float32)
    r3_rel, _ = residual(a, rec3.u, rec3.s.astype(np.float64), rec3.vt)
    log(f"q=3 solve {t_q3 * 1e3:.1f} ms, residual {r3_rel:.2e}")

    extras = {}
    if not args.no_extras:
        del op
        extras = other_paths(args)

    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(a, int(plain[0]), threads)
        log(f"cpu baseline {cpu['value']:.2f}s ({cpu['sample']})")

    traffic = traffic_from_profiles()
    out = {
        "metric": METRIC,
        "value": per_step,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": per_step * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "f32" else "bf16+f32",
        "data": "synthetic",
        "config": {
            "workload": ("randomized_svd on the 100,000 x 20,000 CSR (row nnz ~ Poisson(100), Zipf(1.1) columns, "
                         "log1p(Poisson(2)+1) values, nnz 9,998,163), rank 50, oversample 10, QR normaliser, "
                         "power iterations until the top-50 residual <= 1e-5, U, S, V^T returned in host memory"),
            "rank": RANK, "oversample": OVERSAMPLE, "tol": TOL, "nnz": int(a.nnz),
            "power_iters": sorted(set(int(i[0]) for i in iters)),
        },
        "roofline": {
            "bound": "hbm", "kernel": "spmm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "algorithmic_bytes_per_launch": spmm_bytes, "avg_launch_ms": spmm_avg_ms, "per_direction": per_dir,
        },
        "cpu_baseline": cpu,
        "timing": {"statistic": f"median of {args.steps} solves after {args.warmup} warm-up",
                   "mean_s": wall, "min_s": min(times), "max_s": max(times)},
        "device_resident_solve_s": t_dev,
        "upload_s": t_upload,
        "generate_s": t_gen,
        "cold_first_solve_s": t_cold,
        "residual_fp64": {"per_sigma": r_rel, "norm_relative": r_norm},
        "kernels": kc,
        "ms_per_step_profiled": per_step_prof * 1e3,
        "plain_power_iterations": {"iterations": int(plain[0]), "s": t_plain},
        "q3_solve_s": t_q3,
        "q3_residual_fp64": r3_rel,
    }
    out.update(extras)
    print(json.dumps(out))


def other_paths(args):
    """The other configurations, timed in the same run (medians of 5 solves
    after 1 warm-up): LAS2 at 2,000 x 1,500 (k = 10) and 500,000 x 30,000
    (k = 100, FP32-SpMV/FP64-recurrence, with the SpMV roofline), the 64-mask
    batched randomized SVD, and the adjoint on the 100,000 x 20,000 matrix."""
    import types
    sub = types.SimpleNamespace(steps=5, warmup=1, no_cpu=True, profile_run=False)
    out = {}

    def run(name, fn, wl):
        t0 = time.perf_counter()
        try:
            r = fn(types.SimpleNamespace(**vars(sub), workload=wl))
        except Exception as e:  # an extra must not lose the headline line
            log(f"{name} failed: {e!r}")
            out[name + "_error"] = repr(e)
            return None
        log(f"{name} done in {time.perf_counter() - t0:.1f}s")
        return r

    r = run("las2_cfg1", bench_las2, "las2-cfg1")
    if r:
        out["las2_cfg1_s"] = r["precisions"]["f64"]["s_per_solve"]
        out["las2_cfg1_host_output_s"] = r["precisions"]["f64"]["host_output_s"]
    r = run("las2_cfg3", bench_las2, "las2-cfg3")
    if r:
        out["las2_cfg3_s"] = r["value"]
        out["las2_cfg3_f64_s"] = r["precisions"]["f64"]["s_per_solve"]
        out["las2_cfg3_host_output_s"] = r["precisions"]["f32"]["host_output_s"]
        out["las2_cfg3_spmv_roofline"] = r["roofline"]
        out["las2_cfg3_lanczos_steps"] = r["precisions"]["f32"]["lanczos_steps"]
    r = run("batched_cfg4", bench_batched, "batched")
    if r:
        out["batched_cfg4_s"] = r["value"]
        out["batched_cfg4_one_by_one_s"] = r["one_by_one_s"]
        out["batched_cfg4_roofline"] = r["roofline"]
    r = run("adjoint_cfg5", bench_adjoint, "adjoint")
    if r:
        out["adjoint_cfg5_s"] = r["value"]
        out["adjoint_cfg5"] = {key: r[key] for key in r if key not in ("metric", "n_gpus", "steps", "warmup",
                                                                       "higher_is_better", "scaling", "vs_baseline",
                                                                       "data", "cpu_baseline")}
    return out


if __name__ == "__main__":
    main()
