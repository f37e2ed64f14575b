#!/usr/bin/env python3
"""Benchmark: batch-SOM epoch throughput on one MI355X.

Headline workload (one "step" = one full epoch over the dataset resident in HBM): 64x64 hexagonal map
(4,096 neurons), dim 16, 12,582,912 samples (one per HEALPix nside=1024 pixel) of synthetic
systematics (8 lognormal, 6 Gaussian, sin/cos of the pixel-centre colatitude, standardised),
sigma0 = 32 on the reference's 20-epoch asymptotic schedule, init_rand(seed=0). An epoch is the
reference's whole epoch: BMU search, per-BMU accumulation, neighbourhood update and division, plus
that epoch's quantization/topographic errors (fused into the next BMU pass).

The timed epochs are epochs of the schedule itself: after the warm-up the weights are reset to the
initial weights, and a schedule restarts from them whenever the timed epochs run past its end, so
--steps 20 times exactly one 20-epoch training from init_rand(0).

Legs on the same workload:
  fp32                 plain FP32 MFMA distance GEMM (the baseline of the speed-up),
  bf16_rerank / fp64   split-BF16 MFMA GEMM + exact FP64 rerank: the leg the speed and accuracy bars
                       are claimed on (value),
  bf16_rerank / ref32  the same GEMM + the reference's FP32-difference rerank: the parity leg, whose BMUs
                       equal the reference formula's (agreement with the REF32 oracle).
Accuracy per leg against the oracle: BMU agreement with the FP64 and REF32 oracles at the final weights,
quantisation error, centroid error after one epoch and after the whole schedule (trajectory) against the
FP64 oracle. A second workload (100x100/D32, nside 2048, 10-epoch schedule) is measured the same way.
Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SOM epoch samples/s 64x64 D16, bf16+rerank vs fp32; BMU serve p50/p99 ms"
PEAK_BF16_TFLOPS = 2500.0
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
BARS = {"speedup_vs_fp32": 3.0, "bmu_agreement_vs_fp64": 0.99999, "centroid_rel_err": 1e-5}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed epochs (the schedule restarts past its end)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nside", type=int, default=1024)
    ap.add_argument("--map", type=int, default=64)
    ap.add_argument("--lognormal", type=int, default=8, help="lognormal feature columns")
    ap.add_argument("--gaussian", type=int, default=6, help="Gaussian feature columns (+2 angular)")
    ap.add_argument("--epochs-schedule", type=int, default=20)
    ap.add_argument("--arithmetic", default="fp64", help="rerank arithmetic of the headline bf16_rerank leg")
    ap.add_argument("--no-parity-leg", action="store_true", help="skip the bf16_rerank leg with ref32 rerank")
    ap.add_argument("--no-fp32", action="store_true", help="skip the plain FP32 leg")
    ap.add_argument("--acc-subsample", type=int, default=1 << 20, help="rows of the FP64-oracle agreement")
    ap.add_argument("--ref32-subsample", type=int, default=1 << 17, help="rows of the REF32-oracle agreement")
    ap.add_argument("--traj-subsample", type=int, default=1 << 17, help="rows of the trajectory comparison")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--stream-chunks", type=int, default=8,
                    help="chunks of the streamed-epoch leg (0 skips it)")
    ap.add_argument("--no-pipeline", dest="pipeline", action="store_false",
                    help="skip the covariance / gndm leg")
    ap.add_argument("--c0", action="store_true",
                    help="also time the 10x10/D8/100k CPU-runnable configuration (CPU oracle and GPU)")
    ap.add_argument("--no-c3", dest="c3", action="store_false",
                    help="skip the 100x100/D32 nside-2048 workload")
    ap.add_argument("--c3-steps", type=int, default=10)
    ap.add_argument("--c3-warmup", type=int, default=1)
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one device-to-device copy of the dataset (reference kernel for HBM counters)")
    return ap.parse_args()


def run_schedule(ctx, x_dev, n, W0, topo, m, prec, arith, steps, warmup, E, sigma0):
    """Times `steps` epochs of the E-epoch schedule from W0 (restarting from W0 at each schedule end),
    after `warmup` untimed epochs (then reset to W0). Returns the timing, the per-kernel event timers and
    the weights at the end of a complete schedule (deterministic: every complete run gives the same)."""
    from decontamination_amd import _lib
    from decontamination_amd.runtime import Accumulator, DeviceArray, Model
    pc, ac = _lib.PRECISIONS[prec], _lib.ARITHMETICS[arith]
    mn, dim = W0.shape
    model = Model(ctx, m, m, dim, "hexagonal", topo)
    w0 = DeviceArray.from_host(ctx, W0)
    acc = Accumulator(model)
    errs = DeviceArray.zeros(ctx, (E, 3), np.float64)
    hist = DeviceArray.empty(ctx, (E, mn, dim), np.float32)

    def epoch(k):
        e = k % E
        if e == 0:
            model.set_weights(w0)  # a schedule starts from the initial weights (device copy, async)
        acc.reset()
        acc.epoch_accumulate(x_dev.ptr, None, n, pc, ac, errs.view_ptr((e - 1) * 3) if e > 0 else None)
        acc.epoch_update(e, E, sigma0, hist.view_ptr(e * mn * dim))

    for k in range(warmup):
        epoch(k)
    ctx.synchronize()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for k in range(steps):
        epoch(k)
    ctx.synchronize()
    t1 = time.perf_counter()
    timing = ctx.timing()
    ctx.set_timing(False)
    k = steps
    while k % E != 0:  # finish the schedule untimed
        epoch(k)
        k += 1
    ctx.synchronize()
    W = hist.copy_to_host()[E - 1]
    for a in (errs, hist, w0):
        a.free()
    acc.close()
    model.close()
    return {"seconds": t1 - t0, "timing": timing, "W": W, "steps": steps,
            "samples_per_s": n * steps / (t1 - t0), "ms_per_step": 1e3 * (t1 - t0) / steps}


def streaming_bench(ctx, X, W0, topo, m, sigma0, E, n_chunks, arith, steps=4, warmup=1):
    """The same bf16_rerank epoch with X fed from host memory as n_chunks generator chunks: through the
    uploader ring (copy of chunk k+1 overlapping the kernels of chunk k), and serially (synchronous copy,
    then the kernels), next to the copies alone. Host-to-device rates include PCIe."""
    from decontamination_amd import _lib
    from decontamination_amd.runtime import Accumulator, DeviceArray, Model, Uploader
    pc, ac = _lib.PREC_BF16_RERANK, _lib.ARITHMETICS[arith]
    N, dim = X.shape
    mn = W0.shape[0]
    rows = -(-N // n_chunks)
    chunks = [X[i:i + rows] for i in range(0, N, rows)]
    model = Model(ctx, m, m, dim, "hexagonal", topo)
    model.set_weights(W0)
    acc = Accumulator(model)
    up = Uploader(ctx, 2)
    errs = DeviceArray.zeros(ctx, (warmup + steps + 1, 3), np.float64)
    hist = DeviceArray.empty(ctx, (mn, dim), np.float32)
    serial_buf = DeviceArray.empty(ctx, (rows, dim), np.float32)

    def epoch(e, serial):
        acc.reset()
        for c in chunks:
            if serial:
                _lib.call("somhip_memcpy_h2d", ctx.handle, _lib.vp(serial_buf.ptr), _lib.vp(c), c.nbytes)
                xp = serial_buf.ptr
            else:
                xp, _ = up.put(c)
            acc.epoch_accumulate(xp, None, len(c), pc, ac, errs.view_ptr((e - 1) * 3) if e > 0 else None)
        acc.epoch_update(e, E, sigma0, hist.ptr)

    def timed(fn):
        for e in range(warmup):
            fn(e)
        ctx.synchronize()
        t0 = time.perf_counter()
        for e in range(warmup, warmup + steps):
            fn(e)
        ctx.synchronize()
        return (time.perf_counter() - t0) / steps

    def copies(_e):
        for c in chunks:
            up.put(c)

    copy_s = timed(copies)
    model.set_weights(W0)
    overlap_s = timed(lambda e: epoch(e, False))
    model.set_weights(W0)
    serial_s = timed(lambda e: epoch(e, True))
    for a in (errs, hist, serial_buf):
        a.free()
    acc.close()
    up.close()
    model.close()
    return {"chunks": len(chunks), "chunk_rows": rows, "epochs_timed": steps,
            "copies_only_ms": round(1e3 * copy_s, 3), "h2d_GB_per_s": round(X.nbytes / copy_s / 1e9, 2),
            "overlapped_ms_per_epoch": round(1e3 * overlap_s, 3), "serial_ms_per_epoch": round(1e3 * serial_s, 3),
            "overlapped_samples_per_s": round(N / overlap_s, 1),
            "note": "bf16_rerank epoch with X streamed from pageable host memory as generator chunks; "
                    "PCIe-inclusive, not the headline value"}


def omp_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def pipeline_bench(ctx, X, x_dev, W, m, arith, steps=20):
    """The steps on either side of training, on the same footprint resident in HBM: the weighted
    covariance of the SOM_PCA initialisation (two passes over X) and the galaxy-number-density maps
    from the winners (somhip_gndm), timed per call (wall clock, synchronised), with the oracle's
    restatements of the reference's CPU loops on a prefix as the CPU baseline (Welford: one thread, as
    the reference's loop is sequential; activation maps: parallel over pixels, as its prange kernel)."""
    import ctypes
    from decontamination_amd import _lib
    from decontamination_amd.runtime import DeviceArray, Model
    import oracle
    from oracle import ptr
    N, dim = X.shape
    mn = W.shape[0]
    rng = np.random.default_rng(1)
    rho = rng.poisson(3.0, N) / (0.5 + 0.5 * rng.random(N))  # corrected density: counts / coverage
    r_dev = DeviceArray.from_host(ctx, rho)
    out = np.empty(1 + dim + dim * dim)

    def cov():
        _lib.call("somhip_covariance_chunk", ctx.handle, ctypes.c_void_p(x_dev.ptr), 0, ctypes.c_void_p(r_dev.ptr),
                  N, dim, _lib.vp(out))

    model = Model(ctx, m, m, dim, "hexagonal", np.zeros((mn, 2), np.float32))
    model.set_weights(W)
    win = DeviceArray.empty(ctx, (N,), np.int32)
    model.find_bmus(x_dev.ptr, N, win.ptr, _lib.PREC_BF16_RERANK, _lib.ARITHMETICS[arith])
    maps = DeviceArray.empty(ctx, (3, mn), np.float32)
    gm = DeviceArray.empty(ctx, (N,), np.float32)

    def gndm():
        _lib.call("somhip_gndm", ctx.handle, ctypes.c_void_p(win.ptr), ctypes.c_void_p(r_dev.ptr), N, mn,
                  ctypes.c_void_p(maps.view_ptr(0)), ctypes.c_void_p(maps.view_ptr(mn)),
                  ctypes.c_void_p(maps.view_ptr(2 * mn)), ctypes.c_void_p(gm.ptr))

    def timed(fn):
        fn()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        ctx.synchronize()
        return (time.perf_counter() - t0) / steps

    t_cov = timed(cov)
    t_gndm = timed(gndm)
    for a in (r_dev, win, maps, gm):
        a.free()
    model.close()
    L = oracle.lib()
    p = 1 << 20
    xs = np.ascontiguousarray(X[:p], np.float64)
    sw, mean, m2, dl = np.zeros(1), np.zeros(dim), np.zeros((dim, dim)), np.zeros(dim)
    t0 = time.perf_counter()
    L.orc_welford_update(ptr(sw), ptr(mean), ptr(m2), ptr(dl), ptr(xs), ptr(rho[:p]), p, dim)
    cpu_cov = p / (time.perf_counter() - t0)
    threads = omp_threads()
    q = 1 << 16
    xq = np.ascontiguousarray(X[:q])
    res = np.zeros(mn)
    tmaps = np.zeros((threads, mn))
    Wc = np.ascontiguousarray(W, np.float32)
    t0 = time.perf_counter()
    L.orc_count_bmus_omp(0, ptr(Wc), mn, dim, ptr(xq), ptr(rho[:q]), q, ptr(res), ptr(tmaps), threads)
    L.orc_count_bmus_omp(0, ptr(Wc), mn, dim, ptr(xq), None, q, ptr(res), ptr(tmaps), threads)
    cpu_gndm = q / (time.perf_counter() - t0)
    algo_cov = N * (4 * dim + 8)
    algo_gndm = N * (4 + 8 + 4)
    return {
        "covariance": {"ms": round(1e3 * t_cov, 4), "rows_per_s": round(N / t_cov, 1),
                       "algorithmic_GB_per_s": round(algo_cov / t_cov / 1e9, 1),
                       "note": f"weighted covariance of the {N}x{dim} footprint (SOM_PCA init), 2 passes over X, "
                               "wall clock per call incl. the 2 KB result copy",
                       "cpu_baseline": {"value": round(cpu_cov, 1), "unit": "rows/s", "cores": 1, "kind": "port",
                                        "sample": f"oracle Welford loop (sequential, as the reference's) on the "
                                                  f"first {p} rows"}},
        "gndm": {"ms": round(1e3 * t_gndm, 4), "pixels_per_s": round(N / t_gndm, 1),
                 "algorithmic_GB_per_s": round(algo_gndm / t_gndm / 1e9, 1),
                 "note": "catalog + footprint activation maps from the winners, gnd and gndm = gnd[winners]; the "
                         "reference runs two more BMU searches over the footprint for the maps",
                 "cpu_baseline": {"value": round(cpu_gndm, 1), "unit": "pixels/s", "cores": threads, "kind": "port",
                                  "sample": f"oracle _count_bmus_kernel loop, parallel over pixels (two maps) on "
                                            f"the first {q} pixels"}},
    }


PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")


def profile_figure(filename, key, kernel_symbol):
    """A per-kernel figure `key kernel_symbol: value` from a committed rocprofv3 counter summary, or None."""
    try:
        for line in open(os.path.join(PROFILES, filename)):
            if line.startswith(f"{key} {kernel_symbol}:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def roofline(timing, kernel, flops_per_launch, peak_tflops, traffic=None, mfma_busy=None):
    if kernel not in timing:
        return None
    ms, launches = timing[kernel]
    avg_ms = ms / launches
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    out = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak_tflops, "unit": "TFLOP/s",
           "frac": round(achieved / peak_tflops, 5), "traffic": traffic, "kernel": kernel,
           "avg_ms": round(avg_ms, 4), "launches": launches}
    if mfma_busy is not None:
        out["mfma_busy"] = mfma_busy
    return out


def cpu_baseline(X, W0, topo, sigma0, E, target_s=12.0):
    """The oracle's parallel restatement of the reference CPU epoch (prange + FP32 atomics + errors
    pass) on a prefix of the same workload; extrapolated per sample."""
    import oracle
    from oracle import ptr
    L = oracle.lib()
    mn, dim = W0.shape
    threads = omp_threads()

    def run(p):
        xs = np.ascontiguousarray(X[:p])
        num = np.zeros((mn, dim), np.float32)
        den = np.zeros(mn, np.float32)
        err = np.zeros(3, np.float64)
        t0 = time.perf_counter()
        L.orc_train_epoch_accumulate_omp(ptr(num), ptr(den), ptr(W0), ptr(topo), ptr(xs), None, p, 0, E,
                                         np.float32(sigma0), mn, dim)
        Wn = np.empty((mn, dim), np.float32)
        L.orc_divide(0, ptr(num), ptr(den), mn, dim, ptr(Wn))
        L.orc_compute_errors_omp(ptr(Wn), ptr(topo), mn, dim, ptr(xs), None, p, 1.0, ptr(err))
        return time.perf_counter() - t0

    p = 1024
    t = run(p)
    p2 = int(min(len(X), max(p, p * target_s / max(t, 1e-3))))
    p2 = max(1024, min(p2, 1 << 17))
    t2 = run(p2)
    return {"value": round(p2 / t2, 2), "unit": "samples/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"first {p2} samples of the same 64x64/D16 workload, one epoch (reference loop nest: "
                      f"parallel over samples, FP32 atomic scatter to all {mn} neurons, plus the errors pass), "
                      f"{t2:.2f} s; extrapolated per sample"}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def c0_bench():
    """The reference's CPU-runnable configuration in full: 10x10 square map, D=8, 100,000 samples of a
    seeded 5-Gaussian mixture, 10 epochs from sigma0=5: the oracle's OpenMP restatement of the
    reference CPU epoch (median of 3 runs) beside the GPU path on the same inputs."""
    import decontamination_amd as dam
    import oracle
    from oracle import ptr
    from oracle import som_ref as R
    rng = np.random.default_rng(0)
    means = rng.uniform(-3, 3, (5, 8))
    labels = rng.integers(0, 5, 100_000)
    X = (means[labels] + 0.5 * rng.standard_normal((100_000, 8))).astype(np.float32)
    W0 = rng.uniform(-1, 1, (100, 8)).astype(np.float32)
    m, E, sigma0 = 10, 10, 5.0
    topo = R.topography(m, m, "square")
    L = oracle.lib()

    def cpu_run():
        W = W0.copy()
        t0 = time.perf_counter()
        for e in range(E):
            num = np.zeros((100, 8), np.float32)
            den = np.zeros(100, np.float32)
            L.orc_train_epoch_accumulate_omp(ptr(num), ptr(den), ptr(W), ptr(topo), ptr(X), None, len(X), e, E,
                                             np.float32(sigma0), 100, 8)
            Wn = np.empty((100, 8), np.float32)
            L.orc_divide(0, ptr(num), ptr(den), 100, 8, ptr(Wn))
            err = np.zeros(3)
            L.orc_compute_errors_omp(ptr(Wn), ptr(topo), 100, 8, ptr(X), None, len(X), 2.0, ptr(err))
            W = Wn
        return time.perf_counter() - t0, W

    runs = [cpu_run() for _ in range(3)]
    cpu_s = sorted(r[0] for r in runs)[1]

    def gpu_run():
        som = dam.SOM_Batch(m, m, 8, topology="square", sigma=sigma0)
        som._weights[...] = W0
        t0 = time.perf_counter()
        som.train(X, n_epochs=E, use_best_epoch=False)
        return time.perf_counter() - t0, som.weights

    gpu_run()
    gruns = [gpu_run() for _ in range(3)]
    gpu_s = sorted(r[0] for r in gruns)[1]
    dW = float(np.abs(gruns[0][1] - runs[0][1]).max())
    return {"workload": "10x10 square, D=8, N=100000 (5-Gaussian mixture, seed 0), 10 epochs, sigma0=5",
            "cpu": {"s_total": round(cpu_s, 4), "samples_per_s": round(E * len(X) / cpu_s, 1),
                    "cores": omp_threads(), "cpu_model": cpu_model(), "kind": "port", "runs": 3, "stat": "median"},
            "gpu": {"s_total": round(gpu_s, 4), "samples_per_s": round(E * len(X) / gpu_s, 1),
                    "note": "SOM_Batch.train wall time incl. upload, per-epoch errors and history copy"},
            "max_abs_weight_diff_gpu_vs_cpu": dW}


def centroid_errors(Wg, W64):
    """max-norm relative and element-wise relative centroid errors of Wg against W64."""
    Wg = np.asarray(Wg, np.float64)
    W64 = np.asarray(W64, np.float64)
    err = np.abs(Wg - W64)
    scale = np.abs(W64).max()
    return {"centroid_max_abs_err": float(err.max()),
            "centroid_rel_err_maxnorm": float(err.max() / scale),
            "centroid_max_rel_err": float((err / np.maximum(np.abs(W64), 1e-3 * scale)).max())}


def accuracy(ctx, W, W0, X, topo, m, prec, arith, E, sigma0, n_sub, n_ref, n_traj):
    """Against the oracle, at the final weights W of the schedule:
      * BMU agreement of the GPU winners with the exact FP64 oracle (n_sub rows) and with the REF32 oracle,
        the reference's own distance arithmetic (the first n_ref of those rows),
      * quantisation error vs FP64,
      * one-step centroid error: one more epoch (schedule epoch E-1) on the n_sub rows through the GPU path
        versus the FP64 factorised epoch with the FP64 oracle's BMUs,
      * trajectory centroid error: the whole E-epoch schedule from W0 on n_traj rows through
        SOM_Batch.train versus the FP64 oracle's schedule (weights rounded to FP32 after every epoch, as the
        reference stores them)."""
    import decontamination_amd as dam
    from decontamination_amd import _lib
    from decontamination_amd.runtime import Accumulator, DeviceArray, Model
    from oracle import FP64, REF32
    from oracle import som_ref as R
    t_start = time.perf_counter()
    rng = np.random.default_rng(123)
    N = len(X)
    idx = np.sort(rng.choice(N, size=min(n_sub, N), replace=False))
    Xs = np.ascontiguousarray(X[idx])
    mn, dim = W.shape
    som = dam.SOM_Batch(m, m, dim, precision=prec, arithmetic=arith)
    som._weights[...] = W
    got = som.get_winners(Xs)
    qe, _ = som.compute_errors(Xs)
    c1_64, _, d1_64 = R.top2_omp(W, Xs, FP64)
    nr = min(n_ref, len(Xs))
    c1_32, _, _ = R.top2_omp(W, Xs[:nr], REF32)
    # one more epoch on the subsample through the GPU path
    model = Model(ctx, m, m, dim, "hexagonal", topo)
    model.set_weights(W)
    acc = Accumulator(model)
    xd = DeviceArray.from_host(ctx, Xs)
    acc.reset()
    acc.epoch_accumulate(xd.ptr, None, len(Xs), _lib.PRECISIONS[prec], _lib.ARITHMETICS[arith])
    acc.epoch_update(E - 1, E, sigma0)
    Wg = model.get_weights()
    acc.close()
    xd.free()
    model.close()
    W64, _, _, _ = R.factorised_epoch_fp64(W, topo, Xs, None, R.sigma_at(sigma0, E - 1, E), bmu=c1_64,
                                           out_dtype=np.float64)
    qe64 = float(np.mean(np.sqrt(d1_64)))
    out = {"subsample": len(idx),
           "bmu_agreement_vs_fp64": float(np.mean(got == c1_64)),
           "bmu_disagreements_vs_fp64": int(np.sum(got != c1_64)),
           "ref32_subsample": nr,
           "bmu_agreement_vs_ref32_oracle": float(np.mean(got[:nr] == c1_32)),
           "bmu_disagreements_vs_ref32_oracle": int(np.sum(got[:nr] != c1_32)),
           "qe": qe, "qe_fp64": qe64, "qe_rel_err": abs(qe - qe64) / qe64,
           "one_step": centroid_errors(Wg, W64)}
    if n_traj > 0:
        tidx = np.sort(np.random.default_rng(321).choice(N, size=min(n_traj, N), replace=False))
        Xt = np.ascontiguousarray(X[tidx])
        st = dam.SOM_Batch(m, m, dim, sigma=sigma0, precision=prec, arithmetic=arith)
        st._weights[...] = W0
        st.train(Xt, n_epochs=E, use_best_epoch=False)
        Wt64, _, qt64 = R.trajectory_fp64(W0, topo, Xt, sigma0, E)
        traj = centroid_errors(st.weights, Wt64)
        traj.update({"rows": len(tidx), "epochs": E, "weights_bit_identical": bool(np.array_equal(st.weights, Wt64)),
                     "qe_rel_err": abs(float(st.quantization_errors[-1]) - qt64) / qt64})
        out["trajectory"] = traj
    out["centroid_rel_err_def"] = ("maxnorm: max|W - W_fp64| / max|W_fp64|; max_rel: max over entries of "
                                   "|W - W_fp64| / max(|W_fp64|, 1e-3 max|W_fp64|)")
    out["seconds"] = round(time.perf_counter() - t_start, 1)
    return out


def leg_name(prec, arith):
    return "fp32" if prec == "fp32" else f"{prec}/{arith}"


def workload_bench(ctx, X, m, E, sigma0, steps, warmup, legs, args, acc_rows, label):
    """All legs of one workload (data resident in HBM): timing, roofline, accuracy."""
    import decontamination_amd as dam
    from decontamination_amd.runtime import DeviceArray
    N, dim = X.shape
    som = dam.SOM_Batch(m, m, dim)
    som.init_rand(seed=0)
    W0 = som.weights.copy()
    topo = som._topography
    t = time.perf_counter()
    x_dev = DeviceArray.from_host(ctx, X)
    ctx.synchronize()
    h2d_s = time.perf_counter() - t
    log(f"{label}: uploaded {X.nbytes / 1e9:.2f} GB in {h2d_s:.2f}s")
    if args.calibrate_copy:
        from decontamination_amd import _lib
        tmp = DeviceArray.empty(ctx, X.shape, np.float32)
        _lib.call("somhip_memcpy_d2d", ctx.handle, _lib.vp(tmp.ptr), _lib.vp(x_dev.ptr), X.nbytes)
        ctx.synchronize()
        tmp.free()
        log(f"calibration copy of {X.nbytes} bytes done")
    res = {}
    for prec, arith in legs:
        name = leg_name(prec, arith)
        log(f"{label}: timing {name}: warmup {warmup}, steps {steps} of the {E}-epoch schedule")
        r = run_schedule(ctx, x_dev, N, W0, topo, m, prec, arith, steps, warmup, E, sigma0)
        log(f"{label} {name}: {r['ms_per_step']:.3f} ms/epoch, {r['samples_per_s'] / 1e9:.3f} G samples/s; kernels "
            f"{json.dumps({k: [round(v[0] / v[1], 4), v[1]] for k, v in r['timing'].items()})}")
        res[name] = r
    flops = 2.0 * N * m * m * dim
    for name, r in res.items():
        bf = name.startswith("bf16")
        r["roofline"] = roofline(r["timing"], "bmu_bf16" if bf else "bmu_fp32", flops,
                                 PEAK_BF16_TFLOPS if bf else PEAK_FP32_TFLOPS)
        r["kernels_ms_per_step"] = {k: round(v[0] / r["steps"], 4) for k, v in r["timing"].items()}
        if not args.no_accuracy:
            prec, arith = ("fp32", "ref32") if name == "fp32" else name.split("/")
            n_sub, n_ref, n_traj = acc_rows
            r["accuracy"] = accuracy(ctx, r["W"], W0, X, topo, m, prec, arith, E, sigma0, n_sub, n_ref,
                                     0 if prec == "fp32" else n_traj)
            log(f"{label} accuracy {name}: {json.dumps(r['accuracy'])}")
    return {"res": res, "W0": W0, "topo": topo, "x_dev": x_dev, "h2d_s": h2d_s, "flops": flops}


def leg_summary(r):
    out = {"value": round(r["samples_per_s"], 1), "ms_per_step": round(r["ms_per_step"], 4),
           "roofline": r["roofline"], "kernels_ms_per_step": r["kernels_ms_per_step"]}
    if "accuracy" in r:
        out["accuracy"] = r["accuracy"]
    return out


def claims(res, head_name, parity_name):
    """Which leg carries which claim, and whether it meets the bars."""
    h = res[head_name]
    out = {"leg": head_name}
    if "fp32" in res:
        out["speedup_vs_fp32"] = round(h["samples_per_s"] / res["fp32"]["samples_per_s"], 3)
    a = h.get("accuracy")
    if a:
        out["bmu_agreement_vs_fp64"] = a["bmu_agreement_vs_fp64"]
        out["centroid_max_rel_err_one_step"] = a["one_step"]["centroid_max_rel_err"]
        if "trajectory" in a:
            out["centroid_max_rel_err_trajectory"] = a["trajectory"]["centroid_max_rel_err"]
    meets = {}
    if "speedup_vs_fp32" in out:
        meets["speedup"] = out["speedup_vs_fp32"] >= BARS["speedup_vs_fp32"]
    if a:
        meets["bmu_agreement"] = out["bmu_agreement_vs_fp64"] >= BARS["bmu_agreement_vs_fp64"]
        errs = [out["centroid_max_rel_err_one_step"]] + ([out["centroid_max_rel_err_trajectory"]]
                                                         if "centroid_max_rel_err_trajectory" in out else [])
        meets["centroid_rel_err"] = max(errs) <= BARS["centroid_rel_err"]
    out["bars"] = BARS
    out["meets"] = meets
    if parity_name in res:
        p = res[parity_name]
        par = {"leg": parity_name, "value": round(p["samples_per_s"], 1), "ms_per_step": round(p["ms_per_step"], 4),
               "role": "BMUs of the reference's own FP32-difference distance arithmetic"}
        if "accuracy" in p:
            par["bmu_agreement_vs_ref32_oracle"] = p["accuracy"]["bmu_agreement_vs_ref32_oracle"]
            par["rows"] = p["accuracy"]["ref32_subsample"]
        out["parity_leg"] = par
    return out


def serving_bench(X16, seconds, rate, budget_gb, arith, seed=0, check_rows=256):
    """Three resident trained SOMs (32x32/D16, 64x64/D16, 100x100/D32); BMU and activation-map requests
    with log-uniform row counts in [1e3, 4e6], Poisson arrivals at `rate`/s for `seconds`. After the run, a
    row sample of every winners request is checked against the oracle."""
    import threading
    import decontamination_amd as dam
    from decontamination_amd import synth
    from decontamination_amd.serving import SOMServer, pinned_empty
    from oracle import FP64, REF32
    from oracle import som_ref as R

    t0 = time.perf_counter()
    rng = np.random.default_rng(seed)
    X32 = synth.catalogue_rows(4_200_000, 16, 14, seed=seed + 1)
    specs = [(32, X16), (64, X16), (100, X32)]
    soms = []
    for m, data in specs:
        som = dam.SOM_Batch(m, m, data.shape[1])
        som.init_rand(seed=0)
        som.train(data[: 1 << 20], n_epochs=4)
        soms.append(som)
    pools = {16: pinned_empty(X16.shape, np.float32), 32: pinned_empty(X32.shape, np.float32)}
    pools[16][...] = X16
    pools[32][...] = X32
    outs = [pinned_empty((4_000_000,), np.int32) for _ in range(48)]
    setup_s = time.perf_counter() - t0
    log(f"serving setup {setup_s:.1f}s (3 models trained, pools pinned)")

    arrivals = np.cumsum(rng.exponential(1.0 / rate, int(rate * seconds * 1.5) + 16))
    arrivals = arrivals[arrivals < seconds]
    nreq = len(arrivals)
    models = rng.integers(0, 3, nreq)
    kinds = rng.integers(0, 2, nreq)
    rows = np.exp(rng.uniform(np.log(1e3), np.log(4e6), nreq)).astype(np.int64)
    server = SOMServer(soms, memory_budget=int(budget_gb * (1 << 30)), arithmetic=arith)
    tickets = []
    done = []
    checks = []  # (model, rows of the pool, winners) samples of served results
    lock = threading.Lock()
    stop = threading.Event()
    crng = np.random.default_rng(seed + 7)

    def collector():
        i = 0
        while True:
            with lock:
                have = i < len(tickets)
                t = tickets[i] if have else None
            if not have:
                if stop.is_set():
                    return
                time.sleep(0.0005)
                continue
            res, lat = server.wait(t[0])
            done.append((lat, t[1], t[2], t[3]))
            if t[2] == "winners":
                sel = crng.choice(t[3], size=min(check_rows, t[3]), replace=False)
                checks.append((t[1], t[4] + sel, res[sel].copy()))
            i += 1

    th = threading.Thread(target=collector, daemon=True)
    th.start()
    start = time.perf_counter()
    for k in range(nreq):
        now = time.perf_counter() - start
        if arrivals[k] > now:
            time.sleep(arrivals[k] - now)
        mi = int(models[k])
        pool = pools[soms[mi].dim]
        r = int(min(rows[k], pool.shape[0]))
        off = int(rng.integers(0, pool.shape[0] - r + 1))
        kind = "winners" if kinds[k] == 0 else "activation"
        out = outs[k % len(outs)][:r] if kind == "winners" else None
        t = server.submit(mi, kind, pool[off:off + r], None, out)
        with lock:
            tickets.append((t, mi, kind, r, off))
    stop.set()
    th.join()
    wall = time.perf_counter() - start
    st = server.stats()
    server.close()
    # served winners against the oracle (the server's rerank arithmetic)
    mode = FP64 if arith == "fp64" else REF32
    mism = 0
    nrow = 0
    for mi in range(3):
        sel = [c for c in checks if c[0] == mi]
        if not sel:
            continue
        pool = pools[soms[mi].dim]
        ridx = np.concatenate([c[1] for c in sel])
        got = np.concatenate([c[2] for c in sel])
        exp, _, _ = R.top2_omp(soms[mi].weights, np.ascontiguousarray(pool[ridx]), mode)
        mism += int(np.sum(got != exp))
        nrow += len(ridx)
    lat = np.array([d[0] for d in done])
    tot_rows = int(sum(d[3] for d in done))
    by_model = {}
    for mi, (m, data) in enumerate(specs):
        l = np.array([d[0] for d in done if d[1] == mi])
        if len(l):
            by_model[f"{m}x{m}_d{data.shape[1]}"] = {"p50_ms": round(float(np.percentile(l, 50)), 3),
                                                      "p99_ms": round(float(np.percentile(l, 99)), 3),
                                                      "requests": int(len(l))}
    return {
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "mean_ms": round(float(lat.mean()), 3),
        "max_ms": round(float(lat.max()), 3),
        "requests": int(len(lat)),
        "rate_per_s": rate,
        "seconds": seconds,
        "rows_per_s": round(tot_rows / wall, 1),
        "offered_rows_per_s": round(float(rows.clip(max=4_000_000).sum()) / seconds, 1),
        "launches": st["launches"],
        "tiles": st["tiles"],
        "budget_gb": budget_gb,
        "arithmetic": arith,
        "by_model": by_model,
        "checked_vs_oracle": {"rows": nrow, "mismatches": mism,
                              "note": f"{check_rows} random rows of every winners request vs the "
                                      f"{'FP64' if mode == FP64 else 'REF32'} oracle"},
        "workload": "3 resident SOMs (32x32/D16, 64x64/D16, 100x100/D32, trained 4 epochs on 1M rows), BMU or "
                    "activation-map requests (50/50), rows log-uniform in [1e3, 4e6] sliced from pinned pools of "
                    "synthetic catalogues, Poisson arrivals; latency = submit to result in host memory",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; run with --gpus 1 (training would shard over "
                                   "samples with one exchange of the per-neuron sums per epoch; not "
                                   "implemented)"}), flush=True)
        sys.exit(2)
    from decontamination_amd import synth
    from decontamination_amd.runtime import get_context

    m = args.map
    E = args.epochs_schedule
    sigma0 = m / 2.0
    t = time.perf_counter()
    X = synth.systematics_dataset(args.nside, args.lognormal, args.gaussian, seed=0)
    N, dim = X.shape
    log(f"data {N}x{dim} generated in {time.perf_counter() - t:.1f}s")
    ctx = get_context(0)

    head = ("bf16_rerank", args.arithmetic)
    parity = ("bf16_rerank", "ref32")
    legs = ([] if args.no_fp32 else [("fp32", "ref32")]) + [head]
    if not args.no_parity_leg and args.arithmetic != "ref32":
        legs.append(parity)
    head_name, parity_name = leg_name(*head), leg_name(*parity)
    acc_rows = (args.acc_subsample, args.ref32_subsample, args.traj_subsample)
    wl = workload_bench(ctx, X, m, E, sigma0, args.steps, args.warmup, legs, args, acc_rows, "c1")
    res = wl["res"]
    h = res[head_name]
    default_workload = m == 64 and args.nside == 1024 and dim == 16
    # the committed counter summaries profile the FP64-rerank training pass
    if h["roofline"] is not None and default_workload and args.arithmetic == "fp64":
        tr = profile_figure("som_epoch_64x64_d16_hbm_traffic.md", "traffic_bytes_per_launch", "bmu_bf16_kernel")
        h["roofline"]["traffic"] = None if tr is None else int(tr)
        h["roofline"]["mfma_busy"] = profile_figure("som_epoch_64x64_d16_counters.md", "mfma_busy",
                                                    "bmu_bf16_kernel")
    out = {
        "metric": METRIC,
        "value": round(h["samples_per_s"], 1),
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": f"batch SOM epoch, {m}x{m} hexagonal, dim {dim}, N={N} (HEALPix nside "
                               f"{args.nside}), sigma0={sigma0:g}, timed epochs of the {E}-epoch schedule from "
                               f"init_rand(0); BF16 MFMA distance GEMM + {args.arithmetic} rerank",
                   "map": f"{m}x{m}", "dim": dim, "samples": N, "precision": "bf16_rerank",
                   "rerank_arithmetic": args.arithmetic, "epochs_schedule": E},
        "roofline": h["roofline"],
        "h2d_upload_s": round(wl["h2d_s"], 4),
        "claims": claims(res, head_name, parity_name),
        "legs": {k: leg_summary(r) for k, r in res.items()},
    }
    if args.stream_chunks > 0:
        st = streaming_bench(ctx, X, wl["W0"], wl["topo"], m, sigma0, E, args.stream_chunks, args.arithmetic)
        st["resident_ms_per_epoch"] = round(h["ms_per_step"], 4)
        log(f"streaming: {json.dumps(st)}")
        out["streaming"] = st
    if args.pipeline:
        out["pipeline"] = pipeline_bench(ctx, X, wl["x_dev"], h["W"], m, args.arithmetic)
        log(f"pipeline: {json.dumps(out['pipeline'])}")
    wl["x_dev"].free()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(X, wl["W0"], wl["topo"], sigma0, E)
        log(f"cpu_baseline: {json.dumps(out['cpu_baseline'])}")
    if args.c0:
        out["c0"] = c0_bench()
        log(f"c0: {json.dumps(out['c0'])}")
    if args.c3:
        t = time.perf_counter()
        X3 = synth.systematics_dataset(2048, 16, 14, seed=0)
        log(f"c3 data {X3.shape[0]}x{X3.shape[1]} generated in {time.perf_counter() - t:.1f}s")
        E3, m3 = 10, 100
        w3 = workload_bench(ctx, X3, m3, E3, m3 / 2.0, args.c3_steps, args.c3_warmup, legs, args,
                            (1 << 18, 1 << 16, 1 << 15), "c3")
        w3["x_dev"].free()
        r3 = w3["res"]
        if r3[head_name].get("roofline") and args.arithmetic == "fp64":
            r3[head_name]["roofline"]["mfma_busy"] = profile_figure("som_epoch_100x100_d32_counters.md", "mfma_busy",
                                                                    "bmu_bf16_kernel")
            tr3 = profile_figure("som_epoch_100x100_d32_hbm_traffic.md", "traffic_bytes_per_launch", "bmu_bf16_kernel")
            r3[head_name]["roofline"]["traffic"] = None if tr3 is None else int(tr3)
        out["c3"] = {"workload": f"batch SOM epoch, 100x100 hexagonal, dim 32, N={X3.shape[0]} (HEALPix nside 2048, "
                                 "16 lognormal + 14 Gaussian + 2 angular), sigma0=50, timed epochs of the 10-epoch "
                                 "schedule from init_rand(0), X resident in HBM",
                     "value": round(r3[head_name]["samples_per_s"], 1),
                     "ms_per_step": round(r3[head_name]["ms_per_step"], 4),
                     "roofline": r3[head_name]["roofline"],
                     "claims": claims(r3, head_name, parity_name),
                     "legs": {k: leg_summary(r) for k, r in r3.items()}}
        del X3
    out["serving"] = None
    if args.serve_seconds > 0:
        out["serving"] = serving_bench(X, args.serve_seconds, args.serve_rate, args.serve_budget_gb, args.arithmetic)
        log(f"serving: {json.dumps(out['serving'])}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
