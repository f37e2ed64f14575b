#!/usr/bin/env python3
"""Benchmark: batch-SOM epoch throughput on one MI355X.

Workload (one "step" = one full epoch over the dataset resident in HBM): 64x64 hexagonal map
(4,096 neurons), dim 16, 12,582,912 samples (one per HEALPix nside=1024 pixel) of synthetic
systematics (8 lognormal, 6 Gaussian, sin/cos of the pixel-centre colatitude, standardised),
sigma0 = 32 on the reference's 20-epoch asymptotic schedule, init_rand(seed=0). An epoch is the
reference's whole epoch: BMU search, per-BMU accumulation, neighbourhood update and division, plus
that epoch's quantization/topographic errors (fused into the next BMU pass).

value = samples/s of the BF16 + exact-rerank path; the plain FP32 path is measured on the same
workload and reported beside it with the speed-up. Prints ONE JSON line.
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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=18)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nside", type=int, default=1024)
    ap.add_argument("--map", type=int, default=64)
    ap.add_argument("--lognormal", type=int, default=8, help="lognormal feature columns")
    ap.add_argument("--gaussian", type=int, default=6, help="Gaussian feature columns (+2 angular)")
    ap.add_argument("--acc-subsample", type=int, default=1 << 20)
    ap.add_argument("--epochs-schedule", type=int, default=20)
    ap.add_argument("--arithmetic", default="ref32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--precisions", default="fp32,bf16_rerank")
    ap.add_argument("--no-fp64-rerank", action="store_true",
                    help="skip the extra bf16_rerank run with arithmetic='fp64'")
    ap.add_argument("--stream-chunks", type=int, default=8,
                    help="chunks of the streamed-epoch leg (0 skips it)")
    ap.add_argument("--no-pipeline", dest="pipeline", action="store_false",
                    help="skip the covariance / gndm leg")
    ap.add_argument("--c0", action="store_true",
                    help="also time the 10x10/D8/100k CPU-runnable configuration (CPU oracle and GPU)")
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one device-to-device copy of the dataset (reference kernel for HBM counters)")
    return ap.parse_args()


def run_epochs(ctx, model_cls, acc_cls, dev, x_dev, n, W0, topo, m, prec, arith, steps, warmup, E, sigma0):
    from decontamination_amd import _lib
    pc, ac = _lib.PRECISIONS[prec], _lib.ARITHMETICS[arith]
    mn, dim = W0.shape
    model = model_cls(ctx, m, m, dim, "hexagonal", topo)
    model.set_weights(W0)
    acc = acc_cls(model)
    total = warmup + steps
    errs = dev.zeros(ctx, (total + 1, 3), np.float64)
    hist = dev.empty(ctx, (E, mn, dim), np.float32)

    def epoch(e):
        acc.reset()
        acc.epoch_accumulate(x_dev.ptr, None, n, pc, ac, errs.view_ptr((e - 1) * 3) if e > 0 else None)
        acc.epoch_update(e % E, E, sigma0, hist.view_ptr((e % E) * mn * dim))

    for e in range(warmup):
        epoch(e)
    ctx.synchronize()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for e in range(warmup, total):
        epoch(e)
    ctx.synchronize()
    t1 = time.perf_counter()
    timing = ctx.timing()
    ctx.set_timing(False)
    e3 = errs.copy_to_host()
    W = model.get_weights()
    errs.free()
    hist.free()
    acc.close()
    return {"seconds": t1 - t0, "timing": timing, "errors": e3, "W": W, "model": model}


def streaming_bench(ctx, X, W0, topo, m, sigma0, E, n_chunks, steps=4, warmup=1):
    """The same bf16_rerank epoch with X fed from host memory as n_chunks generator chunks: through the
    uploader ring (copy of chunk k+1 overlapping the kernels of chunk k), and serially (synchronous copy,
    then the kernels), next to the copies alone. Host-to-device rates include PCIe."""
    from decontamination_amd import _lib
    from decontamination_amd.runtime import Accumulator, DeviceArray, Model, Uploader
    pc, ac = _lib.PREC_BF16_RERANK, _lib.RERANK_REF32
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
        acc.epoch_update(e % E, E, sigma0, hist.ptr)

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
    return {"chunks": len(chunks), "chunk_rows": rows, "epochs_timed": steps,
            "copies_only_ms": round(1e3 * copy_s, 3), "h2d_GB_per_s": round(X.nbytes / copy_s / 1e9, 2),
            "overlapped_ms_per_epoch": round(1e3 * overlap_s, 3), "serial_ms_per_epoch": round(1e3 * serial_s, 3),
            "overlapped_samples_per_s": round(N / overlap_s, 1),
            "note": "bf16_rerank epoch with X streamed from pageable host memory as generator chunks; "
                    "PCIe-inclusive, not the headline value"}


def pipeline_bench(ctx, X, x_dev, W, m, steps=20):
    """The steps on either side of training, on the same footprint resident in HBM: the weighted
    covariance of the SOM_PCA initialisation (two passes over X) and the galaxy-number-density maps
    from the winners (somhip_gndm), timed per call (wall clock, synchronised), with the oracle's
    single-thread restatements of the reference's CPU loops on a prefix as the CPU baseline."""
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
    model.find_bmus(x_dev.ptr, N, win.ptr, _lib.PREC_BF16_RERANK, _lib.RERANK_REF32)
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
    # CPU baselines: the reference's Welford loop (one thread, covariance.py:30-84) and its two
    # activation-map passes (_count_bmus_kernel, BMU search per pixel) on prefixes
    L = oracle.lib()
    p = 1 << 20
    xs = np.ascontiguousarray(X[:p], np.float64)
    sw, mean, m2, dl = np.zeros(1), np.zeros(dim), np.zeros((dim, dim)), np.zeros(dim)
    t0 = time.perf_counter()
    L.orc_welford_update(ptr(sw), ptr(mean), ptr(m2), ptr(dl), ptr(xs), ptr(rho[:p]), p, dim)
    cpu_cov = p / (time.perf_counter() - t0)
    q = 1 << 13
    xq = np.ascontiguousarray(X[:q])
    res = np.zeros(mn)
    t0 = time.perf_counter()
    L.orc_count_bmus(0, ptr(W), mn, dim, ptr(xq), ptr(rho[:q]), q, ptr(res))
    L.orc_count_bmus(0, ptr(W), mn, dim, ptr(xq), None, q, ptr(res))
    cpu_gndm = q / (time.perf_counter() - t0)
    algo_cov = N * (4 * dim + 8)
    algo_gndm = N * (4 + 8 + 4)
    return {
        "covariance": {"ms": round(1e3 * t_cov, 4), "rows_per_s": round(N / t_cov, 1),
                       "algorithmic_GB_per_s": round(algo_cov / t_cov / 1e9, 1),
                       "note": f"weighted covariance of the {N}x{dim} footprint (SOM_PCA init), 2 passes over X, "
                               "wall clock per call incl. the 2 KB result copy",
                       "cpu_baseline": {"value": round(cpu_cov, 1), "unit": "rows/s", "cores": 1, "kind": "port",
                                        "sample": f"oracle Welford loop on the first {p} rows"}},
        "gndm": {"ms": round(1e3 * t_gndm, 4), "pixels_per_s": round(N / t_gndm, 1),
                 "algorithmic_GB_per_s": round(algo_gndm / t_gndm / 1e9, 1),
                 "note": "catalog + footprint activation maps from the winners, gnd and gndm = gnd[winners]; the "
                         "reference runs two more BMU searches over the footprint for the maps",
                 "cpu_baseline": {"value": round(cpu_gndm, 1), "unit": "pixels/s", "cores": 1, "kind": "port",
                                  "sample": f"oracle _count_bmus_kernel loop (two maps) on the first {q} pixels"}},
    }


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "som_epoch_64x64_d16_hbm_traffic.md")


def measured_traffic(kernel_symbol):
    """HBM bytes per launch of `kernel_symbol` from the committed FETCH_SIZE/WRITE_SIZE summary (two
    rocprofv3 --pmc passes of this script, scaled by a device copy of known size), or None."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith(f"traffic_bytes_per_launch {kernel_symbol}:"):
                return int(float(line.split(":")[1]))
    except OSError:
        pass
    return None


def roofline(timing, kernel, flops_per_launch, peak_tflops, traffic=None):
    if kernel not in timing:
        return None
    ms, launches = timing[kernel]
    avg_ms = ms / launches
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    return {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak_tflops, "unit": "TFLOP/s",
            "frac": round(achieved / peak_tflops, 5), "traffic": traffic, "kernel": kernel,
            "avg_ms": round(avg_ms, 4), "launches": launches}


def cpu_baseline(X, W0, topo, sigma0, E, target_s=12.0):
    """The oracle's parallel restatement of the reference CPU epoch (prange + FP32 atomics + errors
    pass) on a prefix of the same workload; extrapolated per sample."""
    import ctypes
    import oracle
    from oracle import ptr
    L = oracle.lib()
    mn, dim = W0.shape
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

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
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"workload": "10x10 square, D=8, N=100000 (5-Gaussian mixture, seed 0), 10 epochs, sigma0=5",
            "cpu": {"s_total": round(cpu_s, 4), "samples_per_s": round(E * len(X) / cpu_s, 1), "cores": threads,
                    "cpu_model": cpu_model(), "kind": "port", "runs": 3, "stat": "median"},
            "gpu": {"s_total": round(gpu_s, 4), "samples_per_s": round(E * len(X) / gpu_s, 1),
                    "note": "SOM_Batch.train wall time incl. upload, per-epoch errors and history copy"},
            "max_abs_weight_diff_gpu_vs_cpu": dW}


def accuracy(ctx, res, X, topo, m, prec, arith, E, sigma0, n_sub=1 << 20):
    """Against the FP64 oracle on a random subsample, at the run's final weights W: BMU disagreement,
    quantisation error, and the centroids after one more epoch (schedule epoch E-1) on the subsample
    through the GPU path versus the FP64 factorised epoch on the same rows."""
    import decontamination_amd as dam
    from decontamination_amd import _lib
    from decontamination_amd.runtime import Accumulator, DeviceArray, Model
    from oracle import som_ref as R
    rng = np.random.default_rng(123)
    idx = np.sort(rng.choice(len(X), size=min(n_sub, len(X)), replace=False))
    Xs = np.ascontiguousarray(X[idx])
    W = res["W"]
    mn, dim = W.shape
    som = dam.SOM_Batch(m, m, dim, precision=prec, arithmetic=arith)
    som._weights[...] = W
    got = som.get_winners(Xs)
    qe, _ = som.compute_errors(Xs)
    model = Model(ctx, m, m, dim, "hexagonal", topo)
    model.set_weights(W)
    acc = Accumulator(model)
    xd = DeviceArray.from_host(ctx, Xs)
    acc.reset()
    acc.epoch_accumulate(xd.ptr, None, len(Xs), _lib.PRECISIONS[prec], _lib.ARITHMETICS[arith])
    acc.epoch_update(E - 1, E, sigma0)
    Wg = model.get_weights().astype(np.float64)
    acc.close()
    xd.free()
    model.close()
    W64, c1, _, d1 = R.factorised_epoch_fp64(W, topo, Xs, None, R.sigma_at(sigma0, E - 1, E), out_dtype=np.float64)
    qe64 = float(np.mean(np.sqrt(d1)))
    err = np.abs(Wg - W64)
    scale = np.abs(W64).max()
    dis = float(np.mean(got != c1))
    return {"subsample": len(idx), "bmu_disagreement_vs_fp64": dis, "bmu_agreement_vs_fp64": 1.0 - dis,
            "qe": qe, "qe_fp64": qe64, "qe_rel_err": abs(qe - qe64) / qe64,
            "centroid_max_abs_err": float(err.max()),
            "centroid_rel_err_maxnorm": float(err.max() / scale),
            "centroid_max_rel_err": float((err / np.maximum(np.abs(W64), 1e-3 * scale)).max()),
            "centroid_rel_err_def": "maxnorm: max|W_gpu - W_fp64| / max|W_fp64|; max_rel: max over entries of "
                                    "|W_gpu - W_fp64| / max(|W_fp64|, 1e-3 max|W_fp64|)"}


def serving_bench(X16, seconds, rate, budget_gb, seed=0):
    """Three resident trained SOMs (32x32/D16, 64x64/D16, 100x100/D32); BMU and activation-map requests
    with log-uniform row counts in [1e3, 4e6], Poisson arrivals at `rate`/s for `seconds`."""
    import threading
    import decontamination_amd as dam
    from decontamination_amd import synth
    from decontamination_amd.serving import SOMServer, pinned_empty

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

    # schedule
    arrivals = np.cumsum(rng.exponential(1.0 / rate, int(rate * seconds * 1.5) + 16))
    arrivals = arrivals[arrivals < seconds]
    nreq = len(arrivals)
    models = rng.integers(0, 3, nreq)
    kinds = rng.integers(0, 2, nreq)
    rows = np.exp(rng.uniform(np.log(1e3), np.log(4e6), nreq)).astype(np.int64)
    server = SOMServer(soms, memory_budget=int(budget_gb * (1 << 30)))
    tickets = []
    done = []
    lock = threading.Lock()
    stop = threading.Event()

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
            i += 1

    th = threading.Thread(target=collector, daemon=True)
    th.start()
    start = time.perf_counter()
    for k in range(nreq):
        now = time.perf_counter() - start
        if arrivals[k] > now:
            time.sleep(arrivals[k] - now)
        mi = int(models[k])
        d = soms[mi].dim
        pool = pools[d]
        r = int(min(rows[k], pool.shape[0]))
        off = int(rng.integers(0, pool.shape[0] - r + 1))
        x = pool[off:off + r]
        kind = "winners" if kinds[k] == 0 else "activation"
        out = outs[k % len(outs)][:r] if kind == "winners" else None
        t = server.submit(mi, kind, x, None, out)
        with lock:
            tickets.append((t, mi, kind, r))
    stop.set()
    th.join()
    wall = time.perf_counter() - start
    st = server.stats()
    server.close()
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
        "by_model": by_model,
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
    import decontamination_amd as dam
    from decontamination_amd import synth
    from decontamination_amd.runtime import Accumulator, DeviceArray, Model, get_context

    m = args.map
    E = args.epochs_schedule
    sigma0 = m / 2.0
    t = time.perf_counter()
    X = synth.systematics_dataset(args.nside, args.lognormal, args.gaussian, seed=0)
    N, dim = X.shape
    log(f"data {N}x{dim} generated in {time.perf_counter() - t:.1f}s")
    som = dam.SOM_Batch(m, m, dim)
    som.init_rand(seed=0)
    W0 = som.weights.copy()
    topo = som._topography

    ctx = get_context(0)
    t = time.perf_counter()
    x_dev = DeviceArray.from_host(ctx, X)
    ctx.synchronize()
    h2d_s = time.perf_counter() - t
    log(f"uploaded {X.nbytes / 1e9:.2f} GB in {h2d_s:.2f}s")

    if args.calibrate_copy:
        tmp = DeviceArray.empty(ctx, X.shape, np.float32)
        from decontamination_amd import _lib
        _lib.call("somhip_memcpy_d2d", ctx.handle, _lib.vp(tmp.ptr), _lib.vp(x_dev.ptr), X.nbytes)
        ctx.synchronize()
        tmp.free()
        log(f"calibration copy of {X.nbytes} bytes done")

    results = {}
    for prec in args.precisions.split(","):
        log(f"timing {prec}: warmup {args.warmup}, steps {args.steps}")
        r = run_epochs(ctx, Model, Accumulator, DeviceArray, x_dev, N, W0, topo, m, prec, args.arithmetic,
                       args.steps, args.warmup, E, sigma0)
        r["samples_per_s"] = N * args.steps / r["seconds"]
        r["ms_per_step"] = 1e3 * r["seconds"] / args.steps
        log(f"{prec}: {r['ms_per_step']:.3f} ms/epoch, {r['samples_per_s'] / 1e9:.3f} G samples/s; "
            f"kernels {json.dumps({k: [round(v[0] / v[1], 4), v[1]] for k, v in r['timing'].items()})}")
        results[prec] = r
    extra = None
    if "bf16_rerank" in results and args.arithmetic == "ref32" and not args.no_fp64_rerank:
        # the same path with the rerank in FP64 (arithmetic='fp64'): BMUs of the FP64 formula
        log(f"timing bf16_rerank with fp64 rerank: warmup {args.warmup}, steps {args.steps}")
        extra = run_epochs(ctx, Model, Accumulator, DeviceArray, x_dev, N, W0, topo, m, "bf16_rerank", "fp64",
                           args.steps, args.warmup, E, sigma0)
        extra["samples_per_s"] = N * args.steps / extra["seconds"]
        extra["ms_per_step"] = 1e3 * extra["seconds"] / args.steps
        log(f"bf16_rerank/fp64: {extra['ms_per_step']:.3f} ms/epoch")

    flops = 2.0 * N * m * m * dim
    default_workload = m == 64 and args.nside == 1024 and dim == 16
    head = results.get("bf16_rerank") or next(iter(results.values()))
    head_prec = "bf16_rerank" if "bf16_rerank" in results else next(iter(results))
    out = {
        "metric": METRIC,
        "value": round(head["samples_per_s"], 1),
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16" if head_prec == "bf16_rerank" else "fp32",
        "data": "synthetic",
        "config": {"workload": f"batch SOM epoch, {m}x{m} hexagonal, dim {dim}, N={N} (HEALPix nside "
                               f"{args.nside}), sigma0={sigma0:g} on the {E}-epoch schedule, precision "
                               f"{head_prec}, rerank arithmetic {args.arithmetic}",
                   "map": f"{m}x{m}", "dim": dim, "samples": N, "precision": head_prec},
        "roofline": roofline(head["timing"], "bmu_bf16" if head_prec == "bf16_rerank" else "bmu_fp32", flops,
                             PEAK_BF16_TFLOPS if head_prec == "bf16_rerank" else PEAK_FP32_TFLOPS,
                             measured_traffic("bmu_bf16_kernel") if head_prec == "bf16_rerank" and default_workload
                             else None),
        "h2d_upload_s": round(h2d_s, 4),
    }
    if "fp32" in results and "bf16_rerank" in results:
        f = results["fp32"]
        out["fp32"] = {"value": round(f["samples_per_s"], 1), "ms_per_step": round(f["ms_per_step"], 4),
                       "roofline": roofline(f["timing"], "bmu_fp32", flops, PEAK_FP32_TFLOPS)}
        out["speedup_bf16_rerank_vs_fp32"] = round(head["samples_per_s"] / f["samples_per_s"], 3)
    out["kernels_ms_per_step"] = {k: round(v[0] / args.steps, 4) for k, v in head["timing"].items()}
    if not args.no_accuracy:
        acc = {}
        for prec, r in results.items():
            acc[prec] = accuracy(ctx, r, X, topo, m, prec, args.arithmetic, E, sigma0, args.acc_subsample)
            log(f"accuracy {prec}: {json.dumps(acc[prec])}")
        out["accuracy"] = acc
    if extra is not None:
        out["bf16_rerank_fp64_arithmetic"] = {"value": round(extra["samples_per_s"], 1),
                                              "ms_per_step": round(extra["ms_per_step"], 4)}
        if not args.no_accuracy:
            a64 = accuracy(ctx, extra, X, topo, m, "bf16_rerank", "fp64", E, sigma0, args.acc_subsample)
            log(f"accuracy bf16_rerank/fp64: {json.dumps(a64)}")
            out["bf16_rerank_fp64_arithmetic"]["accuracy"] = a64
    if args.stream_chunks > 0 and "bf16_rerank" in results:
        st = streaming_bench(ctx, X, W0, topo, m, sigma0, E, args.stream_chunks)
        st["resident_ms_per_epoch"] = round(head["ms_per_step"], 4)
        log(f"streaming: {json.dumps(st)}")
        out["streaming"] = st
    if args.pipeline and "bf16_rerank" in results:
        out["pipeline"] = pipeline_bench(ctx, X, x_dev, head["W"], m)
        log(f"pipeline: {json.dumps(out['pipeline'])}")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(X, W0, topo, sigma0, E)
    if args.c0:
        out["c0"] = c0_bench()
        log(f"c0: {json.dumps(out['c0'])}")
    out["serving"] = None
    if args.serve_seconds > 0:
        del x_dev
        out["serving"] = serving_bench(X, args.serve_seconds, args.serve_rate, args.serve_budget_gb)
        log(f"serving: {json.dumps(out['serving'])}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
