#!/usr/bin/env python3
"""Gradient iterations per second at N = 1,000,000 in 2D on one MI355X.

A step is one iteration of the t-SNE hot loop on the device-resident workload: the FFT-interpolated
gradient (attractive CSR term, spread, FFT convolution, interpolation) plus the gains/momentum update,
the step-norm clip and zeroMean (TSNE::run, tsne.cpp:440-531), for the FP32 form.  The FP64 plain GPU
path is timed the same way beside it, and the oracle (the C restatement of the reference CPU path,
FFT by the oracle's own mixed-radix code since FFTW is absent) on a bounded sample of gradient calls.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fit-sne_amd"))

from accuracy import gradient_errors  # noqa: E402

METRIC = "grad iters/sec at N=1M 2D, FP32 vs plain FP64 GPU @ rel err≤1e-4; %HBM peak"
HBM_PEAK_GBS = 8000.0
# the FP32 form's attractive kernel in 2D (bucketed column windows); it dominates the step
ATTR_KERNEL = "k_attr_win<2>"


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def timed_steps(ctx, steps, warmup, prm):
    for _ in range(warmup):
        ctx.step(**prm)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.step(**prm)
    ctx.sync()
    return time.perf_counter() - t0


def kernel_profile(ctx, steps, prm):
    """Per-kernel HIP-event durations over `steps` further iterations on the context's stream."""
    ctx.set_timing(True)
    for _ in range(steps):
        ctx.step(**prm)
    stats = ctx.kernel_stats()
    ctx.set_timing(False)
    return stats


def measured_traffic(kernel, summary="attractive_c2_n1m_2d_hbm_traffic.md"):
    """HBM bytes per dispatch of `kernel` from a committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    under profiles/, each counter scaled by what it read for a plain 64 MiB device copy in the same
    runs; None when the summary is missing or for another kernel."""
    import re

    path = os.path.join(ROOT, "profiles", summary)
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"```json\n(.*?)```", text, re.S)
    if not m:
        return None
    d = json.loads(m.group(1))
    if d.get("kernel") != kernel:
        return None
    kib = (d["fetch_kib_raw"] * d["copy_kib"] / d["copy_fetch_kib_raw"]
           + d["write_kib_raw"] * d["copy_kib"] / d["copy_write_kib_raw"])
    return kib * 1024.0


def per_kernel_roofline(alone, in_step, N, nnz, geo, D=2):
    """HBM roofline of each kernel of the FP32 2D step: algorithmic bytes per launch (each array the kernel
    must read or write, once) over its average duration when the step's kernels run one after another
    (`alone`), with its duration inside the overlapped step beside it.  FFT passes move little HBM data
    (their lines live in LDS); their fraction shows how far they are from being bandwidth bound."""
    n1d, H = geo["n1d"], geo["fft_len"] // 2 + 1
    nodes = n1d ** D
    pts = N * D * 4
    model = {
        "attractive": 8 * (N + 1) + nnz * 8 + 2 * pts,
        "spread_f32": pts + nodes * 4 * 8,
        "rows_fwd": 2 * nodes * 32 + 4 * H * n1d * 8,
        "cols_f32": 4 * H * n1d * 8 + 4 * H * H * 4 + 3 * H * n1d * 8,
        "rows_inv": 3 * H * n1d * 8 + nodes * 16,
        "interp_f32": 7 * pts + nodes * 16,
        "kern_rows": H * n1d * 8,
        "kern_cols": H * n1d * 8 + 4 * H * H * 4,
    }
    out = []
    for name, nbytes in model.items():
        cnt, tot = alone.get(name, (0, 0.0))
        if cnt == 0:
            continue
        ms = tot / cnt
        c2, t2 = in_step.get(name, (0, 0.0))
        gbs = nbytes / (ms * 1e-3) / 1e9
        out.append({"kernel": name, "bound": "hbm", "algorithmic_bytes": int(nbytes), "avg_ms_alone": round(ms, 4),
                    "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                    "avg_ms_in_step": round(t2 / c2, 4) if c2 else None})
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """Threads the CPU baseline uses: the cores this process may run on."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(P, Y32, calls, df=1.0, warmup=1):
    """The oracle's gradient (FP64 C restatement of computeFftGradient*, PARALLEL_FOR threading kept)
    on the same Y and P: `warmup` untimed calls, then the median of `calls` timed ones; returns
    (iters/s, threads, dC, sum_Q, median seconds)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc

    row, col, val = P
    Pu = (row.astype(np.uint32), col.astype(np.uint32), val)
    nthreads = host_threads()
    g = zq = None
    for _ in range(warmup):
        g, zq = orc.gradient(Pu, Y32, nthreads=nthreads, df=df)
    times = []
    for i in range(calls):
        t0 = time.perf_counter()
        g, zq = orc.gradient(Pu, Y32, nthreads=nthreads, df=df)
        times.append(time.perf_counter() - t0)
        log(f"cpu baseline gradient {i + 1}/{calls}: {times[-1]:.3f}s")
    med = float(np.median(times))
    return 1.0 / med, nthreads, g, zq, med


def c4_measure(args, fg, synth, one_by_one=True):
    """C4: 64 independent 2D embeddings stepped together (fitsne_batch_step: every stage of every problem in
    one launch, the whole step one graph launch), against the same 64 problems run one after another with
    the host-planned step (same kernels; results bitwise equal), and optionally each problem's own captured
    step one after another."""
    probs = synth.c4_problems()
    sizes = [Y.shape[0] for (_, Y) in probs]
    prm = [(0.8, max(200.0, n / 12.0), 5.0) for n in sizes]

    def make():
        cs = []
        for (P, Y) in probs:
            c = fg.Context(no_dims=2, precision=fg.FP32)
            c.set_P(*P)
            c.set_Y(Y)
            c.autotune()  # as the batch does for its contexts
            c.N = Y.shape[0]
            cs.append(c)
        return cs

    log("C4 sequential: each problem's steps one after another (host-planned steps)")
    seq = make()
    for c, (m, lr, msn) in zip(seq, prm):
        for _ in range(args.warmup):
            c.step(m, lr, msn)
        c.sync()
    t0 = time.perf_counter()
    for c, (m, lr, msn) in zip(seq, prm):
        for _ in range(args.steps):
            c.step(m, lr, msn)
        c.sync()
    t_seq = time.perf_counter() - t0
    t_cap = None
    if one_by_one:
        log("C4 captured, one problem after another (device-planned steps, no concurrency)")
        cap = make()
        for c, (m, lr, msn) in zip(cap, prm):
            c.set_graph(True)
            for _ in range(args.warmup):
                c.step(m, lr, msn)
            c.sync()
        t0 = time.perf_counter()
        for c, (m, lr, msn) in zip(cap, prm):
            for _ in range(args.steps):
                c.step(m, lr, msn)
            c.sync()
        t_cap = time.perf_counter() - t0
        for c in cap:
            c.close()
    log("C4 batched: one launch per stage for all 64 problems, one graph launch per iteration")
    bat = make()
    batch = fg.Batch(bat)
    for _ in range(args.warmup):
        batch.step(prm)
    bat[0].sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.step(prm)
    bat[0].sync()
    t_bat = time.perf_counter() - t0
    same = all(np.array_equal(a.get_Y(), b.get_Y()) for a, b in zip(seq, bat))
    batch.close()
    for c in seq + bat:
        c.close()
    n = len(sizes)
    out = {
        "value": n * args.steps / t_bat, "unit": "problem-iters/s", "ms_per_step": 1e3 * t_bat / args.steps,
        "workload": "C4: 64 problems, N log-uniform in [2000, 200000], 2D, df=1",
        "N_total": int(sum(sizes)), "N_min": int(min(sizes)), "N_max": int(max(sizes)),
        "sequential": {"value": n * args.steps / t_seq, "unit": "problem-iters/s",
                       "ms_per_round": 1e3 * t_seq / args.steps},
        "speedup_vs_sequential": t_seq / t_bat,
        "bitwise_equal_to_sequential": bool(same),
    }
    if t_cap is not None:
        out["captured_one_by_one"] = {"value": n * args.steps / t_cap, "unit": "problem-iters/s",
                                      "ms_per_round": 1e3 * t_cap / args.steps}
    return out


def bench_c4(args, fg, synth):
    r = c4_measure(args, fg, synth, one_by_one=not args.no_captured)
    out = {
        "metric": "aggregate gradient iterations/s over 64 batched 2D embeddings (C4)",
        "value": r["value"], "unit": "problem-iters/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded Gaussian mixtures, generated in place)",
        "config": {"workload": r["workload"], "N_total": r["N_total"], "N_min": r["N_min"], "N_max": r["N_max"]},
    }
    out.update({k: r[k] for k in ("sequential", "captured_one_by_one", "speedup_vs_sequential",
                                  "bitwise_equal_to_sequential") if k in r})
    print(json.dumps(out), flush=True)


def bench_exact(args, fg):
    """Exact t-SNE (theta = 0) at N = 46,000 in 2D, about the largest N whose N*N fits the reference's
    int arithmetic (tsne.cpp:194, :215-216, :1237-1245): dense P built on the device, then timed steps
    (exact gradient, update, zeroMean).  Roofline of the gradient sweep: it reads the dense P once
    (8 N^2 bytes)."""
    N = args.n if args.n != 1_000_000 else 46_000
    rng = np.random.default_rng(20261202)
    k = 20
    X = rng.normal(0, 4, size=(k, 50))[rng.integers(0, k, N)] + rng.normal(size=(N, 50))
    X -= X.mean(0)
    X /= np.abs(X).max()
    c = fg.Context(no_dims=2, precision=fg.FP64_EXACT)
    log(f"exact: dense affinities N={N} on the device")
    t0 = time.perf_counter()
    c.exact_affinities(X, 30.0, download=False)
    c.sync()
    t_aff = time.perf_counter() - t0
    c.set_Y(rng.normal(0, 10, size=(N, 2)))
    prm = dict(momentum=0.8, learning_rate=max(200.0, N / 12.0), max_step_norm=5.0)
    dt = timed_steps(c, args.steps, args.warmup, prm)
    ms = 1e3 * dt / args.steps
    stats = kernel_profile(c, max(args.profile_steps, 1), prm)
    t0 = time.perf_counter()
    kl = c.kl()
    t_kl = time.perf_counter() - t0
    c.close()
    cnt, tot = stats.get("exact_rows", (0, 0.0))
    rows_ms = tot / max(cnt, 1)
    nbytes = 8.0 * N * N + 8.0 * N * 2 + 8.0 * N * 5
    achieved = nbytes / (rows_ms * 1e-3) / 1e9 if rows_ms > 0 else 0.0
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import orc

        ns = 6000
        log(f"CPU baseline: oracle exact gradient at N={ns} (single thread, as computeExactGradient)")
        Ps, _ = orc.exact_P(X[:ns] / np.abs(X[:ns]).max(), 30.0)
        Ys = rng.normal(0, 10, size=(ns, 2))
        t0 = time.perf_counter()
        orc.exact_gradient(Ps, Ys)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": 1.0 / (t_cpu * (N / ns) ** 2), "unit": "grad iters/s", "cores": 1, "kind": "port",
               "sample": f"one oracle exact gradient (C restatement of computeExactGradient, serial) at N={ns} "
                         f"from the same construction, {t_cpu:.2f} s, scaled by the pair count (N/{ns})^2"}
    out = {
        "metric": "exact (theta=0) gradient iterations/s at N=46000 2D, FP64",
        "value": args.steps / dt, "unit": "grad iters/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Gaussian mixture, generated in place)",
        "config": {"workload": f"exact t-SNE: N={N} points, D=50 input, 2D embedding, perplexity 30, dense P"},
        "roofline": {"bound": "hbm", "kernel": "k_exact_rows<2, true, true>", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes": nbytes, "avg_ms": rows_ms},
        "cpu_baseline": cpu,
        "affinities_s": t_aff,
        "kl_ms": 1e3 * t_kl,
        "kl": kl,
        "kernels_ms": {k: round(v[1] / max(v[0], 1), 4) for k, v in sorted(stats.items(), key=lambda kv: -kv[1][1])},
    }
    print(json.dumps(out), flush=True)


def bench_c3(args, fg, synth):
    """C3: the 1D heavy-tailed gradient (computeFftGradientOneDVariableDf) at N = 2,000,000, df = 0.5,
    late exaggeration 4, span about [-300, 300] (600 boxes, FFT length 3888), FP32 form against the FP64
    plain path on the same Y and P; the attractive kernel's roofline as for C2."""
    N = args.n if args.n != 1_000_000 else 2_000_000
    log(f"generating the C3 workload N={N}")
    P, Y, _ = synth.c3_workload(N)
    nnz = int(P[0][-1])
    df = 0.5
    lr = max(200.0, N / 12.0)
    prm = dict(momentum=0.8, learning_rate=lr, max_step_norm=5.0)
    c32 = fg.Context(no_dims=1, precision=fg.FP32, df=df)
    c32.set_P(*P)
    c32.set_Y(Y)
    c32.set_exaggeration(4.0)
    spread_pts = c32.autotune()
    g32, z32 = c32.gradient()
    geo = c32.last_geometry()
    log(f"geometry: n_boxes={geo['n_boxes']} n1d={geo['n1d']} fft_len={geo['fft_len']} nnz={nnz}")
    res64, g64, z64 = {}, None, None
    if not args.no_fp64:
        c64 = fg.Context(no_dims=1, precision=fg.FP64_PLAIN, df=df)
        c64.set_P(*P)
        c64.set_Y(Y.astype(np.float32).astype(np.float64))
        c64.set_exaggeration(4.0)
        g64, z64 = c64.gradient()
        dt64 = timed_steps(c64, args.steps, args.warmup, prm)
        res64 = {"ms_per_step": 1e3 * dt64 / args.steps, "iters_per_s": args.steps / dt64}
        c64.close()
    dt = timed_steps(c32, args.steps, args.warmup, prm)
    stats = kernel_profile(c32, max(args.profile_steps, 1), prm)
    c32.close()
    cnt, tot = stats.get("attractive", (0, 0.0))
    att_ms = tot / max(cnt, 1)
    att_bytes = 8 * (N + 1) + nnz * (4 + 4) + 2 * N * 4
    achieved = att_bytes / (att_ms * 1e-3) / 1e9 if att_ms > 0 else 0.0
    acc = {}
    if g64 is not None:
        acc["fp32_vs_fp64_gpu_grad_rel"] = rel(g32, g64)
        acc["fp32_vs_fp64_gpu_Z_rel"] = abs(z32 - z64) / abs(z64)
        acc["fp32_vs_fp64_gpu_per_point"] = gradient_errors(g32, g64)
    cpu = None
    if not args.no_cpu:
        log("CPU baseline: one oracle gradient call at N={N} (1D, df = 0.5)".format(N=N))
        v, nthr, _, _, med = cpu_baseline(P, Y.astype(np.float32).astype(np.float64), 3, df=df)
        cpu = {"value": v, "unit": "grad iters/s", "cores": nthr, "kind": "port", "cpu_model": cpu_model(),
               "sample": f"median of 3 gradients ({med:.2f} s each, after one untimed) of the same N={N} 1D "
                         f"workload (same Y, P, df = 0.5) by the oracle's C restatement of "
                         f"computeFftGradientOneDVariableDf, {nthr} threads for the reference's PARALLEL_FOR "
                         f"stages, serial spread/FFT as in the reference"}
    out = {
        "metric": "1D gradient iterations/s at N=2M, df=0.5 (C3), FP32 form",
        "value": args.steps / dt, "unit": "grad iters/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded Gaussian mixture, generated in place)",
        "config": {"workload": f"C3: N={N} 1D, df=0.5, late exaggeration 4, ~{nnz / N:.0f} nnz/row, span ~[-300, 300]",
                   "n_boxes": geo["n_boxes"], "n1d": geo["n1d"], "fft_len": geo["fft_len"],
                   "spread_points_per_workgroup": spread_pts},
        "roofline": {"bound": "hbm", "kernel": "k_attr_win<1>", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic("k_attr_win<1>", "attractive_c3_n2m_1d_hbm_traffic.md"),
                     "algorithmic_bytes": att_bytes, "avg_ms": att_ms},
        "cpu_baseline": cpu,
        "fp64_plain": res64,
        "speedup_vs_fp64": (res64["ms_per_step"] / (1e3 * dt / args.steps)) if res64 else None,
        "accuracy": acc,
        "kernels_ms": {k: round(v[1] / max(v[0], 1), 4) for k, v in sorted(stats.items(), key=lambda kv: -kv[1][1])},
    }
    print(json.dumps(out), flush=True)


def bench_trajectory(args, fg, synth):
    """Step time along a real optimisation: C2's affinities from the PCA start of its D = 50 data (sd 1e-4,
    as fast_tsne.py), early exaggeration 12 to iteration 250, momentum 0.5 then 0.8, 1000 iterations, with
    the points left in the order of their first Y, and again with a spatial re-sort (fitsne_reorder, then
    the autotuner) at the exaggeration switch and every 250 iterations after it.  Reports ms per step
    over 20-step windows starting at iterations 0, 250, 500 and 980, the whole run's time (re-sorts
    included) and the final KL."""
    N = args.n
    log(f"trajectory: generating C2 with its PCA start, N={N}")
    P, _, _, Y0 = synth.c2_workload(N, pca=True)
    lr = max(200.0, N / 12.0)
    windows = [0, 250, 500, 980]
    out_runs = {}
    for mode in ("first_order", "resorted"):
        c = fg.Context(no_dims=2, precision=fg.FP32)
        c.set_P(*P)
        c.set_exaggeration(12.0)
        c.set_Y(Y0)
        c.autotune()
        c.N = N
        win_ms = {}
        resort_s = 0.0
        c.sync()
        t_run = time.perf_counter()
        t_win = None
        for it in range(1000):
            if it in windows:
                c.sync()
                t_win = time.perf_counter()
            c.step(0.5 if it < 250 else 0.8, lr, 5.0)
            if t_win is not None and it + 1 - max(w for w in windows if w <= it) == 20:
                c.sync()
                win_ms[str(max(w for w in windows if w <= it))] = 1e3 * (time.perf_counter() - t_win) / 20
                t_win = None
            if it == 249:
                c.set_exaggeration(1.0)
            if mode == "resorted" and it % 250 == 249:
                c.sync()
                t0 = time.perf_counter()
                c.reorder()
                c.autotune()
                c.sync()
                resort_s += time.perf_counter() - t0
        c.sync()
        total = time.perf_counter() - t_run
        kl = c.kl()
        geo = c.last_geometry()
        c.close()
        out_runs[mode] = {"ms_per_step_at": win_ms, "run_s": total, "resort_s": resort_s, "kl": kl,
                          "n_boxes_end": geo["n_boxes"]}
        log(f"trajectory {mode}: {win_ms}, run {total:.2f} s, KL {kl:.4f}")
    out = {
        "metric": "FP32 gradient step time along a 1000-iteration optimisation at N=1M 2D (C2 affinities, PCA start)",
        "value": 1000.0 / out_runs["resorted"]["run_s"], "unit": "iters/s (whole run, re-sorts included)",
        "n_gpus": 1, "steps": 1000, "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded Gaussian mixture, generated in place)",
        "config": {"workload": f"C2 affinities, N={N}, PCA start sd 1e-4, early exaggeration 12 to 250, "
                               f"lr {lr:.1f}"},
        "runs": out_runs,
    }
    print(json.dumps(out), flush=True)


def normalize_input(X):
    """TSNE::run's input normalisation (tsne.cpp:153-161): zero mean, then divide by max |x|."""
    X = X - X.mean(axis=0)
    X /= np.abs(X).max()
    return X


def bench_c5(args, fg, synth):
    """C5, the learned warm start at scale: 2M (distance, beta) training pairs from two separately seeded
    1M-point mixtures (exact kNN on MFMA, K = 90, the reference's FP64 bisection at perplexity 30 on the
    kNN distances as its list overload passes them, i.e. taken as squared: tsne.cpp:1479), the
    BF16 MFMA network trained on them, then on a 4M-point, D = 100, 200-cluster mixture: exact kNN, the
    bisection cold and warm-started (iterations, |H - log perplexity|), the symmetric P, and the full 2D
    FP32 embedding (1000 iterations, early exaggeration 12 to 250, PCA start)."""
    N = args.n if args.n != 1_000_000 else 4_000_000
    K, perp = 90, 30.0
    logp = np.log(perp)

    def entropy_err(P):
        with np.errstate(divide="ignore", invalid="ignore"):
            H = -np.sum(np.where(P > 0, P * np.log(P), 0.0), axis=1)
        return float(np.abs(H - logp).max())

    dists, betas, t_knn_train = [], [], 0.0
    for seed in (101, 102):
        X, _ = synth.c5_mixture(N=N // 4, D=100, k=200, seed=seed)
        X = normalize_input(X)
        t0 = time.perf_counter()
        _, d, redo = fg.knn_mfma(X, K)
        t_knn_train += time.perf_counter() - t0
        _, b, _ = fg.distances2similarities(d, perp, squared=True)
        dists.append(d)
        betas.append(b)
        log(f"C5 training mixture seed {seed}: {X.shape[0]} points, kNN recomputed {redo}")
        del X
    dtr, btr = np.concatenate(dists), np.concatenate(betas)
    del dists, betas
    model = fg.BetaModel(K=K, seed=11)
    t0 = time.perf_counter()
    loss = model.fit(dtr, btr, epochs=4, batch=1024, lr=2e-3)
    t_fit = time.perf_counter() - t0
    log(f"C5 network trained on {dtr.shape[0]} pairs in {t_fit:.1f} s, loss {loss:.4f}")
    del dtr, btr
    log(f"C5 main mixture: N={N}, D=100, 200 clusters")
    X, _ = synth.c5_mixture(N=N, D=100, k=200, seed=100)
    X = normalize_input(X)
    Y0 = synth.pca_start(X)
    t0 = time.perf_counter()
    idx, dist, redo = fg.knn_mfma(X, K)
    t_knn = time.perf_counter() - t0
    log(f"C5 kNN in {t_knn:.1f} s, {redo} queries recomputed by the FP64 kernel")
    del X
    t0 = time.perf_counter()
    P_cold, b_cold, it_cold = fg.distances2similarities(dist, perp, squared=True)
    t_cold = time.perf_counter() - t0
    t0 = time.perf_counter()
    beta0 = model.predict(dist)
    t_pred = time.perf_counter() - t0
    t0 = time.perf_counter()
    P_warm, b_warm, it_warm = fg.distances2similarities(dist, perp, squared=True, beta0=beta0)
    t_warm = time.perf_counter() - t0
    model.close()
    e_cold, e_warm = entropy_err(P_cold), entropy_err(P_warm)
    log(f"C5 bisection iterations cold {it_cold.mean():.3f} warm {it_warm.mean():.3f}; |H - log perp| "
        f"cold {e_cold:.2e} warm {e_warm:.2e}")
    # symmetric, normalised P from the warm-started rows (tsne.cpp:1730-1828, :321-324)
    row = (np.arange(N + 1, dtype=np.int64) * K).astype(np.uint32)
    sr, sc, sv = fg.symmetrize(row, idx.reshape(-1).astype(np.uint32), P_warm.reshape(-1))
    sv /= sv.sum()
    del P_cold, P_warm, idx, dist
    nnz = int(sr[-1])
    log(f"C5 symmetric P: {nnz} entries ({nnz / N:.1f} per row)")
    c = fg.Context(no_dims=2, precision=fg.FP32)
    c.set_P(sr.astype(np.int64), sc.astype(np.int32), sv)
    c.set_exaggeration(12.0)
    c.set_Y(Y0)
    c.autotune()
    c.N = N
    lr = max(200.0, N / 12.0)
    c.sync()
    t0 = time.perf_counter()
    for it in range(1000):
        c.step(0.5 if it < 250 else 0.8, lr, 5.0)
        if it == 249:
            c.set_exaggeration(1.0)
    c.sync()
    t_emb = time.perf_counter() - t0
    kl = c.kl()
    geo = c.last_geometry()
    c.close()
    log(f"C5 embedding: 1000 iterations in {t_emb:.1f} s, KL {kl:.4f}, {geo['n_boxes']} boxes per dimension")
    out = {
        "metric": "C5 learned warm start: bisection iterations saved at N=4M, D=100, perplexity 30, K=90",
        "value": float(1.0 - it_warm.mean() / it_cold.mean()), "unit": "fraction of bisection iterations saved",
        "n_gpus": 1, "steps": 1000, "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 network (fp32 accumulate), fp64 bisection, fp32 embedding",
        "data": "synthetic (seeded Gaussian mixtures, generated in place)",
        "config": {"workload": f"C5: N={N}, D=100, 200 clusters with sd log-uniform in [0.3, 3], perplexity 30, "
                               f"K=90; network trained on {N // 2} pairs from two separately seeded mixtures"},
        "bisection": {"iters_cold_mean": float(it_cold.mean()), "iters_warm_mean": float(it_warm.mean()),
                      "max_abs_H_minus_log_perp_cold": e_cold, "max_abs_H_minus_log_perp_warm": e_warm,
                      "max_rel_beta_diff": float(np.max(np.abs(b_warm / b_cold - 1.0))),
                      "s_cold": t_cold, "s_warm": t_warm, "s_predict": t_pred},
        "knn": {"s": t_knn, "recomputed_by_fp64": int(redo), "s_training_sets": t_knn_train},
        "network": {"train_pairs": N // 2, "fit_s": t_fit, "final_loss": loss},
        "embedding": {"iterations": 1000, "s": t_emb, "ms_per_step": t_emb, "kl": kl, "nnz": nnz,
                      "n_boxes_end": geo["n_boxes"]},
    }
    print(json.dumps(out), flush=True)


def bench_bh(args, fg, synth):
    """Barnes-Hut (nbody_algo = 1, theta = 0.5) at N = 100,000 in 2D: each step builds the reference's
    SPTree on the host in insertion order, then walks it and sums the edge forces on the GPU (bitwise the
    reference's gradient).  The walk is irregular (pointer chasing over a tree of ~2N cells), so it has no
    meaningful bandwidth roofline; the line reports the step and its two halves."""
    N = args.n if args.n != 1_000_000 else 100_000
    P, Y, _ = synth.small_mixture_workload(N=N, D=2, k=20, per_point=30, span=60.0, seed=20261202)
    theta = 0.5
    c = fg.Context(no_dims=2, precision=fg.FP64_BH)
    c.set_theta(theta)
    c.set_P(*P)
    c.set_Y(Y)
    prm = dict(momentum=0.8, learning_rate=max(200.0, N / 12.0), max_step_norm=5.0)
    dt = timed_steps(c, args.steps, args.warmup, prm)
    stats = kernel_profile(c, max(args.profile_steps, 1), prm)
    c.close()
    gpu_ms = sum(v[1] for v in stats.values()) / max(args.profile_steps, 1)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import orc

        log(f"CPU baseline: oracle Barnes-Hut gradient at N={N} (serial, the reference's SPTree)")
        t0 = time.perf_counter()
        orc.bh_gradient(P, Y, theta)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": 1.0 / t_cpu, "unit": "grad iters/s", "cores": 1, "kind": "port",
               "sample": f"one gradient of the same N={N} workload by the oracle's restatement of "
                         f"computeGradient (tree build and walk, one thread), {t_cpu:.2f} s"}
    out = {
        "metric": "Barnes-Hut (theta=0.5) gradient iterations/s at N=100000 2D, FP64",
        "value": args.steps / dt, "unit": "grad iters/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded Gaussian mixture, generated in place)",
        "config": {"workload": f"Barnes-Hut t-SNE: N={N} 2D, theta={theta}, ~60 nnz/row symmetric P"},
        "roofline": None,
        "gpu_ms_per_step": gpu_ms,
        "host_tree_ms_per_step": 1e3 * dt / args.steps - gpu_ms,
        "cpu_baseline": cpu,
        "kernels_ms": {k: round(v[1] / max(v[0], 1), 4) for k, v in sorted(stats.items(), key=lambda kv: -kv[1][1])},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 100 timed steps: a 0.3 ms step timed over 20 varied by several percent from run to run
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--cpu-calls", type=int, default=9, help="timed oracle calls (after one untimed) for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-c4", action="store_true", help="leave the batched C4 figure out of the default line")
    ap.add_argument("--no-captured", action="store_true",
                    help="--config c4: leave out the run of each problem's captured step one after another")
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--config", default="c2", choices=["c2", "c3", "c4", "c5", "exact", "bh", "trajectory"],
                    help="c2 (default): the headline step at N=1M 2D; c3: 1D df=0.5 at N=2M; c4: 64 batched "
                         "embeddings; exact: theta = 0 at N=46000 2D; bh: Barnes-Hut at N=100000 2D")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one plain 64 MiB device copy (torch) as the reference kernel for "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE scaling")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: runs on several GPUs are not supported (the gradient path is "
              "single-GPU)", file=sys.stderr)
        sys.exit(2)

    import fitsne_gpu as fg
    import synth

    if args.copy_calibration:
        import torch

        # source written first and then pushed out of L2 and the MALL by 448 MiB of later writes
        src = torch.ones(512 * 2 ** 20 // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty(64 * 2 ** 20 // 4, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        dst.copy_(src[: dst.numel()])
        torch.cuda.synchronize()
        del src, dst
        log("copy calibration: one copy kernel reading 64 MiB (cold) and writing 64 MiB")

    if args.config == "c4":
        return bench_c4(args, fg, synth)
    if args.config == "c3":
        return bench_c3(args, fg, synth)
    if args.config == "exact":
        return bench_exact(args, fg)
    if args.config == "bh":
        return bench_bh(args, fg, synth)
    if args.config == "trajectory":
        return bench_trajectory(args, fg, synth)
    if args.config == "c5":
        return bench_c5(args, fg, synth)

    N = args.n
    log(f"generating the C2 workload N={N}")
    P, Y, _ = synth.c2_workload(N)
    nnz = int(P[0][-1])
    lr = max(200.0, N / 12.0)
    prm = dict(momentum=0.8, learning_rate=lr, max_step_norm=5.0)

    log("FP32 context: upload P, Y")
    c32 = fg.Context(no_dims=2, precision=fg.FP32)
    c32.set_P(*P)
    c32.set_Y(Y)
    spread_pts = c32.autotune()
    log(f"autotuner: {spread_pts} points per spread workgroup")
    g32, z32 = c32.gradient()
    geo = c32.last_geometry()
    log(f"geometry: n_boxes={geo['n_boxes']} n1d={geo['n1d']} fft_len={geo['fft_len']} nnz={geo['nnz']}")

    res64 = {}
    g64 = z64 = None
    if not args.no_fp64:
        log("FP64 plain context")
        c64 = fg.Context(no_dims=2, precision=fg.FP64_PLAIN)
        c64.set_P(*P)
        c64.set_Y(Y.astype(np.float32).astype(np.float64))  # the Y the FP32 context holds
        g64, z64 = c64.gradient()
        dt64 = timed_steps(c64, args.steps, args.warmup, prm)
        res64 = {"ms_per_step": 1e3 * dt64 / args.steps, "iters_per_s": args.steps / dt64}
        log(f"FP64 plain: {res64['ms_per_step']:.3f} ms/step")
        c64.close()

    log("FP32 timed steps")
    dt = timed_steps(c32, args.steps, args.warmup, prm)
    ms = 1e3 * dt / args.steps
    value = args.steps / dt
    log(f"FP32: {ms:.3f} ms/step, {value:.1f} iters/s")
    stats = kernel_profile(c32, max(args.profile_steps, 1), prm)
    geo_end = c32.last_geometry()
    # the same kernels one after another on one stream: each kernel's time alone
    c32.set_overlap(False)
    stats_alone = kernel_profile(c32, max(args.profile_steps, 1), prm)
    geo_alone = c32.last_geometry()
    c32.close()

    cnt, tot = stats.get("attractive", (0, 0.0))
    att_ms = tot / max(cnt, 1)
    att_bytes = 8 * (N + 1) + nnz * (4 + 4) + 2 * N * 2 * 4
    achieved = att_bytes / (att_ms * 1e-3) / 1e9 if att_ms > 0 else 0.0
    kernels_ms = {k: round(v[1] / max(v[0], 1), 4) for k, v in sorted(stats.items(), key=lambda kv: -kv[1][1])}
    per_kernel = per_kernel_roofline(stats_alone, stats, N, nnz, geo_alone)

    acc = {}
    if g64 is not None:
        acc["fp32_vs_fp64_gpu_grad_rel"] = rel(g32, g64)
        acc["fp32_vs_fp64_gpu_Z_rel"] = abs(z32 - z64) / abs(z64)
        acc["fp32_vs_fp64_gpu_per_point"] = gradient_errors(g32, g64)
    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline: 1 + {args.cpu_calls} oracle gradient calls at N={N}")
        Y32 = Y.astype(np.float32).astype(np.float64)
        v, nthr, gref, zref, med = cpu_baseline(P, Y32, args.cpu_calls)
        cpu = {"value": v, "unit": "grad iters/s", "cores": nthr, "kind": "port", "cpu_model": cpu_model(),
               "sample": f"median of {args.cpu_calls} full gradient evaluations ({med:.3f} s each) after one "
                         f"untimed call, on the same N={N} workload (same Y, P), by the oracle's C restatement "
                         f"of computeFftGradient: {nthr} threads for the reference's PARALLEL_FOR stages, "
                         f"serial spread/FFT as in the reference, own FFT (no FFTW on the box)"}
        acc["fp32_vs_oracle_grad_rel"] = rel(g32, gref)
        acc["fp32_vs_oracle_Z_rel"] = abs(z32 - zref) / abs(zref)
        acc["fp32_vs_oracle_per_point"] = gradient_errors(g32, gref)
        if g64 is not None:
            acc["fp64_vs_oracle_grad_rel_at_fp32_rounded_Y"] = rel(g64, gref)

    c4 = None
    if not args.no_c4:
        c4 = c4_measure(args, fg, synth, one_by_one=False)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "grad iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian mixture, generated in place)",
        "config": {"workload": "C2: N=1,000,000 2D, df=1, symmetric P ~90 nnz/row, Y span ~[-80,80]",
                   "N": N, "nnz": nnz, "n_boxes_per_dim": geo["n_boxes"], "n1d": geo["n1d"],
                   "fft_len": geo["fft_len"], "n_boxes_end": geo_end["n_boxes"], "learning_rate": lr,
                   "momentum": 0.8, "max_step_norm": 5.0, "parallelism": "single GPU",
                   "spread_points_per_workgroup": spread_pts},
        "roofline": {"bound": "hbm", "kernel": ATTR_KERNEL, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(ATTR_KERNEL),
                     "algorithmic_bytes": att_bytes, "avg_ms": att_ms, "per_kernel": per_kernel},
        "cpu_baseline": cpu,
        "fp64_plain": res64,
        "speedup_vs_fp64": (res64["ms_per_step"] / ms) if res64 else None,
        "accuracy": acc,
        "kernels_ms": kernels_ms,
        "c4_batched": c4,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
