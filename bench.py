#!/usr/bin/env python3
"""Benchmark of the LL iteration path on one MI355X.

Workload: the large single exponent -- seeded prime p = 77,005,277 at FFT
length N = 4,194,304 words (18.36 bits/word), s_0 = 4.  One step is one LL
iteration (square sweep + carry sweep, hipGraph-replayed).  Prints one JSON
line: ms per iteration over exactly --steps iterations after --warmup
untimed ones, the roofline of the dominant kernel (algorithmic bytes
16 N per sweep launch over its average duration, from events set at each
profiled kernel's own start and end), and a
CPU baseline: the oracle's IBDWT restatement of the same (p, N) over a
bounded sample, on every host core the process may use (OMP_NUM_THREADS,
else the affinity mask; "cores" in the line) and on one core ("one_core"),
each residue checked against the GPU's.

    python bench.py [--gpus 1] [--steps 10000] [--warmup 200]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
# LL_OPT_NO_GRAPH under --no-graph: kernels launched one by one (for
# profilers that do not follow hipGraph replay); the kernels are the same
CTX_FLAGS = 0
FP64_PEAK_TFLOPS = 78.6


def algorithmic(n):
    """Bytes and flops of one iteration at FFT length n (words)."""
    import math

    m = n // 2
    return 32.0 * n, 10.0 * m * math.log2(m) + 20.0 * n


def _cpu_leg(p, n, budget_s):
    """One timed run of the oracle IBDWT from s_0 = 4: (ms/iter, iterations,
    seconds, residue words)."""
    import oracle as O

    c = O.IBDWT(p, n)
    c.load(O.initial_digits(p, n, 0), 0)
    t0 = time.perf_counter()
    c.iterate(1)
    one = time.perf_counter() - t0
    extra = max(2, min(50, int(budget_s / max(one, 1e-3)) - 1))
    t0 = time.perf_counter()
    c.iterate(extra)
    dt = time.perf_counter() - t0
    d, off = c.read()
    words = O.digits_to_residue(p, n, d, off)
    c.close()
    return dt * 1000.0 / extra, 1 + extra, dt, words


def cpu_baseline(p, n, budget_s):
    """Oracle IBDWT on the same workload, on every host core this process may
    use (its FFT passes, pair squaring and rounding run in parallel; the carry
    chain is sequential) and on one core; each residue is checked against the
    GPU's after the same number of iterations."""
    import cudalucas_amd as L
    import oracle as O

    omp = (os.environ.get("OMP_NUM_THREADS") or "").split(",")[0].strip()  # "16" or a list "16,8"
    want = (int(omp) if omp.isdigit() else 0) or len(os.sched_getaffinity(0))
    legs = []
    for threads in ([want, 1] if want > 1 else [1]):
        used = O.set_threads(threads)
        try:
            ms, iters, dt, words = _cpu_leg(p, n, budget_s)
        finally:
            O.set_threads(1)
        with L.LLContext(p, n) as g:
            g.load_initial(0)
            g.iterate(iters)
            gw, _ = g.residue()
        legs.append({
            "value": round(ms, 3),
            "unit": "ms/iter",
            "cores": used,
            "kind": "port",
            "sample": "iterations 2..%d of p=%d at N=%d from s_0=4 (oracle IBDWT, %d host thread%s, %.1f s)"
            % (iters, p, n, used, "" if used == 1 else "s", dt),
            "residue_matches_gpu": bool((words == gw).all()),
        })
    out = legs[0]
    if len(legs) > 1:
        out["one_core"] = legs[1]
    return out


def batched_leg(steps=4, step_iters=512, plain_iters=200, plain=True):
    """Batched workload: 256 seeded exponents log-uniform in [1e5, 5e6] at their
    chosen FFT lengths.  Exponents/hour, with finished tests replaced by new
    ones of the same mix: batched = 3600 * 256 / (t_step * mean(p - 2));
    plain (one exponent at a time on the same GPU) =
    3600 / mean((p - 2) * t_iter(p)), left out when `plain` is false."""
    import numpy as np

    import cudalucas_amd as L
    from cudalucas_amd import tuner as T
    from cudalucas_amd import workloads as W

    ps = W.log_uniform_primes(W.BATCH_SEED, W.BATCH_COUNT, *W.BATCH_RANGE)
    lens = [L.choose_fft_length(p)[0] for p in ps]
    tests = [(p, n, p % 1000, p - 2) for p, n in zip(ps, lens)]
    tuned = os.path.exists(T.BATCH_TABLE)
    if tuned:  # schedules picked by the batched autotuner
        T.install(T.load(T.BATCH_TABLE))
    try:
        with L.LLBatch(tests) as b:
            b.step(64)
            total_ms, worst = 0.0, 0.0
            for _ in range(steps):
                e = b.step(step_iters)
                total_ms += b.last_step_ms()
                worst = max(worst, float(e.max()))
            groups = b.groups()
    finally:
        if tuned:
            T.install([])
    t_step = total_ms * 1e-3 / (steps * step_iters)
    mean_iters = float(np.mean([p - 2 for p in ps]))
    batched = 3600.0 * len(ps) / (t_step * mean_iters)
    # algorithmic bytes of one batch iteration: 32 N per test (two read+write
    # sweeps of its N-word FP64 signal; LDS-resident tests counted alike)
    b_alg = 32.0 * float(sum(lens))
    out = {
        "workload": "256 seeded primes log-uniform in [1e5, 5e6], FFT lengths %dK-%dK (%d schedule groups)"
        % (min(lens) // 1024, max(lens) // 1024, groups),
        "exps_per_hour_batched": round(batched, 1),
        "batch_iteration_ms": round(t_step * 1e3, 4),
        "hbm_gbs_alg": round(b_alg / t_step / 1e9, 1),
        "sample": "%d timed batch iterations after 64 warm-up" % (steps * step_iters),
        "schedules": "batched autotuner table" if tuned else "built-in preference",
        "max_roundoff": round(worst, 4),
    }
    if not plain:
        return out
    per = []
    for p, n in zip(ps, lens):
        with L.LLContext(p, n) as c:
            c.load_initial(p % 1000)
            c.iterate(32)  # builds and replays the context's graph once
            c.iterate(plain_iters)
            per.append((p - 2) * c.last_iterate_ms() * 1e-3 / plain_iters)
    plain_rate = 3600.0 / float(np.mean(per))
    out["exps_per_hour_plain"] = round(plain_rate, 1)
    out["speedup"] = round(batched / plain_rate, 2)
    out["sample"] += "; plain: %d iterations per exponent after 32 warm-up, all 256" % plain_iters
    return out


def smooth_leg(iters=2000, warmup=100):
    """The smooth-length sweep: each 3/5/7-smooth length with its seeded
    exponent at ~18.5 bits/word, ms/iteration over `iters` hipGraph-replayed
    iterations (wall clock and hipEvent GPU time) and HBM GB/s at the
    algorithmic 32 N bytes per iteration."""
    import cudalucas_amd as L
    from cudalucas_amd import workloads as W

    rows = []
    for n in W.SMOOTH_LENGTHS:
        p = W.smooth_prime(n)
        with L.LLContext(p, n, flags=CTX_FLAGS) as c:
            info = c.info()
            c.load_initial(0)
            c.iterate(warmup)
            t0 = time.perf_counter()
            m = c.iterate(iters)
            wall = time.perf_counter() - t0
            gpu = c.last_iterate_ms() / iters
        b_alg, _ = algorithmic(n)
        rows.append({
            "n": n, "p": p, "bits_per_word": round(p / n, 3),
            "schedule": "W=%d H=%d R=%d" % (info["width"], info["height"], info["lines"]),
            "ms_per_iter": round(wall * 1000.0 / iters, 5), "gpu_ms_per_iter": round(gpu, 5),
            "hbm_gbs_alg": round(b_alg / (gpu * 1e-3) / 1e9, 1),
            "hbm_frac_alg": round(b_alg / (gpu * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "max_roundoff": round(float(m), 4),
        })
    return rows


def lengths_leg(spec, iters=500, warmup=100):
    """Extra lengths ("8192K,3136K,..."): the densest exponent the
    reference's policy assigns to each, ms/iteration (hipEvent GPU time) and
    HBM GB/s at 32 N bytes per iteration, with the schedule and kernel path
    (specialised sweeps or the generic runtime-plan kernels)."""
    import cudalucas_amd as L
    from cudalucas_amd import workloads as W

    rows = []
    for tok in spec.split(","):
        n = L.fft_from_str(tok.strip())
        p = W.densest_prime(n)
        with L.LLContext(p, n, flags=CTX_FLAGS) as c:
            info = c.info()
            c.load_initial(0)
            c.iterate(warmup)
            m = c.iterate(iters)
            gpu = c.last_iterate_ms() / iters
        b_alg, _ = algorithmic(n)
        rows.append({"n": n, "p": p, "bits_per_word": round(p / n, 3), "path": info["path"],
                     "schedule": "W=%d H=%d R=%d" % (info["width"], info["height"], info["lines"]),
                     "gpu_ms_per_iter": round(gpu, 5), "hbm_gbs_alg": round(b_alg / (gpu * 1e-3) / 1e9, 1),
                     "hbm_frac_alg": round(b_alg / (gpu * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "max_roundoff": round(float(m), 4)})
    return rows


def small_leg():
    """The plain single exponent at N = 4096 (LDS-resident): all p - 2
    iterations in one call, microseconds per iteration."""
    import cudalucas_amd as L
    from cudalucas_amd import workloads as W

    p, n = W.SMALL_P, W.SMALL_N
    with L.LLContext(p, n) as c:
        c.load_initial(0)
        c.iterate(1000)
        c.load_initial(0)
        t0 = time.perf_counter()
        m = c.iterate(p - 2)
        wall = time.perf_counter() - t0
        gpu = c.last_iterate_ms() / (p - 2)
        _, r = c.residue()
    return {"p": p, "n": n, "iterations": p - 2, "us_per_iter": round(wall * 1e6 / (p - 2), 3),
            "gpu_us_per_iter": round(gpu * 1e3, 3), "iter_per_s": round((p - 2) / wall, 1),
            "max_roundoff": round(float(m), 4), "res64": "0x%016x" % r}


def load_traffic(name):
    path = os.path.join(HERE, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--profile-iters", type=int, default=200)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-only", action="store_true",
                    help="only the 64 MiB streaming-copy calibration kernel")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="batched leg without the one-exponent-at-a-time comparison")
    ap.add_argument("--no-smooth", action="store_true", help="skip the smooth-length and N=4096 legs")
    ap.add_argument("--lengths", default="", help="extra lengths to time, e.g. 8192K,3136K (densest exponents)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch each iteration's kernels one by one instead of replaying a hipGraph")
    args = ap.parse_args()
    global CTX_FLAGS
    if args.no_graph:
        CTX_FLAGS = 2  # LL_OPT_NO_GRAPH
    if args.gpus != 1:
        print("bench.py: measures one GPU; the LL path of one exponent runs on a single "
              "MI355X, and several GPUs would each take their own exponents", file=sys.stderr)
        sys.exit(2)

    import cudalucas_amd as L
    from cudalucas_amd import workloads as W

    copy_bytes = 64 << 20
    copy_ms = L.bench_copy(copy_bytes, reps=50)
    copy_gbs = 2.0 * copy_bytes / (copy_ms * 1e-3) / 1e9
    if args.copy_only:
        print(json.dumps({"copy_bytes": copy_bytes, "copy_ms": copy_ms, "copy_gbs": round(copy_gbs, 1)}))
        return

    from cudalucas_amd import tuner as T

    p, n = W.LARGE_P, W.LARGE_N
    single_tuned = os.path.exists(T.SINGLE_TABLE)
    if single_tuned:  # schedules picked by the single-exponent autotuner
        T.install(T.load(T.SINGLE_TABLE))
    ctx = L.LLContext(p, n, flags=CTX_FLAGS)
    if single_tuned:
        T.install([])
    info = ctx.info()
    ctx.load_initial(0)
    if args.warmup:
        ctx.iterate(args.warmup)
    t0 = time.perf_counter()
    maxerr, errs = ctx.iterate(args.steps, per_iter=True)
    wall = time.perf_counter() - t0
    gpu_ms = ctx.last_iterate_ms()
    sq_ms, ca_ms = ctx.profile_kernels(args.profile_iters)
    words, res64 = ctx.residue()
    done_iters = args.warmup + args.steps + args.profile_iters
    ctx.close()

    b_alg, f_alg = algorithmic(n)
    ms_per_step = wall * 1000.0 / args.steps
    gpu_ms_per_iter = gpu_ms / args.steps
    dom_name, dom_ms = ("carry_sweep", ca_ms) if ca_ms >= sq_ms else ("square_sweep", sq_ms)
    per_launch_bytes = 16.0 * n  # one read + one write of the N-word signal
    achieved = per_launch_bytes / (dom_ms * 1e-3) / 1e9
    traffic = load_traffic("traffic_4m.json")
    roofline = {
        "bound": "hbm",
        "kernel": dom_name,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic.get(dom_name) if traffic else None,
        "avg_launch_ms": {"square_sweep": round(sq_ms, 5), "carry_sweep": round(ca_ms, 5)},
        "copy_kernel_gbs": round(copy_gbs, 1),
    }
    out = {
        "metric": "ms/iter at 4M FFT (p~77M) + batched exps/hour vs plain; HBM GB/s",
        "value": round(ms_per_step, 5),
        "unit": "ms/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "large single exponent: seeded prime p=%d, FFT length N=%d words "
            "(%.2f bits/word), s_0=4" % (p, n, p / n),
            "p": p,
            "n": n,
            "schedule": "%s W=%d H=%d R=%d (%s)" % (info["path"], info["width"], info["height"], info["lines"],
                                                     "autotuned" if single_tuned else "built-in preference"),
        },
        "gpu_ms_per_iter": round(gpu_ms_per_iter, 5),
        "iter_per_s": round(1000.0 / ms_per_step, 1),
        "hbm_gbs_alg": round(b_alg / (gpu_ms_per_iter * 1e-3) / 1e9, 1),
        "hbm_frac_alg": round(b_alg / (gpu_ms_per_iter * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "fp64_util": round(f_alg / (gpu_ms_per_iter * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
        "max_roundoff": round(float(maxerr), 5),
        "res64": "0x%016x" % res64,
        "res64_iteration": done_iters,
        "roofline": roofline,
    }
    if not args.no_smooth:
        if single_tuned:
            T.install(T.load(T.SINGLE_TABLE))
        try:
            out["smooth"] = smooth_leg()
        finally:
            if single_tuned:
                T.install([])
        out["small"] = small_leg()
    if args.lengths:
        if single_tuned:
            T.install(T.load(T.SINGLE_TABLE))
        try:
            out["lengths"] = lengths_leg(args.lengths)
        finally:
            if single_tuned:
                T.install([])
    if not args.no_batched:
        out["batched"] = batched_leg(plain=not args.no_plain)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(p, n, args.cpu_budget)
        except Exception as e:  # the GPU measurement above stands on its own
            out["cpu_baseline"] = {"value": None, "unit": "ms/iter", "error": repr(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
