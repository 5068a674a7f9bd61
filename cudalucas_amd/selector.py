"""Learned FFT-length selection (roundoff surrogate), trained on the GPU.

The reference chooses an FFT length from fixed power-law bounds
(choose_fft_length, CUDALucas.cu:1027-1068) and discovers a length that is
too short only when the careful round-off test fails (CUDALucas.cu:2589-2648:
any of the first 1000 iterations above 0.35, or their average above 0.25,
restarts at the next length).  Here an FP32 MLP (hand-written HIP kernels,
ll_mlp_*) predicts the maximum roundoff of (p, n) from seven features, and
ll_selector_choose takes the fastest length predicted to stay under 0.35 over
a whole test; the round-off test stays as the fallback check.

Pipeline (every GPU step goes through libcudalucas_amd.so):

1. ``trace_plan``: p log-uniform seeded primes in [1e5, 1.2e8]; n uniform
   over the candidate lengths of p (supported multiples of 1K words in
   [0.62, 1.25] x the reference policy's length, at most 28 bits/word).
2. ``collect_traces``: each (p, n) runs from s_0 = 4 (random offset) in the
   batched engine -- 32 warm-up iterations (s_k has few bits until k ~ log2 p),
   then 8 windows of 16 iterations, each window's maximum roundoff recorded.
3. ``fit``: target log2(max over the 128 iterations), clipped to
   [2^-24, 0.5]; standardised features; MLP 7-64-64-64-1 trained with Adam
   on the GPU; margin = 99.5 % quantile of held-out residuals near the limit;
   growth_scale = how much faster the maximum grew from 16 to 128 iterations
   than the Gaussian model says (never below 1).
4. ``save_model`` writes the text model that ``ll_selector_load`` reads.
5. ``evaluate``: selector choice -> careful round-off test in the batched
   engine -> fallback to the next length on failure; residues at the end of
   the test compared with the same iterations at the reference policy's
   length; projected one-at-a-time time from measured ms/iteration per
   length (against the reference policy on its default list and on the list
   its own cufftbench rule would keep), and measured batched throughput at
   the reference's and at the selected lengths.

    python -m cudalucas_amd.selector traces --out traces.npz
    python -m cudalucas_amd.selector train --traces traces.npz --model model.txt
    python -m cudalucas_amd.selector times --out times.txt
    python -m cudalucas_amd.selector evaluate --model model.txt --times times.txt
"""
import argparse
import ctypes
import hashlib
import json
import math
import os
import sys
import time

import numpy as np

from . import LLBatch, LLContext, LLError, MLP, _ptr, choose_fft_length, lib
from . import workloads as W

_HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(_HERE, "data", "roundoff_mlp.txt")
TIMES_PATH = os.path.join(_HERE, "data", "mi355x_fft_times.txt")
BATCH_TIMES_PATH = os.path.join(_HERE, "data", "mi355x_batched_times.txt")

FEATURES = 7
SIZES = [FEATURES, 64, 64, 64, 1]
WARMUP = 32          # iterations before the first window
WINDOWS = 8          # windows per trace
WINDOW_ITERS = 16    # iterations per window
LIMIT = 0.35         # careful round-off test: per-iteration maximum
LIMIT_AVG = 0.25     # careful round-off test: average over the test
CHECK_ITERS = 1000   # careful round-off test length
CLIP_LO = 2.0 ** -24


# ------------------------------------------------------------- features

def features(p, n):
    f = np.zeros(FEATURES, dtype=np.float32)
    k = lib().ll_selector_features(p, n, _ptr(f))
    if k != FEATURES:
        raise LLError(k, "ll_selector_features(%d, %d)" % (p, n))
    return f


def feature_matrix(ps, ns):
    return np.stack([features(int(p), int(n)) for p, n in zip(ps, ns)])


def candidates(p):
    k = lib().ll_selector_candidates(p, None, 0)
    if k < 0:
        raise LLError(k, "ll_selector_candidates(%d)" % p)
    out = np.zeros(k, dtype=np.uint32)
    lib().ll_selector_candidates(p, _ptr(out), k)
    return [int(x) for x in out]


# ---------------------------------------------------------------- traces

def trace_plan(count=W.SELECTOR_TRAIN_COUNT, seed=W.SELECTOR_TRAIN_SEED, lo=W.SELECTOR_TRAIN_RANGE[0], hi=W.SELECTOR_TRAIN_RANGE[1]):
    """[(p, n, offset)]: seeded log-uniform primes, a uniform candidate n."""
    ps = W.log_uniform_primes(seed, count, lo, hi)
    rng = W.SplitMix64(seed ^ 0x5E1EC7)
    plan = []
    for p in ps:
        c = candidates(p)
        n = c[rng.next() % len(c)]
        plan.append((p, n, rng.next() % p))
    return plan


def _chunks(items, words_budget, max_tests):
    """Consecutive chunks bounded by total words and test count."""
    chunk, words = [], 0
    for it in items:
        if chunk and (words + it[1] > words_budget or len(chunk) >= max_tests):
            yield chunk
            chunk, words = [], 0
        chunk.append(it)
        words += it[1]
    if chunk:
        yield chunk


def collect_traces(plan, words_budget=1 << 27, max_tests=256, log=None):
    """Window maxima [len(plan), WINDOWS] of every (p, n, offset)."""
    order = sorted(range(len(plan)), key=lambda i: (plan[i][1], plan[i][0]))
    items = [(i, plan[i][1]) for i in order]
    out = np.zeros((len(plan), WINDOWS), dtype=np.float32)
    done, t0 = 0, time.time()
    total = WARMUP + WINDOWS * WINDOW_ITERS
    for chunk in _chunks(items, words_budget, max_tests):
        tests = [(plan[i][0], plan[i][1], plan[i][2], total) for i, _ in chunk]
        with LLBatch(tests) as b:
            b.step(WARMUP)
            for w in range(WINDOWS):
                e = b.step(WINDOW_ITERS)
                for k, (i, _) in enumerate(chunk):
                    out[i, w] = e[k]
        done += len(chunk)
        if log:
            log("traces %d/%d (n up to %dK) %.1f s" % (done, len(plan), chunk[-1][1] // 1024, time.time() - t0))
    return out


def targets(wmax):
    """log2 of the 128-iteration maximum, clipped; failures (>= 0.5, NaN,
    carry overflow) saturate at log2(0.5)."""
    m = np.nan_to_num(np.asarray(wmax, dtype=np.float64).max(axis=1), nan=0.5, posinf=0.5)
    return np.log2(np.clip(m, CLIP_LO, 0.5))


# ----------------------------------------------------------------- model

def growth(n, iters, window):
    """log2 growth of a Gaussian maximum from n*window to n*iters samples."""
    return 0.5 * np.log2(np.log(np.asarray(n, float) * iters) / np.log(np.asarray(n, float) * window))


def fit(ps, ns, wmax, seed=7, epochs=(200, 100, 100), lrs=(2e-3, 5e-4, 1e-4), batch=256, log=None):
    """Trains the roundoff MLP on the GPU; returns (model dict, stats)."""
    ps, ns = np.asarray(ps), np.asarray(ns)
    X = feature_matrix(ps, ns).astype(np.float64)
    y = targets(wmax)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(y))
    nval = max(1, len(y) // 10)
    val, tr = perm[:nval], perm[nval:]
    mean = X[tr].mean(axis=0)
    scale = X[tr].std(axis=0)
    scale[scale < 1e-6] = 1.0
    tmean, tscale = float(y[tr].mean()), float(y[tr].std() or 1.0)
    Xn = ((X - mean) / scale).astype(np.float32)
    yn = ((y - tmean) / tscale).astype(np.float32)
    losses = []
    with MLP(SIZES) as m:
        for ep, lr in zip(epochs, lrs):
            l = m.train(Xn[tr], yn[tr], batch, ep, lr, seed)
            losses.extend(float(v) for v in l)
            if log:
                log("train: %d epochs at lr %g -> loss %.5f" % (ep, lr, l[-1]))
        params = m.get_params()
        pred = m.forward(Xn[val]) * tscale + tmean
    r = y[val] - pred
    near = (y[val] > math.log2(0.02)) & (y[val] < math.log2(0.45))
    margin = float(max(0.0, np.quantile(r[near], 0.995))) if near.any() else 0.5
    # check of the extrapolation: first window alone -> all 8 windows
    wm = np.nan_to_num(np.asarray(wmax, dtype=np.float64), nan=0.5, posinf=0.5)
    ok = (wm.max(axis=1) < 0.4) & (wm[:, 0] > 1e-6)
    obs = np.log2(wm[ok].max(axis=1) / wm[ok, 0])
    pre = growth(ns[ok], WINDOWS * WINDOW_ITERS, WINDOW_ITERS)
    gscale = float(max(1.0, obs.mean() / pre.mean())) if len(obs) else 1.0
    stats = {
        "samples": int(len(y)),
        "train": int(len(tr)),
        "validation": int(len(val)),
        "final_train_loss": losses[-1],
        "val_rmse_log2": float(np.sqrt(np.mean(r ** 2))),
        "val_rmse_log2_near_limit": float(np.sqrt(np.mean(r[near] ** 2))) if near.any() else None,
        "val_near_limit": int(near.sum()),
        "margin_log2": margin,
        "saturated_fraction": float(np.mean(y >= -1.0 - 1e-9)),
        "growth_16_to_128_observed_mean": float(obs.mean()) if len(obs) else None,
        "growth_16_to_128_predicted_mean": float(pre.mean()) if len(obs) else None,
        "growth_scale": gscale,
        "loss_curve": [losses[0], losses[len(losses) // 2], losses[-1]],
    }
    model = {
        "sizes": SIZES,
        "feature_mean": mean.tolist(),
        "feature_scale": scale.tolist(),
        "target": [tmean, tscale],
        "window": WINDOWS * WINDOW_ITERS,
        "margin": margin,
        "growth_scale": gscale,
        "limit": LIMIT,
        "params": [float(v) for v in params],
    }
    return model, stats


def save_model(model, path, comment=""):
    with open(path, "w") as f:
        f.write("cudalucas_amd roundoff-mlp 1\n")
        for line in comment.splitlines():
            f.write("# %s\n" % line)
        f.write("sizes %d %s\n" % (len(model["sizes"]), " ".join(str(s) for s in model["sizes"])))
        f.write("feature_mean %s\n" % " ".join("%.9g" % v for v in model["feature_mean"]))
        f.write("feature_scale %s\n" % " ".join("%.9g" % v for v in model["feature_scale"]))
        f.write("target %.9g %.9g\n" % tuple(model["target"]))
        f.write("window %d\n" % model["window"])
        f.write("margin %.6f\n" % model["margin"])
        f.write("growth_scale %.6f\n" % model.get("growth_scale", 1.0))
        f.write("limit %.4f\n" % model["limit"])
        ps = model["params"]
        f.write("params %d\n" % len(ps))
        for i in range(0, len(ps), 8):
            f.write(" ".join("%.9g" % v for v in ps[i:i + 8]) + "\n")


def read_model(path):
    """Parses the text model (the same grammar ll_selector_load reads)."""
    model = {}
    with open(path) as f:
        head = f.readline().split()
        if head != ["cudalucas_amd", "roundoff-mlp", "1"]:
            raise ValueError("%s: not a roundoff-mlp v1 file" % path)
        toks = []
        for line in f:
            if line.lstrip().startswith("#"):
                continue
            toks.extend(line.split())
    i = 0
    while i < len(toks):
        k = toks[i]
        i += 1
        if k == "sizes":
            c = int(toks[i])
            model["sizes"] = [int(x) for x in toks[i + 1:i + 1 + c]]
            i += 1 + c
        elif k in ("feature_mean", "feature_scale"):
            model[k] = [float(x) for x in toks[i:i + FEATURES]]
            i += FEATURES
        elif k == "target":
            model[k] = [float(toks[i]), float(toks[i + 1])]
            i += 2
        elif k in ("window", "margin", "limit", "growth_scale"):
            model[k] = float(toks[i])
            i += 1
        elif k == "params":
            c = int(toks[i])
            model[k] = [float(x) for x in toks[i + 1:i + 1 + c]]
            i += 1 + c
        else:
            raise ValueError("unknown key %s" % k)
    return model


# --------------------------------------------------------------- timings

def bench_lengths(lengths, iters=128, log=None):
    """Measured ms per plain LL iteration at each length (the analog of the
    reference's -cufftbench, CUDALucas.cu:2413-2587)."""
    out = {}
    for n in lengths:
        p = W.next_prime(int(n * 17))
        with LLContext(p, n) as c:
            c.load_initial(1)
            c.iterate(32)
            c.iterate(iters)
            out[n] = c.last_iterate_ms() / iters
        if log:
            log("length %dK: %.5f ms/iter" % (n // 1024, out[n]))
    return out


def save_times(times, path):
    with open(path, "w") as f:
        f.write("# FFT length (K words), ms per LL iteration (plain engine, hipEvents)\n")
        for n in sorted(times):
            f.write("%d %.6f\n" % (n // 1024, times[n]))


def read_times(path):
    out = {}
    with open(path) as f:
        for line in f:
            if line.strip() and not line.startswith("#"):
                k, ms = line.split()
                out[int(k) * 1024] = float(ms)
    return out


def supported_lengths(max_k=16384):
    """Multiples of 1K words with a specialised schedule (the selector's
    candidates; lengths only the generic kernels run are left out)."""
    from . import fft_schedule

    out = []
    for k in range(1, max_k + 1):
        s = fft_schedule(k * 1024)
        if s is not None and s[0] != "generic":
            out.append(k * 1024)
    return out


def timed_lengths(max_k=16384):
    """Lengths the timing table covers: the selector's candidates and every
    length of the reference's default list up to max_k K words (an
    exponent's reference-policy length may be one only the generic kernels
    run)."""
    from . import fft_lengths

    return sorted(set(supported_lengths(max_k)) | {n for n in fft_lengths() if n <= max_k * 1024})


# -------------------------------------------------------------- selector

class Selector:
    """ll_selector_* on the GPU: loads a text model, predicts, chooses."""

    def __init__(self, path=MODEL_PATH, device=0, times=None):
        h = ctypes.c_void_p()
        rc = lib().ll_selector_load(device, path.encode(), ctypes.byref(h))
        self._h = h.value
        if rc:
            msg = lib().ll_selector_last_error(self._h).decode() if self._h else "ll_selector_load"
            self.close()
            raise LLError(rc, msg)
        if times:
            self.set_times(times)

    def _check(self, rc):
        if rc < 0:
            raise LLError(rc, lib().ll_selector_last_error(self._h).decode())
        return rc

    def close(self):
        if getattr(self, "_h", None):
            lib().ll_selector_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    def set_times(self, times):
        ns = np.asarray(sorted(times), dtype=np.uint32)
        ms = np.asarray([times[int(n)] for n in ns], dtype=np.float32)
        self._check(lib().ll_selector_set_times(self._h, len(ns), _ptr(ns), _ptr(ms)))

    def predict(self, p, ns):
        """Predicted full-test maximum roundoff (not log2) at each length."""
        ns = np.ascontiguousarray(ns, dtype=np.uint32)
        out = np.zeros(len(ns), dtype=np.float32)
        self._check(lib().ll_selector_predict(self._h, p, len(ns), _ptr(ns), _ptr(out)))
        return np.exp2(out.astype(np.float64))

    def choose(self, p, above=0):
        e = ctypes.c_float()
        n = self._check(lib().ll_selector_choose(self._h, p, above, ctypes.byref(e)))
        return n, e.value


# ------------------------------------------------------------ evaluation

def careful_test(errs, limit=LIMIT, limit_avg=LIMIT_AVG):
    """The reference's round-off test verdict (CUDALucas.cu:2606-2633)."""
    errs = np.asarray(errs, dtype=np.float64)
    return bool(np.all(errs <= limit) and errs.mean() <= limit_avg)


def _digest(words):
    return hashlib.sha256(np.ascontiguousarray(words, dtype="<u4").tobytes()).hexdigest()[:32]


def _run_checked(tests, iters, words_budget=1 << 27, max_tests=256):
    """[(max, mean, res64, digest)] per (p, n, offset) after `iters`
    iterations in the batched engine (per-iteration errors kept)."""
    out = [None] * len(tests)
    order = sorted(range(len(tests)), key=lambda i: tests[i][1])
    items = [(i, tests[i][1]) for i in order]
    for chunk in _chunks(items, words_budget, max_tests):
        batch = [(tests[i][0], tests[i][1], tests[i][2], iters) for i, _ in chunk]
        with LLBatch(batch) as b:
            left = iters
            errs = [[] for _ in chunk]
            while left:
                k = min(left, LLBatch.MAX_STEP)
                b.step(k)
                for j in range(len(chunk)):
                    errs[j].append(b.step_errors(j))
                left -= k
            for j, (i, _) in enumerate(chunk):
                e = np.concatenate(errs[j])
                w, r = b.residue(j)
                out[i] = (e, r, _digest(w))
    return out


def tuned_reference_lengths(times):
    """The length list the reference's own tuning would leave on this GPU:
    cufftbench's rule (CUDALucas.cu:2505-2522) keeps each length faster than
    every longer one, and every power of two (in K), and init_ffts uses the
    kept lengths (CUDALucas.cu:947-1025)."""
    keep, best = [], float("inf")
    for n in sorted(times, reverse=True):
        k = n // 1024
        if times[n] < best:
            best = times[n]
            keep.append(n)
        elif k & (k - 1) == 0:
            keep.append(n)
    return sorted(keep)


def reference_choices(primes, lengths=None):
    """choose_fft_length over the default list, or over `lengths`."""
    L = lib()
    if lengths:
        arr = np.asarray(lengths, dtype=np.uint32)
        if L.ll_fft_length_list_set(len(arr), _ptr(arr)):
            raise LLError(-1, "ll_fft_length_list_set")
    try:
        return [choose_fft_length(p)[0] for p in primes]
    finally:
        if lengths:
            L.ll_fft_length_list_set(0, None)


def _select_and_check(selector, primes, offs, times, check_iters, log=None, tag=""):
    """First choices, the round-off check, fallbacks until every exponent
    has an accepted length.  Returns (first, accepted, fallbacks, results,
    wasted ms) with results[i] = (max err, mean err, res64, digest)."""
    cur = [selector.choose(p)[0] for p in primes]
    first = list(cur)
    accepted = [None] * len(primes)
    result = [None] * len(primes)
    fallbacks = [0] * len(primes)
    wasted_ms = 0.0
    pending = list(range(len(primes)))
    rounds = 0
    while pending and rounds < 8:
        rounds += 1
        res = _run_checked([(primes[i], cur[i], offs[i]) for i in pending], check_iters)
        nxt = []
        for i, (e, r, dg) in zip(pending, res):
            if careful_test(e):
                accepted[i] = cur[i]
                result[i] = (float(e.max()), float(e.mean()), r, dg)
            else:
                fallbacks[i] += 1
                wasted_ms += check_iters * times[cur[i]]
                cur[i] = selector.choose(primes[i], above=cur[i])[0]
                nxt.append(i)
        pending = nxt
        if log:
            log("%s round %d: %d exponents failed the round-off test" % (tag, rounds, len(nxt)))
    if pending:
        raise RuntimeError("selector did not settle for %d exponents" % len(pending))
    return first, accepted, fallbacks, result, wasted_ms


def _batch_iteration_ms(tests, iters, words_budget=1 << 27, max_tests=256):
    """ms of one batched iteration over all (p, n, offset) tests (chunks run
    one after another)."""
    order = sorted(range(len(tests)), key=lambda i: tests[i][1])
    total = 0.0
    for chunk in _chunks([(i, tests[i][1]) for i in order], words_budget, max_tests):
        batch = [(tests[i][0], tests[i][1], tests[i][2], 64 + iters) for i, _ in chunk]
        with LLBatch(batch) as b:
            b.step(64)
            b.step(iters)
            total += b.last_step_ms() / iters
    return total


def evaluate(model_path, primes, times, batch_times=None, schedules=None, check_iters=CHECK_ITERS,
             batch_iters=512, seed=W.SELECTOR_EVAL_SEED, log=None):
    """The learned selector against the reference policy on `primes`.

    Two uses of the same model, each through the round-off check with
    fallback, each with residues compared with the reference-length run:

    * "timed": the fastest length predicted safe by the plain engine's
      measured ms/iteration (one exponent at a time), projected full-test
      hours against the reference policy on its default list and on the list
      its own cufftbench rule keeps;
    * "batched": the cheapest length predicted safe by the batched engine's
      per-test cost table (`batch_times`; without one, the shortest safe
      length), with the batched iteration time measured at the reference's
      and at the selected lengths (`schedules`: the tuned batched schedule
      table, installed for both)."""
    from . import tuner

    rng = W.SplitMix64(seed ^ 0x0FF5E7)
    offs = [rng.next() % p for p in primes]
    nref = reference_choices(primes)
    ntuned = reference_choices(primes, tuned_reference_lengths(times))
    if schedules:
        tuner.install(schedules)
    try:
        with Selector(model_path, times=times) as st:
            t_first, t_acc, t_fb, t_res, t_waste = _select_and_check(st, primes, offs, times, check_iters, log, "timed")
        with Selector(model_path, times=batch_times) as sb:
            b_first, b_acc, b_fb, b_res, _ = _select_and_check(sb, primes, offs, times, check_iters, log, "batched")
        ref = _run_checked([(p, n, o) for p, n, o in zip(primes, nref, offs)], check_iters)
        b_ref = _batch_iteration_ms([(p, n, o) for p, n, o in zip(primes, nref, offs)], batch_iters)
        b_sel = _batch_iteration_ms([(p, n, o) for p, n, o in zip(primes, b_acc, offs)], batch_iters)
    finally:
        if schedules:
            tuner.install([])
    ref_fail = sum(0 if careful_test(e) else 1 for e, _, _ in ref)
    mism_t = [p for p, r, q in zip(primes, t_res, ref) if r[3] != q[2]]
    mism_b = [p for p, r, q in zip(primes, b_res, ref) if r[3] != q[2]]

    def hours(lengths, waste_ms=0.0):
        return (sum((p - 2) * times[n] for p, n in zip(primes, lengths)) + waste_ms) * 1e-3 / 3600.0

    h_ref, h_tuned, h_timed = hours(nref), hours(ntuned), hours(t_acc, t_waste)
    mean_iters = float(np.mean([p - 2 for p in primes]))
    eph = lambda ms: 3600.0 * len(primes) / (ms * 1e-3 * mean_iters)
    plain_ref_eph = len(primes) / h_ref
    if log:
        log("batched iteration: reference lengths %.4f ms, selected %.4f ms" % (b_ref, b_sel))
    return {
        "exponents": len(primes),
        "range": [min(primes), max(primes)],
        "check": "careful round-off test: %d iterations, max <= %.2f and mean <= %.2f" % (check_iters, LIMIT, LIMIT_AVG),
        "reference_length_check_failures": ref_fail,
        "timed": {
            "shorter_than_reference": sum(a < b for a, b in zip(t_acc, nref)),
            "longer_than_reference": sum(a > b for a, b in zip(t_acc, nref)),
            "fallbacks": int(sum(t_fb)),
            "residue_mismatches_vs_reference_length": mism_t,
            "max_roundoff": max(r[0] for r in t_res),
            "projected_hours": h_timed,
        },
        "batched": {
            "cost_table": "batched per-test ms" if batch_times else "none (shortest safe length)",
            "shorter_than_reference": sum(a < b for a, b in zip(b_acc, nref)),
            "longer_than_reference": sum(a > b for a, b in zip(b_acc, nref)),
            "fallbacks": int(sum(b_fb)),
            "residue_mismatches_vs_reference_length": mism_b,
            "max_roundoff": max(r[0] for r in b_res),
            "words_fraction_of_reference": float(sum(b_acc)) / float(sum(nref)),
            "schedules": "tuned batched table" if schedules else "built-in preference",
        },
        "plain_projected_hours": {"reference_policy": h_ref, "reference_policy_tuned_list": h_tuned,
                                  "selector_timed": h_timed},
        "time_saved_fraction_plain": 1.0 - h_timed / h_ref,
        "time_saved_fraction_plain_vs_tuned_list": 1.0 - h_timed / h_tuned,
        "batched_iteration_ms": {"reference_lengths": b_ref, "selected_lengths": b_sel},
        "exps_per_hour": {"plain_reference_lengths": plain_ref_eph, "plain_selector": len(primes) / h_timed,
                          "batched_reference_lengths": eph(b_ref), "batched_selected_lengths": eph(b_sel)},
        "time_saved_fraction_batched": 1.0 - b_sel / b_ref,
        "batched_selected_vs_plain_reference": eph(b_sel) / plain_ref_eph,
        "per_exponent": [
            {"p": p, "offset": o, "n_reference": a, "n_tuned_list": tl, "n_timed_first": tf, "n_timed": ta,
             "n_batched_first": bf, "n_batched": ba, "fallbacks_timed": fa, "fallbacks_batched": fb,
             "max_err_timed": rt[0], "max_err_batched": rb[0], "res64": "%016x" % rt[2]}
            for p, o, a, tl, tf, ta, bf, ba, fa, fb, rt, rb in zip(primes, offs, nref, ntuned, t_first, t_acc, b_first,
                                                                  b_acc, t_fb, b_fb, t_res, b_res)
        ],
    }


# ------------------------------------------------------------------- cli

def _log(msg):
    print(msg, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m cudalucas_amd.selector")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("traces")
    a.add_argument("--out", required=True)
    a.add_argument("--count", type=int, default=W.SELECTOR_TRAIN_COUNT)
    a = sub.add_parser("train")
    a.add_argument("--traces", required=True)
    a.add_argument("--model", default=MODEL_PATH)
    a.add_argument("--stats")
    a = sub.add_parser("times")
    a.add_argument("--out", default=TIMES_PATH)
    a.add_argument("--iters", type=int, default=128)
    a = sub.add_parser("evaluate")
    a.add_argument("--model", default=MODEL_PATH)
    a.add_argument("--times", default=TIMES_PATH)
    a.add_argument("--batch-times", default=BATCH_TIMES_PATH)
    a.add_argument("--schedules", default=None)
    a.add_argument("--count", type=int, default=W.SELECTOR_EVAL_COUNT)
    a.add_argument("--out")
    args = ap.parse_args(argv)
    if args.cmd == "traces":
        t0 = time.time()
        plan = trace_plan(args.count)
        wmax = collect_traces(plan, log=_log)
        np.savez_compressed(args.out, p=np.array([x[0] for x in plan], dtype=np.uint32),
                            n=np.array([x[1] for x in plan], dtype=np.uint32),
                            offset=np.array([x[2] for x in plan], dtype=np.uint32), window_max=wmax)
        _log("traces: %d in %.1f s -> %s" % (len(plan), time.time() - t0, args.out))
    elif args.cmd == "train":
        d = np.load(args.traces)
        model, stats = fit(d["p"], d["n"], d["window_max"], log=_log)
        save_model(model, args.model, comment="trained on %d (p, n) roundoff traces, p log-uniform in "
                   "[1e5, 1.2e8] (seed %d)" % (stats["samples"], W.SELECTOR_TRAIN_SEED))
        _log(json.dumps(stats))
        if args.stats:
            with open(args.stats, "w") as f:
                json.dump(stats, f, indent=1)
    elif args.cmd == "times":
        save_times(bench_lengths(timed_lengths(), args.iters, log=_log), args.out)
    elif args.cmd == "evaluate":
        times = read_times(args.times)
        primes = W.log_uniform_primes(W.SELECTOR_EVAL_SEED, args.count, *W.SELECTOR_EVAL_RANGE)
        from . import tuner

        btimes = read_times(args.batch_times) if args.batch_times and os.path.exists(args.batch_times) else None
        sched = args.schedules or tuner.BATCH_TABLE
        sched = tuner.load(sched) if os.path.exists(sched) else None
        rep = evaluate(args.model, primes, times, batch_times=btimes, schedules=sched, log=_log)
        summary = {k: v for k, v in rep.items() if k != "per_exponent"}
        _log(json.dumps(summary))
        if args.out:
            with open(args.out, "w") as f:
                json.dump(rep, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
