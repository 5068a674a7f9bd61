"""Schedule autotuner: picks the W x H x R (or LDS-resident) schedule of each
FFT length by timing every candidate on the GPU.

The reference tunes thread counts per FFT length with -threadbench and
reads them back from "<device> threads.txt" (CUDALucas.cu:812-945,
2143-2276).  Here the tunable is the kernel schedule (ll_fft_schedules), and
it is tuned for the way the length will run:

* ``single``: one exponent per length, one iteration per call of the
  hipGraph (ll_tune_length), as CUDALucasAMD runs;
* ``batched``: ``copies`` exponents of the length advanced together by the
  batched engine (ll_batch_*), where workgroups of different tests share the
  GPU and the best schedule can differ.

The table is the text file ``K path W H R ms`` that CUDALucasAMD reads as
"<device> schedules.txt"; ``install`` hands it to the library
(ll_schedule_table_set) for ll_create and ll_batch_create.

    python -m cudalucas_amd.tuner --lengths 8K-256K --mode batched --out table.txt
"""
import argparse
import ctypes
import os
import sys
import time

import numpy as np

from . import LLBatch, LLError, _ptr, fft_supported, lib
from . import workloads as W

_HERE = os.path.dirname(os.path.abspath(__file__))
BATCH_TABLE = os.path.join(_HERE, "data", "mi355x_schedules_batched.txt")
SINGLE_TABLE = os.path.join(_HERE, "data", "mi355x_schedules_single.txt")


class Schedule(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("path", ctypes.c_uint32), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32), ("lines", ctypes.c_uint32), ("ms", ctypes.c_float)]

    def key(self):
        return (self.n, self.path, self.width, self.height, self.lines)

    def __repr__(self):
        kind = "resident" if self.path == 0 else "sweeps %dx%dx%d" % (self.width, self.height, self.lines)
        return "%dK %s %.5f ms" % (self.n // 1024, kind, self.ms)


def schedules(n):
    k = lib().ll_fft_schedules(n, None, 0)
    arr = (Schedule * max(k, 1))()
    if k > 0:
        lib().ll_fft_schedules(n, ctypes.cast(arr, ctypes.c_void_p), k)
    return list(arr[:k]) if k > 0 else []


def install(table):
    """Preferred schedules for ll_create / ll_batch_create ([] clears)."""
    arr = (Schedule * max(len(table), 1))(*table)
    rc = lib().ll_schedule_table_set(len(table), ctypes.cast(arr, ctypes.c_void_p) if table else None)
    if rc:
        raise LLError(rc, "ll_schedule_table_set")


def tune_single(n, iters=256, device=0):
    """Every schedule of length n, fastest first (ll_tune_length)."""
    k = lib().ll_fft_schedules(n, None, 0)
    arr = (Schedule * max(k, 1))()
    got = lib().ll_tune_length(device, 0, n, iters, ctypes.cast(arr, ctypes.c_void_p), k)
    if got < 0:
        raise LLError(got, "ll_tune_length(%d)" % n)
    return list(arr[:got])


def tune_batched(n, copies=16, iters=256, seed=7):
    """Every schedule of length n timed in the batched engine with `copies`
    exponents of that length (about 17 bits/word); fastest first."""
    rng = W.SplitMix64(seed + n)
    ps = [W.next_prime(int(n * 17.0) + (rng.next() % 1000)) for _ in range(copies)]
    out = []
    for s in schedules(n):
        install([s])
        tests = [(p, n, rng.next() % p, 64 + iters) for p in ps]
        try:
            with LLBatch(tests, flags=0) as b:
                b.step(64)
                b.step(iters)
                ms = b.last_step_ms() / iters
        finally:
            install([])
        r = Schedule(*s.key(), ms)
        out.append(r)
    out.sort(key=lambda r: r.ms)
    return out


def parse_lengths(spec):
    """'8K-256K' (every supported multiple of 1K), or '4096K,5M'."""
    if "-" in spec:
        a, b = spec.split("-")
        lo, hi = _words(a), _words(b)
        return [k * 1024 for k in range(max(1, lo // 1024), hi // 1024 + 1) if fft_supported(k * 1024)]
    return [_words(x) for x in spec.split(",")]


def _words(s):
    s = s.strip().upper()
    if s.endswith("K"):
        return int(s[:-1]) * 1024
    if s.endswith("M"):
        return int(s[:-1]) * 1024 * 1024
    return int(s)


def save(table, path, comment=""):
    with open(path, "w") as f:
        f.write("# %s\n# K path W H R ms/iteration\n" % comment)
        for s in sorted(table, key=lambda s: s.n):
            f.write("%5d %u %u %u %u %10.5f\n" % (s.n // 1024, s.path, s.width, s.height, s.lines, s.ms))


def load(path):
    out = {}
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            k, path_, w, h, r, ms = line.split()
            out[int(k) * 1024] = Schedule(int(k) * 1024, int(path_), int(w), int(h), int(r), float(ms))
    return [out[n] for n in sorted(out)]


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m cudalucas_amd.tuner")
    ap.add_argument("--lengths", required=True, help="e.g. 8K-256K or 4096K,5M")
    ap.add_argument("--mode", choices=["single", "batched"], default="single")
    ap.add_argument("--copies", type=int, default=16)
    ap.add_argument("--iters", type=int, default=256)
    ap.add_argument("--out", required=True)
    ap.add_argument("--times-out", help="per-test ms of each length's best schedule (selector cost table)")
    args = ap.parse_args(argv)
    best = []
    t0 = time.time()
    for n in parse_lengths(args.lengths):
        res = tune_single(n, args.iters) if args.mode == "single" else tune_batched(n, args.copies, args.iters)
        if not res:
            continue
        print("%dK: %s" % (n // 1024, "; ".join(repr(r) for r in res)), flush=True)
        best.append(res[0])
    save(best, args.out, "%s schedules, %s, %.0f s" % (args.mode, args.lengths, time.time() - t0))
    if args.times_out:
        per = args.copies if args.mode == "batched" else 1
        with open(args.times_out, "w") as f:
            f.write("# FFT length (K words), ms per test-iteration (%s engine, best schedule)\n" % args.mode)
            for s in best:
                f.write("%d %.6f\n" % (s.n // 1024, s.ms / per))
    print("wrote %s (%d lengths)" % (args.out, len(best)), flush=True)


if __name__ == "__main__":
    sys.exit(main())
