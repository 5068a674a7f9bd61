"""Writes tests/golden/gmp_residues.json: exact LL residues from GMP.

For each (p, iterations) it records Res64 and the SHA-256 of the full
residue as ceil(p/32) little-endian uint32 words (s_0 = 4, unshifted).  The
large exponents take GMP minutes to hours on one core, so the values are
computed once here and committed as data.

    python tests/golden/make_gmp_fixtures.py [--quick] [--workers K]

Jobs: the smooth-length sweep (iterations 200 and 2000), the large single
exponent (every 100 iterations up to 1000), the two-sweep schedule cases,
one length-coverage case per default FFT length above 16K words (the
densest exponent the reference's policy assigns to it, workloads.densest_prime),
and the learned selector's 512 evaluation primes after its 1000-iteration
round-off test.
Jobs run in K worker processes (default 1); the file is rewritten as each
job ends, so an interrupted run keeps what it finished.
"""
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from cudalucas_amd import workloads as W  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gmp_residues.json")

# exponents of the two-sweep schedule tests large enough that GMP takes
# more than a few seconds for SWEEP_ITERS iterations
SWEEP_ITERS = 200
SWEEP_FIXTURE_PRIMES = [9914521, 10616897, 19535569, 29097221, 48495761, 67895203]


def entry(p, iters, words):
    return {
        "p": p,
        "iterations": iters,
        "res64": "0x%016x" % O.res64(words),
        "sha256": hashlib.sha256(np.ascontiguousarray(words, dtype="<u4").tobytes()).hexdigest(),
    }


# default FFT lengths (words) of the reference's list above this get a
# length-coverage case
LENGTH_CASES_ABOVE = 16384
# iterations of the selector's round-off test (cudalucas_amd/selector.py CHECK_ITERS)
SELECTOR_CHECK_ITERS = 1000


def run_job(job):
    p, marks, quick = job
    state, done, rows = None, 0, []
    t0 = time.time()
    for m in marks:
        state = O.gmp_ll(p, m - done, state)
        done = m
        rows.append(entry(p, m, state))
        print("p=%d iter=%d res64=%s (%.0fs)" % (p, m, rows[-1]["res64"], time.time() - t0), flush=True)
        if quick:
            break
    return p, rows


def main():
    quick = "--quick" in sys.argv
    workers = int(sys.argv[sys.argv.index("--workers") + 1]) if "--workers" in sys.argv else 1
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    jobs = []
    # smooth-length sweep exponents: prefixes of 200 and 2000 iterations
    for n in W.SMOOTH_LENGTHS:
        jobs.append((W.smooth_prime(n), [200, W.SMOOTH_ITERS]))
    # large single exponent: checkpoints every 100 iterations up to 1000
    jobs.append((W.LARGE_P, list(range(100, W.LARGE_CHECK_ITER + 1, 100))))
    # two-sweep schedule cases of tests/test_gpu_parity.py above 5e6
    for p in SWEEP_FIXTURE_PRIMES:
        jobs.append((p, [SWEEP_ITERS]))
    # one case per default length above 1M words, smallest first
    for n in O.default_fft_lengths():
        if n > LENGTH_CASES_ABOVE:
            p = W.densest_prime(n)
            jobs.append((p, [W.length_check_iters(p)]))
    # the learned selector's evaluation primes: Res64 after the round-off
    # test (1000 iterations from s_0 = 4)
    for p in W.log_uniform_primes(W.SELECTOR_EVAL_SEED, W.SELECTOR_EVAL_COUNT, *W.SELECTOR_EVAL_RANGE):
        jobs.append((p, [SELECTOR_CHECK_ITERS]))
    todo = []
    for p, marks in jobs:
        have = {e["iterations"] for e in data.get(str(p), [])}
        if not all(m in have for m in marks):
            todo.append((p, marks, quick))
    # cheapest first (GMP's cost grows a little faster than p * iterations)
    todo.sort(key=lambda j: j[0] * max(j[1]))

    def save(p, rows):
        old = {e["iterations"]: e for e in data.get(str(p), [])}
        old.update({e["iterations"]: e for e in rows})
        data[str(p)] = [old[k] for k in sorted(old)]
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)

    if workers <= 1:
        for job in todo:
            save(*run_job(job))
    else:
        import multiprocessing as mp

        with mp.Pool(workers) as pool:
            for p, rows in pool.imap_unordered(run_job, todo):
                save(p, rows)


if __name__ == "__main__":
    main()
