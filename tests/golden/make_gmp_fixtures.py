"""Writes tests/golden/gmp_residues.json: exact LL residues from GMP.

For each (p, iterations) it records Res64 and the SHA-256 of the full
residue as ceil(p/32) little-endian uint32 words (s_0 = 4, unshifted).  The
large exponents take GMP minutes to hours on one core, so the values are
computed once here and committed as data.

    python tests/golden/make_gmp_fixtures.py [--quick]
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


def main():
    quick = "--quick" in sys.argv
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    jobs = []
    # smooth-length sweep exponents: prefix of 200 iterations
    for n in W.SMOOTH_LENGTHS:
        jobs.append((W.smooth_prime(n), [200]))
    # large single exponent: checkpoints every 100 iterations up to 1000
    jobs.append((W.LARGE_P, list(range(100, W.LARGE_CHECK_ITER + 1, 100))))
    # two-sweep schedule cases of tests/test_gpu_parity.py above 5e6
    for p in SWEEP_FIXTURE_PRIMES:
        jobs.append((p, [SWEEP_ITERS]))
    for p, marks in jobs:
        key = str(p)
        have = {e["iterations"] for e in data.get(key, [])}
        if all(m in have for m in marks):
            continue
        state, done, rows = None, 0, []
        t0 = time.time()
        for m in marks:
            state = O.gmp_ll(p, m - done, state)
            done = m
            rows.append(entry(p, m, state))
            print("p=%d iter=%d res64=%s (%.0fs)" % (p, m, rows[-1]["res64"], time.time() - t0), flush=True)
            if quick:
                break
        data[key] = rows
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
