"""The batched workload run to completion (256 seeded exponents log-uniform in
[1e5, 5e6] at the reference policy's lengths, autotuned batched schedules),
plus the nine Mersenne-prime exponents of that range, whose final residue
must be exactly 0.

Every test runs all p - 2 iterations in the batched engine.  Checks: every
Mersenne-prime exponent ends at residue 0 and every other one does not;
no iteration of any test exceeds the 0.35 roundoff limit (so there is no
undetected failure to look for); and for the smallest exponents the full
final residue equals the plain single-exponent engine's, run to completion
beside it.  The whole run takes about half an hour of GPU time, so by default
only the exponents below 8e5 run (with the Mersenne primes 110503 ...
756839); the four parts of the whole run are asked for by name:

    LL_BATCH_COMPLETION=low  python -m pytest -s tests/test_gpu_batch_completion.py   # p < 2e6
    LL_BATCH_COMPLETION=mid  python -m pytest -s tests/test_gpu_batch_completion.py   # 2e6 <= p < 3.5e6
    LL_BATCH_COMPLETION=high python -m pytest -s tests/test_gpu_batch_completion.py   # 3.5e6 <= p < 4.3e6
    LL_BATCH_COMPLETION=top  python -m pytest -s tests/test_gpu_batch_completion.py   # p >= 4.3e6
"""
import hashlib
import json
import os
import time

import numpy as np
import pytest

import cudalucas_amd as L
from cudalucas_amd import tuner as T
from cudalucas_amd import workloads as W
from conftest import announce

pytestmark = pytest.mark.gpu

MERSENNE_PRIME_EXPONENTS = [110503, 132049, 216091, 756839, 859433, 1257787, 1398269, 2976221, 3021377]
PARTS = {"default": (0, 800_000), "low": (0, 2_000_000), "mid": (2_000_000, 3_500_000),
         "high": (3_500_000, 4_300_000), "top": (4_300_000, 10_000_000)}
PLAIN_CHECKS = 24  # smallest exponents also run to completion in the plain engine
PART = os.environ.get("LL_BATCH_COMPLETION", "") or "default"


def digest(words):
    return hashlib.sha256(np.ascontiguousarray(words, dtype="<u4").tobytes()).hexdigest()[:32]


@pytest.mark.skipif(PART not in PARTS, reason="LL_BATCH_COMPLETION must be one of %s" % "|".join(PARTS))
def test_batch_to_completion():
    ps = W.log_uniform_primes(W.BATCH_SEED, W.BATCH_COUNT, *W.BATCH_RANGE)
    ps = sorted(set(ps) | set(MERSENNE_PRIME_EXPONENTS))
    lo, hi = PARTS[PART]
    ps = [p for p in ps if lo <= p < hi]
    tests = [(p, L.choose_fft_length(p)[0], p % 1000, p - 2) for p in ps]
    announce("batch completion part %s: %d exponents, %d..%d" % (PART, len(ps), ps[0], ps[-1]))
    if os.path.exists(T.BATCH_TABLE):
        T.install(T.load(T.BATCH_TABLE))
    t0 = time.time()
    rows = []
    try:
        with L.LLBatch(tests) as b:
            worst = np.zeros(len(tests), dtype=np.float32)
            step = 0
            while any(b.remaining()):
                worst = np.maximum(worst, b.step(L.LLBatch.MAX_STEP))
                step += 1
                if step % 40 == 0:
                    left = sum(1 for r in b.remaining() if r)
                    print("step %d: %d tests running, %.0f s" % (step, left, time.time() - t0), flush=True)
            wall = time.time() - t0
            for i, (p, n, off, iters) in enumerate(tests):
                done, rem, m = b.status(i)
                w, r = b.residue(i)
                rows.append({"p": p, "n": n, "iterations": int(done), "max_roundoff": float(m),
                             "res64": "%016x" % r, "digest": digest(w), "zero": bool(not w.any())})
    finally:
        T.install([])
    print("part %s: %d exponents in %.1f s" % (PART, len(rows), wall), flush=True)
    # the plain engine on the smallest exponents, to completion
    plain = {}
    for row in rows[:PLAIN_CHECKS] if PART in ("low", "default") else []:
        p, n = row["p"], row["n"]
        with L.LLContext(p, n) as c:
            c.load_initial(p % 1000)
            m = c.iterate(p - 2)
            w, r = c.residue()
        plain[p] = digest(w)
    out = os.environ.get("LL_BATCH_REPORT")
    if out:
        with open(out, "w") as f:
            json.dump({"part": PART, "wall_s": wall, "rows": rows, "plain_digests": plain}, f, indent=1)
    for row in rows:
        assert row["iterations"] == row["p"] - 2
        assert row["max_roundoff"] < 0.35, row
        assert row["zero"] == (row["p"] in MERSENNE_PRIME_EXPONENTS), row
        if row["p"] in plain:
            assert plain[row["p"]] == row["digest"], row
