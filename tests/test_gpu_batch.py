"""Batched engine (ll_batch_*) on the GPU.

Every residue is checked bit-exactly: against GMP where an exact run is
cheap, and against the plain single-exponent engine (itself pinned to GMP
and the reference's Res64 tables) for a sample of the seeded 256-exponent
batch workload.
"""
import numpy as np
import pytest

import cudalucas_amd as L
import oracle as O
from cudalucas_amd import workloads as W
from conftest import announce

pytestmark = pytest.mark.gpu


def test_mixed_batch_runs_to_completion_vs_gmp():
    tests = [
        (44497, 4096, 17, 44497 - 2),       # Mersenne prime: residue 0
        (21701, 2048, 5, 21701 - 2),        # Mersenne prime: residue 0
        (W.SMALL_P, 4096, 77, W.SMALL_P - 2),     # full composite test
        (86243, 8192, 3, 3000),
        (1257787, 131072, 11, 1500),        # two-sweep group (256 x 256)
        (2553659, 131072, 12, 1500),
        (4723783, 262144, 13, 1200),        # second two-sweep group
    ]
    announce("batch mixed %d tests" % len(tests))
    with L.LLBatch(tests) as b:
        assert b.groups() == 5
        worst = b.run()
        assert not any(b.remaining())
        for i, (p, n, off, iters) in enumerate(tests):
            done, rem, m = b.status(i)
            assert done == iters and rem == 0
            w, r = b.residue(i)
            assert (w == O.gmp_ll(p, iters)).all(), "test %d p=%d" % (i, p)
            assert worst[i] < 0.35
        # finished tests are frozen
        before = [b.residue(i)[1] for i in range(len(tests))]
        b.step(64)
        assert [b.residue(i)[1] for i in range(len(tests))] == before


def batch_sample(k):
    ps = W.log_uniform_primes(W.BATCH_SEED, W.BATCH_COUNT, *W.BATCH_RANGE)
    idx = np.linspace(0, len(ps) - 1, k).astype(int)
    out = []
    for j in idx:
        p = ps[j]
        n, _ = L.choose_fft_length(p)
        out.append((p, n))
    return out


def test_seeded_batch_matches_plain_engine():
    sample = batch_sample(12)
    iters = 1500
    tests = [(p, n, (p * 7) % p, iters) for p, n in sample]
    announce("batch workload sample %s" % [(p, n) for p, n in sample])
    with L.LLBatch(tests) as b:
        b.step(1000)
        b.step(500)
        res = [b.residue(i)[0] for i in range(len(tests))]
    for (p, n, off, it), w in zip(tests, res):
        with L.LLContext(p, n) as c:
            c.load_initial(off)
            c.iterate(it)
            wp, _ = c.residue()
        assert (w == wp).all(), "p=%d n=%d" % (p, n)


def test_batch_forced_sweeps_small_lengths():
    """Small lengths forced onto the two-sweep path in one batch group."""
    tests = [(2553659, 131072, 1, 700), (1257787, 131072, 2, 900), (1414741, 131072, 3, 800)]
    with L.LLBatch(tests, flags=L.LL_OPT_FORCE_SWEEPS) as b:
        assert b.groups() == 1
        b.run(max_step=256)
        for i, (p, n, off, it) in enumerate(tests):
            assert (b.residue(i)[0] == O.gmp_ll(p, it)).all()


def test_large_multi_group_batch_has_no_failed_iterations():
    """A batch of several large two-sweep groups (256K-1M words, 8 tests
    each, the shape of the round-off check's batches): every iteration's
    error stays far below the limit, so no carry wait timed out (those
    report 3.0), and residues equal the plain engine's.  Regression test
    for carry launches of different groups sharing the GPU."""
    rng = W.SplitMix64(2024)
    lengths = [262144, 524288, 1048576]
    tests = []
    for n in lengths:
        for _ in range(8):
            p = W.next_prime(int(n * (14.0 + (rng.next() % 1000) / 250.0)))
            tests.append((p, n, rng.next() % p, 600))
    announce("large multi-group batch %d tests" % len(tests))
    with L.LLBatch(tests) as b:
        assert b.groups() == len(lengths)
        errs = [[] for _ in tests]
        for k in (512, 88):
            b.step(k)
            for i in range(len(tests)):
                errs[i].append(b.step_errors(i))
        res = [b.residue(i)[0] for i in range(len(tests))]
    for i, (p, n, off, it) in enumerate(tests):
        e = np.concatenate(errs[i])
        assert len(e) == it and e.max() < 0.35, "p=%d n=%d max %.3f" % (p, n, e.max())
    for i in (0, 8, 16):  # one per group against the plain engine
        p, n, off, it = tests[i]
        with L.LLContext(p, n) as c:
            c.load_initial(off)
            c.iterate(it)
            wp, _ = c.residue()
        assert (res[i] == wp).all(), "p=%d n=%d" % (p, n)
