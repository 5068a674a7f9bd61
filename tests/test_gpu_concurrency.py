"""Two-sweep contexts iterating at the same time on one GPU, from two
threads of one process and from two processes.

A carry sweep waits on other workgroups of its own launch; two such
launches side by side could hold each other's workgroup slots.  Every call
that runs carry sweeps holds the GPU's lock (an in-process mutex and an
flock on a per-GPU lock file), so the calls interleave and each residue is
exact: checked against the committed GMP fixtures (200 iterations), with no
iteration reporting a timed-out carry wait (error 3.0)."""
import hashlib
import os
import subprocess
import sys
import threading

import numpy as np
import pytest

import cudalucas_amd as L
from conftest import ROOT, announce

pytestmark = pytest.mark.gpu

CASES = [(19535569, 1048576), (9914521, 524288), (10616897, 589824)]
ITERS = 200
CHUNK = 10


def sha(words):
    return hashlib.sha256(np.ascontiguousarray(words, dtype="<u4").tobytes()).hexdigest()


def run_case(p, n):
    """(Res64 hex, SHA-256, max roundoff) after ITERS iterations from s_0 = 4,
    in calls of CHUNK iterations (so that calls of several contexts
    interleave)."""
    with L.LLContext(p, n, flags=L.LL_OPT_FORCE_SWEEPS) as c:
        c.load_initial(0)
        worst = 0.0
        for _ in range(ITERS // CHUNK):
            worst = max(worst, c.iterate(CHUNK))
        w, r = c.residue()
    return "0x%016x" % r, sha(w), worst


def check(p, result, gmp_fixtures):
    res, digest, worst = result
    fx = gmp_fixtures[(p, ITERS)]
    assert res == fx["res64"] and digest == fx["sha256"], p
    assert worst < 0.35, (p, worst)  # 2.0 / 3.0 would flag an overflow / a timed-out wait


def test_threads_iterate_concurrently(gmp_fixtures):
    announce("concurrent threads: %s" % CASES)
    results, errors = {}, []

    def worker(p, n):
        try:
            results[p] = run_case(p, n)
        except Exception as e:  # reported below, not lost in the thread
            errors.append((p, repr(e)))

    threads = [threading.Thread(target=worker, args=c) for c in CASES]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=240)
    assert not errors, errors
    for p, _ in CASES:
        check(p, results[p], gmp_fixtures)


def test_processes_iterate_concurrently(gmp_fixtures):
    announce("concurrent processes: %s" % CASES)
    worker = os.path.join(ROOT, "tests", "concurrent_worker.py")
    procs = [subprocess.Popen([sys.executable, worker, str(p), str(n), str(ITERS), str(CHUNK)],
                              stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, cwd=ROOT)
             for p, n in CASES]
    outs = []
    try:
        for pr in procs:
            out, err = pr.communicate(timeout=300)
            assert pr.returncode == 0, err[-2000:]
            outs.append(out.split())
    finally:
        for pr in procs:
            if pr.poll() is None:
                pr.kill()
                pr.wait()
    for (p, _), o in zip(CASES, outs):
        check(p, (o[0], o[1], float(o[2])), gmp_fixtures)
