"""GPU parity of the HIP LL engine against exact arithmetic.

Every case runs through the C-ABI (libcudalucas_amd.so) on the GPU and is
compared bit-exactly with GMP (the oracle), with the reference's own Res64
self-test tables (CUDALucas.cu:2673-2729), or with committed GMP fixtures
for the large exponents.  Floating point only enters through the reported
roundoff, which must stay under the reference's 0.35 limit
(CUDALucas.cu:2616, 2883).
"""
import hashlib
import json
import os

import numpy as np
import pytest

import cudalucas_amd as L
import oracle as O
from cudalucas_amd import workloads as W
from conftest import GOLDEN, announce

pytestmark = pytest.mark.gpu

ROUNDOFF_LIMIT = 0.35


def sha(words):
    return hashlib.sha256(np.ascontiguousarray(words, dtype="<u4").tobytes()).hexdigest()


def random_state(p, seed):
    rng = np.random.default_rng(seed)
    w = rng.integers(0, 1 << 32, O.nwords(p), dtype=np.uint64).astype(np.uint32)
    if p % 32:
        w[-1] &= (1 << (p % 32)) - 1
    return w


def rotl(words, p, k):
    v = O.words_to_int(words)
    k %= p
    return O.int_to_words(((v << k) | (v >> (p - k))) & ((1 << p) - 1), p)


# ----------------------------------------------------------- small N (LDS)

@pytest.mark.parametrize("p,n,iters,offset", [
    (W.SMALL_P, 4096, 500, 0),
    (W.SMALL_P, 4096, 500, 12345),
    (44497, 4096, 300, 777),
    (22133, 1024, 1000, 5),
    (86243, 8192, 400, 99),
    (169409, 8192, 300, 3),
    (86243, 16384, 300, 11),  # 5.3 bits/word
])
def test_resident_vs_gmp(p, n, iters, offset):
    announce("resident p=%d n=%d iters=%d offset=%d" % (p, n, iters, offset))
    w, r, m = L.run_ll(p, n, iters, offset)
    g = O.gmp_ll(p, iters)
    assert r == O.res64(g)
    assert (w == g).all()
    assert m < ROUNDOFF_LIMIT


# every LDS-resident instantiation (M = n / 2 from 256 to 8192 points, the
# 7-smooth ones included) at about 16.5 bits/word
RESIDENT_LENGTHS = [512, 1024, 2048, 3072, 4096, 5120, 6144, 7168, 8192, 10240, 12288, 14336, 16384]


@pytest.mark.parametrize("n", RESIDENT_LENGTHS)
def test_resident_lengths_vs_gmp(n):
    p = W.next_prime(int(n * 16.5))
    announce("resident length p=%d n=%d" % (p, n))
    with L.LLContext(p, n) as c:
        assert c.info()["path"] == "resident"
    w, r, m = L.run_ll(p, n, 300, 7)
    g = O.gmp_ll(p, 300)
    assert r == O.res64(g)
    assert m < ROUNDOFF_LIMIT


def test_small_exponent_full_run_vs_gmp():
    p, n = W.SMALL_P, W.SMALL_N
    announce("small exponent full run p=%d n=%d" % (p, n))
    with L.LLContext(p, n) as c:
        c.load_initial(0)
        m, errs = c.iterate(p - 2, per_iter=True)
        w, r = c.residue()
    g = O.gmp_ll(p, p - 2)
    assert (w == g).all()
    assert errs.shape == (p - 2,) and float(errs.max()) == pytest.approx(m)
    assert m < ROUNDOFF_LIMIT
    assert (errs[40:] > 0).all()  # every iteration reports its own maximum


def test_mersenne_prime_44497_is_zero():
    p, n = W.SMALL_PRIME_P, 4096
    announce("M44497 full run")
    w, r, m = L.run_ll(p, n, p - 2, offset=31337)
    assert r == 0 and not w.any()
    assert m < ROUNDOFF_LIMIT


# ----------------------------------------------------- two HBM sweeps path

SWEEP_CASES = [
    # (p, n, W, H, R)
    (756839, 131072, 256, 256, 4),   # 5.8 bits/word: carry-in must cross the whole chain
    (2553659, 131072, 256, 256, 4),
    (1257787, 131072, 256, 256, 4),
    (4723783, 262144, 512, 256, 4),
    (4723783, 262144, 256, 512, 4),
    (9914521, 524288, 1024, 256, 4),
    (9914521, 524288, 512, 512, 8),
    (19535569, 1048576, 256, 2048, 4),
    (19535569, 1048576, 1024, 512, 2),
    (19535569, 1048576, 2048, 256, 2),
    (19535569, 1048576, 1024, 512, 8),   # 512-thread carry workgroups
    (4723783, 262144, 256, 512, 8),
    (29097221, 1572864, 768, 1024, 8),   # 3 * 2^18 complex points
    (48495761, 2621440, 640, 2048, 8),   # 5 * 2^18
    (67895203, 3670016, 896, 2048, 8),   # 7 * 2^18
    (4723783, 262144, 512, 256, 16),     # 16-line carry workgroups (256-byte column runs)
    (10616897, 589824, 72, 4096, 8),     # 4096-point columns, 512-thread square workgroups
]


@pytest.mark.parametrize("p,n,w,h,r", SWEEP_CASES)
def test_sweeps_vs_gmp(p, n, w, h, r, gmp_fixtures):
    """200 iterations of every schedule against GMP: live for small p, the
    committed GMP fixture (Res64 + SHA-256 of the residue) above 5e6."""
    iters = 200
    announce("sweeps p=%d n=%d W=%d H=%d R=%d" % (p, n, w, h, r))
    with L.LLContext(p, n, width=w, height=h, lines=r, flags=L.LL_OPT_FORCE_SWEEPS) as c:
        info = c.info()
        assert (info["path"], info["width"], info["height"], info["lines"]) == ("sweeps", w, h, r)
        c.load_initial(4242)
        m = c.iterate(iters)
        words, res = c.residue()
    fx = gmp_fixtures.get((p, iters))
    if fx is not None:
        assert "0x%016x" % res == fx["res64"] and sha(words) == fx["sha256"]
    else:
        assert p < 5_000_000, "missing GMP fixture for p=%d" % p
        assert (words == O.gmp_ll(p, iters)).all()
    assert m < ROUNDOFF_LIMIT


@pytest.mark.parametrize("p,n,flags", [
    (42839, 4096, 0),
    (2553659, 131072, L.LL_OPT_FORCE_SWEEPS),
    (2553659, 131072, L.LL_OPT_FORCE_SWEEPS | L.LL_OPT_NO_GRAPH),
    (9914521, 524288, 0),
])
def test_random_state_single_steps(p, n, flags):
    """Random full-size states (not the low-entropy s_k prefix): one and
    three iterations from a shifted random value vs GMP."""
    s = random_state(p, seed=p)
    off = (p * 7919) % p or 1
    shifted = rotl(s, p, off)
    digits = L.balance_digits(p, n, L.unpack_bits(p, n, shifted))
    announce("random state p=%d n=%d flags=%d" % (p, n, flags))
    with L.LLContext(p, n, flags=flags) as c:
        c.load_digits(digits, off)
        w0, _ = c.residue()
        assert (w0 == s).all()  # load/unload round trip
        m1 = c.iterate(1)
        w1, _ = c.residue()
        m3 = c.iterate(2)
        w3, _ = c.residue()
    e1 = O.gmp_square_sub(p, s, 2)
    e3 = O.gmp_ll(p, 2, start=e1)
    assert (w1 == e1).all()
    assert (w3 == e3).all()
    assert max(m1, m3) < ROUNDOFF_LIMIT


# ------------------------------------------------- reference golden vectors

def _selftest_cases():
    with open(os.path.join(GOLDEN, "reference_selftest.json")) as f:
        d = json.load(f)
    out = []
    for p, r in d["tests"] + d["testl"]:
        n, _ = L.choose_fft_length(p) if os.path.exists(L.LIB_PATH) else (0, 0)
        if n > 0:
            out.append((p, int(r, 16), n))
    return out


@pytest.mark.parametrize("p,res64,n", _selftest_cases())
def test_reference_selftest_res64(p, res64, n):
    """The reference's -r self-test: Res64 after 10,000 iterations."""
    announce("selftest p=%d n=%d" % (p, n))
    with L.LLContext(p, n) as c:
        c.load_initial(p // 3)
        m = c.iterate(10000)
        _, r = c.residue()
    assert r == res64, "p=%d n=%d got %016x want %016x (maxerr %.4f)" % (p, n, r, res64, m)
    assert m < 0.4


# ---------------------------------------------------- large configurations

@pytest.mark.parametrize("schedule", [None, (512, 4096, 8)])
def test_77m_at_4m_words_vs_gmp_fixture(gmp_fixtures, schedule):
    """The built-in schedule and the autotuned one (the bench's)."""
    p, n = W.LARGE_P, W.LARGE_N
    marks = sorted(k[1] for k in gmp_fixtures if k[0] == p)
    assert marks, "missing GMP fixture for p=%d" % p
    announce("77M at 4M words p=%d n=%d marks=%s schedule=%s" % (p, n, marks, schedule))
    kw = {} if schedule is None else dict(width=schedule[0], height=schedule[1], lines=schedule[2],
                                           flags=L.LL_OPT_FORCE_SWEEPS)
    with L.LLContext(p, n, **kw) as c:
        c.load_initial(0)
        done, worst = 0, 0.0
        for mk in marks:
            worst = max(worst, c.iterate(mk - done))
            done = mk
            w, r = c.residue()
            fx = gmp_fixtures[(p, mk)]
            assert "0x%016x" % r == fx["res64"], "iteration %d" % mk
            assert sha(w) == fx["sha256"], "iteration %d" % mk
    print("77M at 4M words: max roundoff over %d iterations: %.4f" % (done, worst))
    assert worst < ROUNDOFF_LIMIT


def _check_roundoff(p, n, errs, ref_sampled_limit=ROUNDOFF_LIMIT):
    """The roundoff limits of a run of per-iteration maxima `errs`.

    The reference samples the error every ErrorIterations = 100 iterations
    (CUDALucas.cu:1895-1900, 2935-2936) against error_limit = 0.35 at its
    default ini (CUDALucas.cu:2883), and treats 0.40 as an error that must
    not occur at all (check_residue, CUDALucas.cu:2765).  For an exponent
    the reference's own length policy assigns to n (CUDALucas.cu:1027-1068)
    every iteration must stay under 0.35.  Beyond that density the
    reference's own sampled check applies, and no iteration may reach 0.40."""
    errs = np.asarray(errs)
    sampled = errs[99::100]
    assert sampled.max() < ref_sampled_limit, (p, n, float(sampled.max()))
    if L.choose_fft_length(p)[0] == n:
        assert errs.max() < ROUNDOFF_LIMIT, (p, n, float(errs.max()))
    else:
        assert errs.max() < 0.40, (p, n, float(errs.max()))


@pytest.mark.parametrize("n", W.SMOOTH_LENGTHS)
def test_smooth_lengths_vs_gmp_fixture(n, gmp_fixtures):
    """The smooth-length sweep: 2,000 iterations at ~18.5 bits/word, the
    residue at 200 and 2,000 iterations equal to GMP's and, at 2,000, to the
    power-of-two length's run of the same exponent.  These densities lie
    beyond the reference's length policy (it assigns each of these
    exponents a longer length), so the roundoff is held to the reference's
    sampled check (_check_roundoff)."""
    p = W.smooth_prime(n)
    iters = W.SMOOTH_ITERS
    announce("smooth length p=%d n=%d iters=%d" % (p, n, iters))
    with L.LLContext(p, n) as c:
        c.load_initial(0)
        _, e1 = c.iterate(200, per_iter=True)
        w, r = c.residue()
        fx = gmp_fixtures[(p, 200)]
        assert "0x%016x" % r == fx["res64"] and sha(w) == fx["sha256"]
        _, e2 = c.iterate(iters - 200, per_iter=True)
        w, r = c.residue()
    errs = np.concatenate([e1, e2])
    fx = gmp_fixtures[(p, iters)]
    assert "0x%016x" % r == fx["res64"]
    assert sha(w) == fx["sha256"]
    n2 = W.SMOOTH_POW2[n]
    announce("smooth length p=%d power-of-two n=%d iters=%d" % (p, n2, iters))
    with L.LLContext(p, n2) as c:
        c.load_initial(0)
        m2 = c.iterate(iters)
        w2, _ = c.residue()
    assert (w2 == w).all()
    print("smooth length n=%d p=%d bits/word=%.3f max roundoff %.4f (every 100th: %.4f; power of two %d: %.4f)"
          % (n, p, p / n, errs.max(), errs[99::100].max(), n2, m2))
    _check_roundoff(p, n, errs)
    assert m2 < ROUNDOFF_LIMIT


@pytest.mark.parametrize("n", [n for n in W.SMOOTH_LENGTHS if n in O.default_fft_lengths()])
def test_smooth_lengths_at_reference_density(n):
    """The smooth lengths of the reference's list at the densest exponent
    its policy assigns to them: 2,000 iterations with every iteration's
    roundoff under 0.35, and the residue equal to the power-of-two length's
    (GMP pins both lengths' first 200 iterations in the length-coverage
    cases)."""
    p = W.densest_prime(n)
    assert L.choose_fft_length(p)[0] == n
    iters = W.SMOOTH_ITERS
    announce("smooth length at reference density p=%d n=%d" % (p, n))
    with L.LLContext(p, n) as c:
        c.load_initial(0)
        m, errs = c.iterate(iters, per_iter=True)
        w, _ = c.residue()
    n2 = W.SMOOTH_POW2[n]
    with L.LLContext(p, n2) as c:
        c.load_initial(0)
        c.iterate(iters)
        w2, _ = c.residue()
    assert (w == w2).all()
    print("reference density n=%d p=%d bits/word=%.3f max roundoff %.4f" % (n, p, p / n, m))
    _check_roundoff(p, n, errs)


SHORT_LINE_LENGTHS_K = [18, 20, 32, 36, 42, 48, 50, 56, 60, 64, 70, 80, 84, 96, 112, 120, 126, 144, 160,
                        180, 224]


@pytest.mark.parametrize("nk", SHORT_LINE_LENGTHS_K)
def test_short_line_lengths_vs_gmp(nk):
    """Lengths whose line FFT carries the 3/5/7 factors (W = 72..448)."""
    n = nk * 1024
    p = W.next_prime(int(n * 18.2))
    iters = 300
    announce("short-line length n=%d p=%d" % (n, p))
    with L.LLContext(p, n) as c:
        info = c.info()
        c.load_initial(p // 5)
        m = c.iterate(iters)
        w, _ = c.residue()
    assert (w == O.gmp_ll(p, iters)).all(), info
    assert m < ROUNDOFF_LIMIT


# ---------------------------------------------------- length coverage

def _length_cases():
    """Every default length this build supports above 16K words (the
    two-sweep lengths) with its committed GMP fixture: the densest exponent
    the reference's policy assigns to the length."""
    with open(os.path.join(GOLDEN, "gmp_residues.json")) as f:
        fx = json.load(f)
    out = []
    for n in O.default_fft_lengths():
        if n <= 16384 or not (os.path.exists(L.LIB_PATH) and L.fft_supported(n)):
            continue
        p = W.densest_prime(n)
        out.append(pytest.param(p, n, W.length_check_iters(p), id="%dK" % (n // 1024),
                                marks=() if str(p) in fx else pytest.mark.skip(reason="no GMP fixture yet")))
    return out


@pytest.mark.parametrize("p,n,iters", _length_cases())
def test_length_coverage_vs_gmp_fixture(p, n, iters, gmp_fixtures):
    announce("length coverage p=%d n=%d iters=%d" % (p, n, iters))
    with L.LLContext(p, n) as c:
        info = c.info()
        c.load_initial(0)
        m = c.iterate(iters)
        w, r = c.residue()
    fx = gmp_fixtures[(p, iters)]
    assert "0x%016x" % r == fx["res64"], info
    assert sha(w) == fx["sha256"], info
    print("length %dK p=%d bits/word=%.3f schedule W=%d H=%d R=%d max roundoff %.4f"
          % (n // 1024, p, p / n, info["width"], info["height"], info["lines"], m))
    assert m < ROUNDOFF_LIMIT


def _tuned_cases():
    """Lengths whose autotuned schedule (the single-exponent table that
    bench.py and the driver install) differs from the built-in preference
    the coverage cases above run: the tuned schedule against the same GMP
    fixture."""
    from cudalucas_amd import tuner as T

    with open(os.path.join(GOLDEN, "gmp_residues.json")) as f:
        fx = json.load(f)
    if not (os.path.exists(L.LIB_PATH) and os.path.exists(T.SINGLE_TABLE)):
        return []
    table = {s.n: s for s in T.load(T.SINGLE_TABLE)}
    out = []
    for n in O.default_fft_lengths():
        e = table.get(n)
        if n <= 16384 or e is None or e.path != 1 or not L.fft_supported(n):
            continue
        if L.fft_schedule(n)[1:] == (e.width, e.height, e.lines):
            continue
        p = W.densest_prime(n)
        if str(p) not in fx:
            continue
        out.append(pytest.param(p, n, W.length_check_iters(p), e.width, e.height, e.lines,
                                id="%dK-%dx%dx%d" % (n // 1024, e.width, e.height, e.lines)))
    return out


@pytest.mark.parametrize("p,n,iters,w,h,r", _tuned_cases())
def test_tuned_schedule_vs_gmp_fixture(p, n, iters, w, h, r, gmp_fixtures):
    announce("tuned schedule p=%d n=%d W=%d H=%d R=%d iters=%d" % (p, n, w, h, r, iters))
    with L.LLContext(p, n, width=w, height=h, lines=r, flags=L.LL_OPT_FORCE_SWEEPS) as c:
        info = c.info()
        assert (info["width"], info["height"], info["lines"]) == (w, h, r)
        c.load_initial(0)
        m = c.iterate(iters)
        wd, rr = c.residue()
    fx = gmp_fixtures[(p, iters)]
    assert "0x%016x" % rr == fx["res64"], info
    assert sha(wd) == fx["sha256"], info
    assert m < ROUNDOFF_LIMIT


# --------------------------------------------------- per-call overhead

def test_single_iteration_call_is_cheap():
    """ll_iterate(ctx, 1) at 4M words costs its kernels plus one
    synchronisation: the per-iteration call the reference makes
    (lucas_square, CUDALucas.cu:1869-1986) must not pay fixed host costs."""
    import time

    p, n = W.LARGE_P, W.LARGE_N
    announce("single-iteration call cost p=%d n=%d" % (p, n))
    with L.LLContext(p, n) as c:
        c.load_initial(0)
        c.iterate(64)
        walls, gpus = [], []
        for _ in range(60):
            t0 = time.perf_counter()
            c.iterate(1)
            walls.append(time.perf_counter() - t0)
            gpus.append(c.last_iterate_ms() * 1e-3)
        t0 = time.perf_counter()
        c.iterate(20)
        wall20 = time.perf_counter() - t0
        gpu20 = c.last_iterate_ms() * 1e-3
    walls, gpus = np.array(walls[10:]), np.array(gpus[10:])
    print("ll_iterate(1): median wall %.1f us, GPU %.1f us; ll_iterate(20): wall %.1f us/iter, GPU %.1f us/iter"
          % (np.median(walls) * 1e6, np.median(gpus) * 1e6, wall20 / 20 * 1e6, gpu20 / 20 * 1e6))
    assert np.median(walls) < 1e-4
    assert wall20 < 1.15 * gpu20 + 1e-4


# ------------------------------------------------------- generic kernels

GENERIC_CASES = [
    # (p, n, W, H, R): shapes the specialised kernels also run, forced onto
    # the runtime-plan kernels, and shapes only they run
    (2553659, 131072, 256, 256, 4),
    (4723783, 262144, 512, 256, 8),
    (9914521, 524288, 1024, 256, 2),
    (19535569, 1048576, 1024, 512, 8),
    (29097221, 1572864, 768, 1024, 8),   # 3 * 2^18
    (4723783, 262144, 256, 512, 1),      # one line per carry block
    (3021377, 165888, 36, 2304, 8),      # 162K: W = 9 * 4, H = 9 * 256
    (4723783, 264600, 210, 630, 2),      # 2^3 3^3 5^2 7^2 / 2: every radix
]


@pytest.mark.parametrize("p,n,w,h,r", GENERIC_CASES)
def test_generic_kernels_vs_gmp(p, n, w, h, r, gmp_fixtures):
    iters = 200
    announce("generic p=%d n=%d W=%d H=%d R=%d" % (p, n, w, h, r))
    with L.LLContext(p, n, width=w, height=h, lines=r, flags=L.LL_OPT_GENERIC) as c:
        info = c.info()
        assert (info["width"], info["height"], info["lines"]) == (w, h, r)
        c.load_initial(4242)
        w0, _ = c.residue()
        m = c.iterate(iters)
        words, res = c.residue()
    assert (w0 == O.gmp_ll(p, 0)).all()  # load / unload round trip
    fx = gmp_fixtures.get((p, iters))
    if fx is not None:
        assert "0x%016x" % res == fx["res64"] and sha(words) == fx["sha256"]
    else:
        assert (words == O.gmp_ll(p, iters)).all()
    assert m < ROUNDOFF_LIMIT
