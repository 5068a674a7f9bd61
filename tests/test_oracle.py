"""Pins the CPU oracle before it is trusted as the checker.

- GMP LL against the reference's own Res64 self-test tables
  (CUDALucas.cu:2673-2729) for every entry cheap enough on one core, and
  against the closed form (Mersenne-prime exponents end at residue 0).
- The oracle's FFT against a brute-force DFT and its inverse.
- The IBDWT restatement (lucas_square) against GMP, from s_0 = 4 and from
  random states, at power-of-two and 3/5/7-smooth lengths; a planted error
  in the restatement must be caught.
- The checkpoint checksum against the published CRC-32/MPEG-2 check value
  (same polynomial, MSB first, no reflection; init 0 instead of ~0 gives
  0x89a1897f for "123456789").
- Codec restatements against exact Python integers.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN


def py_ll(p, iters, s=4):
    m = (1 << p) - 1
    for _ in range(iters):
        s = (s * s - 2) % m
    return s


def selftest_table():
    with open(os.path.join(GOLDEN, "reference_selftest.json")) as f:
        d = json.load(f)
    return [(p, int(r, 16)) for p, r in d["tests"] + d["testl"]]


@pytest.mark.parametrize("p,res64", [(p, r) for p, r in selftest_table() if p < 400000])
def test_gmp_reproduces_reference_selftest(p, res64):
    w = O.gmp_ll(p, 10000)
    assert O.res64(w) == res64


def test_gmp_matches_python_int():
    for p in (61, 89, 107, 127, 521, 607, 1279, 9689, 11213):
        for it in (1, 5, 50):
            assert O.words_to_int(O.gmp_ll(p, it)) == py_ll(p, it)


@pytest.mark.parametrize("p", [521, 607, 1279, 2203, 2281, 3217, 4253, 4423, 9689, 9941, 11213, 19937, 21701, 23209, 44497])
def test_gmp_mersenne_primes_end_at_zero(p):
    assert not O.gmp_ll(p, p - 2).any()


@pytest.mark.parametrize("p", [523, 1283, 9697, 44483])
def test_gmp_composite_mersenne_nonzero(p):
    assert O.gmp_ll(p, p - 2).any()


def test_seeded_primes_agree_with_product_definition():
    from cudalucas_amd import workloads as W

    assert O.seeded_prime(W.SMALL_SEED, 40000, 50000) == W.SMALL_P
    assert O.seeded_prime(W.LARGE_SEED, 77_000_000, 77_100_000) == W.LARGE_P
    for n in W.SMOOTH_LENGTHS:
        assert O.is_prime(W.smooth_prime(n))


@pytest.mark.parametrize("m", [1, 2, 3, 4, 5, 6, 7, 8, 12, 14, 15, 16, 21, 35, 49, 64, 96, 105, 128])
def test_fft_against_bruteforce_dft(m):
    rng = np.random.default_rng(m)
    z = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    k = np.arange(m)
    dft = np.exp(2j * np.pi * np.outer(k, k) / m) @ z  # sign +1 (INVERSE)
    got = O.fft_inverse(z)
    assert np.abs(got - dft).max() < 1e-12 * max(1, m)
    # inverse(conj) round trip: conj(F(conj(F(z)))) = m z
    back = np.conj(O.fft_inverse(np.conj(got))) / m
    assert np.abs(back - z).max() < 1e-13 * max(1, m)


def test_fft_rejects_non_smooth():
    with pytest.raises(ValueError):
        O.fft_inverse(np.ones(11))


@pytest.mark.parametrize("p,n,iters", [
    (42839, 4096, 300),
    (44497, 4096, 200),
    (22133, 1024, 300),
    (86243, 8192, 100),
    (21701, 1536, 200),   # 3 * 2^9
    (43541, 2560, 150),   # 5 * 2^9
    (60937, 3584, 120),   # 7 * 2^9
    (104459, 6144, 60),   # 3 * 2^11
])
def test_ibdwt_matches_gmp(p, n, iters):
    assert O.is_prime(p)
    off = (p * 31) % p
    c = O.IBDWT(p, n)
    c.load(O.initial_digits(p, n, off), off)
    m, errs = c.iterate(iters, per_iter=True)
    d, o = c.read()
    assert (O.digits_to_residue(p, n, d, o) == O.gmp_ll(p, iters)).all()
    assert m < 0.35 and errs.shape == (iters,)


def test_ibdwt_random_state_matches_gmp():
    p, n = 44497, 4096
    rng = np.random.default_rng(3)
    s = rng.integers(0, 1 << 32, O.nwords(p), dtype=np.uint64).astype(np.uint32)
    s[-1] &= (1 << (p % 32)) - 1
    off = 1000
    v = O.words_to_int(s)
    shifted = O.int_to_words(((v << off) | (v >> (p - off))) & ((1 << p) - 1), p)
    d = O.balance_digits(p, n, O.unpack_bits(p, n, shifted))
    c = O.IBDWT(p, n)
    c.load(d, off)
    c.iterate(3)
    d2, o2 = c.read()
    expect = O.gmp_ll(p, 3, start=s)
    assert (O.digits_to_residue(p, n, d2, o2) == expect).all()


def test_ibdwt_threads_change_nothing():
    """The IBDWT's data-parallel loops give bit-identical digits and
    per-iteration errors for any thread count, and still match GMP."""
    p, n = 1257787, 65536
    runs = []
    for threads in (1, 4):
        assert O.set_threads(threads) in (1, threads)
        try:
            c = O.IBDWT(p, n)
            c.load(O.initial_digits(p, n, 5), 5)
            _, errs = c.iterate(40, per_iter=True)
            d, off = c.read()
            c.close()
        finally:
            O.set_threads(1)
        runs.append((d.copy(), off, np.asarray(errs).copy()))
    (d1, o1, e1), (d4, o4, e4) = runs
    assert o1 == o4 and (d1 == d4).all() and (e1 == e4).all()
    assert (O.digits_to_residue(p, n, d4, o4) == O.gmp_ll(p, 40)).all()


def test_planted_error_is_caught():
    """A plausible slip (forgetting the -2) must fail the GMP comparison."""
    p, n, iters = 44497, 4096, 50
    c = O.IBDWT(p, n)
    c.load(O.initial_digits(p, n, 0), 0)
    c.iterate(iters)
    d, o = c.read()
    good = O.digits_to_residue(p, n, d, o)
    assert (good == O.gmp_ll(p, iters)).all()
    # the same state one squaring without the -2: GMP disagrees
    wrong = O.gmp_square_sub(p, O.gmp_ll(p, iters - 1), 0)
    assert not (good == wrong).all()


def test_checkpoint_crc_check_value():
    assert O.checkpoint_crc(b"123456789") == 0x89A1897F
    assert O.checkpoint_crc(b"") == 0


def test_codec_roundtrip_and_residue():
    rng = np.random.default_rng(7)
    for p, n in [(44497, 4096), (1257787, 65536), (2976221, 163840)]:
        v = int.from_bytes(rng.bytes(O.nwords(p) * 4), "little") % ((1 << p) - 1)
        w = O.int_to_words(v, p)
        d = O.unpack_bits(p, n, w)
        widths = O.digit_widths(p, n)
        assert int(widths.sum()) == p
        assert (O.pack_bits(p, n, d) == w).all()
        b = O.balance_digits(p, n, d)
        assert np.abs(b).max() <= (1 << int(widths.max() - 1))
        off = int(rng.integers(0, p))
        unshifted = ((v >> off) | (v << (p - off))) & ((1 << p) - 1)
        r = O.digits_to_residue(p, n, b, off)
        assert O.words_to_int(r) == unshifted % ((1 << p) - 1)
        assert O.find_residue(p, n, b, off) == unshifted & ((1 << 64) - 1)
        s, z = O.standardize_digits(p, n, b)
        assert not z and (O.pack_bits(p, n, s) == w).all()


def test_initial_digits_value():
    for p, n, off in [(44497, 4096, 0), (44497, 4096, 44495), (86243, 8192, 12345)]:
        d = O.initial_digits(p, n, off)
        assert O.words_to_int(O.digits_to_residue(p, n, d, off)) == 4


def test_fft_bounds_reference_examples():
    # choose_fft_length's power laws (CUDALucas.cu:1034-1050), worked by
    # hand: 77M needs more than 4M words, 4096K ends near 75.8M.
    lb, ub = O.fft_bounds(77_000_000)
    assert lb == 4216441 and lb > 4194304
    lens = O.default_fft_lengths()
    assert len(lens) == 162 and lens[0] == 1024 and lens[-1] == 65536 * 1024
    assert all(a < b for a, b in zip(lens, lens[1:]))
