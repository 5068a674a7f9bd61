"""Seeded synthetic workloads for the LL iteration path.

A "seeded prime in [lo, hi)" is the first prime >= lo + (splitmix64(seed) mod
(hi - lo)); "log-uniform in [a, b]" draws exp(U(ln a, ln b)) from successive
splitmix64 outputs.  These definitions are the ones the parity tests, the
fixtures under tests/golden/ and bench.py share.
"""
import math

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / (1 << 53))


def is_prime(n):
    if n < 2:
        return False
    small = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)
    for b in small:
        if n % b == 0:
            return n == b
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in small:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def next_prime(x):
    while not is_prime(x):
        x += 1
    return x


def seeded_prime(seed, lo, hi):
    rng = SplitMix64(seed)
    return next_prime(lo + (rng.next() % (hi - lo) if hi > lo else 0))


def log_uniform_primes(seed, count, lo, hi):
    rng = SplitMix64(seed)
    out = []
    la, lb = math.log(lo), math.log(hi)
    while len(out) < count:
        x = int(math.exp(la + (lb - la) * rng.uniform()))
        p = next_prime(max(x, 3))
        if p <= hi:
            out.append(p)
    return out


# Plain LL test, single exponent: seeded prime in [40000, 50000], N = 4096,
# all p-2 iterations; 44497 (a Mersenne-prime exponent) ends at residue 0.
SMALL_SEED = 1001
SMALL_N = 4096
SMALL_P = seeded_prime(SMALL_SEED, 40000, 50000)
SMALL_PRIME_P = 44497

# Large single exponent: seeded prime near 77,000,000 at N = 4,194,304.
LARGE_SEED = 2002
LARGE_N = 4194304
LARGE_P = seeded_prime(LARGE_SEED, 77_000_000, 77_100_000)
LARGE_TIMED_ITERS = 10_000
LARGE_CHECK_ITER = 1000
# Reference self-test exponent in the same FFT class (CUDALucas.cu:2705):
# Res64 after 10,000 iterations.
LARGE_REF_P = 75_846_319
LARGE_REF_RES64 = 0x626AB6B09B607BA9

# Smooth-length sweep: p at ~18.5 bits/word for 3*2^k, 5*2^k, 7*2^k lengths.
SMOOTH_SEED = 3003
SMOOTH_LENGTHS = [1572864, 2621440, 3670016, 5242880, 6291456]
SMOOTH_ITERS = 2000
# the power-of-two length each smooth length's residue is compared with
SMOOTH_POW2 = {1572864: 2097152, 2621440: 4194304, 3670016: 4194304, 5242880: 8388608,
               6291456: 8388608}


def smooth_prime(n, seed=SMOOTH_SEED, bits_per_word=18.5):
    target = int(bits_per_word * n)
    return seeded_prime(seed + n, target - 1000, target)


# Length coverage: for each default FFT length, the densest exponent the
# reference's policy assigns to it (the largest p whose lower bound
# 0.999 * 1024 * 3.57822505975293e-5 * p^1.0221799779697 stays within the
# length, CUDALucas.cu:1034-1050), capped at the supported 1.1e9
# (CHANGELOG:12): a seeded prime in the 100,000 exponents below that.
LENGTH_SEED = 6006
MAX_EXPONENT = 1_100_000_000


def densest_exponent_bound(n):
    ff1 = 0.999 * 1024.0 * 0.0000357822505975293
    e1 = 1.022179977969700
    return min(int((n / ff1) ** (1.0 / e1)), MAX_EXPONENT)


def densest_prime(n, seed=LENGTH_SEED):
    hi = densest_exponent_bound(n) - 1000
    return seeded_prime(seed + n, hi - 100_000, hi)


def length_check_iters(p):
    """GMP-pinned prefix used for the length-coverage cases: 200 iterations,
    100 above 2e8 (GMP takes seconds per iteration there)."""
    return 200 if p <= 200_000_000 else 100


# Batched: 256 exponents log-uniform in [1e5, 5e6].
BATCH_SEED = 4004
BATCH_COUNT = 256
BATCH_RANGE = (100_000, 5_000_000)

# Learned selector: training p log-uniform in [1e5, 1.2e8]; evaluation on
# 512 seeded primes in [1e5, 2e7].
SELECTOR_TRAIN_SEED = 5005
SELECTOR_EVAL_SEED = 5006
SELECTOR_EVAL_COUNT = 512
SELECTOR_EVAL_RANGE = (100_000, 20_000_000)
SELECTOR_TRAIN_COUNT = 20_000
SELECTOR_TRAIN_RANGE = (100_000, 120_000_000)
