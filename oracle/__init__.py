"""CPU oracle for the Lucas-Lehmer path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this package, and only as the checker.  The product
(``cudalucas_amd``) never imports it.

``liboracle.so`` is built from ``ll_oracle.c`` by ``oracle/Makefile``; its
functions restate tdulcet/CUDALucas v2.06 (file:line cited in the C source).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u32, u64, i32, dbl = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_double
        P = ctypes.c_void_p
        L.orc_splitmix64.argtypes = [ctypes.POINTER(u64)]
        L.orc_splitmix64.restype = u64
        L.orc_is_prime.argtypes = [u64]
        L.orc_is_prime.restype = ctypes.c_int
        L.orc_seeded_prime.argtypes = [u64, u64, u64]
        L.orc_seeded_prime.restype = u32
        L.orc_gmp_ll.argtypes = [u32, P, u64, P]
        L.orc_gmp_square_sub.argtypes = [u32, P, u32, P]
        L.orc_gmp_mulmod.argtypes = [u32, P, P, P]
        L.orc_digit_width.argtypes = [u32, u32, u32]
        L.orc_unpack_bits.argtypes = [u32, u32, P, P]
        L.orc_balance_digits.argtypes = [u32, u32, P]
        L.orc_standardize_digits.argtypes = [u32, u32, P]
        L.orc_pack_bits.argtypes = [u32, u32, P, P]
        L.orc_find_residue.argtypes = [u32, u32, P, u32]
        L.orc_find_residue.restype = u64
        L.orc_digits_to_residue.argtypes = [u32, u32, P, u32, P]
        L.orc_initial_digits.argtypes = [u32, u32, u32, P]
        L.orc_checkpoint_crc.argtypes = [P, ctypes.c_size_t]
        L.orc_checkpoint_crc.restype = u32
        L.orc_default_fft_count.restype = ctypes.c_int
        L.orc_default_fft_length.argtypes = [ctypes.c_int]
        L.orc_fft_bounds.argtypes = [u32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
        L.orc_fft_inverse.argtypes = [u32, P]
        L.orc_ibdwt_create.argtypes = [u32, u32]
        L.orc_ibdwt_create.restype = P
        L.orc_ibdwt_destroy.argtypes = [P]
        L.orc_ibdwt_load.argtypes = [P, P, u32]
        L.orc_ibdwt_iterate.argtypes = [P, u64, P]
        L.orc_ibdwt_iterate.restype = dbl
        L.orc_ibdwt_read.argtypes = [P, P, P]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


_PARSE_REF_PATH = os.path.join(_HERE, "_ref", "libparse_ref.so")
_parse_ref = None


class RefAssignment(ctypes.Structure):
    """struct ASSIGNMENT of the reference's parse.c (parse.c:230-236)."""
    _fields_ = [("exponent", ctypes.c_int), ("fft_length", ctypes.c_int),
                ("hex_key", ctypes.c_char * 192), ("comment", ctypes.c_char * 192)]


def parse_ref():
    """The reference's own parse.c (worktodo and ini reading) as compiled in
    place by oracle/Makefile into oracle/_ref/, or None where the reference
    sources were not present to build it."""
    global _parse_ref
    if _parse_ref is None and os.path.exists(_PARSE_REF_PATH):
        L = ctypes.CDLL(_PARSE_REF_PATH)
        P, i32, cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p
        L.parse_worktodo_line.argtypes = [P, ctypes.POINTER(RefAssignment), P, ctypes.POINTER(P)]
        L.get_next_assignment.argtypes = [cp, ctypes.POINTER(i32), ctypes.POINTER(i32), P]
        L.clear_assignment.argtypes = [cp, i32]
        L.valid_assignment.argtypes = [i32, i32]
        L.isprime.argtypes = [ctypes.c_uint]
        L.IniGetInt.argtypes = [cp, cp, ctypes.POINTER(i32), i32]
        L.IniGetStr.argtypes = [cp, cp, P, cp]
        L.IniGetInt2.argtypes = [cp, cp, ctypes.POINTER(i32), ctypes.POINTER(i32), i32]
        L.fopen_and_lock.argtypes = [cp, cp]
        L.fopen_and_lock.restype = P
        L.unlock_and_fclose.argtypes = [P]
        _parse_ref = L
    return _parse_ref


def set_threads(n):
    """Threads for the IBDWT's data-parallel loops; returns the count in
    effect.  Results are the same for every count."""
    return int(lib().orc_set_threads(n))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def nwords(p):
    return (p + 31) // 32


def words_to_int(words):
    return int.from_bytes(np.ascontiguousarray(words, dtype="<u4").tobytes(), "little")


def int_to_words(value, p):
    return np.frombuffer(value.to_bytes(4 * nwords(p), "little"), dtype="<u4").copy()


def seeded_prime(seed, lo, hi):
    return int(lib().orc_seeded_prime(seed, lo, hi))


def is_prime(n):
    return bool(lib().orc_is_prime(n))


def gmp_ll(p, iters, start=None):
    """Exact s_iters (s_0 = 4, or `start` words) mod 2^p-1 as uint32 words."""
    out = np.zeros(nwords(p), dtype=np.uint32)
    st = None if start is None else np.ascontiguousarray(start, dtype=np.uint32)
    lib().orc_gmp_ll(p, None if st is None else _ptr(st), iters, _ptr(out))
    return out


def gmp_square_sub(p, words, sub):
    w = np.ascontiguousarray(words, dtype=np.uint32)
    out = np.zeros(nwords(p), dtype=np.uint32)
    lib().orc_gmp_square_sub(p, _ptr(w), sub, _ptr(out))
    return out


def gmp_mulmod(p, a, b):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    out = np.zeros(nwords(p), dtype=np.uint32)
    lib().orc_gmp_mulmod(p, _ptr(a), _ptr(b), _ptr(out))
    return out


def res64(words):
    return int(words[0]) | (int(words[1]) << 32)


def digit_widths(p, n):
    b = p % n
    j = np.arange(n, dtype=np.uint64)
    f = (n - (np.uint64(b) * j) % np.uint64(n)) % np.uint64(n)
    return (p // n + (f < b)).astype(np.int32)


def unpack_bits(p, n, packed):
    packed = np.ascontiguousarray(packed, dtype=np.uint32)
    d = np.zeros(n, dtype=np.int32)
    lib().orc_unpack_bits(p, n, _ptr(packed), _ptr(d))
    return d


def balance_digits(p, n, digits):
    d = np.array(digits, dtype=np.int32)
    lib().orc_balance_digits(p, n, _ptr(d))
    return d


def standardize_digits(p, n, digits):
    d = np.array(digits, dtype=np.int32)
    z = lib().orc_standardize_digits(p, n, _ptr(d))
    return d, bool(z)


def pack_bits(p, n, digits):
    d = np.ascontiguousarray(digits, dtype=np.int32)
    out = np.zeros(nwords(p) + 1, dtype=np.uint32)
    lib().orc_pack_bits(p, n, _ptr(d), _ptr(out))
    return out[: nwords(p)]


def find_residue(p, n, digits, offset):
    d = np.ascontiguousarray(digits, dtype=np.int32)
    return int(lib().orc_find_residue(p, n, _ptr(d), offset))


def digits_to_residue(p, n, digits, offset):
    d = np.ascontiguousarray(digits, dtype=np.int32)
    out = np.zeros(nwords(p), dtype=np.uint32)
    lib().orc_digits_to_residue(p, n, _ptr(d), offset, _ptr(out))
    return out


def initial_digits(p, n, offset):
    d = np.zeros(n, dtype=np.int32)
    lib().orc_initial_digits(p, n, offset, _ptr(d))
    return d


def checkpoint_crc(data):
    buf = np.frombuffer(bytes(data), dtype=np.uint8)
    return int(lib().orc_checkpoint_crc(_ptr(buf), buf.size))


def default_fft_lengths():
    L = lib()
    return [L.orc_default_fft_length(i) for i in range(L.orc_default_fft_count())]


def fft_bounds(p):
    lb, ub = ctypes.c_int32(), ctypes.c_int32()
    lib().orc_fft_bounds(p, ctypes.byref(lb), ctypes.byref(ub))
    return lb.value, ub.value


def fft_inverse(z):
    """Unnormalised DFT with sign +1 of a complex vector (7-smooth length)."""
    a = np.ascontiguousarray(np.asarray(z, dtype=np.complex128)).copy()
    rc = lib().orc_fft_inverse(a.size, _ptr(a))
    if rc:
        raise ValueError("length not 7-smooth")
    return a


class IBDWT:
    """CPU restatement of the reference's LL iteration (lucas_square)."""

    def __init__(self, p, n):
        self.p, self.n = p, n
        self._h = lib().orc_ibdwt_create(p, n)
        if not self._h:
            raise ValueError("bad (p, n)")

    def close(self):
        if self._h:
            lib().orc_ibdwt_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def load(self, digits, offset):
        d = np.ascontiguousarray(digits, dtype=np.int32)
        lib().orc_ibdwt_load(self._h, _ptr(d), offset)

    def iterate(self, count, per_iter=False):
        errs = np.zeros(count, dtype=np.float64) if per_iter else None
        m = lib().orc_ibdwt_iterate(self._h, count, None if errs is None else _ptr(errs))
        return (m, errs) if per_iter else m

    def read(self):
        d = np.zeros(self.n, dtype=np.int32)
        off = ctypes.c_uint32()
        lib().orc_ibdwt_read(self._h, _ptr(d), ctypes.byref(off))
        return d, off.value
