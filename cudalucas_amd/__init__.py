"""MI355X-native Lucas-Lehmer engine: Python view of the C-ABI.

The product is ``libcudalucas_amd.so`` (HIP kernels for gfx950 behind the
C-ABI in ``include/ll_api.h``) and the CUDALucas-compatible driver binary
``CUDALucasAMD``.  This module binds the C-ABI with ctypes for tests,
``bench.py`` and Python callers.  There is no fallback: if the library is
missing or cannot reach a GPU, calls raise.
"""
import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcudalucas_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "ll_api.h")

LL_OK = 0
LL_ERR_ARG = -1
LL_ERR_UNSUPPORTED = -2
LL_ERR_HIP = -3
LL_ERR_STATE = -4
LL_ERR_ALLOC = -5
LL_ERR_ROUNDOFF = -6

LL_OPT_FORCE_SWEEPS = 1
LL_OPT_NO_GRAPH = 2
LL_OPT_GENERIC = 4

# ll_info.path: LDS-resident, two sweeps, two sweeps on the generic kernels
_PATHS = {0: "resident", 1: "sweeps", 2: "generic"}

_lib = None


class LLError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("LL error %d: %s" % (code, msg))
        self.code = code


class _Options(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("threads", ctypes.c_int32 * 2),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("lines", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
    ]


class _Info(ctypes.Structure):
    _fields_ = [
        ("p", ctypes.c_uint32),
        ("n", ctypes.c_uint32),
        ("path", ctypes.c_uint32),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("lines", ctypes.c_uint32),
        ("stream", ctypes.c_void_p),
        ("iterations", ctypes.c_uint64),
    ]


def build():
    """Compile the HIP library and the driver (make, in-tree)."""
    import subprocess

    subprocess.run(["make", "-s", "-j8", "-C", os.path.join(_HERE, "csrc")], check=True)


def exported_symbols():
    """Function names declared in include/ll_api.h."""
    with open(HEADER_PATH) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(ll_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libcudalucas_amd.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    u32, u64, i32, f32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_float
    P = ctypes.c_void_p
    L.ll_create.argtypes = [u32, u32, ctypes.POINTER(_Options), ctypes.POINTER(P)]
    L.ll_destroy.argtypes = [P]
    L.ll_destroy.restype = None
    L.ll_load_digits.argtypes = [P, P, u32]
    L.ll_iterate.argtypes = [P, u32, P, ctypes.POINTER(f32)]
    L.ll_iterate_noalloc.argtypes = [P, u32, P, ctypes.POINTER(f32)]
    L.ll_read_digits.argtypes = [P, P, ctypes.POINTER(u32)]
    L.ll_reset_error.argtypes = [P]
    L.ll_error_max.argtypes = [P]
    L.ll_error_max.restype = f32
    L.ll_get_info.argtypes = [P, ctypes.POINTER(_Info)]
    L.ll_last_error.argtypes = [P]
    L.ll_last_error.restype = ctypes.c_char_p
    L.ll_debug_inject.argtypes = [P, u32, u64, f32]
    L.ll_last_iterate_ms.argtypes = [P]
    L.ll_last_iterate_ms.restype = ctypes.c_double
    L.ll_profile_kernels.argtypes = [P, u32, P]
    L.ll_bench_copy.argtypes = [i32, u64, u32, ctypes.POINTER(f32)]
    L.ll_batch_create.argtypes = [i32, u32, P, u32, ctypes.POINTER(P)]
    L.ll_batch_destroy.argtypes = [P]
    L.ll_batch_destroy.restype = None
    L.ll_batch_step.argtypes = [P, u32, P]
    L.ll_batch_status.argtypes = [P, u32, ctypes.POINTER(u64), ctypes.POINTER(u32), ctypes.POINTER(f32)]
    L.ll_batch_read_digits.argtypes = [P, u32, P, ctypes.POINTER(u32)]
    L.ll_batch_last_step_ms.argtypes = [P]
    L.ll_batch_last_step_ms.restype = ctypes.c_double
    L.ll_batch_groups.argtypes = [P]
    L.ll_batch_last_error.argtypes = [P]
    L.ll_batch_last_error.restype = ctypes.c_char_p
    L.ll_digit_width.argtypes = [u32, u32, u32]
    L.ll_unpack_bits.argtypes = [u32, u32, P, P]
    L.ll_unpack_bits.restype = None
    L.ll_balance_digits.argtypes = [u32, u32, P]
    L.ll_balance_digits.restype = None
    L.ll_standardize_digits.argtypes = [u32, u32, P]
    L.ll_pack_bits.argtypes = [u32, u32, P, P]
    L.ll_pack_bits.restype = None
    L.ll_residue_words.argtypes = [u32, u32, P, u32, P]
    L.ll_res64.argtypes = [u32, u32, P, u32]
    L.ll_res64.restype = u64
    L.ll_initial_digits.argtypes = [u32, u32, u32, P]
    L.ll_initial_digits.restype = None
    L.ll_checkpoint_crc.argtypes = [P, ctypes.c_size_t]
    L.ll_checkpoint_crc.restype = u32
    L.ll_fft_length_count.restype = ctypes.c_int
    L.ll_fft_length_at.argtypes = [ctypes.c_int]
    L.ll_fft_supported.argtypes = [u32]
    L.ll_choose_fft_length.argtypes = [u32, ctypes.POINTER(ctypes.c_int)]
    L.ll_fft_bounds.argtypes = [u32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.ll_fft_bounds.restype = None
    L.ll_fft_from_str.argtypes = [ctypes.c_char_p]
    L.ll_fft_schedule.argtypes = [u32, ctypes.POINTER(_Options), ctypes.POINTER(_Info)]
    L.ll_fft_length_list_set.argtypes = [u32, P]
    L.ll_fft_schedules.argtypes = [u32, P, u32]
    L.ll_tune_length.argtypes = [i32, u32, u32, u32, P, u32]
    L.ll_schedule_table_set.argtypes = [u32, P]
    L.ll_get_device_info.argtypes = [i32, P]
    L.ll_batch_step_errors.argtypes = [P, u32, P, u32]
    L.ll_mlp_create.argtypes = [i32, u32, P, ctypes.POINTER(P)]
    L.ll_mlp_destroy.argtypes = [P]
    L.ll_mlp_destroy.restype = None
    L.ll_mlp_last_error.argtypes = [P]
    L.ll_mlp_last_error.restype = ctypes.c_char_p
    L.ll_mlp_param_count.argtypes = [P]
    L.ll_mlp_param_count.restype = u32
    L.ll_mlp_set_params.argtypes = [P, P]
    L.ll_mlp_get_params.argtypes = [P, P]
    L.ll_mlp_forward.argtypes = [P, P, u32, P]
    L.ll_mlp_gradient.argtypes = [P, P, P, u32, P, P]
    L.ll_mlp_train.argtypes = [P, P, P, u32, u32, u32, f32, u64, P]
    L.ll_selector_features.argtypes = [u32, u32, P]
    L.ll_selector_candidates.argtypes = [u32, P, u32]
    L.ll_selector_load.argtypes = [i32, ctypes.c_char_p, ctypes.POINTER(P)]
    L.ll_selector_destroy.argtypes = [P]
    L.ll_selector_destroy.restype = None
    L.ll_selector_last_error.argtypes = [P]
    L.ll_selector_last_error.restype = ctypes.c_char_p
    L.ll_selector_set_times.argtypes = [P, u32, P, P]
    L.ll_selector_predict.argtypes = [P, u32, u32, P, P]
    L.ll_selector_choose.argtypes = [P, u32, u32, ctypes.POINTER(f32)]
    L.ll_work_parse_line.argtypes = [ctypes.c_char_p, ctypes.POINTER(Assignment)]
    L.ll_is_prime.argtypes = [u32]
    L.ll_valid_assignment.argtypes = [u32, u32]
    L.ll_work_next.argtypes = [ctypes.c_char_p, ctypes.POINTER(Assignment)]
    L.ll_work_clear.argtypes = [ctypes.c_char_p, u32]
    L.ll_format_report.argtypes = [ctypes.c_char_p, ctypes.POINTER(Report), ctypes.c_char_p, ctypes.c_size_t]
    L.ll_fopen_locked.argtypes = [ctypes.c_char_p, ctypes.c_char_p, u32]
    L.ll_fopen_locked.restype = P
    L.ll_fclose_locked.argtypes = [P]
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def nwords(p):
    return (p + 31) // 32


# ------------------------------------------------------------------ codec

def digit_width(p, n, j):
    return lib().ll_digit_width(p, n, j)


def unpack_bits(p, n, packed):
    packed = np.ascontiguousarray(packed, dtype=np.uint32)
    d = np.zeros(n, dtype=np.int32)
    lib().ll_unpack_bits(p, n, _ptr(packed), _ptr(d))
    return d


def balance_digits(p, n, digits):
    d = np.array(digits, dtype=np.int32)
    lib().ll_balance_digits(p, n, _ptr(d))
    return d


def standardize_digits(p, n, digits):
    d = np.array(digits, dtype=np.int32)
    z = lib().ll_standardize_digits(p, n, _ptr(d))
    return d, bool(z)


def pack_bits(p, n, digits):
    d = np.ascontiguousarray(digits, dtype=np.int32)
    out = np.zeros(nwords(p), dtype=np.uint32)
    lib().ll_pack_bits(p, n, _ptr(d), _ptr(out))
    return out


def residue_words(p, n, digits, offset):
    d = np.ascontiguousarray(digits, dtype=np.int32)
    out = np.zeros(nwords(p) + 2, dtype=np.uint32)
    rc = lib().ll_residue_words(p, n, _ptr(d), offset, _ptr(out))
    if rc:
        raise LLError(rc, "ll_residue_words")
    return out[: nwords(p)]


def res64(p, n, digits, offset):
    d = np.ascontiguousarray(digits, dtype=np.int32)
    return int(lib().ll_res64(p, n, _ptr(d), offset))


def initial_digits(p, n, offset):
    d = np.zeros(n, dtype=np.int32)
    lib().ll_initial_digits(p, n, offset, _ptr(d))
    return d


def checkpoint_crc(data):
    buf = np.frombuffer(bytes(data), dtype=np.uint8)
    return int(lib().ll_checkpoint_crc(_ptr(buf), buf.size))


def fft_lengths():
    L = lib()
    return [L.ll_fft_length_at(i) for i in range(L.ll_fft_length_count())]


def fft_supported(n):
    return bool(lib().ll_fft_supported(n))


def choose_fft_length(p, index=-1):
    idx = ctypes.c_int(index)
    n = lib().ll_choose_fft_length(p, ctypes.byref(idx))
    return n, idx.value


def fft_bounds(p):
    lb, ub = ctypes.c_int32(), ctypes.c_int32()
    lib().ll_fft_bounds(p, ctypes.byref(lb), ctypes.byref(ub))
    return lb.value, ub.value


def fft_schedule(n, width=0, height=0, lines=0, flags=0):
    """(path, W, H, R) ll_create would use for length n, or None."""
    o = _Options(0, (ctypes.c_int32 * 2)(0, 0), width, height, lines, flags)
    info = _Info()
    if lib().ll_fft_schedule(n, ctypes.byref(o), ctypes.byref(info)):
        return None
    return (_PATHS.get(info.path, "sweeps"), info.width, info.height, info.lines)


def fft_from_str(s):
    return lib().ll_fft_from_str(s.encode())


# ---------------------------------------------------------------- worktodo

WORK_OK, WORK_EOF, WORK_LONG_LINE, WORK_NO_TEST_EQUAL = 0, 1, 2, 3
WORK_INVALID_FORMAT, WORK_INVALID_DATA, WORK_BLANK, WORK_COMMENT = 4, 5, 6, 7


class Assignment(ctypes.Structure):
    """ll_assignment: one parsed work line (struct ASSIGNMENT, parse.c:230-236)."""
    _fields_ = [("exponent", ctypes.c_uint32), ("fft_length", ctypes.c_uint32),
                ("key", ctypes.c_char * 192), ("comment", ctypes.c_char * 192)]


class Report(ctypes.Structure):
    """ll_report: the fields of one report line (process_output,
    CUDALucas.cu:1633-1701)."""
    _fields_ = [("p", ctypes.c_uint32), ("n", ctypes.c_uint32), ("iteration", ctypes.c_uint32),
                ("offset", ctypes.c_uint32), ("residue", ctypes.c_uint64), ("max_err", ctypes.c_float),
                ("interval_iterations", ctypes.c_uint32), ("interval_us", ctypes.c_uint64),
                ("eta_s", ctypes.c_uint64), ("now", ctypes.c_int64), ("program", ctypes.c_char_p)]


def format_report(fmt, **fields):
    """ll_format_report: one report line by an OutputString format."""
    prog = fields.pop("program", "CUDALucasAMD")
    r = Report(program=prog.encode(), **fields)
    need = lib().ll_format_report(fmt.encode(), ctypes.byref(r), None, 0)
    buf = ctypes.create_string_buffer(need + 1)
    lib().ll_format_report(fmt.encode(), ctypes.byref(r), buf, need + 1)
    return buf.value.decode()


def work_parse_line(line):
    """parse_worktodo_line (parse.c:241-484): (status, Assignment)."""
    a = Assignment()
    st = lib().ll_work_parse_line(line.encode(), ctypes.byref(a))
    return st, a


def work_next(path):
    """get_next_assignment (parse.c:486-554): (0, Assignment), or (1 = cannot
    open, 2 = no valid line, None)."""
    a = Assignment()
    rc = lib().ll_work_next(os.fsencode(path), ctypes.byref(a))
    return rc, (a if rc == 0 else None)


def work_clear(path, p):
    """clear_assignment (parse.c:556-678): 0, or 3/4/5/6 as the reference."""
    return lib().ll_work_clear(os.fsencode(path), p)


# ----------------------------------------------------------------- engine

class LLContext:
    """One exponent at one FFT length on one GPU (ll_create .. ll_destroy)."""

    def __init__(self, p, n, device=0, width=0, height=0, lines=0, flags=0, threads=(0, 0)):
        L = lib()
        opts = _Options()
        opts.device = device
        opts.threads[0], opts.threads[1] = threads
        opts.width, opts.height, opts.lines, opts.flags = width, height, lines, flags
        h = ctypes.c_void_p()
        rc = L.ll_create(p, n, ctypes.byref(opts), ctypes.byref(h))
        self._h = h.value
        if rc:
            msg = L.ll_last_error(self._h).decode() if self._h else "ll_create"
            self.close()
            raise LLError(rc, msg)
        self.p, self.n = p, n

    def _check(self, rc):
        if rc:
            raise LLError(rc, lib().ll_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().ll_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    def info(self):
        i = _Info()
        self._check(lib().ll_get_info(self._h, ctypes.byref(i)))
        return {
            "p": i.p, "n": i.n, "path": _PATHS.get(i.path, "sweeps"),
            "width": i.width, "height": i.height, "lines": i.lines,
            "stream": i.stream, "iterations": i.iterations,
        }

    def load_digits(self, digits, offset):
        d = np.ascontiguousarray(digits, dtype=np.int32)
        assert d.size == self.n
        self._check(lib().ll_load_digits(self._h, _ptr(d), offset))

    def load_initial(self, offset=0):
        self.load_digits(initial_digits(self.p, self.n, offset), offset)

    def iterate(self, count, per_iter=False):
        errs = np.zeros(count, dtype=np.float32) if per_iter else None
        m = ctypes.c_float()
        self._check(lib().ll_iterate(self._h, count, None if errs is None else _ptr(errs), ctypes.byref(m)))
        return (m.value, errs) if per_iter else m.value

    def iterate_noalloc(self, count, per_iter=False):
        """ll_iterate_noalloc: the same iterations with the shuttle and error
        slots in the device's static scratch array."""
        errs = np.zeros(count, dtype=np.float32) if per_iter else None
        m = ctypes.c_float()
        self._check(lib().ll_iterate_noalloc(self._h, count, None if errs is None else _ptr(errs),
                                             ctypes.byref(m)))
        return (m.value, errs) if per_iter else m.value

    def read_digits(self):
        d = np.zeros(self.n, dtype=np.int32)
        off = ctypes.c_uint32()
        self._check(lib().ll_read_digits(self._h, _ptr(d), ctypes.byref(off)))
        return d, off.value

    def residue(self):
        """(unshifted residue words, Res64) of the current state."""
        d, off = self.read_digits()
        w = residue_words(self.p, self.n, d, off)
        return w, int(w[0]) | (int(w[1]) << 32)

    def last_iterate_ms(self):
        """GPU time of the kernels of the last iterate() (hipEvents)."""
        return lib().ll_last_iterate_ms(self._h)

    def profile_kernels(self, count):
        """Average per-launch ms of (square sweep, carry sweep)."""
        out = np.zeros(2, dtype=np.float32)
        self._check(lib().ll_profile_kernels(self._h, count, _ptr(out)))
        return float(out[0]), float(out[1])

    def error_max(self):
        return lib().ll_error_max(self._h)

    def reset_error(self):
        self._check(lib().ll_reset_error(self._h))

    def inject(self, kind, after=0, value=0.0):
        self._check(lib().ll_debug_inject(self._h, kind, after, value))


class _BatchTest(ctypes.Structure):
    _fields_ = [("p", ctypes.c_uint32), ("n", ctypes.c_uint32), ("offset", ctypes.c_uint32),
                ("iterations", ctypes.c_uint32)]


class LLBatch:
    """Many LL tests advanced together (ll_batch_create .. ll_batch_destroy).

    tests: iterable of (p, n, offset, iterations)."""

    MAX_STEP = 1024

    def __init__(self, tests, device=0, flags=0):
        tests = list(tests)
        arr = (_BatchTest * len(tests))(*[_BatchTest(*t) for t in tests])
        h = ctypes.c_void_p()
        rc = lib().ll_batch_create(device, len(tests), ctypes.cast(arr, ctypes.c_void_p), flags, ctypes.byref(h))
        self._h = h.value
        self.tests = tests
        if rc:
            msg = lib().ll_batch_last_error(self._h).decode() if self._h else "ll_batch_create"
            self.close()
            raise LLError(rc, msg)

    def _check(self, rc):
        if rc:
            raise LLError(rc, lib().ll_batch_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().ll_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    def step(self, iterations):
        errs = np.zeros(len(self.tests), dtype=np.float32)
        self._check(lib().ll_batch_step(self._h, iterations, _ptr(errs)))
        return errs

    def last_step_ms(self):
        return lib().ll_batch_last_step_ms(self._h)

    def groups(self):
        return lib().ll_batch_groups(self._h)

    def status(self, i):
        done, rem, m = ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_float()
        self._check(lib().ll_batch_status(self._h, i, ctypes.byref(done), ctypes.byref(rem), ctypes.byref(m)))
        return done.value, rem.value, m.value

    def remaining(self):
        return [self.status(i)[1] for i in range(len(self.tests))]

    def residue(self, i):
        p, n = self.tests[i][0], self.tests[i][1]
        d = np.zeros(n, dtype=np.int32)
        off = ctypes.c_uint32()
        self._check(lib().ll_batch_read_digits(self._h, i, _ptr(d), ctypes.byref(off)))
        w = residue_words(p, n, d, off.value)
        return w, int(w[0]) | (int(w[1]) << 32)

    def step_errors(self, i):
        """Per-iteration roundoff of test i over the last step."""
        k = lib().ll_batch_step_errors(self._h, i, None, 0)
        self._check(min(k, 0))
        out = np.zeros(k, dtype=np.float32)
        if k:
            self._check(min(lib().ll_batch_step_errors(self._h, i, _ptr(out), k), 0))
        return out

    def run(self, max_step=MAX_STEP):
        """Steps until every test has finished; returns per-test max roundoff."""
        worst = np.zeros(len(self.tests), dtype=np.float32)
        while any(self.remaining()):
            worst = np.maximum(worst, self.step(max_step))
        return worst


class MLP:
    """FP32 MLP on the GPU (ll_mlp_*): tanh hidden layers, linear scalar
    output, mean-squared-error loss, Adam.  sizes = [inputs, hidden..., 1]."""

    def __init__(self, sizes, device=0):
        self.sizes = [int(x) for x in sizes]
        arr = np.asarray(self.sizes, dtype=np.uint32)
        h = ctypes.c_void_p()
        rc = lib().ll_mlp_create(device, len(self.sizes) - 1, _ptr(arr), ctypes.byref(h))
        self._h = h.value
        if rc:
            msg = lib().ll_mlp_last_error(self._h).decode() if self._h else "ll_mlp_create"
            self.close()
            raise LLError(rc, msg)
        self.count = lib().ll_mlp_param_count(self._h)

    def _check(self, rc):
        if rc:
            raise LLError(rc, lib().ll_mlp_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().ll_mlp_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    def get_params(self):
        out = np.zeros(self.count, dtype=np.float32)
        self._check(lib().ll_mlp_get_params(self._h, _ptr(out)))
        return out

    def set_params(self, params):
        params = np.ascontiguousarray(params, dtype=np.float32)
        assert params.shape == (self.count,)
        self._check(lib().ll_mlp_set_params(self._h, _ptr(params)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(len(x), dtype=np.float32)
        self._check(lib().ll_mlp_forward(self._h, _ptr(x), len(x), _ptr(y)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.count, dtype=np.float32)
        loss = ctypes.c_float()
        self._check(lib().ll_mlp_gradient(self._h, _ptr(x), _ptr(t), len(t), _ptr(g), ctypes.byref(loss)))
        return loss.value, g

    def train(self, x, t, batch, epochs, lr, seed=1):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        losses = np.zeros(epochs, dtype=np.float32)
        self._check(lib().ll_mlp_train(self._h, _ptr(x), _ptr(t), len(t), batch, epochs, lr, seed, _ptr(losses)))
        return losses


def bench_copy(nbytes, reps=20, device=0):
    """Average ms of a plain device copy of nbytes (read nbytes + write nbytes)."""
    ms = ctypes.c_float()
    rc = lib().ll_bench_copy(device, nbytes, reps, ctypes.byref(ms))
    if rc:
        raise LLError(rc, "ll_bench_copy")
    return ms.value


def run_ll(p, n, iterations, offset=0, **kw):
    """s_iterations (s_0 = 4) at FFT length n: (residue words, Res64, max error)."""
    with LLContext(p, n, **kw) as c:
        c.load_initial(offset)
        m = c.iterate(iterations) if iterations else 0.0
        w, r = c.residue()
    return w, r, m
