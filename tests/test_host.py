"""Host-side logic of the product, checked on the CPU (no GPU calls).

- libcudalucas_amd.so loads and exports every entry point include/ll_api.h
  declares.
- The product's codec, Res64 and checkpoint checksum agree with the oracle's
  restatements of the reference (and hence, through the oracle's pins, with
  exact integers and the CRC check value).
- Checkpoint files: byte layout and CRC of write_checkpoint
  (CUDALucas.cu:1707-1861), rotation c -> t, backup fallback when the current
  file is corrupt, the -s save folder name.
- FFT-length policy against the oracle's restatement of choose_fft_length.
- The driver's option/ini precedence (parse_args / process_options).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import cudalucas_amd as L
import oracle as O

DRIVER = os.path.join(os.path.dirname(L.LIB_PATH), "CUDALucasAMD")


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = L.exported_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def random_value(p, rng):
    v = int.from_bytes(rng.bytes(O.nwords(p) * 4), "little") % ((1 << p) - 1)
    return v


@pytest.mark.parametrize("p,n", [(44497, 4096), (86243, 4096), (1257787, 65536), (2976221, 163840),
                                 (77005277, 4194304)])
def test_codec_matches_oracle(p, n):
    rng = np.random.default_rng(p)
    v = random_value(p, rng)
    w = O.int_to_words(v, p)
    assert (L.unpack_bits(p, n, w) == O.unpack_bits(p, n, w)).all()
    d = L.balance_digits(p, n, L.unpack_bits(p, n, w))
    assert (d == O.balance_digits(p, n, O.unpack_bits(p, n, w))).all()
    assert all(L.digit_width(p, n, j) == O.lib().orc_digit_width(p, n, j) for j in range(0, n, max(1, n // 257)))
    off = int(rng.integers(0, p))
    assert (L.residue_words(p, n, d, off) == O.digits_to_residue(p, n, d, off)).all()
    assert L.res64(p, n, d, off) == O.find_residue(p, n, d, off)
    s, z = L.standardize_digits(p, n, d)
    assert not z and (L.pack_bits(p, n, s) == w).all()
    assert (L.initial_digits(p, n, off) == O.initial_digits(p, n, off)).all()


def test_standardize_handles_unbalanced_and_zero():
    p, n = 44497, 4096
    rng = np.random.default_rng(5)
    d = rng.integers(-(1 << 20), 1 << 20, n).astype(np.int32)  # far outside balanced range
    assert (L.residue_words(p, n, d, 0) == O.digits_to_residue(p, n, d, 0)).all()
    # 2^p - 1 == 0
    ones = np.array([(1 << L.digit_width(p, n, j)) - 1 for j in range(n)], dtype=np.int32)
    s, z = L.standardize_digits(p, n, ones)
    assert z and not s.any()


def test_crc_matches_oracle_and_check_value():
    rng = np.random.default_rng(1)
    for size in (0, 1, 9, 1000, 4099):
        b = rng.bytes(size)
        assert L.checkpoint_crc(b) == O.checkpoint_crc(b)
    assert L.checkpoint_crc(b"123456789") == 0x89A1897F


class Ckpt(ctypes.Structure):
    _fields_ = [("p", ctypes.c_uint32), ("n", ctypes.c_uint32), ("next_iter", ctypes.c_uint32),
                ("offset", ctypes.c_uint32), ("total_us", ctypes.c_uint64), ("time_adj_us", ctypes.c_uint64),
                ("iter_adj", ctypes.c_uint32)]


def encode(p, n, digits, nxt, off, total_us=0):
    lib = L.lib()
    lib.ll_checkpoint_words.restype = ctypes.c_size_t
    nw = lib.ll_checkpoint_words(p)
    words = np.zeros(nw, dtype=np.uint32)
    m = Ckpt(p, n, nxt, off, total_us, 0, 0)
    d = np.ascontiguousarray(digits, dtype=np.int32)
    rc = lib.ll_checkpoint_encode(ctypes.byref(m), d.ctypes.data_as(ctypes.c_void_p),
                                  words.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    return words


def test_checkpoint_layout_matches_reference_format():
    p, n = 86243, 8192
    rng = np.random.default_rng(2)
    v = random_value(p, rng)
    w = O.int_to_words(v, p)
    d = L.balance_digits(p, n, L.unpack_bits(p, n, w))
    words = encode(p, n, d, 12346, 777, total_us=123 << 15)
    end = O.nwords(p)
    assert words.size == end + 10
    assert (words[:end] == w).all()  # packed standardized residue (pack_bits)
    assert list(words[end:end + 9]) == [p, n, 12346, 777, 123, 0, 0, 0, 0]
    assert words[end + 9] == O.checkpoint_crc(words[:end + 9].tobytes())
    # decode
    m = Ckpt()
    assert L.lib().ll_checkpoint_decode(p, words.ctypes.data_as(ctypes.c_void_p), ctypes.byref(m)) == 0
    assert (m.n, m.next_iter, m.offset, m.total_us) == (n, 12346, 777, 123 << 15)
    # digits at another length carry the same value
    d2 = np.zeros(16384, dtype=np.int32)
    L.lib().ll_checkpoint_digits(p, 16384, words.ctypes.data_as(ctypes.c_void_p), d2.ctypes.data_as(ctypes.c_void_p))
    assert O.words_to_int(L.residue_words(p, 16384, d2, 0)) == v
    # a flipped bit fails the CRC
    bad = words.copy()
    bad[5] ^= 1
    assert L.lib().ll_checkpoint_decode(p, bad.ctypes.data_as(ctypes.c_void_p), ctypes.byref(m)) != 0


def test_checkpoint_files_rotation_and_backup(tmp_path):
    lib = L.lib()
    p, n = 44497, 4096
    d0 = L.initial_digits(p, n, 5)
    w1 = encode(p, n, d0, 10001, 5)
    w2 = encode(p, n, d0, 20001, 5)
    P = ctypes.c_void_p
    dirb = str(tmp_path).encode()
    assert lib.ll_checkpoint_write(dirb, p, w1.ctypes.data_as(P), None, ctypes.c_uint64(0)) == 0
    assert lib.ll_checkpoint_write(dirb, p, w2.ctypes.data_as(P), str(tmp_path / "save").encode(),
                                   ctypes.c_uint64(0xABCDEF)) == 0
    assert (tmp_path / ("c%d" % p)).read_bytes() == w2.tobytes()
    assert (tmp_path / ("t%d" % p)).read_bytes() == w1.tobytes()
    assert (tmp_path / "save" / ("s%d.20000.0000000000abcdef" % p)).exists()
    out = np.zeros_like(w1)
    assert lib.ll_checkpoint_read(dirb, p, out.ctypes.data_as(P)) == 0 and (out == w2).all()
    # corrupt c: the backup (validated by CRC) is used
    raw = bytearray((tmp_path / ("c%d" % p)).read_bytes())
    raw[17] ^= 0x40
    (tmp_path / ("c%d" % p)).write_bytes(bytes(raw))
    assert lib.ll_checkpoint_read(dirb, p, out.ctypes.data_as(P)) == 1 and (out == w1).all()
    # both corrupt / truncated: no checkpoint
    (tmp_path / ("t%d" % p)).write_bytes(w1.tobytes()[:-4])
    assert lib.ll_checkpoint_read(dirb, p, out.ctypes.data_as(P)) < 0
    lib.ll_checkpoint_remove(dirb, p)
    assert not (tmp_path / ("c%d" % p)).exists() and not (tmp_path / ("t%d" % p)).exists()


def test_fft_length_policy_matches_oracle():
    assert L.fft_lengths() == O.default_fft_lengths()
    for p in (22133, 42839, 86243, 1000003, 20996011, 57885161, 75846319, 77005277, 116391683):
        assert L.fft_bounds(p) == O.fft_bounds(p)
        n, idx = L.choose_fft_length(p)
        lb, ub = O.fft_bounds(p)
        assert n >= lb and L.fft_supported(n)
        # the first supported list length at or above the lower bound
        first = next(x for x in O.default_fft_lengths() if x >= lb and L.fft_supported(x))
        assert n == first
    # manual override (an index beyond the list is a length in words) is
    # clamped to [lb, ub] like the reference: 4M is below lb(77M) = 4216441,
    # so it is raised to the next list length, 4320K (every list length has
    # kernels in this build)
    n, _ = L.choose_fft_length(77005277, 4194304)
    assert n == 4423680
    assert all(L.fft_supported(x) for x in O.default_fft_lengths())
    n, _ = L.choose_fft_length(75846319, 4194304)
    assert n == 4194304
    n, _ = L.choose_fft_length(44497, 4096)
    assert n == 4096


def test_fft_from_str():
    assert L.fft_from_str("4096K") == 4194304
    assert L.fft_from_str("4M") == 4194304
    assert L.fft_from_str("1536k") == 1572864
    assert L.fft_from_str("65536") == 65536
    assert L.fft_from_str("4Kx") == -1 and L.fft_from_str("K") == -1


def test_supported_lengths_cover_the_configs():
    from cudalucas_amd import workloads as W

    for n in [W.SMALL_N, W.LARGE_N] + W.SMOOTH_LENGTHS:
        assert L.fft_supported(n)


def run_driver(args, cwd):
    return subprocess.run([DRIVER] + args, cwd=cwd, capture_output=True, text=True, timeout=60)


def test_driver_version_help_and_bad_options(tmp_path):
    r = run_driver(["-v"], tmp_path)
    assert "CUDALucasAMD" in r.stdout and r.returncode == 2
    r = run_driver(["-h"], tmp_path)
    assert "-threads" in r.stderr
    r = run_driver(["-f"], tmp_path)
    assert r.returncode == 2 and "can't parse -f" in r.stderr
    r = run_driver(["-f", "12Q", "44497"], tmp_path)
    assert r.returncode == 2 and "can't parse fft length" in r.stderr


def test_driver_ini_precedence(tmp_path):
    (tmp_path / "CUDALucas.ini").write_text(
        "CheckpointIterations=50000\nReportIterations=2000\nErrorLimit=60\nThreads=512 256\n"
        "FFTLength=4M\nWorkFile=todo.txt\nResultsFile=out.txt\nSaveAllCheckpoints=1\nSaveFolder=keep\n")
    r = run_driver(["--print-config"], tmp_path)
    cfg = dict(kv.split("=") for kv in r.stdout.split())
    assert cfg["checkpoint"] == "50000" and cfg["report"] == "2000"
    assert cfg["error_limit"] == "47"  # clamped to [1, 47] (CUDALucas.cu:3519-3520)
    assert cfg["threads"] == "512,256" and cfg["fft"] == "4194304"
    assert cfg["workfile"] == "todo.txt" and cfg["results"] == "out.txt" and cfg["save_all"] == "1"
    # command line beats the ini file
    r = run_driver(["-c", "200000", "-x", "5000", "-f", "2048K", "-threads", "128", "64", "--print-config"], tmp_path)
    cfg = dict(kv.split("=") for kv in r.stdout.split())
    assert cfg["checkpoint"] == "200000" and cfg["report"] == "5000"
    assert cfg["fft"] == "2097152" and cfg["threads"] == "128,64"
    # intervals are rounded to k * 10^m, k in {1, 2, 5}
    r = run_driver(["-x", "3000", "--print-config"], tmp_path)
    assert "report=2000" in r.stdout


def test_fft_length_list_can_be_replaced_and_restored():
    """init_ffts merging "<device> fft.txt" (CUDALucas.cu:947-1025) goes
    through ll_fft_length_list_set; choose_fft_length then works on it."""
    lib = L.lib()
    default = L.fft_lengths()
    custom = np.array([1024 * k for k in (1, 2, 4, 8, 20, 64, 128, 256, 512, 1024, 4096, 8192)], dtype=np.uint32)
    assert lib.ll_fft_length_list_set(len(custom), custom.ctypes.data) == 0
    try:
        assert L.fft_lengths() == [int(x) for x in custom]
        n, idx = L.choose_fft_length(500009)  # lb ~ 24.5K words: next listed length is 64K
        assert n == 65536 and L.fft_lengths()[idx] == n
        bad = np.array([2048, 1024], dtype=np.uint32)
        assert lib.ll_fft_length_list_set(2, bad.ctypes.data) == L.LL_ERR_ARG  # not ascending
    finally:
        assert lib.ll_fft_length_list_set(0, None) == 0
    assert L.fft_lengths() == default


def test_schedule_table_overrides_builtin_preference():
    from cudalucas_amd import tuner as T

    n = 262144
    every = T.schedules(n)
    assert len(every) >= 4 and all(s.n == n for s in every)
    default = L.fft_schedule(n)
    other = next(s for s in every if (s.width, s.height, s.lines) != default[1:])
    T.install([other])
    try:
        assert L.fft_schedule(n) == ("sweeps", other.width, other.height, other.lines)
        # an explicit schedule in the options still wins
        assert L.fft_schedule(n, width=default[1], height=default[2], lines=default[3]) == default
        assert L.fft_schedule(524288) == L.fft_schedule(524288)  # other lengths untouched
    finally:
        T.install([])
    assert L.fft_schedule(n) == default


def test_tuner_table_round_trip(tmp_path):
    from cudalucas_amd import tuner as T

    t = [T.Schedule(8192, 0, 4096, 1, 1, 0.01), T.Schedule(131072, 1, 256, 256, 8, 0.02)]
    T.save(t, str(tmp_path / "s.txt"), "test")
    back = T.load(str(tmp_path / "s.txt"))
    assert [b.key() for b in back] == [x.key() for x in t]
    assert T.parse_lengths("4096K,5M") == [4194304, 5242880]
    assert all(L.fft_supported(n) for n in T.parse_lengths("8K-64K"))


def test_driver_bench_options_parse(tmp_path):
    r = run_driver(["-h"], tmp_path)
    assert "-cufftbench" in r.stderr and "-threadbench" in r.stderr and "--selector" in r.stderr
    assert "-threadbench 0 fft passes" in r.stderr
    r = run_driver(["-cufftbench", "1"], tmp_path)
    assert r.returncode == 2 and "can't parse -cufftbench" in r.stderr
    r = run_driver(["-memtest", "10", "1"], tmp_path)
    assert r.returncode == 2 and "-memtest is not supported" in r.stderr


@pytest.mark.skipif(O.parse_ref() is None, reason="reference parse.c not built (oracle/_ref absent)")
def test_driver_ini_reading_matches_reference_readers(tmp_path):
    """process_options' ini reads (IniGetInt / IniGetStr, parse.c:683-771)
    against the reference's own readers on seeded random ini files:
    duplicate keys, values that do not parse (passed over for a later line),
    empty values, CRLF endings, near-miss keys."""
    import random
    R = O.parse_ref()
    rng = random.Random(3388)
    ints = {"DeviceNumber": ("device", 0), "ErrorReset": ("error_reset", 85),
            "SaveAllCheckpoints": ("save_all", 0), "RoundoffTest": ("roundoff", 0)}
    strs = {"WorkFile": ("workfile", "worktodo.txt"), "ResultsFile": ("results", "results.txt"),
            "SaveFolder": ("save_folder", "savefiles")}
    for case in range(60):
        lines = []
        for _ in range(rng.randrange(0, 14)):
            key = rng.choice(list(ints) + list(strs) + ["DeviceNumbers", "workfile", " WorkFile", "Polite", "Threads"])
            if key in strs or key.lower() == "workfile":
                val = rng.choice(["", "w%d.txt" % rng.randrange(100), "dir/r.txt", "x"])
            elif key == "Threads":
                val = rng.choice(["256 128", "512", "64 abc", "", "1024 256", " 128  64"])
            else:
                val = rng.choice(["", str(rng.randrange(0, 200)), "abc", " 7", "12abc", "0x10", "-0"])
            lines.append("%s%s%s%s" % (key, rng.choice(["=", "=", " ="]), val, rng.choice(["\n", "\r\n"])))
        ini = tmp_path / ("c%d.ini" % case)
        ini.write_text("".join(lines), newline="")
        r = run_driver(["-i", ini.name, "--print-config"], tmp_path)
        assert r.returncode == 0, r.stderr
        cfg = dict(kv.split("=", 1) for kv in r.stdout.split())
        for key, (name, dflt) in ints.items():
            v = ctypes.c_int(0)
            R.IniGetInt(os.fsencode(ini), key.encode(), ctypes.byref(v), dflt)
            assert int(cfg[name]) == v.value, (case, key, lines)
        for key, (name, dflt) in strs.items():
            buf = ctypes.create_string_buffer(192)
            R.IniGetStr(os.fsencode(ini), key.encode(), buf, dflt.encode())
            assert cfg[name] == buf.value.decode(), (case, key, lines)
        # Threads (IniGetInt2): the same pair where the reference finds one;
        # without one this build keeps its automatic choice (0,0) where the
        # reference falls back to 256
        t0, t1 = ctypes.c_int(0), ctypes.c_int(0)
        if R.IniGetInt2(os.fsencode(ini), b"Threads", ctypes.byref(t0), ctypes.byref(t1), 256):
            assert cfg["threads"] == "%d,%d" % (t0.value, t1.value), (case, lines)
        else:
            assert cfg["threads"] == "0,0", (case, lines)


def test_locked_files_wait_time_out_and_exclude_the_reference(tmp_path, capfd):
    """fopen_and_lock / unlock_and_fclose (parse.c:775-894): "<path>.lck" is
    the lock; a holder is waited for, and the reference's own lock (where it
    is built) excludes this build's and the other way round."""
    import threading
    import time
    lib = L.lib()
    path = os.fsencode(tmp_path / "results.txt")
    lck = tmp_path / "results.txt.lck"
    f = lib.ll_fopen_locked(path, b"a", 0)
    assert f and lck.exists()
    t0 = time.monotonic()
    assert not lib.ll_fopen_locked(path, b"a", 60)  # held: gives up after about 60 ms
    assert time.monotonic() - t0 >= 0.05
    # a waiter without a limit gets the file once the holder lets go
    threading.Timer(0.15, lambda: lib.ll_fclose_locked(f)).start()
    g = lib.ll_fopen_locked(path, b"a", 0)
    assert g and lck.exists()
    assert lib.ll_fclose_locked(g) == 0 and not lck.exists()
    libc = ctypes.CDLL(None)
    libc.fflush.argtypes = [ctypes.c_void_p]
    libc.fflush(None)  # the messages go through C stdio
    out = capfd.readouterr()
    assert "results.txt is locked, waiting ..." in out.err and "Locked " in out.out
    R = O.parse_ref()
    if R is None:
        return
    h = R.fopen_and_lock(path, b"a")
    assert h and lck.exists()
    assert not lib.ll_fopen_locked(path, b"a", 60)
    R.unlock_and_fclose(h)
    assert not lck.exists()
    f = lib.ll_fopen_locked(path, b"a", 60)
    assert f
    lib.ll_fclose_locked(f)
    assert not lck.exists()


INI_FORMAT = "|  %M %D  %h%m%s  | %P %i  %r  | %f  %x5 %q4 %c2s  | %T%d2%%  |"


@pytest.mark.parametrize("j,residue,err,secs,eta,row", [
    (18900000, 0x219381f578a2abf9, 0.28516, 539.92, (2, 11, 5, 11),
     "|  Dec 15  12:46:01  |  M57885161  18900000  0x219381f578a2abf9  |  3136K  0.28516   5.3992  539.92s  |  2:11:05:11  32.65%  |"),
    (19000000, 0x41e63fccc43ad9a5, 0.29297, 539.93, (2, 10, 50, 19),
     "|  Dec 15  12:55:01  |  M57885161  19000000  0x41e63fccc43ad9a5  |  3136K  0.29297   5.3993  539.93s  |  2:10:50:19  32.82%  |"),
    (19100000, 0x6e8b11826f5f2620, 0.29688, 539.98, (2, 10, 37, 4),
     "|  Dec 15  13:04:01  |  M57885161  19100000  0x6e8b11826f5f2620  |  3136K  0.29688   5.3998  539.98s  |  2:10:37:04  32.99%  |"),
])
def test_report_table_reproduces_the_documented_rows(j, residue, err, secs, eta, row):
    """The OutputString example and its three rows in CUDALucas.ini:113-144.
    The rows were written with an ETA column one character narrower than
    process_output's %12s (CUDALucas.cu:1460, 1678), so that one column is
    compared with its leading blank taken off; every other character must
    match."""
    import time
    clock = row.split("|")[1].split()
    hh, mm, ss = (int(x) for x in clock[2].split(":"))
    now = int(time.mktime((2012, 12, int(clock[1]), hh, mm, ss, 0, 0, -1)))
    d, h, m, s = eta
    line = L.format_report(INI_FORMAT, p=57885161, n=3136 * 1024, iteration=j, residue=residue, max_err=err,
                           interval_iterations=100000, interval_us=int(round(secs * 1e6)),
                           eta_s=((d * 24 + h) * 60 + m) * 60 + s, now=now)
    assert line.replace("|   %d:" % d, "|  %d:" % d) == row


def test_report_codes_widths_and_cuts():
    """Each code of encode_output_options / process_output
    (CUDALucas.cu:3189-3372, 1633-1701) with the widths the reference prints."""
    f = dict(p=86243, n=8192, iteration=20000, offset=12345, residue=0xABCDEF0123456789, max_err=0.1875,
             interval_iterations=10000, interval_us=2500000, eta_s=3725, now=0, program="CUDALucasAMD v1.0")
    assert L.format_report("%p|%P|%f|%F|%i|%o", **f) == "    86243|    M86243|    8K|    8192|    20000|    12345"
    assert L.format_report("%r %R %N %%", **f) == "0xabcdef0123456789 0xABCDEF0123456789 CUDALucasAMD v1.0 %"
    assert L.format_report("%x5|%x3|%x0|%x9|%x", **f) == "0.18750|0.187|0|0.18750|0.18750"
    assert L.format_report("%q4|%c2|%Q1|%d2", **f) == "  0.2500|   2.50|4000.0|  23.19"
    assert L.format_report("[%C][%T]", **f) == "[        0:02][     1:02:05]"
    assert L.format_report("[%T]", **dict(f, eta_s=2 * 86400 + 5)) == "[  2:00:00:05]"
    assert L.format_report("a%jb%", **f) == "ab"  # unknown code prints nothing; a lone % at the end too
    assert L.format_report("%x5|", **f) == "0.18750|" and L.format_report("%xz|", **f) == "0.18750|"
    long = "%r" * 40
    assert L.format_report(long, **f) == "0xabcdef0123456789" * 40  # no fixed line buffer


def test_report_table_options_from_ini(tmp_path):
    """OutputHInterval: missing = 0 (the one-line report), negative = 15
    (CUDALucas.cu:3459, 3488)."""
    r = run_driver(["--print-config"], tmp_path)
    assert "output_interval=0" in r.stdout
    (tmp_path / "CUDALucas.ini").write_text("OutputHInterval=-1\n")
    assert "output_interval=15" in run_driver(["--print-config"], tmp_path).stdout
    (tmp_path / "CUDALucas.ini").write_text("OutputHInterval=19\nOutputString=%p %i\n")
    assert "output_interval=19" in run_driver(["--print-config"], tmp_path).stdout


def test_c_abi_header_builds_as_c99_and_links(tmp_path):
    """include/ll_api.h is plain C: a C99 client (the language of the
    reference's host code) compiles with -pedantic -Werror, links against the
    library and calls host-only entry points (no GPU needed)."""
    import shutil
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    src = tmp_path / "client.c"
    src.write_text(r'''
#include <stdio.h>
#include "ll_api.h"
int main(void) {
  ll_assignment a;
  ll_report r = {0};
  char line[64];
  int st = ll_work_parse_line("Test=N/A,77005277,4M", &a);
  r.p = 77005277; r.n = 4194304; r.iteration = 10000; r.residue = 0x1234u;
  ll_format_report("%P %i %f %r", &r, line, sizeof line);
  printf("%d %u %u %s|%d %d|%s\n", st, a.exponent, a.fft_length, a.key, ll_is_prime(77005277),
         ll_fft_from_str("4096K"), line);
  return 0;
}
''')
    lib_dir = os.path.dirname(L.LIB_PATH)
    inc = os.path.join(os.path.dirname(lib_dir), "include")
    exe = tmp_path / "client"
    b = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I", inc, "-o", str(exe),
                        str(src), "-L", lib_dir, "-lcudalucas_amd", "-Wl,-rpath," + lib_dir],
                       capture_output=True, text=True, timeout=120)
    assert b.returncode == 0, b.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "0 77005277 4194304 N/A|1 4194304| M77005277     10000  4096K 0x0000000000001234"
