"""The worktodo workflow (parse.c) of the product, checked on the CPU.

Where oracle/_ref/libparse_ref.so is built (the reference's own parse.c,
compiled in place by oracle/Makefile), the product's functions are compared
with the reference's on the same input:

- ll_work_parse_line with parse_worktodo_line (parse.c:241-484) on
  hand-written lines and on seeded random lines drawn from the work-line
  grammar (key, exponent, FFT length, trial-factoring depth, P-1 flag,
  comments, blanks, malformed fields);
- ll_work_next with get_next_assignment (parse.c:486-554) and ll_work_clear
  with clear_assignment (parse.c:556-678) on seeded random worktodo files:
  same return code, same assignment, byte-identical rewritten file;
- ll_valid_assignment / ll_is_prime with valid_assignment / isprime.

Known answers read off the reference's rules run without it.  Two
deliberate differences are kept out of the comparisons and checked on their
own: trailing blanks of the key and the comment are trimmed, and a '/'
inside the last field of a line with a // comment stays in that field (the
reference ends that field at its first '/', parse.c:349-350).  Lines longer
than 190 characters are one rejected line here, where the reference reads
their remainder as a further line.
"""
import ctypes
import os
import random
import subprocess

import pytest

import cudalucas_amd as L
import oracle as O

REF = O.parse_ref()
needs_ref = pytest.mark.skipif(REF is None, reason="reference parse.c not built (oracle/_ref absent)")
DRIVER = os.path.join(os.path.dirname(L.LIB_PATH), "CUDALucasAMD")

_libc = ctypes.CDLL(None)
_libc.fopen.restype = ctypes.c_void_p
_libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
_libc.fclose.argtypes = [ctypes.c_void_p]
_libc.fflush.argtypes = [ctypes.c_void_p]


@pytest.fixture(autouse=True)
def _flush_reference_output():
    """The reference prints its reasons through C stdio: flush them while
    pytest still captures the test's output."""
    yield
    _libc.fflush(None)


def ref_parse_lines(lines, tmp_path):
    """parse_worktodo_line over a file holding the lines, one call per line."""
    path = tmp_path / "lines.txt"
    path.write_bytes(b"".join(l.encode() + b"\n" for l in lines))
    f = _libc.fopen(os.fsencode(path), b"r")
    assert f
    out = []
    try:
        for _ in lines:
            a = O.RefAssignment()
            copy = ctypes.create_string_buffer(192)
            tail = ctypes.c_void_p()
            st = REF.parse_worktodo_line(f, ctypes.byref(a), ctypes.cast(copy, ctypes.c_void_p), ctypes.byref(tail))
            out.append((st, a.exponent, a.fft_length, a.hex_key.decode(), a.comment.decode()))
    finally:
        _libc.fclose(f)
    return out


def same_parse(ours, ref):
    st, a = ours
    if st != ref[0]:
        return False
    if st != L.WORK_OK:
        return True
    return (a.exponent, a.fft_length, a.key.decode(), a.comment.decode()) == \
        (ref[1], ref[2], ref[3].rstrip(), ref[4].rstrip())


HAND_LINES = [
    "Test=77005277", "Test=N/A,77005277,75,1", "DoubleCheck=0123456789ABCDEF0123456789ABCDEF,75846319,74,1 # ok",
    "test=M77005277", "Test=77005277,4096K", "Test=77005277,4M", "Test=77005277,4 M", "Test=77005277,40 96K",
    "Test=1,2,3,4,5,6", "Test=", "Test=,77005277", "Test=99", "Test=77005277,2K", "  # comment", "// c",
    "\\\\ c", "", "   ", "Factor=1,2", "Test=N/A,77005277 // note here ", "Test=77005277,-5",
    "Test=0x1234,77005277", "Test=7" + "0" * 30, "Test=77005277,", "Test=abc,44497\r", "Test=77005277,0K",
    "Test=77005277,4096Kxyz", "Test=1aK,44497", "Test=  N/A  ,  44497  ,  69 ", "Test=44497\\\\done",
    "DOUBLECHECK=44497#x", "Test=44497,3000000000", "Test=44497,-5K", "Test=K,44497", "Test=44497,+8K",
    "Test=44497 75", "Test=N/A,N/A,44497", "Test=44497,4096K,8192K",
]


@needs_ref
def test_parse_line_matches_reference_on_hand_written_lines(tmp_path):
    ref = ref_parse_lines(HAND_LINES, tmp_path)
    for line, r in zip(HAND_LINES, ref):
        assert same_parse(L.work_parse_line(line), r), (line, L.work_parse_line(line), r)


def random_field(rng):
    kind = rng.random()
    if kind < 0.25:
        return str(rng.choice([rng.randrange(100, 1 << 27), 44497, 86243, 77005277, 75846319, 7499, 7507, 99, 150]))
    if kind < 0.38:
        return rng.choice(["N/A", "n/a", "".join(rng.choice("0123456789ABCDEF") for _ in range(32)), "abc", "1aK"])
    if kind < 0.52:
        return rng.choice(["%dK" % rng.choice([2, 4, 1024, 2048, 4096, 6144]), "%dM" % rng.choice([1, 2, 4]),
                           "4096 K", "40 96K", "4096k", "3m", "M77005277", "K", "0K", "-5K"])
    if kind < 0.7:
        return str(rng.randrange(0, 90))
    if kind < 0.8:
        return rng.choice(["", " ", "0x12", "-5", "3000000000", "7" + "0" * 25, "+44497", "44497xyz"])
    return " " * rng.randrange(0, 3) + str(rng.randrange(100, 1 << 31)) + " " * rng.randrange(0, 3)


def random_line(rng):
    r = rng.random()
    if r < 0.05:
        return " " * rng.randrange(0, 4)
    if r < 0.1:
        return rng.choice(["#", "//", "\\\\"]) + " a comment"
    prefix = rng.choice(["Test=", "DoubleCheck=", "test=", "doublecheck=", "Test= ", "  Test=", "Factor=", "Tset="])
    fields = [random_field(rng) for _ in range(rng.choice([1, 2, 2, 3, 4, 4, 5, 6]))]
    line = prefix + ",".join(fields)
    c = rng.random()
    if c < 0.15:
        line += " # note %d  " % rng.randrange(100)
    elif c < 0.25:
        line += "//note"
    elif c < 0.3:
        line += " \\\\ done"
    return line[:190]


def slash_in_last_field_before_comment(line):
    cut = line.find("//")
    if cut < 0:
        return False
    body = line[:cut]
    return "/" in body[body.rfind(",") + 1:] if "," in body else "/" in body[body.find("=") + 1:]


@needs_ref
def test_parse_line_matches_reference_on_random_lines(tmp_path):
    rng = random.Random(20261203)
    lines = [random_line(rng) for _ in range(4000)]
    compared = [l for l in lines if not slash_in_last_field_before_comment(l)]
    assert len(compared) > 3500
    ref = ref_parse_lines(compared, tmp_path)
    statuses = set()
    for line, r in zip(compared, ref):
        statuses.add(r[0])
        assert same_parse(L.work_parse_line(line), r), (line, L.work_parse_line(line), r)
    # the corpus reaches every status a single line can have
    assert statuses >= {L.WORK_OK, L.WORK_NO_TEST_EQUAL, L.WORK_INVALID_FORMAT, L.WORK_INVALID_DATA,
                        L.WORK_BLANK, L.WORK_COMMENT}


def test_parse_line_known_answers():
    """Values read off parse.c's rules; they hold without the reference."""
    cases = {
        "Test=77005277": (L.WORK_OK, 77005277, 0, "", ""),
        "DoubleCheck=0123456789ABCDEF0123456789ABCDEF,75846319,74,1 # ok":
            (L.WORK_OK, 75846319, 0, "0123456789ABCDEF0123456789ABCDEF", "ok"),
        "Test=N/A,77005277,4096K,75,1": (L.WORK_OK, 77005277, 4194304, "N/A", ""),
        "Test=77005277,4M": (L.WORK_OK, 77005277, 4194304, "", ""),
        "Test=abc,44497\r": (L.WORK_OK, 44497, 0, "abc", ""),
    }
    for line, want in cases.items():
        st, a = L.work_parse_line(line)
        assert (st, a.exponent, a.fft_length, a.key.decode(), a.comment.decode()) == want, line
    status = {
        "": L.WORK_BLANK, "  # c": L.WORK_COMMENT, "// c": L.WORK_COMMENT, "\\\\ c": L.WORK_COMMENT,
        "Factor=1,2": L.WORK_NO_TEST_EQUAL, "Test=1,2,3,4,5,6": L.WORK_INVALID_FORMAT,
        "Test=,44497": L.WORK_INVALID_FORMAT, "Test=44497,4096 K": L.WORK_INVALID_FORMAT,
        "Test=M77005277": L.WORK_INVALID_FORMAT, "Test=99": L.WORK_INVALID_DATA,
        "Test=44497,2K": L.WORK_INVALID_DATA, "Test=44497,3000000000": L.WORK_INVALID_DATA,
        "Test=44497," + "1" * 175: L.WORK_INVALID_DATA, "Test=44497," + "1" * 180: L.WORK_LONG_LINE,
    }
    for line, want in status.items():
        assert L.work_parse_line(line)[0] == want, line
    # deliberate difference: '/' in the last field before a // comment
    st, a = L.work_parse_line("Test=44497,N/A // c")
    assert (st, a.exponent, a.key, a.comment) == (L.WORK_OK, 44497, b"N/A", b"c")


def random_workfile(rng, n_lines):
    lines = []
    for _ in range(n_lines):
        r = rng.random()
        if r < 0.15:
            lines.append("")
        elif r < 0.3:
            lines.append(rng.choice(["# queued", "// next", "\\\\ x"]))
        elif r < 0.4:
            lines.append(random_line(rng))
        else:
            key = rng.choice(["", "N/A,", "".join(rng.choice("0123456789abcdef") for _ in range(32)) + ","])
            p = rng.choice([44497, 86243, 110503, 132049, 216091, 756839, 859433, 1257787, 1398269,
                            77005277, 75846319, 7499, 7501, 44499, 500])
            tail = rng.choice(["", ",69,0", ",75,1", ",4096K,74,1", ",2M"])
            lines.append("%s=%s%d%s" % (rng.choice(["Test", "DoubleCheck"]), key, p, tail))
    lines = [l for l in lines if not slash_in_last_field_before_comment(l)]
    return "".join(l + ("\r\n" if rng.random() < 0.2 else "\n") for l in lines)


@needs_ref
def test_next_assignment_matches_reference(tmp_path):
    rng = random.Random(77005277)
    for case in range(150):
        text = random_workfile(rng, rng.randrange(0, 12))
        path = tmp_path / "worktodo.txt"
        path.write_text(text, newline="")
        q, n = ctypes.c_int(0), ctypes.c_int(0)
        key = ctypes.create_string_buffer(192)
        rc_ref = REF.get_next_assignment(os.fsencode(path), ctypes.byref(q), ctypes.byref(n),
                                         ctypes.cast(key, ctypes.c_void_p))
        rc, a = L.work_next(path)
        assert rc == rc_ref, (case, text)
        if rc == 0:
            assert (a.exponent, a.fft_length, a.key.decode()) == (q.value, n.value, key.value.decode().rstrip()), text
    rc, a = L.work_next(tmp_path / "missing.txt")
    assert rc == 1 and REF.get_next_assignment(os.fsencode(tmp_path / "missing.txt"), ctypes.byref(q),
                                               ctypes.byref(n), None) == 1


@needs_ref
def test_clear_assignment_matches_reference(tmp_path, monkeypatch):
    rng = random.Random(44497)
    ref_dir, our_dir = tmp_path / "ref", tmp_path / "ours"
    ref_dir.mkdir()
    our_dir.mkdir()
    monkeypatch.chdir(ref_dir)  # clear_assignment writes __worktodo__.tmp in the working directory
    found = 0
    for case in range(200):
        text = random_workfile(rng, rng.randrange(1, 14))
        parsed = [L.work_parse_line(l) for l in text.splitlines()]
        listed = [a.exponent for st, a in parsed if st == L.WORK_OK]
        p = rng.choice(listed) if listed and rng.random() < 0.85 else rng.choice([44497, 1257787, 12345])
        (ref_dir / "worktodo.txt").write_text(text, newline="")
        (our_dir / "worktodo.txt").write_text(text, newline="")
        rc_ref = REF.clear_assignment(b"worktodo.txt", p)
        rc = L.work_clear(our_dir / "worktodo.txt", p)
        assert rc == rc_ref, (case, p, text)
        assert (our_dir / "worktodo.txt").read_bytes() == (ref_dir / "worktodo.txt").read_bytes(), (case, p, text)
        assert not (our_dir / "worktodo.txt.tmp").exists()
        found += rc == 0
        if (ref_dir / "__worktodo__.tmp").exists():
            (ref_dir / "__worktodo__.tmp").unlink()
    assert found > 100
    assert L.work_clear(our_dir / "missing.txt", 44497) == 3 == REF.clear_assignment(b"missing.txt", 44497)


def test_clear_assignment_takes_the_comments_above_it(tmp_path):
    """clear_assignment's rule (parse.c:592-660): the lines between the
    previous assignment and the finished one go with it; with no assignment
    above, the lines after the file's first blank line do, and those before
    it stay."""
    path = tmp_path / "worktodo.txt"
    path.write_text("# header\n\n# for 44497\nTest=44497\n# for 86243\nTest=86243\n# tail\n")
    assert L.work_clear(path, 86243) == 0
    assert path.read_text() == "# header\n\n# for 44497\nTest=44497\n# tail\n"
    assert L.work_clear(path, 44497) == 0
    assert path.read_text() == "# header\n\n# tail\n"
    path.write_text("# header\nTest=44497\n")
    assert L.work_clear(path, 44497) == 0
    assert path.read_text() == "# header\n"
    assert L.work_clear(path, 44497) == 5 and path.read_text() == "# header\n"


@needs_ref
def test_valid_assignment_and_isprime_match_reference():
    rng = random.Random(3)
    values = list(range(0, 2000)) + [7499, 7500, 7501, 7507, 44497, 2147483647, 2147483646, 4294967291,
                                     4294967295, 65537 * 65537 - 2] + [rng.randrange(1, 1 << 32) for _ in range(3000)]
    for v in values:
        assert L.lib().ll_is_prime(v) == REF.isprime(v), v
    for p in [7499, 7500, 7501, 7507, 44497, 44499, 77005277, 77005279] + [rng.randrange(7000, 1 << 27) for _ in range(300)]:
        for n in [0, 1024, 4096 * 1024, 1000, 4097]:
            assert L.lib().ll_valid_assignment(p, n) == REF.valid_assignment(p, n), (p, n)


def test_driver_worktodo_without_valid_line_and_composite_exponent(tmp_path):
    (tmp_path / "worktodo.txt").write_text("Test=N/A,44499,69,0\nFactor=44497,1,2\nTest=99\n")
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "Warning: exponent is not prime!" in r.stderr
    assert r.stderr.count("Warning: ignoring line") == 3 and "No valid assignment found." in r.stderr
    assert (tmp_path / "worktodo.txt").read_text().startswith("Test=N/A,44499")  # nothing cleared
    r = subprocess.run([DRIVER, "7000"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "exponents < 7500 are not supported" in r.stderr
