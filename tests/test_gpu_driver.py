"""The CUDALucasAMD driver on the GPU: CUDALucas's command-line behaviour.

Results are compared with GMP (the oracle).  Recovery paths are forced from
the host: roundoff errors through --inject-error (ll_debug_inject), an
interrupt through SIGINT, a resume through --stop-after and through a
checkpoint written in the reference's file format by the oracle.
"""
import os
import re
import signal
import subprocess
import time

import numpy as np
import pytest

import cudalucas_amd as L
import oracle as O
from cudalucas_amd import workloads as W
from conftest import announce

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(L.LIB_PATH), "CUDALucasAMD")


def drive(args, cwd, timeout=300):
    announce("driver " + " ".join(str(a) for a in args))
    r = subprocess.run([DRIVER] + [str(a) for a in args], cwd=cwd, capture_output=True, text=True,
                       timeout=timeout)
    return r


def result_line(cwd):
    return (cwd / "results.txt").read_text().strip().splitlines()[-1]


def gmp_res64(p, iters):
    return O.res64(O.gmp_ll(p, iters))


def test_prime_exponent_reports_prime(tmp_path):
    r = drive(["-f", "4096", "-x", "10000", "-c", "10000", "--offset", "99", 44497], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "M( 44497 )P" in r.stdout
    assert result_line(tmp_path).startswith("M( 44497 )P, offset = ")
    # a prime keeps its checkpoints (the reference removes them only for a
    # nonzero residue, CUDALucas.cu:3175): the last one holds s_(k-1),
    # shifted by its offset, for its next iteration k
    p = 44497
    words = np.frombuffer((tmp_path / ("c%d" % p)).read_bytes(), dtype=np.uint32)
    end = O.nwords(p)
    assert int(words[end]) == p and int(words[end + 1]) == 4096
    assert int(words[end + 9]) == O.checkpoint_crc(words[:end + 9].tobytes())
    nxt, off = int(words[end + 2]), int(words[end + 3])
    assert 30000 < nxt < p - 2
    value = O.words_to_int(words[:end])
    unshifted = ((value >> off) | (value << (p - off))) & ((1 << p) - 1)
    assert O.int_to_words(unshifted, p).tolist() == O.gmp_ll(p, nxt - 1).tolist()


def test_composite_run_matches_gmp_and_reports(tmp_path):
    p = W.SMALL_P
    r = drive(["-f", "4096", "-x", "10000", "-c", "10000", "--offset", "123", p], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    want = gmp_res64(p, p - 2)
    assert "M( %d )C, 0x%016x" % (p, want) in result_line(tmp_path)
    # report lines at every ReportIterations with the Res64 of that iteration
    reps = re.findall(r"Iteration (\d+) / %d, 0x([0-9a-f]{16})" % p, r.stdout)
    assert [int(i) for i, _ in reps] == list(range(10000, p - 2, 10000))
    for it, res in reps[:2]:
        assert int(res, 16) == gmp_res64(p, int(it))
    # a finished composite test removes its checkpoints
    assert not (tmp_path / ("c%d" % p)).exists()


def test_report_table_rows_and_header(tmp_path):
    """process_output (CUDALucas.cu:1633-1701) in a run: OutputString rows
    with the Res64 of their iteration, the header above rows 1 and 3 of 4
    with OutputHInterval=3."""
    p = W.SMALL_P
    (tmp_path / "CUDALucas.ini").write_text(
        "OutputHInterval=3\nOutputHeader=HEAD\nOutputString=| %P %i %r | %f %x5 | %d2%% |\n")
    r = drive(["-x", "10000", "--offset", "77", p], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [l for l in r.stdout.splitlines() if l == "HEAD" or l.startswith("| ")]
    rows = [l.split() for l in lines if l != "HEAD"]
    assert [int(x[2]) for x in rows] == list(range(10000, p - 2, 10000))
    assert [i for i, l in enumerate(lines) if l == "HEAD"] == [0, 3]
    for x in rows:
        assert x[1] == "M%d" % p and int(x[3], 16) == gmp_res64(p, int(x[2]))
        assert x[5] == "%dK" % (L.choose_fft_length(p)[0] // 1024) and float(x[6]) < 0.4 and x[8].endswith("%")


def test_stop_and_resume_from_own_checkpoint(tmp_path):
    p = W.SMALL_P
    r1 = drive(["-f", "4096", "-x", "5000", "-c", "5000", "--offset", "7", "--stop-after", "12000", p], tmp_path)
    assert r1.returncode == 0, r1.stdout + r1.stderr
    assert (tmp_path / ("c%d" % p)).exists()
    r2 = drive(["-f", "4096", "-x", "5000", "-c", "5000", p], tmp_path)
    assert r2.returncode == 0, r2.stdout + r2.stderr
    assert "Continuing M%d @ iteration 12001" % p in r2.stdout
    assert "M( %d )C, 0x%016x" % (p, gmp_res64(p, p - 2)) in result_line(tmp_path)


def test_resume_from_reference_format_checkpoint(tmp_path):
    """A c<p> file laid out as CUDALucas writes it (CUDALucas.cu:1707-1849),
    built by the oracle's restatement of pack_bits and the CRC from the exact
    GMP state after 20,000 iterations, shift offset 0, FFT length 4096."""
    p, n, done = W.SMALL_P, 4096, 20000
    s = O.gmp_ll(p, done)
    end = O.nwords(p)
    words = np.zeros(end + 10, dtype=np.uint32)
    digits = O.unpack_bits(p, n, s)
    words[:end] = O.pack_bits(p, n, digits)
    words[end:end + 9] = [p, n, done + 1, 0, 0, 0, 0, 0, 0]
    words[end + 9] = O.checkpoint_crc(words[:end + 9].tobytes())
    (tmp_path / ("c%d" % p)).write_bytes(words.tobytes())
    r = drive(["-x", "10000", p], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Continuing M%d @ iteration %d" % (p, done + 1) in r.stdout
    assert "M( %d )C, 0x%016x" % (p, gmp_res64(p, p - 2)) in result_line(tmp_path)


def test_roundoff_error_retry_then_larger_fft(tmp_path):
    # 8K is inside [lb, ub] = [5705, 14336] for this exponent, so 10K and
    # then 14K may follow
    p = 120223
    # once: retried from the last report state and the error goes away
    r = drive(["-f", "8192", "-x", "5000", "--offset", "5", "--inject-error", "7001:0.45:1", p], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Round off error at iteration = 7001" in r.stdout
    assert "Restarting from last checkpoint to see if the error is repeatable." in r.stdout
    assert "Looks like the error went away, continuing." in r.stdout
    assert "M( %d )C, 0x%016x" % (p, gmp_res64(p, p - 2)) in result_line(tmp_path)
    # twice: a larger FFT until the next report, then back (CUDALucas.cu:3043-3088)
    (tmp_path / "results.txt").unlink()
    r = drive(["-f", "8192", "-x", "5000", "--offset", "5", "--inject-error", "7001:0.45:2", p], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "The error persists." in r.stdout and "Trying a larger fft until the next checkpoint." in r.stdout
    assert "Resetting fft." in r.stdout
    assert "M( %d )C, 0x%016x" % (p, gmp_res64(p, p - 2)) in result_line(tmp_path)
    # three times: a larger FFT did not help either
    r = drive(["-f", "8192", "-x", "5000", "--offset", "5", "--inject-error", "7001:0.45:3", p], tmp_path)
    assert r.returncode == 2 and "The error won't go away. I give up." in r.stdout


def test_roundoff_error_at_the_upper_bound_quits(tmp_path):
    """At the largest length the policy allows (ub(42839) = 5491 words, so
    no list length above 4K), a roundoff error ends the run
    (CUDALucas.cu:3047-3051) instead of retrying the same length."""
    p = W.SMALL_P
    r = drive(["-f", "4096", "-x", "5000", "--offset", "5", "--inject-error", "7001:0.45:2", p], tmp_path)
    assert r.returncode == 2, r.stdout + r.stderr
    assert "Round off error at iteration = 7001" in r.stdout
    assert "Something is wrong! Quitting." in r.stdout
    assert "Trying a larger fft" not in r.stdout


def test_selftest_short_table(tmp_path):
    r = drive(["-r", "0", "--offset", "1000"], tmp_path, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr
    assert r.stdout.count("This residue is correct.") == 20


def test_worktodo_and_results(tmp_path):
    """get_next_assignment / clear_assignment (parse.c:486-678) around check:
    a composite exponent's line is skipped with the reference's warning and
    stays; a length on the work line (8K, inside 86243's bounds) holds for
    that test only."""
    aid = "0123456789ABCDEF0123456789ABCDEF"
    (tmp_path / "worktodo.txt").write_text(
        "Test=N/A,%d,69,0\n# comment\nTest=N/A,44499,69,0\nDoubleCheck=%s,86243,8K,68,1\n" % (W.SMALL_P, aid))
    assert L.choose_fft_length(86243)[0] == 4096  # the default the work line overrides
    r = drive(["-x", "10000"], tmp_path)
    assert r.returncode == 2  # stops when no assignment is left
    assert "Warning: exponent is not prime!" in r.stderr and "No valid assignment found." in r.stderr
    lines = (tmp_path / "results.txt").read_text().strip().splitlines()
    assert lines[0].startswith("M( %d )C, 0x%016x" % (W.SMALL_P, gmp_res64(W.SMALL_P, W.SMALL_P - 2)))
    assert lines[0].split("n = ")[1].startswith("%dK" % (L.choose_fft_length(W.SMALL_P)[0] // 1024))
    assert lines[1].startswith("M( 86243 )P") and lines[1].endswith("AID: " + aid)
    assert lines[1].split("n = ")[1].startswith("8K")
    assert (tmp_path / "worktodo.txt").read_text() == "# comment\nTest=N/A,44499,69,0\n"


def test_sigint_writes_checkpoint(tmp_path):
    p = 1257787
    announce("driver SIGINT p=%d" % p)
    with open(tmp_path / "out.txt", "w") as out, open(tmp_path / "err.txt", "w") as err:
        proc = subprocess.Popen([DRIVER, "-x", "1000", "--offset", "3", str(p)], cwd=tmp_path,
                                stdout=out, stderr=err, stdin=subprocess.DEVNULL)
        try:
            time.sleep(4.0)
            proc.send_signal(signal.SIGINT)
            rc = proc.wait(timeout=60)
        finally:
            if proc.poll() is None:
                proc.kill()
                proc.wait()
    outs, errs = (tmp_path / "out.txt").read_text(), (tmp_path / "err.txt").read_text()
    assert rc == 0, outs[-2000:] + errs
    assert "SIGINT caught" in errs
    assert "Estimated time spent so far" in outs
    words = np.frombuffer((tmp_path / ("c%d" % p)).read_bytes(), dtype=np.uint32)
    end = O.nwords(p)
    nxt, off = int(words[end + 2]), int(words[end + 3])
    assert 1 < nxt < p - 2
    # the saved state is exactly s_(nxt-1), shifted by `off`: compared with
    # an independent engine run from s_0 = 4 (GMP for ~10^5 iterations at
    # this size would take minutes; the engine is pinned to GMP elsewhere)
    value = O.words_to_int(words[:end])
    unshifted = ((value >> off) | (value << (p - off))) & ((1 << p) - 1)
    w, _, _ = L.run_ll(p, 131072, nxt - 1, offset=3)
    assert O.int_to_words(unshifted, p).tolist() == w.tolist()


def test_cufftbench_and_threadbench_write_tuning_files(tmp_path):
    """-cufftbench (CUDALucas.cu:2413-2587) and -threadbench
    (CUDALucas.cu:2143-2276) write "<device> fft.txt" / "<device>
    schedules.txt"; a later run reads both back."""
    r = drive(["-cufftbench", 16, 40, 1], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    fft = [f for f in os.listdir(tmp_path) if f.endswith(" fft.txt")]
    assert len(fft) == 1 and not fft[0].startswith(" ")  # named for the device (or its architecture)
    rows = [l.split() for l in (tmp_path / fft[0]).read_text().splitlines() if l[:5].strip().isdigit()]
    assert rows and all(L.fft_supported(int(k) * 1024) and float(ms) > 0 for k, _, ms in rows)
    assert "16" in [k for k, _, _ in rows]  # powers of two are always kept
    r = drive(["-threadbench", 256, 256, 1, 8], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    sch = [f for f in os.listdir(tmp_path) if f.endswith(" schedules.txt")]
    assert len(sch) == 1 and not sch[0].startswith(" ")
    k, path, w, h, rr, ms = (tmp_path / sch[0]).read_text().split()[:6]
    assert (int(k), int(path)) == (256, 1) and int(w) * int(h) * 2 == 262144
    # the files are read back: a run still matches GMP
    p = 1257787
    r = drive(["-x", "1000", "--stop-after", 2000, "--offset", 5, p], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    rep = re.findall(r"Iteration (\d+) / %d, 0x([0-9a-f]{16})" % p, r.stdout)
    assert rep and int(rep[0][1], 16) == gmp_res64(p, int(rep[0][0]))


def test_kernel_bench_shares(tmp_path):
    """-threadbench 0 K passes (kernel_percentages, CUDALucas.cu:2002-2141):
    per-iteration time and each kernel's share, for the two sweeps at 4096K
    and the LDS-resident kernel at 4K."""
    r = drive(["-threadbench", "0", "4096", "1", "0"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = {}
    for line in r.stdout.splitlines():
        name, _, rest = line.partition(":")
        if name in ("Total time", "Square sweep", "Carry sweep"):
            rows[name] = [float(x) for x in rest.split(",")]
    assert set(rows) == {"Total time", "Square sweep", "Carry sweep"}, r.stdout
    assert abs(rows["Square sweep"][2] + rows["Carry sweep"][2] - 100.0) < 0.05
    assert all(v > 0 for row in rows.values() for v in row[:2])
    r = drive(["-threadbench", "0", "4", "1"], tmp_path)
    assert r.returncode == 0 and "Resident kernel:" in r.stdout, r.stdout + r.stderr


def test_info_prints_device(tmp_path):
    r = drive(["-info"], tmp_path)  # no exponent and no worktodo: exits 2 after the listing
    assert "Device Name" in r.stdout and "multiProcessorCount" in r.stdout, r.stdout + r.stderr


def test_learned_selector_run_matches_gmp(tmp_path):
    from cudalucas_amd import selector as S

    if not os.path.exists(S.MODEL_PATH):
        pytest.skip("no trained model committed")
    p = 2000003
    r = drive(["--selector", S.MODEL_PATH, "--selector-times", S.TIMES_PATH, "-x", "1000", "--stop-after", 3000,
               "--offset", 17, p], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Learned length selection:" in r.stdout
    rep = re.findall(r"Iteration (\d+) / %d, 0x([0-9a-f]{16})" % p, r.stdout)
    assert rep
    for it, res in rep:
        assert int(res, 16) == gmp_res64(p, int(it))
