"""The host library's text handling under AddressSanitizer and UBSan.

tests/fuzz_host_text.cpp drives ll_work_parse_line, ll_work_next,
ll_work_clear, ll_format_report and ll_fopen_locked with seeded random
lines, worktodo files and format strings, and the digit codec and
checkpoint files through round trips on random small states (each round
trip must give its input back); it is built here with g++ and
-fsanitize=address,undefined together with cudalucas_amd/csrc/ll_host.cpp
alone (host code only, no HIP), and any sanitizer report fails the run.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_text_handling_is_clean_under_sanitizers(tmp_path):
    exe = tmp_path / "fuzz_host_text"
    build = subprocess.run(
        ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "cudalucas_amd", "csrc"),
         "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-o", str(exe),
         os.path.join(ROOT, "tests", "fuzz_host_text.cpp"), os.path.join(ROOT, "cudalucas_amd", "csrc", "ll_host.cpp")],
        capture_output=True, text=True, timeout=300)
    if build.returncode != 0 and "asan" in build.stderr.lower():
        pytest.skip("sanitizer runtime not available: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(exe), str(tmp_path), "40000", "120"], capture_output=True, text=True, timeout=600,
                         env=env)
    assert run.returncode == 0 and "fuzz done" in run.stdout, run.stderr[-3000:]
    assert "runtime error" not in run.stderr and "AddressSanitizer" not in run.stderr
