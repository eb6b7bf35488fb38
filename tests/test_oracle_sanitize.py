"""The CPU oracle under AddressSanitizer, LeakSanitizer and
UndefinedBehaviorSanitizer: tests/cpp/oracle_sanitize.c drives every oracle
entry point over empty, single-row and ragged inputs with every storage kind,
nulls, NaN, -0.0, empty strings, every aggregation kind and both null
policies, and checks invariants of the results.  The memory-checking
counterpart of the reference's compute-sanitizer runs
(ci/run_cudf_memcheck_ctests.sh:12-20), for the test oracle; it needs no GPU.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build(tmp_path, flags):
    exe = str(tmp_path / "oracle_sanitize")
    cmd = ["gcc", "-std=c11", "-g", "-O1", "-fno-omit-frame-pointer", *flags,
           os.path.join(ROOT, "oracle", "groupby_oracle.c"), os.path.join(ROOT, "tests", "cpp", "oracle_sanitize.c"),
           "-lm", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "sanitize" in r.stderr:
        pytest.skip(f"compiler without sanitizer runtime: {r.stderr[-300:]}")
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_clean_under_address_and_undefined_sanitizers(tmp_path):
    exe = _build(tmp_path, ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed checks" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_sanitizer_build_reports_a_planted_overflow(tmp_path):
    """The instrumented build must catch what it is there to catch: a copy of
    the oracle whose row-to-group array is one element short faults under
    ASan with the same driver."""
    src = open(os.path.join(ROOT, "oracle", "groupby_oracle.c")).read()
    line = "int64_t* group_of_row = (int64_t*)malloc(sizeof(int64_t) * (n > 0 ? n : 1));"
    assert line in src
    planted = src.replace(line, "int64_t* group_of_row = (int64_t*)malloc(sizeof(int64_t) * (n > 1 ? n - 1 : 1));", 1)
    path = tmp_path / "planted_oracle.c"
    path.write_text(planted)
    exe = str(tmp_path / "planted")
    cmd = ["gcc", "-std=c11", "-g", "-O1", "-fsanitize=address", str(path),
           os.path.join(ROOT, "tests", "cpp", "oracle_sanitize.c"), "-lm", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "sanitize" in r.stderr:
        pytest.skip("compiler without sanitizer runtime")
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr
