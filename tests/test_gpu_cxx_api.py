"""Runs the C++ API test program (tests/cpp/test_cxx_api.cpp): the reference's
group-by known answers through the cudf:: facade headers, linked against the
C-ABI library."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_cxx_api")


def test_cxx_api_known_answers():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (python -m cudf_amd.build)")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout
