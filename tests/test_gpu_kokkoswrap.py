"""The C++ KokkosWrap::LocalMatrix shim over the C-ABI (include/KokkosWrap.hpp),
exercised by tests/cpp/test_kokkoswrap.cpp (built by __graft_entry__.build())."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "cpp", "test_kokkoswrap")


def test_kokkoswrap_shim():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built; run __graft_entry__.build()")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
