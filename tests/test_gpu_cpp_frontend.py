"""The C++ drop-in GIR class (host/gir.h) on the GPU: runs the differential
self-test tests/cpp/gir_selftest (built by __graft_entry__.build())."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "tests", "cpp", "gir_selftest")


def test_cpp_gir_selftest():
    assert os.path.exists(BIN), "build the C++ front end first (__graft_entry__.build())"
    print("[case] cpp gir_selftest 16 iterations", flush=True)
    r = subprocess.run([BIN, "16"], capture_output=True, text=True, timeout=110)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
