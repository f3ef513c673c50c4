"""The drop-in C++ API on the GPU: tests/cpp/test_dropin.cpp is user code written against
include/cmpmc/ (the reference's names and signatures) and linked to libcmpmc_amd.so. It
follows the reference's tests/test_parallel.cpp and tests/test_engine_energy_closure.cpp:
bit-identical results for any number of generator threads, run_events equal to a serial
transport_charges/transport loop over the same streams, hit_offsets with n + 1 entries,
ledger closure, keep_hits = false, and exceptions reaching the caller."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_dropin")


def test_dropin_cpp_program():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: build it with cmpmc_amd/build.py (or __graft_entry__.build())")
    print("case: tests/cpp/test_dropin 203 events", flush=True)
    r = subprocess.run([BIN, "203"], capture_output=True, text=True, timeout=300)
    print(r.stdout[-4000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, "drop-in C++ checks failed"
    assert "# 0 check(s) failed" in r.stdout
