"""Runs the C++ host test programs built from tests/cpp (see its Makefile).

test_semantics (CPU): the reference-compatible API layer and the oracle
SerialSimulator's plugin event semantics.
test_hipsimulator (GPU): HipSimulator against the oracle SerialSimulator on
the same user cells and plugins.
"""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "bin"


def run(name, timeout):
    exe = BIN / name
    assert exe.exists(), f"{exe} missing: build with __graft_entry__.build()"
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=timeout)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "0 failures" in r.stdout


def test_cpp_semantics():
    run("test_semantics", 300)


@pytest.mark.gpu
def test_cpp_hipsimulator():
    run("test_hipsimulator", 600)
