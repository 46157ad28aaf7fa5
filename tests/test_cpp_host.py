"""Runs the C++ host test programs built from tests/cpp (see its Makefile).

test_semantics (CPU): the reference-compatible API layer and the oracle
SerialSimulator's plugin event semantics.
test_hipsimulator (GPU): HipSimulator against the oracle SerialSimulator on
the same user cells and plugins.
test_generic (GPU): cells without a hand-written kernel family, their own
update() compiled by hipcc (the generic kernel family), against the oracle.
test_soa (GPU): struct-of-arrays cells and their static updateLineX (the
SoA kernel family) and array-of-structs line cells, against the oracle and
the reference's closed forms.
"""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "bin"


def run(name, timeout, cwd):
    """Runs one program in `cwd` (the writer tests create files there)."""
    exe = BIN / name
    assert exe.exists(), f"{exe} missing: build with __graft_entry__.build()"
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=timeout,
                       cwd=str(cwd))
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "0 failures" in r.stdout


def test_cpp_semantics(tmp_path):
    run("test_semantics", 300, tmp_path)


@pytest.mark.gpu
def test_cpp_hipsimulator(tmp_path):
    run("test_hipsimulator", 600, tmp_path)


@pytest.mark.gpu
def test_cpp_generic(tmp_path):
    run("test_generic", 600, tmp_path)


@pytest.mark.gpu
def test_cpp_soa(tmp_path):
    run("test_soa", 600, tmp_path)
