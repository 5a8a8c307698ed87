"""Runs the C++ HipBackend test program (tests/cpp/test_hip_backend.cpp,
built into the package's lib/ by build()): its CPU-only cases everywhere,
its device cases on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "indigenous-cpu-gpu-optimization-engine-for-indian-industry_amd", "lib", "test_hip_backend")


def _run(args):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    p = subprocess.run([EXE, *args], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(p.stdout)
    return p


def test_cpp_backend_without_device():
    p = _run(["--no-gpu"])
    assert p.returncode == 0, p.stdout


@pytest.mark.gpu
def test_cpp_backend_on_device():
    p = _run([])
    assert p.returncode == 0, p.stdout
    assert "PASSED" in p.stdout
