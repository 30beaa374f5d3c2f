"""Runs the C++ test programs built against include/Seis3D.hpp (tests/cpp).

Like the reference's CTest programs they run their CPU checks first and exit 77 (skip) when no
GPU is visible; on a GPU box the full program must pass (exit 0).
"""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "bin"


def run(name, timeout):
    exe = BIN / name
    if not exe.exists():
        pytest.fail(f"{exe} not built (run __graft_entry__.build())")
    return subprocess.run([str(exe)], capture_output=True, text=True, timeout=timeout)


def test_dropin_session_host_checks():
    """The host checks of the drop-in caller run everywhere; its device part may skip with 77."""
    r = run("dropin_session", 300)
    assert r.returncode in (0, 77), r.stderr + r.stdout[-2000:]
    assert "HOST dt 847 us" in r.stdout


@pytest.mark.gpu
def test_dropin_session_gpu(gpu):
    r = run("dropin_session", 300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stderr + r.stdout[-2000:]
    assert "PASS" in r.stdout and "MIRROR" in r.stdout


def test_io_regression_cpu_checks():
    """SEG-Y gather round trip and .s3dm inspection run everywhere; the GPU part may skip with 77."""
    r = run("seis3d_io_regression", 300)
    assert r.returncode in (0, 77), r.stderr + r.stdout[-2000:]
    assert "GATHER 150 receivers" in r.stdout and "MODEL 40x32x24" in r.stdout


@pytest.mark.gpu
def test_io_regression_gpu(gpu):
    r = run("seis3d_io_regression", 300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stderr + r.stdout[-2000:]
    assert "PASS" in r.stdout


def test_2d_record_regression():
    """Host-only: the 2D record and shot-package checks must pass everywhere."""
    r = run("seis2d_record_regression", 300)
    assert r.returncode == 0, r.stderr + r.stdout[-2000:]
    assert "PASS" in r.stdout
