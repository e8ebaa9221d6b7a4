"""The C++ host surface (include/mini/**): mini::temporal and
mini::spatial::mi355x::FiniteElement, exercised by compiled drivers.

The drivers are built in-tree by __graft_entry__.build() (tests/cpp/Makefile
-> tests/cpp/_build/).  The CPU test builds its driver itself when missing;
the GPU test requires the prebuilt driver, which links libminicfd.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")


def _run(binary):
    r = subprocess.run([binary], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    return r


def test_runge_kutta_bitwise_cpp():
    """RungeKutta<1|2|3> of the C++ surface against closed-form combinations (test/temporal/rk.cpp:21-72)."""
    binary = os.path.join(CPP, "_build", "test_temporal")
    if not os.path.exists(binary):
        subprocess.run(["make", "-C", CPP, os.path.join("_build", "test_temporal")], check=True)
    r = _run(binary)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS") == 18


@pytest.mark.gpu
def test_system_contract_cpp_on_gpu():
    """FiniteElement driven by RungeKutta<3> through the System contract equals the fused device step bit for bit."""
    binary = os.path.join(CPP, "_build", "test_mi355x_system")
    assert os.path.exists(binary), "the C++ drivers are built by __graft_entry__.build()"
    r = _run(binary)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS") == 4
