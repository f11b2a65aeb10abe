"""The C++20 PhysicalGpuHashAggregate (slothdb_amd/host/) through its own test
program, tests/cpp/test_operator.cpp: result types and error paths on the CPU,
the reference's known answers, GetData chunking, TopN, row limit, filter and
NULL keys on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "slothdb_amd", "lib", "test_operator")


def _run(*args):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is missing: build with `python -m slothdb_amd.build`")
    r = subprocess.run([EXE, *args], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=110)
    print(r.stdout)
    assert r.returncode == 0, r.stdout


def test_cpp_operator_planning_cpu():
    _run("--cpu")


@pytest.mark.gpu
def test_cpp_operator_gpu():
    _run()


def test_fixed_point_double_sums_cpu():
    # the exact double SUM arithmetic of the device header, on the host:
    # round-half-even scaling, order-independent words, one rounding at the
    # end, IEEE specials, precision-hazard detection, NaN images for MIN/MAX
    exe = os.path.join(ROOT, "slothdb_amd", "lib", "test_fixed_point")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is missing: build with `python -m slothdb_amd.build`")
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=110)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
