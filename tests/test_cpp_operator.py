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
