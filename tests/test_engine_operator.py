"""PhysicalGpuHashAggregate (slothdb_amd/host/engine/) compiled against the
reference engine's own PhysicalOperator, DataChunk, Vector, Value and string_t,
through tests/cpp/test_engine_operator.cpp: planning, CapturePredicates and
GpuAggregateSupports on the CPU; the known answers of
test/unit/execution/test_aggregation.cpp, long strings, NULLs, a captured
predicate, multi-chunk output and a TopN hint on the GPU.

The program is built from the reference's sources (oracle/Makefile, target
`engine`) only where the reference tree is present; without it these tests
skip."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "test_engine_operator")


def _run(*args):
    if not os.path.exists(EXE):
        pytest.skip(f"{EXE} not built (needs the reference sources; `make -C oracle engine`)")
    r = subprocess.run([EXE, *args], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=110)
    print(r.stdout)
    assert r.returncode == 0, r.stdout


def test_engine_operator_planning_cpu():
    _run("--cpu")


@pytest.mark.gpu
def test_engine_operator_gpu():
    _run()
