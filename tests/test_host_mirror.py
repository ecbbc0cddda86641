"""The C++ host mirror (include/pgfem_assembler.hpp) driven from a C++ program."""
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_host_mirror")


def _run(mode):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    r = subprocess.run([EXE, mode], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.rstrip().endswith("0 failure(s)")


def test_host_mirror_cpu_side():
    """ISV layout vs the harness's qp counts, flatten / unflatten, copy_data."""
    _run("cpu")


@pytest.mark.gpu
def test_host_mirror_assembles_like_the_oracle():
    """AssemblerGPU::assemble (all schedules) vs the oracle: K blocks, P_int, ISVs."""
    _run("gpu")
