"""Newton histories with the GPU assembly equal the oracle's (FP64 and mixed)."""
import dataclasses

import numpy as np
import pytest

from pigfem_amd import AssemblerGPU, RVEProblem, small_config
from pigfem_amd.newton import NewtonDriver, gpu_assemble_fn
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,strain", [("op", 0.02), ("bl", 0.005)])
def test_newton_iteration_counts_identical(kind, strain):
    cfg = dataclasses.replace(small_config(kind, n=30, n_incl=3, seed=4), strain=strain)
    p = RVEProblem(cfg)
    o = Oracle(p)
    runs = {}
    for name, fn in (("oracle", lambda U, isv: o.assemble(U, isv)),
                     ("fp64", gpu_assemble_fn(AssemblerGPU(p, "fp64"))),
                     ("mixed", gpu_assemble_fn(AssemblerGPU(p, "mixed")))):
        d = NewtonDriver(p, fn)
        d.solve()
        runs[name] = d
        print(name, d.iteration_counts(), flush=True)
    ref = runs["oracle"].iteration_counts()
    assert runs["fp64"].iteration_counts() == ref
    assert runs["mixed"].iteration_counts() == ref
    nU = np.linalg.norm(runs["oracle"].U)
    assert np.linalg.norm(runs["fp64"].U - runs["oracle"].U) <= 1e-10 * nU
    assert np.linalg.norm(runs["mixed"].U - runs["oracle"].U) <= 1e-6 * nU
