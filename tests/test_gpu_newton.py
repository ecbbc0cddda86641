"""Newton histories with the GPU assembly against the oracle's.

Identical iteration counts are asserted on load paths whose Newton history is
robust: the oracle's own history is unchanged when its tangent is perturbed by
random relative noise of 1e-6 (checked on the CPU in tests/test_newton.py).  On
a path where the oracle needs 20 of its 25 allowed iterations in one step, a
1e-10 relative perturbation of K already changes the history, so only the
FP64 drop-in (same arithmetic and summation order as the reference) can repeat
it; the other schedules are reported there, not asserted.
"""
import dataclasses

import numpy as np
import pytest

from pigfem_amd import AssemblerGPU, RVEProblem, config, small_config
from pigfem_amd.newton import NewtonDriver, gpu_assemble_fn
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _problem(kind, strain):
    cfg = dataclasses.replace(small_config(kind, n=30, n_incl=3, seed=4), strain=strain)
    return RVEProblem(cfg)


def _histories(p, modes):
    o = Oracle(p)
    runs = {}
    for name in ("oracle",) + modes:
        fn = (lambda U, isv: o.assemble(U, isv)) if name == "oracle" else gpu_assemble_fn(AssemblerGPU(p, name))
        d = NewtonDriver(p, fn)
        d.solve()
        runs[name] = d
        print(name, d.iteration_counts(), flush=True)
    return runs


@pytest.mark.parametrize("kind,strain", [("op", 0.01), ("op", 0.005), ("bl", 0.005)])
def test_newton_iteration_counts_identical(kind, strain):
    p = _problem(kind, strain)
    runs = _histories(p, ("fp64", "binned64", "mixed", "mixed_cut64"))
    ref = runs["oracle"].iteration_counts()
    nU = np.linalg.norm(runs["oracle"].U)
    for mode, tol in (("fp64", 1e-10), ("binned64", 1e-10), ("mixed", 1e-6), ("mixed_cut64", 1e-6)):
        assert runs[mode].iteration_counts() == ref, mode
        assert np.linalg.norm(runs[mode].U - runs["oracle"].U) <= tol * nU, mode
        assert np.abs(runs[mode].isv - runs["oracle"].isv).max() <= tol * max(np.abs(runs["oracle"].isv).max(), 1e-30)


def test_fp64_repeats_a_sensitive_newton_history():
    """OP at 2 % strain: one step converges in 20 iterations on the oracle and
    fails (25) under a 1e-10 perturbation of K; the FP64 schedule repeats it."""
    p = _problem("op", 0.02)
    runs = _histories(p, ("fp64", "mixed"))
    assert runs["fp64"].iteration_counts() == runs["oracle"].iteration_counts()


def test_newton_c1_ortiz_pandolfi_to_two_percent():
    """The 100 x 100 RVE with 16 fibers and Ortiz-Pandolfi interfaces, 10 load
    steps to 2 % strain: identical iteration counts for every schedule."""
    p = RVEProblem(dataclasses.replace(config("C1"), cohesive=2, name="C1-OP"))
    runs = _histories(p, ("fp64", "binned64", "mixed", "mixed_cut64"))
    ref = runs["oracle"].iteration_counts()
    for mode in ("fp64", "binned64", "mixed", "mixed_cut64"):
        assert runs[mode].iteration_counts() == ref, mode
