"""GPU assembly of a batch of independent RVEs in one launch sequence: parity
with the oracle on the merged problem, and each instance's slice equal, bit for
bit, to the instance assembled on its own by the same schedule."""
import numpy as np
import pytest

from pigfem_amd import AssemblerGPU, PMatrix, PVector, RVEProblem
from oracle.oracle import Oracle
from tests.helpers import rel
from tests.test_batch import _instance_u, small_batch

pytestmark = pytest.mark.gpu


def _run(asm, p, U, ic):
    K, P = PMatrix(p), PVector(p)
    it = np.zeros(p.n_cqp)
    asm.committed_isvs[:] = ic
    asm.assemble(K, P, 0.0, it, U, True, True)
    return K.values, P.values, it


@pytest.mark.parametrize("mode", ["fp64", "binned64", "mixed"])
def test_batch_parity_and_instance_slices(mode):
    cfgs = small_batch()
    B = RVEProblem.batch(cfgs)
    singles = [RVEProblem(c) for c in cfgs]
    U, Us = _instance_u(B, singles)
    ic = np.random.default_rng(4).uniform(0, 5e-5, B.n_cqp)
    Kr, Pr, Ir = Oracle(B).assemble(U, ic)
    K, P, It = _run(AssemblerGPU(B, mode), B, U, ic)
    print(mode, "relK", rel(K, Kr), "relP", rel(P, Pr), flush=True)
    assert rel(K, Kr) < (1e-6 if mode == "mixed" else 1e-12) and rel(P, Pr) < 1e-12
    assert np.abs(It - Ir).max() <= 1e-12 * np.abs(Ir).max()
    for i, p in enumerate(singles):
        o, q = B.inst_off[i], B.inst_off[i + 1]
        Ki, Pi, Ii = _run(AssemblerGPU(p, mode), p, Us[i], ic[o[5]: q[5]])
        kr, pr = B.instance_ranges(i)
        assert np.array_equal(np.concatenate([K[a:b] for a, b in kr]), Ki)
        assert np.array_equal(np.concatenate([P[a:b] for a, b in pr]), Pi)
        assert np.array_equal(It[o[5]: q[5]], Ii)


@pytest.mark.parametrize("mode", ["binned64", "mixed"])
def test_batch_structured_path_is_bitwise_the_generic_one(mode):
    """The instances of a batch have different grid strides: each structured
    node carries its own.  With the structured path off (group variant bit 7)
    every node takes the generic kernel; the results agree bit for bit."""
    cfgs = small_batch()
    B = RVEProblem.batch(cfgs)
    U, _ = _instance_u(B, [RVEProblem(c) for c in cfgs])
    ic = np.random.default_rng(4).uniform(0, 5e-5, B.n_cqp)
    on = AssemblerGPU(B, mode)
    assert on.grid_nodes() > 0.5 * B.n_nodes
    off = AssemblerGPU(B, mode)
    off.set_tiles(0, 0, 128)
    a, b = _run(on, B, U, ic), _run(off, B, U, ic)
    print(mode, "grid nodes", on.grid_nodes(), "of", B.n_nodes, flush=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
