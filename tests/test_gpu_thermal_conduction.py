"""GPU parity of the heat-conduction assembler (AssemblerThermal; pg_therm_*)
against the oracle (pinned in test_thermal.py): K within 1e-12, P within 1e-12,
on tri3 meshes cut by one, two or three inclusions (polygon children) and on a
tet4 mesh; func-only / jac-only calls; device API equal to the host API bit for
bit; repeatable."""
import dataclasses

import numpy as np
import pytest

from oracle.oracle import Oracle
from pigfem_amd import AssemblerGPU, AssemblerGPUThermal, DeviceBuffer, PMatrix, PVector, PgError, RVEProblem, \
    small_config, synchronize
from tests.helpers import rel

pytestmark = pytest.mark.gpu

CASES = [("c0", 24, 3), ("tet", 8, 1), ("multi", 24, 3)]


@pytest.mark.parametrize("kind,n,ni", CASES, ids=[c[0] for c in CASES])
def test_conduction_parity_vs_oracle(kind, n, ni):
    p = RVEProblem(dataclasses.replace(small_config(kind, n=n, n_incl=ni), thermal=True))
    o = Oracle(p)
    T = np.random.default_rng(3).uniform(-1, 1, p.n_dof)
    Kr, Pr, _ = o.assemble(T)
    print(f"conduction case {kind} n={n} elems={p.n_elem} cut={p.n_cut}", flush=True)
    asm = AssemblerGPUThermal(p)
    K, P = PMatrix(p), PVector(p)
    asm.assemble(K, P, 0.0, None, T)
    assert rel(K.values, Kr) < 1e-12 and rel(P.values, Pr) < 1e-12
    # repeatable; func-only and jac-only leave the other output alone
    K2, P2 = PMatrix(p), PVector(p)
    K2.values[:] = 5.0
    asm.assemble(K2, P2, 0.0, None, T, True, False)
    assert np.all(K2.values == 5.0) and np.array_equal(P2.values, P.values)
    P2.values[:] = 3.0
    asm.assemble(K2, P2, 0.0, None, T, False, True)
    assert np.all(P2.values == 3.0) and np.array_equal(K2.values, K.values)
    # device API
    dT, dK, dP = DeviceBuffer.from_numpy(T), DeviceBuffer.zeros(p.nnz_total), DeviceBuffer.zeros(p.n_dof)
    asm.assemble_device(dT, dK, dP)
    synchronize()
    assert np.array_equal(dK.download(p.nnz_total), K.values) and np.array_equal(dP.download(p.n_dof), P.values)


def test_conduction_refusals():
    s = RVEProblem(small_config("c0", n=12, n_incl=2))
    with pytest.raises(ValueError):
        AssemblerGPUThermal(s)                      # a structural problem
    t = RVEProblem(dataclasses.replace(small_config("c0", n=12, n_incl=2), thermal=True))
    with pytest.raises(PgError):
        AssemblerGPU(t)                             # the structural assembler refuses a scalar field
