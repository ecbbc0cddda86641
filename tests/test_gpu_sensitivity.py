"""GPU parity of the material-parameter sensitivity dP/dd (pg_asm_sensitivity_dPdd)
against the oracle, whose restatement test_sensitivity.py pins: LEI parameters
(E, nu, lambda, mu, K) with the reference's child-local B and with each child's
own B, Ortiz-Pandolfi sigma_c and delta_c with random committed delta_max and
history sensitivities, OP_NU; 1e-12 relative.  Meshes with a bilinear cohesive
law are refused, as the reference stops there."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from pigfem_amd import AssemblerGPU, PgError, RVEProblem, _ffi, small_config
from pigfem_amd.sensitivity import SensitivityMaterialParameterGPU, lei_dmat_sensitivity
from tests.helpers import rel

pytestmark = pytest.mark.gpu

CASES = [("c0", 24, 3), ("op", 24, 3), ("tet", 8, 1), ("tet-op", 6, 3), ("multi-op", 24, 3)]


@pytest.mark.parametrize("kind,n,ni", CASES, ids=[c[0] for c in CASES])
def test_sensitivity_parity_vs_oracle(kind, n, ni):
    cfg = small_config(kind, n=n, n_incl=ni)
    p = RVEProblem(cfg)
    o = Oracle(p)
    asm = AssemblerGPU(p, "fp64")
    U = p.random_u(3)
    rng = np.random.default_rng(4)
    isv = np.where(rng.uniform(size=p.n_cqp) < 0.5, 0.0, 5e-3) if p.n_cqp else None
    isvs = rng.uniform(-1, 1, p.n_cqp) if p.n_cqp else None
    print(f"sensitivity case {kind} n={n} elems={p.n_elem} cut={p.n_cut}", flush=True)
    params = [(0, "E"), (0, "nu"), (1, "lambda"), (1, "shear modulus"), (0, "bulk modulus")]
    if p.n_cqp:
        params += [(2, "sigma_c"), (2, "delta_c")]
    for mat, name in params:
        E, nu = (cfg.E_matrix, cfg.nu_matrix) if mat == 0 else (cfg.E_fiber, cfg.nu_fiber)
        for own in (False, True):
            sp = SensitivityMaterialParameterGPU(asm, mat, name, E=E, nu=nu, own_child_b=own)
            dP = sp.assemble_dPdd(U, isv, isvs)
            ref = o.sens_dPdd(sp.prm, U, isv, isvs, sp.flags)
            assert np.abs(ref).max() > 0
            assert rel(dP, ref) < 1e-12, (mat, name, own, rel(dP, ref))
    # repeatable, and an assembly between two calls changes nothing
    sp = SensitivityMaterialParameterGPU(asm, 0, "E", E=cfg.E_matrix, nu=cfg.nu_matrix)
    a = sp.assemble_dPdd(U, isv, isvs)
    from pigfem_amd import PMatrix, PVector
    K, P = PMatrix(p), PVector(p)
    asm.assemble(K, P, 0.0, np.zeros(p.n_cqp), U, True, True)
    assert np.array_equal(a, sp.assemble_dPdd(U, isv, isvs))


def test_sensitivity_op_nu_and_refusals():
    cfg = small_config("op", n=24, n_incl=3)
    cfg.cohesive = _ffi.PGH_COH_OP_NU
    p = RVEProblem(cfg)
    o = Oracle(p)
    asm = AssemblerGPU(p, "fp64")
    U = p.random_u(5)
    for name in ("sigma_c", "delta_c"):
        sp = SensitivityMaterialParameterGPU(asm, 2, name)
        assert rel(sp.assemble_dPdd(U), o.sens_dPdd(sp.prm, U)) < 1e-12
    b = RVEProblem(small_config("bl", n=16, n_incl=2))
    ab = AssemblerGPU(b, "fp64")
    sp = SensitivityMaterialParameterGPU(ab, 0, "E", E=3.5e3, nu=0.35)
    with pytest.raises(PgError):
        sp.assemble_dPdd(b.random_u(0))
    with pytest.raises(ValueError):
        SensitivityMaterialParameterGPU(ab, 2, "sigma_c")
    with pytest.raises(ValueError):
        SensitivityMaterialParameterGPU(asm, 0, "thickness", E=3.5e3, nu=0.35)
    # the dD/dd the mirror forms is the reference's
    S = lei_dmat_sensitivity(3.5e3, 0.35, 2, "E")
    assert np.allclose(S * 3.5e3, np.array([[1 - 0.35, 0.35, 0], [0.35, 1 - 0.35, 0], [0, 0, 0.15]])
                       * 3.5e3 / (1.35 * 0.3), rtol=1e-14)
