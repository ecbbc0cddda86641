"""GPU parity of the thermal pre-load (AssemblerSmallStrainStructural_ThermalPreLoad).

The HIP assembly with a temperature change, through the C ABI, against the
oracle (whose thermal strain is pinned in test_thermal_preload.py): every
schedule, the node-group kernel forms (LDS, global-memory fallback, list form,
structured grid), one or both materials expanding; K as without the change,
P within 1e-12.  The host mirror's parameter and load-step rules follow the
reference (src/AssemblerSmallStrainStructural_ThermalPreLoad.cpp:197-267).
"""
import numpy as np
import pytest

from pigfem_amd import AssemblerGPU, AssemblerGPUThermalPreLoad, PMatrix, PVector, PgError, RVEProblem, small_config
from oracle.oracle import Oracle
from tests.helpers import rel, u_scale

pytestmark = pytest.mark.gpu

CASES = [("c0", 24, 3), ("op", 24, 3), ("tet", 8, 1), ("tet-op", 6, 3), ("multi-op", 24, 3)]
MODES = ["fp64", "binned64", "mixed", "mixed_cut64"]


def _run(asm, p, U, ic):
    K, P = PMatrix(p), PVector(p)
    it = np.zeros(p.n_cqp)
    asm.committed_isvs[:] = ic
    asm.assemble(K, P, 0.0, it, U, True, True)
    return K.values, P.values, it


@pytest.mark.parametrize("kind,n,ni", CASES, ids=[c[0] for c in CASES])
def test_thermal_parity_vs_oracle(kind, n, ni):
    cfg = small_config(kind, n=n, n_incl=ni)
    p = RVEProblem(cfg)
    p.desc.mats[0].thermal_exp, p.desc.mats[0].thermal_exp_set = 2.5e-5, 1   # matrix
    p.desc.mats[1].thermal_exp, p.desc.mats[1].thermal_exp_set = 6.0e-6, 1   # fiber
    U = p.random_u(5, u_scale(cfg))
    ic = np.random.default_rng(2).uniform(0, 5e-4, p.n_cqp) if p.n_cqp else np.zeros(0)
    o = Oracle(p)
    dT = -120.0
    Kr, Pr, Ir = o.assemble(U, ic, temp_change=dT)
    K0, P0, _ = o.assemble(U, ic)
    assert rel(P0, Pr) > 1e-3          # the change matters at this U
    print(f"thermal case {kind} n={n} elems={p.n_elem} cut={p.n_cut}", flush=True)
    variants = (0, 8, 128) if p.dim == 3 else (0, 2, 8, 128)
    for mode in MODES:
        for v in variants if mode != "fp64" else (0,):
            asm = AssemblerGPU(p, mode)
            asm.set_tiles(0, 0, v)
            asm.set_temperature_change(dT)
            K, P, It = _run(asm, p, U, ic)
            tol = 1e-6 if mode.startswith("mixed") else 1e-12
            assert rel(K, Kr) < tol, (mode, v, rel(K, Kr))
            assert rel(P, Pr) < 1e-12, (mode, v, rel(P, Pr))
            if p.n_cqp:
                assert np.abs(It - Ir).max() <= 1e-12 * max(np.abs(Ir).max(), 1e-30)
            # back to no change: the plain structural result
            asm.set_temperature_change(0.0)
            K, P, _ = _run(asm, p, U, ic)
            assert rel(P, P0) < 1e-12 and rel(K, K0) < tol


def test_thermal_preload_parameter_rules():
    """set_parameter / get_parameter / assemble_new_load_step of the reference:
    the current change is delta_T * time fraction until relaxation completes,
    then delta_T; setting delta_T resets the current change to 0."""
    p = RVEProblem(small_config("c0", n=16, n_incl=2))
    p.desc.mats[0].thermal_exp, p.desc.mats[0].thermal_exp_set = 1e-5, 1
    o = Oracle(p)
    U = p.random_u(1)
    asm = AssemblerGPUThermalPreLoad(p, "fp64")
    with pytest.raises(ValueError):
        asm.assemble_new_load_step(0.5)          # no temperature set yet
    with pytest.raises(ValueError):
        asm.set_parameter("bogus", 1.0)
    asm.set_parameter("delta_t", 200.0)
    assert asm.get_parameter("DELTA_T") == 200.0 and asm.get_parameter("current_temp") == 0.0
    for frac, expect in ((0.25, 50.0), (1.0, 200.0)):
        asm.assemble_new_load_step(frac)
        assert asm.get_parameter("CURRENT TEMP") == expect
        _, P, _ = _run(asm, p, U, np.zeros(0))
        _, Pr, _ = o.assemble(U, temp_change=expect)
        assert rel(P, Pr) < 1e-12
    asm.set_parameter("COMPLETE", 1.0)
    asm.assemble_new_load_step(0.1)
    assert asm.get_parameter("current_temp") == 200.0 and asm.get_parameter("converged") == 1.0
    asm.clear()
    _, P, _ = _run(asm, p, U, np.zeros(0))
    _, Pr, _ = o.assemble(U)
    assert rel(P, Pr) < 1e-12
    with pytest.raises(PgError):
        asm.set_temperature_change(float("nan"))
