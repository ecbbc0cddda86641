"""GPU parity at the edges of the input space: meshes without a cut element
(no enrichment nodes, no cut kernel, no overlap), the smallest meshes (one or
two cells, fewer nodes than one node group), a cohesive law with no cohesive
element in the mesh, and a batch of one instance.  Bar as in the main parity
tests: FP64 schedules 1e-12, mixed 1e-6 on K, 1e-12 on P."""
import numpy as np
import pytest

from pigfem_amd import AssemblerGPU, PMatrix, PVector, RVEConfig, RVEProblem, _ffi, small_config
from oracle.oracle import Oracle
from tests.helpers import rel

pytestmark = pytest.mark.gpu

MODES = ["fp64", "binned64", "mixed", "mixed_cut64"]


def _check(p, seed=3, scale=1e-3):
    o = Oracle(p)
    U = p.random_u(seed, scale)
    ic = np.zeros(p.n_cqp)
    Kr, Pr, Ir = o.assemble(U, ic)
    print(f"edge case: dim={p.dim} elems={p.n_elem} cut={p.n_cut} cqp={p.n_cqp}", flush=True)
    for mode in MODES:
        asm = AssemblerGPU(p, mode)
        K, P = PMatrix(p), PVector(p)
        it = np.zeros(p.n_cqp)
        asm.assemble(K, P, 0.0, it, U, True, True)
        tol = 1e-12 if mode in ("fp64", "binned64") else 1e-6
        assert rel(K.values, Kr) < tol, mode
        assert rel(P.values, Pr) < 1e-12, mode
        assert np.array_equal(it, Ir), mode


@pytest.mark.parametrize("dim,n", [(2, 1), (2, 3), (2, 40), (3, 1), (3, 2), (3, 7)])
def test_mesh_without_cut_elements(dim, n):
    p = RVEProblem(RVEConfig(dim=dim, n=n, n_incl=0, r_min=0.1, r_max=0.1, name=f"plain-{dim}d-{n}"))
    assert p.n_cut == 0 and p.n_enr == 0
    _check(p)


def test_cohesive_law_without_cohesive_elements():
    cfg = RVEConfig(dim=2, n=12, n_incl=0, r_min=0.1, r_max=0.1, cohesive=_ffi.PGH_COH_OP, name="op-no-incl")
    p = RVEProblem(cfg)
    assert p.n_cqp == 0
    _check(p)


def test_single_inclusion_small_mesh():
    p = RVEProblem(small_config("op", n=10, n_incl=1))
    assert p.n_cut > 0
    _check(p, scale=1e-3)


def test_batch_of_one_instance():
    cfg = small_config("bl", n=16, n_incl=2)
    p = RVEProblem.batch([cfg])
    q = RVEProblem(cfg)
    assert p.n_elem == q.n_elem and p.n_dof == q.n_dof
    _check(p, scale=1e-5)


@pytest.mark.parametrize("mode", MODES)
def test_inverted_child_is_reported_by_both_apis(mode):
    """A cut element whose first child lists two nodes swapped has a negative
    child Jacobian: the reference aborts in store_shape_functions
    (src/Mesh.cpp:1561); here the host API returns PG_E_MESH and the device API
    reports it through pg_asm_check_error.  A good problem assembled next on
    the device reports nothing (the flag belongs to one assembly)."""
    from pigfem_amd import DeviceBuffer, PgError
    p = RVEProblem(small_config("c0", n=12, n_incl=1))
    rec = p.desc.cut[0]
    c = rec.child_node  # child 0's nodes lead the list
    c[1], c[2] = c[2], c[1]
    assert p.desc.cut[0].child_node[1] == c[1]
    U = p.random_u(1)
    asm = AssemblerGPU(p, mode)
    K, P = PMatrix(p), PVector(p)
    print(f"inverted child: {mode}", flush=True)
    with pytest.raises(PgError, match=r"\(-4\)"):
        asm.assemble(K, P, 0.0, np.zeros(0), U, True, True)
    bufs = [DeviceBuffer.from_numpy(U), None, None, DeviceBuffer.zeros(p.nnz_total), DeviceBuffer.zeros(p.n_dof)]
    asm.assemble_device(*bufs)
    with pytest.raises(PgError, match=r"\(-4\)"):
        asm.check_error()
    q = RVEProblem(small_config("c0", n=12, n_incl=1))
    good = AssemblerGPU(q, mode)
    good.assemble_device(*bufs)
    good.check_error()
    # and the handle that failed recovers once its flag is cleared by the next
    # assembly of a good description: a second handle on the bad one fails again
    asm.assemble_device(*bufs)
    with pytest.raises(PgError, match=r"\(-4\)"):
        asm.check_error()
