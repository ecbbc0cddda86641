"""GPU parity: the HIP assembly through the C ABI against the CPU oracle.

Bar: pattern / dofs / scatter indices bit-exact (the GPU writes into the harness
pattern, which equals the oracle's insertion pattern); FP64 schedule within 1e-12
relative (Frobenius over the four K blocks, 2-norm of P); mixed schedule within
1e-6 on K (FP32 element tangents) and 1e-12 on P and the damage history (kept FP64).
"""
import numpy as np
import pytest

from pigfem_amd import AssemblerGPU, DeviceBuffer, PMatrix, PVector, RVEProblem, _ffi, config, small_config
from pigfem_amd._ffi import lib
from oracle.oracle import Oracle
from tests.helpers import full_matrix, rel, u_scale

pytestmark = pytest.mark.gpu

# multi*: 120 circles packed 0.05 h apart, so that tri3 elements are cut by two
# or three inclusions (polygon children); tet-*-q: spheres anywhere, so that tet4
# interfaces are CohTri3 and CohQuad4
CASES = [("c0", 24, 3), ("bl", 24, 3), ("op", 24, 3), ("tet", 8, 1), ("tet-op", 6, 3), ("multi", 24, 3),
         ("multi-op", 24, 3), ("multi-bl", 24, 3), ("tet-op-q", 8, 1), ("tet-bl-q", 8, 1)]


def _run(asm, p, U, ic, func=True, jac=True):
    K, P = PMatrix(p), PVector(p)
    it = np.zeros(p.n_cqp)
    asm.committed_isvs[:] = ic
    asm.assemble(K, P, 0.0, it, U, func, jac)
    return K.values, P.values, it


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request):
    kind, n, ni = request.param
    cfg = small_config(kind, n=n, n_incl=ni)
    p = RVEProblem(cfg)
    o = Oracle(p)
    for b in range(4):
        assert np.array_equal(o.rp[b], p.rp[b]) and np.array_equal(o.ci[b], p.ci[b])
    U = p.random_u(7, u_scale(cfg))
    ic = np.random.default_rng(1).uniform(0, 5e-4, p.n_cqp) if p.n_cqp else np.zeros(0)
    Kr, Pr, Ir = o.assemble(U, ic)
    print(f"case {kind} n={n} elems={p.n_elem} cut={p.n_cut}", flush=True)
    return p, U, ic, Kr, Pr, Ir


@pytest.mark.parametrize("mode", ["fp64", "binned64", "mixed", "mixed_cut64"])
def test_parity_vs_oracle(case, mode):
    p, U, ic, Kr, Pr, Ir = case
    asm = AssemblerGPU(p, mode)
    K, P, It = _run(asm, p, U, ic)
    tol = 1e-6 if mode.startswith("mixed") else 1e-12
    assert rel(K, Kr) < tol, rel(K, Kr)
    assert rel(P, Pr) < 1e-12, rel(P, Pr)
    if p.n_cqp:
        assert np.abs(It - Ir).max() <= 1e-12 * max(np.abs(Ir).max(), 1e-30)


@pytest.mark.parametrize("mode", ["fp64", "binned64", "mixed", "mixed_cut64"])
def test_bitwise_repeatable(case, mode):
    p, U, ic, *_ = case
    asm = AssemblerGPU(p, mode)
    a = _run(asm, p, U, ic)
    b = _run(asm, p, U, ic)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("mode", ["fp64", "binned64", "mixed", "mixed_cut64"])
def test_func_only_and_jac_only(case, mode):
    p, U, ic, Kr, Pr, Ir = case
    asm = AssemblerGPU(p, mode)
    tol = 1e-12 if mode in ("fp64", "binned64") else 1e-6
    K, P = PMatrix(p), PVector(p)
    K.values[:] = 7.0
    it = np.zeros(p.n_cqp)
    asm.committed_isvs[:] = ic
    asm.assemble(K, P, 0.0, it, U, True, False)
    assert np.all(K.values == 7.0) and rel(P.values, Pr) < 1e-12
    assert np.abs(it - Ir).max(initial=0.0) <= 1e-12 * max(np.abs(Ir).max(initial=0.0), 1e-30)
    P.values[:] = 3.0
    asm.assemble(K, P, 0.0, it, U, False, True)
    assert np.all(P.values == 3.0) and rel(K.values, Kr) < tol
    assert asm.assemble(K, P, 0.0, it, U, False, False) == 0


def test_c1_full_config_fp64_and_mixed():
    """Config 1 (20k tri3, 16 fibers, perfect bond), one assembly at U ~ U[-1e-3, 1e-3]."""
    p = RVEProblem(config("C1"))
    o = Oracle(p)
    U = p.random_u(0)
    Kr, Pr, _ = o.assemble(U)
    for mode, tol in (("fp64", 1e-12), ("binned64", 1e-12), ("mixed", 1e-6)):
        K, P, _ = _run(AssemblerGPU(p, mode), p, U, np.zeros(0))
        print(mode, "relK", rel(K, Kr), "relP", rel(P, Pr), flush=True)
        assert rel(K, Kr) < tol and rel(P, Pr) < 1e-12


@pytest.mark.parametrize("mode", ["fp64", "binned64", "mixed", "mixed_cut64"])
def test_tangent_keeps_rigid_translations(case, mode):
    """A rigid translation (regular-node dofs equal, enrichment dofs zero) is a
    zero-energy mode of the exact tangent: the parent gradients sum to zero and
    the cohesive fold couples only enrichment nodes.  Every schedule keeps it to
    FP64 rounding on the rows of regular nodes -- the FP32 ones because their
    diagonal blocks are taken as minus the FP64 sum of the row's other (parent)
    blocks; element-wise FP32 rounding alone leaves about 1e-8 of the row scale,
    which costs Newton iterations on large meshes.  Enrichment rows hold it to
    the tangent's own precision."""
    p, U, ic, Kr, Pr, Ir = case
    K, _, _ = _run(AssemblerGPU(p, mode), p, U, ic)
    A = full_matrix(p, K)
    Ar = full_matrix(p, Kr)
    reg = p.node_dof[: p.n_nodes].ravel()
    enr = p.node_dof[p.n_nodes:].ravel()
    s_reg = abs(Ar[reg]).max()  # entry scale of each row class (cohesive
    s_enr = abs(Ar[enr]).max() if len(enr) else 0.0  # springs sit on enrichment rows)
    Af = full_matrix(p, Kr.astype(np.float32).astype(np.float64))
    for comp in range(p.dim):
        V = np.zeros(p.n_dof)
        V[p.node_dof[: p.n_nodes, comp]] = 1.0
        assert np.abs((Ar @ V)[reg]).max() <= 1e-12 * s_reg  # the premise, on the oracle
        # and the check has teeth: the oracle's tangent rounded to FP32 entry by entry fails it
        assert np.abs((Af @ V)[reg]).max() > 1e-12 * s_reg
        r = A @ V
        assert np.abs(r[reg]).max() <= 1e-12 * s_reg, mode
        if len(enr):
            tol = 1e-12 if mode in ("fp64", "binned64", "mixed_cut64") else 1e-6
            assert np.abs(r[enr]).max() <= tol * s_enr, mode


def test_unknown_flag_bits_are_rejected():
    p = RVEProblem(small_config("c0", n=8, n_incl=1))
    asm = AssemblerGPU(p, "mixed")
    dU = DeviceBuffer.from_numpy(p.random_u(0))
    dK, dP = DeviceBuffer.zeros(p.nnz_total), DeviceBuffer.zeros(p.n_dof)
    rc = lib().pg_asm_assemble_device(asm._h, dU.ptr, None, None, 0.0,
                                      _ffi.PG_FUNC | _ffi.PG_JAC | (1 << 31), dK.ptr, dP.ptr, None)
    assert rc == _ffi.PG_E_ARG


def test_device_api_matches_host_api():
    from pigfem_amd import DeviceBuffer, synchronize
    p = RVEProblem(small_config("op", n=24, n_incl=3))
    U = p.random_u(4)
    ic = np.zeros(p.n_cqp)
    asm = AssemblerGPU(p, "mixed")
    Kh, Ph, Ih = _run(asm, p, U, ic)
    dU = DeviceBuffer.from_numpy(U)
    dic = DeviceBuffer.zeros(p.n_cqp)
    dit = DeviceBuffer.zeros(p.n_cqp)
    dK = DeviceBuffer.zeros(p.nnz_total)
    dP = DeviceBuffer.zeros(p.n_dof)
    asm.assemble_device(dU, dic, dit, dK, dP)
    synchronize()
    assert np.array_equal(dK.download(p.nnz_total), Kh) and np.array_equal(dP.download(p.n_dof), Ph)
    assert np.array_equal(dit.download(p.n_cqp), Ih)


@pytest.mark.parametrize("kind,n,ni", [("op", 24, 3), ("tet", 8, 1), ("tet-op", 6, 3), ("multi-op", 24, 3),
                                      ("tet-op-q", 8, 1)])
def test_every_schedule_variant(kind, n, ni):
    """All node-group sizes, group-kernel forms (and the global-memory fallback
    that groups overflowing the LDS tables take) and cut-kernel variants give
    the same result, bit for bit (same arithmetic, same summation order).  The
    one exception: with FP32 tangents the structured tet4 nodes read their
    element blocks from the grid's table (rounded once from exact values), so
    the variants that take that kernel form their own bit-identical class, whose
    K agrees with the others' to FP32 rounding and whose P is bit-identical."""
    p = RVEProblem(small_config(kind, n=n, n_incl=ni))
    o = Oracle(p)
    U = p.random_u(9)
    ic = np.zeros(p.n_cqp)
    Kr, Pr, _ = o.assemble(U, ic)
    groups = (8, 16, 32) if p.dim == 2 else (2, 4, 8)
    for mode in ("binned64", "mixed", "mixed_cut64"):
        # bit 3: every group on the global-memory fallback path; bit 4: enrichment groups with TV = 1;
        # bits 5-6: another group size for the groups after the cut kernels; bit 7: the nodes of the
        # structured grid through the generic kernel; bits 8-9: consecutive groups per block of the structured kernel;
        # bit 10: FP32 structured tet4 nodes through the kernel that evaluates each incidence's blocks
        variants = ((*range(8), 8, 11, 16, 18, 32, 64, 96, 97, 100, 128, 129, 130, 132, 133, 228, 256, 516, 768, 897) if p.dim == 2
                    else (0, 1, 4, 5, 8, 13, 16, 20, 32, 64, 97, 101, 128, 133, 256, 517, 768, 1024, 1541))
        refs = {}
        if kind in ("op", "multi-op", "tet"):
            assert AssemblerGPU(p, mode).grid_nodes() > (0.6 if kind == "op" else 0.05) * p.n_nodes
        for g, c, v in [(g, c, v) for g in groups for c in range(4) for v in variants]:
            asm = AssemblerGPU(p, mode)
            asm.set_tiles(g, c, v)
            K, P, _ = _run(asm, p, U, ic)
            assert rel(K, Kr) < (1e-12 if mode == "binned64" else 1e-6)
            assert rel(P, Pr) < 1e-12
            tab = p.dim == 3 and mode != "binned64" and not v & (8 | 128 | 1024)
            if tab not in refs:
                refs[tab] = (K, P)
            else:
                assert np.array_equal(refs[tab][0], K) and np.array_equal(refs[tab][1], P), (g, c, v)
        if len(refs) == 2:
            assert rel(refs[True][0], refs[False][0]) < 1e-7
            assert np.array_equal(refs[True][1], refs[False][1])
