"""GPU parity of the VTK writers' fields (pg_asm_plot_fields) against the
oracle's restatement (pinned in test_vtk.py): the nodal displacements with the
enrichment-node interpolation, and per cell the volume, strain and stress
integrals, within 1e-12; the file written from GPU data equals the file written
from the oracle's fields."""
import numpy as np
import pytest

from oracle.oracle import plot_fields
from pigfem_amd import AssemblerGPU, RVEProblem, small_config
from pigfem_amd.vtk import write_vtk
from tests.helpers import rel, u_scale

pytestmark = pytest.mark.gpu

CASES = [("c0", 24, 3), ("op", 24, 3), ("tet", 8, 1), ("tet-op", 6, 3), ("multi-op", 24, 3)]


@pytest.mark.parametrize("kind,n,ni", CASES, ids=[c[0] for c in CASES])
def test_plot_fields_parity(kind, n, ni, tmp_path):
    cfg = small_config(kind, n=n, n_incl=ni)
    p = RVEProblem(cfg)
    U = p.random_u(4, u_scale(cfg))
    print(f"vtk case {kind} n={n} elems={p.n_elem} cut={p.n_cut}", flush=True)
    nodes_r, cells_r = plot_fields(p, U)
    asm = AssemblerGPU(p, "mixed")
    nodes, cells = asm.plot_fields(U)
    assert nodes.shape == nodes_r.shape and cells.shape == cells_r.shape
    assert rel(nodes, nodes_r) < 1e-14
    for j in range(cells.shape[1]):
        assert rel(cells[:, j], cells_r[:, j]) < 1e-12, j
    # a second call (the cell list is built once) gives the same rows
    n2, c2 = asm.plot_fields(U)
    assert np.array_equal(n2, nodes) and np.array_equal(c2, cells)
    fa, fb = tmp_path / "gpu.vtk", tmp_path / "ref.vtk"
    write_vtk(str(fa), p, nodes, cells, binary=(kind == "op"))
    write_vtk(str(fb), p, nodes_r, cells_r, binary=(kind == "op"))
    a, b = fa.read_bytes(), fb.read_bytes()
    assert len(a) == len(b) or kind != "op"
    if kind != "op":
        la, lb = a.decode().split("\n"), b.decode().split("\n")
        assert len(la) == len(lb)
        for x, y in zip(la, lb):
            if x != y:   # the same number but for the last printed digit
                fx, fy = [float(t) for t in x.split()], [float(t) for t in y.split()]
                assert np.allclose(fx, fy, rtol=1e-5, atol=0)
