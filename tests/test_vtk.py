"""VTK output (Writer_VTK + Writer_ASCII_VTK / Writer_Binary_VTK): pins of the
oracle's restatement of the gathered fields and of the file writer.

What fixes the answers:
  * a linear displacement u = A x (enrichment dofs 0) has the strain
    sym(A) everywhere, so every cell's average strain is sym(A) and its average
    stress D_c sym(A); the cell volumes sum to the domain's (1 in 2D, 6 in 3D:
    the tet4 weights sum to 1);
  * an enrichment node's plotted displacement is the field's value there: A x
    for the linear field, and the kinked field's own value when its enrichment
    dofs carry the kink (src/Problem.cpp:538-570);
  * the files parse back to the arrays written (counts, connectivity, types,
    big-endian binary), with the reference's float conversions.
"""
import numpy as np
import pytest

from oracle.oracle import plot_fields
from pigfem_amd import RVEProblem, small_config
from pigfem_amd.vtk import cell_averages, von_mises, vtk_cells, write_vtk


def linear_u(p, A):
    U = np.zeros(p.n_dof)
    x = p.coords[: p.n_nodes]
    for c in range(p.dim):
        U[p.node_dof[: p.n_nodes, c]] = x @ A[c]
    return U


@pytest.mark.parametrize("kind,n,ni", [("c0", 12, 2), ("op", 16, 3), ("tet", 6, 1)])
def test_linear_field_cells_and_nodes(kind, n, ni):
    p = RVEProblem(small_config(kind, n=n, n_incl=ni))
    rng = np.random.default_rng(1)
    A = rng.uniform(-1e-3, 1e-3, (p.dim, p.dim))
    U = linear_u(p, A)
    nodes, cells = plot_fields(p, U)
    conn, types, mats = vtk_cells(p)
    assert cells.shape[0] == len(conn) == p.n_elem - p.n_cut + sum(p.cut[c].n_child for c in range(p.n_cut))
    vol_total = 1.0 if p.dim == 2 else 6.0
    assert abs(cells[:, 0].sum() - vol_total) < 1e-12
    if p.dim == 2:
        eps = np.array([A[0, 0], A[1, 1], A[0, 1] + A[1, 0]])
    else:
        eps = np.array([A[0, 0], A[1, 1], A[2, 2], A[1, 2] + A[2, 1], A[0, 2] + A[2, 0], A[0, 1] + A[1, 0]])
    ns = len(eps)
    avg = cells[:, 1:1 + ns] / cells[:, :1]
    assert np.abs(avg - eps).max() < 1e-12 * np.abs(eps).max()
    for c, m in enumerate(mats):
        D = np.array(p.desc.mats[m].D[: ns * ns]).reshape(ns, ns)
        s = cells[c, 1 + ns:] / cells[c, 0]
        assert np.abs(s - D @ eps).max() < 1e-11 * np.abs(D @ eps).max()
    x = p.coords
    assert np.abs(nodes - x @ A.T).max() < 1e-15
    # cells list their nodes in the child order, the child's own nodes
    for c0 in conn:
        assert len(c0) >= 3 and max(c0) < p.n_nodes + p.n_enr


def test_enrichment_node_takes_the_kinked_field_value():
    """A field kinked along each inclusion's chords, carried by enrichment dofs:
    the plotted value at an enrichment node is the field's value there."""
    p = RVEProblem(small_config("c0", n=12, n_incl=2))
    rng = np.random.default_rng(3)
    U = np.zeros(p.n_dof)
    x = p.coords
    a = rng.uniform(-1, 1, (2, 2))
    U[p.node_dof[: p.n_nodes].ravel()] = (x[: p.n_nodes] @ a.T).ravel()
    # an enrichment dof adds the kink: the field value there is the parent
    # interpolation along the edge plus the dof
    kink = rng.uniform(-1, 1, (p.n_enr, 2))
    U[p.node_dof[p.n_nodes:].ravel()] = kink.ravel()
    nodes, _ = plot_fields(p, U)
    exp = x[p.n_nodes:] @ a.T + kink
    assert np.abs(nodes[p.n_nodes:] - exp).max() < 1e-14
    # the edge fraction puts the enrichment node on its edge
    for k in range(p.n_enr):
        n1, n2 = p.enr_edge[k]
        xe = x[n1] + p.enr_alpha[k] * (x[n2] - x[n1])
        assert np.abs(xe - x[p.n_nodes + k]).max() < 1e-15


@pytest.mark.parametrize("binary", [False, True])
def test_writer_round_trip(tmp_path, binary):
    p = RVEProblem(small_config("op", n=10, n_incl=2))
    U = p.random_u(2)
    nodes, cells = plot_fields(p, U)
    f = tmp_path / ("out.vtk" if not binary else "outb.vtk")
    write_vtk(str(f), p, nodes, cells, curr_t=0.25, binary=binary)
    conn, types, mats = vtk_cells(p)
    nall, nc = p.n_nodes + p.n_enr, len(conn)
    strain, stress = cell_averages(cells, 3)
    raw = f.read_bytes()
    head = b"# vtk DataFile Version 3.0\nIGFEM 2D simulation: t=0.25\n"
    assert raw.startswith(head)
    if not binary:
        lines = raw.decode().split("\n")
        i = lines.index(f"POINTS {nall} float")
        pts = np.array([[float(v) for v in l.split()] for l in lines[i + 1:i + 1 + nall]])
        assert np.abs(pts[:, :2] - p.coords).max() < 1e-6
        i = lines.index(f"CELLS {nc} {nc + sum(len(c) for c in conn)}")
        for k in range(nc):
            v = [int(t) for t in lines[i + 1 + k].split()]
            assert v[0] == len(conn[k]) and v[1:] == conn[k]
        i = lines.index(f"CELL_TYPES {nc}")
        assert [int(t) for t in lines[i + 1:i + 1 + nc]] == types
        i = lines.index("VECTORS displacement float ")
        d = np.array([[float(v) for v in l.split()] for l in lines[i + 1:i + 1 + nall]])
        assert np.abs(d[:, :2] - nodes).max() <= 1e-5 * np.abs(nodes).max() and np.all(d[:, 2] == 0)
        i = lines.index("SCALARS e12 float")
        e12 = np.array([float(t) for t in lines[i + 2:i + 2 + nc]])
        assert np.abs(e12 - strain[:, 2]).max() <= 1e-5 * np.abs(strain[:, 2]).max()
        i = lines.index("SCALARS von_mises float")
        vm = np.array([float(t) for t in lines[i + 2:i + 2 + nc]])
        s = stress.astype(np.float64)
        vm_ref = np.sqrt(s[:, 0] ** 2 - s[:, 0] * s[:, 1] + s[:, 1] ** 2 + 3 * s[:, 2] ** 2)
        assert np.abs(vm - vm_ref).max() <= 1e-5 * vm_ref.max()
        assert lines[lines.index(f"CELL_DATA {nc}") + 1] == "SCALARS materials int"
    else:
        def after(tag, n, fmt):
            k = raw.index(tag) + len(tag)
            return np.frombuffer(raw[k:k + 4 * n], dtype=fmt)
        pts = after(f"POINTS {nall} float\n".encode(), 3 * nall, ">f4").reshape(nall, 3)
        assert np.array_equal(pts[:, :2], p.coords.astype(np.float32))
        types_b = after(f"\nCELL_TYPES {nc}\n".encode(), nc, ">i4")
        assert list(types_b) == types
        k = raw.index(f"\nCELLS {nc} ".encode())
        k = raw.index(b"\n", k + 1) + 1
        flat = np.frombuffer(raw[k:k + 4 * (nc + sum(len(c) for c in conn))], dtype=">u4")
        assert list(flat) == [v for c in conn for v in [len(c)] + c]
        d = after(b"VECTORS displacement float\n", 3 * nall, ">f4").reshape(nall, 3)
        assert np.array_equal(d[:, :2], nodes.astype(np.float32))
        m = after(b"SCALARS materials int\nLOOKUP_TABLE default\n", nc, ">u4")
        assert list(m) == mats
        vm = after(b"SCALARS von_mises float\nLOOKUP_TABLE default\n", nc, ">f4")
        assert np.array_equal(vm, np.array([von_mises(s) for s in stress], np.float32))


def test_von_mises_closed_forms():
    assert von_mises(np.array([3.0, 0.0, 0.0], np.float32)) == np.float32(3.0)     # uniaxial
    assert von_mises(np.array([0.0, 0.0, 1.0], np.float32)) == np.float32(np.sqrt(3.0))  # pure shear
    s6 = np.array([2.0, 2.0, 2.0, 0.0, 0.0, 0.0], np.float32)
    assert von_mises(s6) == 0.0                                                   # hydrostatic
    s6 = np.array([0, 0, 0, 1.0, 0, 0], np.float32)
    assert von_mises(s6) == np.float32(np.sqrt(3.0))
