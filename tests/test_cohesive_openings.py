"""Cohesive openings output (WriterCohesiveOpenings, src/WriterCohesiveOpenings.cpp:44-197).

CPU: the restatement in oracle/oracle.py is pinned by invariants that a wrong
frame, index or sign breaks (a rigid translation opens nothing; a uniform jump v
of the upper nodes gives openings of norm |v| / delta_c in every edge frame, with
the normal part v . n; each edge is a chord of its own inclusion's circle), and
the writer's sort and format are checked on a hand-made input.  GPU: the device
kernel against the restatement."""

import numpy as np
import pytest

from pigfem_amd import AssemblerGPU, PgError, RVEProblem, small_config, write_cohesive_openings
from oracle.oracle import cohesive_openings


@pytest.fixture(scope="module", params=["op", "multi-op"])
def op_problem(request):
    # multi-op: elements cut by two or three inclusions, whose cohesive edges
    # belong to different inclusions
    return RVEProblem(small_config(request.param, n=24, n_incl=3))


def _upper_nodes(p):
    up = set()
    for c in range(p.n_cut):
        r = p.cut_record(c)
        gl = [int(x) for x in p.conn[p.cut_elem[c]]] + [int(r.enr[k]) for k in range(r.n_enr)]
        for k in range(r.n_coh):
            up.update(gl[r.coh_nodes[k][n + 2]] for n in range(2))
    return sorted(up)


def test_rigid_translation_opens_nothing(op_problem):
    p = op_problem
    U = np.empty(p.n_dof)
    U[p.node_dof[:, 0]] = 3e-4
    U[p.node_dof[:, 1]] = -1e-4
    inc, rows = cohesive_openings(p, U)
    assert rows.shape == (p.n_cqp // 2, 3) and len(rows) > 0
    assert np.abs(rows[:, 1:]).max() < 1e-12


def test_uniform_jump_keeps_its_norm(op_problem):
    p = op_problem
    v = np.array([2e-4, -5e-4])
    U = np.zeros(p.n_dof)
    for a in _upper_nodes(p):
        U[p.node_dof[a]] = v
    inc, rows = cohesive_openings(p, U)
    dc = p.desc.mats[p.cut_record(0).coh_mat[0]].delta_c
    norms = np.hypot(rows[:, 1], rows[:, 2]) * dc
    assert np.allclose(norms, np.linalg.norm(v), rtol=1e-12)
    # the normal opening is v . n with n the unit normal (-t_y, t_x) of the edge
    for c in range(p.n_cut):
        r = p.cut_record(c)
        gl = [int(x) for x in p.conn[p.cut_elem[c]]] + [int(r.enr[k]) for k in range(r.n_enr)]
        for k in range(r.n_coh):
            x0, x1 = p.coords[gl[r.coh_nodes[k][0]]], p.coords[gl[r.coh_nodes[k][1]]]
            t = (x1 - x0) / np.linalg.norm(x1 - x0)
            n = np.array([-t[1], t[0]])
            assert abs(rows[r.isv_off // 2 + k, 2] * dc - v @ n) < 1e-15


def test_midpoints_lie_on_their_inclusion(op_problem):
    p = op_problem
    inc, rows = cohesive_openings(p, np.zeros(p.n_dof))
    assert inc.min() >= 0 and inc.max() < len(p.incl_radii)
    for c in range(p.n_cut):
        r = p.cut_record(c)
        gl = [int(x) for x in p.conn[p.cut_elem[c]]] + [int(r.enr[k]) for k in range(r.n_enr)]
        for k in range(r.n_coh):
            row = r.isv_off // 2 + k
            xm = 0.5 * (p.coords[gl[r.coh_nodes[k][0]]] + p.coords[gl[r.coh_nodes[k][1]]])
            d = xm - p.incl_centers[inc[row]]
            rad = p.incl_radii[inc[row]]
            # a chord of the circle: both ends on it, the midpoint within it
            for n in range(2):
                x = p.coords[gl[r.coh_nodes[k][n]]]
                assert abs(np.linalg.norm(x - p.incl_centers[inc[row]]) - rad) <= 1e-6 * rad
            assert np.linalg.norm(d) <= rad * (1 + 1e-6)
            assert abs(np.arctan2(d[1], d[0]) - rows[row, 0]) < 1e-15


def test_writer_sorts_and_formats(tmp_path):
    inc = np.array([1, 0, 1, 0])
    rows = np.array([[0.5, 1.0, 2.0], [-1.0, 0.25, 3.0], [-0.5, 1e-7, -2.0], [-1.0, 0.125, 4.0]])
    f = tmp_path / "openings.csv"
    write_cohesive_openings(str(f), inc, rows)
    assert f.read_text() == ("0,-1,0.125,4\n0,-1,0.25,3\n1,-0.5,1e-07,-2\n1,0.5,1,2\n")


@pytest.mark.gpu
def test_gpu_openings_match_restatement(op_problem, tmp_path):
    p = op_problem
    U = p.random_u(5, 1e-3)
    inc_r, rows_r = cohesive_openings(p, U)
    asm = AssemblerGPU(p, "mixed")
    inc, rows = asm.cohesive_openings(U)
    print(f"openings: {len(rows)} cohesive edges", flush=True)
    assert np.array_equal(inc, inc_r)
    assert np.abs(rows - rows_r).max() <= 1e-12 * np.abs(rows_r).max()
    write_cohesive_openings(str(tmp_path / "a.csv"), inc, rows)
    assert len((tmp_path / "a.csv").read_text().splitlines()) == len(rows)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["bl", "tet-op"])
def test_gpu_openings_reject_what_the_reference_rejects(kind):
    # the reference stops on laws other than Ortiz-Pandolfi and on 3D meshes
    p = RVEProblem(small_config(kind, n=12 if kind == "bl" else 6, n_incl=2))
    asm = AssemblerGPU(p, "mixed")
    with pytest.raises(PgError):
        asm.cohesive_openings(p.random_u(1))
