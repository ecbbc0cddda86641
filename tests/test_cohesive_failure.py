"""Cohesive failure lengths (WriterCohesiveFailureLengths,
src/WriterCohesiveFailureLengths.cpp:64-258).

CPU: the restatement in oracle/oracle.py is pinned by closed forms that a wrong
frame, law or clamp breaks: the laws without their own failure measure give 0; a
uniform jump v of the upper nodes gives sqrt((v.n)^2 + beta2 (v.t)^2) / delta_c on
every edge (delta_c1 for the bilinear law); a zero opening gives the clamped
1e-20 * sqrt(1 + beta2) / delta_c; the edge lengths are chords of their
inclusions, so they sum to just below the circumferences.  The writer is checked
on a hand-made input in both of its modes.  GPU: the device kernel against the
restatement."""

import numpy as np
import pytest

from pigfem_amd import AssemblerGPU, PgError, RVEProblem, _ffi, small_config, write_cohesive_failure_lengths
from oracle.oracle import cohesive_failure


def _problem(kind, law=None):
    p = RVEProblem(small_config(kind, n=24 if kind != "tet-op" else 6, n_incl=3 if kind != "tet-op" else 2))
    if law is not None:
        for i in range(p.desc.n_mat):
            if p.desc.mats[i].type != _ffi.PG_MAT_LEI:
                p.desc.mats[i].type = law
                p.desc.mats[i].beta2 = 2.25  # unequal weights tell the normal from the tangent
    return p


def _edges(p):
    for c in range(p.n_cut):
        r = p.cut_record(c)
        gl = [int(x) for x in p.conn[p.cut_elem[c]]] + [int(r.enr[k]) for k in range(r.n_enr)]
        for k in range(r.n_coh):
            yield c, r.isv_off // 2 + k, r.coh_mat[k], [gl[r.coh_nodes[k][n]] for n in range(4)]


def test_base_laws_report_no_failure():
    for kind in ("op", "bl"):
        p = _problem(kind)
        rows = cohesive_failure(p, p.random_u(3, 1e-3))
        assert len(rows) > 0 and np.all(rows[:, 1] == 0.0) and np.all(rows[:, 0] > 0)


@pytest.mark.parametrize("law", [_ffi.PG_MAT_OP_NU, _ffi.PG_MAT_BL_NU])
def test_uniform_jump_closed_form(law):
    p = _problem("op" if law == _ffi.PG_MAT_OP_NU else "bl", law)
    v = np.array([2e-4, -5e-4])
    U = np.zeros(p.n_dof)
    for _, _, _, nodes in _edges(p):
        for a in nodes[2:]:
            U[p.node_dof[a]] = v
    rows = cohesive_failure(p, U)
    for _, row, mi, nodes in _edges(p):
        m = p.desc.mats[mi]
        assert m.beta2 != 1.0
        x0, x1 = p.coords[nodes[0]], p.coords[nodes[1]]
        t = (x1 - x0) / np.linalg.norm(x1 - x0)
        n = np.array([-t[1], t[0]])
        dc = m.delta_c if law == _ffi.PG_MAT_OP_NU else m.delta_c1
        want = np.sqrt((v @ n) ** 2 + m.beta2 * (v @ t) ** 2) / dc
        assert abs(rows[row, 1] - want) <= 1e-12 * want


def test_zero_opening_is_clamped():
    p = _problem("op", _ffi.PG_MAT_OP_NU)
    rows = cohesive_failure(p, np.zeros(p.n_dof))
    for _, row, mi, _ in _edges(p):
        m = p.desc.mats[mi]
        want = 1e-20 * np.sqrt(1 + m.beta2) / m.delta_c
        assert abs(rows[row, 1] - want) <= 1e-12 * want


def test_lengths_are_chords_of_their_inclusions():
    p = _problem("op")
    rows = cohesive_failure(p, np.zeros(p.n_dof))
    incl = p.edge_inclusions()
    tot = np.zeros(len(p.incl_radii))
    for c, row, _, _ in _edges(p):
        tot[incl[row]] += rows[row, 0]
    for i, r in enumerate(p.incl_radii):
        if tot[i] > 0:
            assert 0.9 * 2 * np.pi * r < tot[i] <= 2 * np.pi * r * (1 + 1e-6)


def test_writer_modes(tmp_path):
    rows = np.array([[0.5, 0.2], [0.25, 1.5], [0.125, 2.5], [1e-7, 6.0]])
    f = tmp_path / "a.csv"
    write_cohesive_failure_lengths(str(f), rows, 0)
    write_cohesive_failure_lengths(str(f), rows[:, ::-1] * [1, 0], 1)
    assert f.read_text() == "0.8750000999999999\n0.3750001,0.1250001,1e-07,1e-07,1e-07\n0,0,0,0,0\n"
    g = tmp_path / "b.csv"
    write_cohesive_failure_lengths(str(g), rows, 0, individual=True)
    write_cohesive_failure_lengths(str(g), rows, 3, individual=True)
    assert g.read_text() == "0.5,0.25,0.125,1e-07,\n0.2,1.5,2.5,6,\n0.2,1.5,2.5,6,\n"


@pytest.mark.gpu
@pytest.mark.parametrize("kind,law", [("op", None), ("op", _ffi.PG_MAT_OP_NU), ("bl", _ffi.PG_MAT_BL_NU)])
def test_gpu_failure_matches_restatement(kind, law):
    p = _problem(kind, law)
    U = p.random_u(7, 1e-3)
    want = cohesive_failure(p, U)
    print(f"failure: {kind} law {law}, {len(want)} cohesive edges", flush=True)
    got = AssemblerGPU(p, "mixed").cohesive_failure(U)
    assert np.abs(got[:, 0] - want[:, 0]).max() <= 1e-14 * want[:, 0].max()
    assert np.abs(got[:, 1] - want[:, 1]).max() <= 1e-12 * max(want[:, 1].max(), 1e-300)


@pytest.mark.gpu
def test_gpu_failure_rejects_3d():
    p = _problem("tet-op")
    with pytest.raises(PgError):
        AssemblerGPU(p, "mixed").cohesive_failure(p.random_u(1))
