"""Heat conduction (AssemblerThermal, src/AssemblerThermal.cpp): harness and
oracle pins for the scalar field (one dof per node, ProblemLinearThermal).

What fixes the answers independently of the oracle's code:
  * an uncut element's conductance is kappa |e| g_a . g_b (the tri3 weights
    sum to the reference area 1/2, the tet4 weights to 1, so 6 |e| in 3D, the
    reference's tet4 quirk);
  * a constant temperature is a null vector of K, and P = K T (the flux of
    the linear law);
  * a field that is linear on every child and continuous (kinked along the
    interface chord) is reproduced exactly by the IGFEM dofs: T^T K_e T equals
    sum over children kappa_c |c| |grad T_c|^2 (src/Mesh.cpp:1596-1614).
"""
import dataclasses

import numpy as np
import pytest
import scipy.sparse as sp

from oracle.oracle import Oracle
from pigfem_amd import PgError, RVEProblem, _ffi, small_config

from helpers import full_matrix, rel


def thermal(kind="c0", n=12, n_incl=2, **kw):
    return RVEProblem(dataclasses.replace(small_config(kind, n=n, n_incl=n_incl), thermal=True, **kw))


def test_scalar_dofs_and_pattern():
    p = thermal()
    nall = p.n_nodes + p.n_enr
    assert p.nndof == 1 and p.node_dof.shape == (nall, 1)
    assert p.n_free + p.n_pres == nall
    assert np.array_equal(np.sort(p.node_dof.ravel()), np.arange(nall))
    assert p.n_pres == 2 * (p.cfg.n + 1)                 # T on LEFT and RIGHT
    x = p.coords[: p.n_nodes, 0]
    pres = p.node_dof[: p.n_nodes, 0] >= p.n_free
    assert np.array_equal(pres, (x == 0.0) | (x == 1.0))
    o = Oracle(p)
    for b in range(4):
        assert np.array_equal(o.rp[b], p.rp[b]) and np.array_equal(o.ci[b], p.ci[b])
    assert p.npat_map.shape == (p.n_pairs,)
    assert p.desc.mats[0].type == _ffi.PG_MAT_LT
    with pytest.raises(PgError):
        RVEProblem(dataclasses.replace(small_config("op", n=12, n_incl=2), thermal=True))


def _simplex_area(X):
    d = X.shape[1]
    M = np.hstack([np.ones((d + 1, 1)), X])
    return abs(np.linalg.det(M)) / (2.0 if d == 2 else 6.0)


def _grads(X):
    d = X.shape[1]
    M = np.hstack([np.ones((d + 1, 1)), X])
    return np.linalg.inv(M)[1:].T        # row a: grad N_a


@pytest.mark.parametrize("dim", [2, 3])
def test_uncut_conductance_closed_form(dim):
    p = thermal("c0" if dim == 2 else "tet", n=6 if dim == 2 else 3, n_incl=0, kappa_matrix=0.7)
    assert p.n_cut == 0
    o = Oracle(p)
    K, _, _ = o.assemble(np.zeros(p.n_dof))
    rows, cols, vals = [], [], []
    for e in range(p.n_elem):
        nodes = p.conn[e]
        X = p.coords[nodes]
        G = _grads(X)
        vol = _simplex_area(X) * (1.0 if dim == 2 else 6.0)   # tet4 weights sum to 1 (Q3)
        Ke = 0.7 * vol * G @ G.T
        d = p.node_dof[nodes, 0]
        rows += list(np.repeat(d, dim + 1))
        cols += list(np.tile(d, dim + 1))
        vals += list(Ke.ravel())
    ref = sp.csr_matrix((vals, (rows, cols)), shape=(p.n_dof, p.n_dof)).toarray()
    A = full_matrix(p, K).toarray()
    assert np.abs(A - ref).max() < 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("kind,n,ni", [("c0", 16, 3), ("tet", 6, 1), ("multi", 24, 3)])
def test_constant_null_space_and_P_equals_KT(kind, n, ni):
    p = thermal(kind, n=n, n_incl=ni)
    o = Oracle(p)
    one = np.zeros(p.n_dof)
    one[p.node_dof[: p.n_nodes, 0]] = 1.0
    K, P1, _ = o.assemble(one)
    A = full_matrix(p, K)
    assert np.abs(A @ one).max() < 1e-12 * np.abs(K).max()
    assert np.abs(P1).max() < 1e-12 * np.abs(K).max()
    T = np.random.default_rng(2).uniform(-1, 1, p.n_dof)
    _, P, _ = o.assemble(T)
    assert rel(P, A @ T) < 1e-13


@pytest.mark.parametrize("kind", ["c0", "tet"])
def test_kinked_patch_test_on_cut_elements(kind):
    """Each cut element alone: T linear on every child, continuous; T^T K_e T =
    sum_c kappa_c |c| |grad T_c|^2."""
    p = thermal(kind, n=12 if kind == "c0" else 6, n_incl=2 if kind == "c0" else 1)
    dim = p.dim
    rng = np.random.default_rng(5)
    kap = [p.desc.mats[0].D[0], p.desc.mats[1].D[0]]
    checked = 0
    for ci in range(0, p.n_cut, max(1, p.n_cut // 10)):
        e = int(p.cut_elem[ci])
        rec = p.cut[ci]
        local = [int(x) for x in p.conn[e]] + [int(rec.enr[k]) for k in range(rec.n_enr)]
        X = p.coords[local]
        nl = len(local)
        if dim == 2:
            a = rng.uniform(-1, 1, 2)
            kinks = []
            for k in range(rec.n_child):
                if rec.child_mat[k] != 1:
                    continue
                cn = _ffi.child_nodes(rec, dim, k)
                en = [l for l in cn if l > dim]
                t = X[en[0]] - X[en[1]]
                nrm = np.array([-t[1], t[0]]) / np.linalg.norm(t)
                if nrm @ (X[cn].mean(axis=0) - X[en[0]]) < 0:
                    nrm = -nrm
                kinks.append((X[en[0]], nrm, rng.uniform(-1, 1)))
            V = np.array([a @ x + sum(c * max(0.0, float(nv @ (x - x0))) for x0, nv, c in kinks) for x in X])
        else:
            V = rng.uniform(-1, 1, nl)
        M = np.hstack([np.ones((dim + 1, 1)), X[:dim + 1]])
        Tl = V.copy()
        for j in range(dim + 1, nl):
            lam = np.linalg.solve(M.T, np.concatenate([[1.0], X[j]]))
            Tl[j] = V[j] - lam @ V[:dim + 1]
        T = np.zeros(p.n_dof)
        for l in range(nl):
            T[p.node_dof[local[l], 0]] = Tl[l]
        energy = 0.0
        for k in range(rec.n_child):
            cn = _ffi.child_nodes(rec, dim, k)
            Xc = X[cn]
            if dim == 2:
                # child area (shoelace) and the gradient of V on it (least squares
                # over its nodes: V is linear on the child)
                x, y = Xc[:, 0], Xc[:, 1]
                area = 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))
                Mc = np.hstack([np.ones((len(cn), 1)), Xc])
                g = np.linalg.lstsq(Mc, V[cn], rcond=None)[0][1:]
            else:
                area = 6.0 * _simplex_area(Xc)
                g = _grads(Xc).T @ V[cn]
            energy += kap[rec.child_mat[k]] * area * (g @ g)
        o = Oracle(p, e, e + 1)
        K, _, _ = o.assemble(T)
        A = full_matrix(p, K, o.rp, o.ci)
        assert abs(T @ (A @ T) - energy) < 1e-12 * energy
        checked += 1
    assert checked >= 2
