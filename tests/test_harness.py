"""Host set-up: mesh, detection, enrichment, dofs, AIJ pattern and scatter indices.

The pattern, dof map and scatter indices must be bit-exact with the reference's:
the harness pattern is compared with the oracle's own pattern, built the way
PETSc's MatSetValue builds it (insert every (i, j) of every element dof block),
and the cut tables with hand-enumerated cases of src/tri3.cpp:323-356.
"""
import dataclasses

import numpy as np
import pytest

from pigfem_amd import RVEProblem, RVEConfig, config, small_config, _ffi
from oracle.oracle import Oracle


def _same_pattern(p, o):
    return all(np.array_equal(o.rp[b], p.rp[b]) and np.array_equal(o.ci[b], p.ci[b]) for b in range(4))


@pytest.mark.parametrize("kind", ["c0", "bl", "op", "tet"])
def test_pattern_matches_insertion_pattern(kind):
    p = RVEProblem(small_config(kind, n=16 if kind != "tet" else 7, n_incl=3 if kind != "tet" else 1))
    assert _same_pattern(p, Oracle(p))


def test_pattern_matches_insertion_pattern_c1():
    p = RVEProblem(config("C1"))
    assert p.n_elem == 20000 and p.n_nodes == 10201 and p.n_cut > 0
    assert _same_pattern(p, Oracle(p))


def test_node_detection_brute_force():
    p = RVEProblem(config("C1"))
    x = p.coords[: p.n_nodes]
    want = np.full(p.n_nodes, -1)
    for j in range(len(p.incl_radii) - 1, -1, -1):  # last inclusion wins, as detect_node loops
        c, r = p.incl_centers[j], p.incl_radii[j]
        inside = (np.sqrt(((x - c) ** 2).sum(1)) - r) / r < -1e-12
        want[(want == -1) & inside] = j
    assert np.array_equal(p.node_detect, want)


def test_inclusions_respect_gap_and_box():
    cfg = config("C1")
    p = RVEProblem(cfg)
    c, r = p.incl_centers, p.incl_radii
    assert len(r) == cfg.n_incl
    assert np.all(r >= cfg.r_min) and np.all(r <= cfg.r_max)
    assert np.all(c - r[:, None] >= 0) and np.all(c + r[:, None] <= 1)
    d = np.sqrt(((c[:, None] - c[None]) ** 2).sum(-1)) - r[:, None] - r[None]
    np.fill_diagonal(d, np.inf)
    assert d.min() >= cfg.gap_h / cfg.n


@pytest.mark.parametrize("kind", ["c0", "op", "tet"])
def test_enrichment_nodes_on_interfaces(kind):
    cfg = small_config(kind, n=16 if kind != "tet" else 7, n_incl=3 if kind != "tet" else 1)
    p = RVEProblem(cfg)
    coh = cfg.cohesive != _ffi.PGH_COH_NONE
    # every edge whose end nodes differ in detection holds exactly one enrichment node
    edges = set()
    tri = [(0, 1), (1, 2), (2, 0)] if p.dim == 2 else [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
    for e in range(p.n_elem):
        for a, b in tri:
            na, nb = p.conn[e, a], p.conn[e, b]
            if p.node_detect[na] != p.node_detect[nb]:
                edges.add((min(na, nb), max(na, nb)))
    n_init = p.n_enr // 2 if coh else p.n_enr
    assert n_init == len(edges)
    h = 1.0 / cfg.n
    xe = p.coords[p.n_nodes:p.n_nodes + n_init]
    # the intersection lies on some inclusion surface (Steffensen |d alpha| < 1e-7)
    dist = np.abs(np.sqrt(((xe[:, None, :] - p.incl_centers[None]) ** 2).sum(-1)) - p.incl_radii[None]).min(1)
    assert dist.max() < 1e-6 * h
    if coh:  # mirrors sit on the originals
        assert np.array_equal(p.coords[p.n_nodes + n_init:], xe)


def test_hand_enumerated_cut_tables():
    """2x2 tri3 mesh, one circle around the centre node 4: all 8 elements are cut
    with node 4 as the separate (fiber) node.  Expected children from the table of
    Tri3::detection (src/tri3.cpp:323-356) with the separate node inside, i.e. the
    'else' branches, then transformDetectionCohToReg (src/elem.cpp:547-559)."""
    base = RVEConfig(dim=2, n=2, n_incl=1, r_min=0.3, r_max=0.3, gap_h=0, seed=0)
    for coh in (_ffi.PGH_COH_NONE, _ffi.PGH_COH_BL):
        # force the circle to (0.5, 0.5): r = 0.3 and a centre draw in [r, 1-r]; search seeds
        for seed in range(2000):
            p = RVEProblem(dataclasses.replace(base, seed=seed, cohesive=coh))
            if np.allclose(p.incl_centers[0], [0.5, 0.5], atol=0.08):
                break
        assert p.node_detect.tolist() == [-1, -1, -1, -1, 0, -1, -1, -1, -1]
        assert p.n_cut == 8
        # enrichment numbering: std::set order of (lo, hi, inclusion): edges (0,4) (1,4) (2,4) (3,4) (4,5) (4,6) (4,7) (4,8)
        eid = {(0, 4): 0, (1, 4): 1, (2, 4): 2, (3, 4): 3, (4, 5): 4, (4, 6): 5, (4, 7): 6, (4, 8): 7}
        tri_edges = [(0, 1), (1, 2), (2, 0)]
        for ci in range(8):
            e = p.cut_elem[ci]
            r = p.cut_record(ci)
            nodes = list(p.conn[e])
            sep = nodes.index(4)
            cut_edges = {0: (2, 0), 1: (0, 1), 2: (1, 2)}[sep]
            for s, le in enumerate(cut_edges):
                a, b = tri_edges[le]
                ga, gb = nodes[a], nodes[b]
                assert r.enr[s] == p.n_nodes + eid[(min(ga, gb), max(ga, gb))]
                if coh:
                    assert r.enr[s + 2] == p.n_nodes + 8 + eid[(min(ga, gb), max(ga, gb))]
            if coh:
                want = {0: [[0, 6, 5], [1, 2, 3, 4]], 1: [[1, 6, 5], [2, 0, 3, 4]], 2: [[2, 6, 5], [0, 1, 3, 4]]}[sep]
                want_coh = [4, 3, 6, 5]
                assert r.n_coh == 1 and list(r.coh_nodes[0][:4]) == want_coh and r.coh_nn == 4
            else:
                want = {0: [[0, 4, 3], [1, 2, 3, 4]], 1: [[1, 4, 3], [2, 0, 3, 4]], 2: [[2, 4, 3], [0, 1, 3, 4]]}[sep]
                assert r.n_coh == 0
            got = [_ffi.child_nodes(r, 2, k) for k in range(r.n_child)]
            assert got == want
            assert list(r.child_mat[:2]) == [1, 0]      # separate (fiber) child first
            assert r.n_vqp == 8


def test_dof_numbering_free_first():
    p = RVEProblem(config("C1"))
    n = p.cfg.n
    nd = p.node_dof
    x_idx = np.arange(p.n_nodes) % (n + 1)
    pres = np.zeros(nd.shape, bool)
    pres[: p.n_nodes, 0] = (x_idx == 0) | (x_idx == n)
    pres[0, :] = True
    assert np.all((nd >= p.n_free) == pres)
    free = nd[~pres]
    assert np.array_equal(np.sort(free), np.arange(p.n_free))
    # node-major ordering within the free and the prescribed sets
    assert np.all(np.diff(nd[~pres]) > 0) and np.all(np.diff(nd[pres]) > 0)
    assert p.bc_values[nd[n, 0] - p.n_free] == pytest.approx(p.cfg.strain)


@pytest.mark.parametrize("kind", ["bl", "tet"])
def test_scatter_indices_decode_to_their_dofs(kind):
    p = RVEProblem(small_config(kind, n=12 if kind != "tet" else 6, n_incl=2 if kind != "tet" else 1))
    d = p.dim
    off = np.cumsum([0] + p.nnz)
    for q in range(p.n_pairs):
        a, b = p.npat_row[q], p.npat_col[q]
        for i in range(d):
            for j in range(d):
                pos = p.npat_map[(q * d + i) * d + j]
                blk = np.searchsorted(off, pos, side="right") - 1
                r, c = p.node_dof[a, i], p.node_dof[b, j]
                want_blk = (0 if r < p.n_free else 2) + (0 if c < p.n_free else 1)
                assert blk == want_blk
                row = r if r < p.n_free else r - p.n_free
                loc = pos - off[blk]
                assert p.rp[blk][row] <= loc < p.rp[blk][row + 1]
                assert p.ci[blk][loc] == (c if c < p.n_free else c - p.n_free)


def test_incidence_is_ascending_and_complete():
    p = RVEProblem(small_config("op", n=16, n_incl=3))
    for a in range(0, p.n_nodes + p.n_enr, 7):
        lst = p.n2e[p.n2e_ptr[a]:p.n2e_ptr[a + 1]]
        assert np.all(np.diff(lst) > 0)
    # every (element, local node) appears once
    assert p.n2e_ptr[-1] == p.n_elem * 3 + sum(p.cut_record(i).n_enr for i in range(p.n_cut))


def test_rejects_unsupported():
    # spheres packed closer than a tet: elements between two inclusions (the
    # reference's own cases are incomplete, src/tet4.cpp:566-610) are refused
    with pytest.raises(Exception, match="tet4 element cut by"):
        RVEProblem(RVEConfig(dim=3, n=4, n_incl=40, r_min=0.1, r_max=0.15, gap_h=0.01,
                             place=_ffi.PGH_PLACE_BEST_EFFORT))
    # a large sphere cuts tets one node against three (CohTri3) and two against
    # two (CohQuad4: 4 lower + 4 upper nodes, 4 qps)
    p = RVEProblem(RVEConfig(dim=3, n=8, n_incl=1, r_min=0.3, r_max=0.3, cohesive=_ffi.PGH_COH_OP))
    nn = [p.cut_record(i).coh_nn for i in range(p.n_cut)]
    assert 6 in nn and 8 in nn
    assert p.n_cqp == 4 * sum(p.cut_record(i).n_coh for i in range(p.n_cut))
    # node-centred spheres cut one node against three only: CohTri3 interfaces
    p = RVEProblem(small_config("tet-op", n=6, n_incl=2))
    assert p.n_cut > 0 and p.n_cqp == 4 * sum(p.cut_record(i).n_coh for i in range(p.n_cut))
    assert all(p.cut_record(i).coh_nn == 6 for i in range(p.n_cut))


@pytest.mark.parametrize("kind", ["op", "tet"])
def test_build_pattern_from_external_mesh(kind):
    """pgh_build_pattern (the integration path: mesh set up by PIGFEM) rebuilds
    exactly the pattern, scatter indices and incidence of the generator."""
    import ctypes as C
    from pigfem_amd import _ffi
    p = RVEProblem(small_config(kind, n=24 if kind == "op" else 8, n_incl=3 if kind == "op" else 1))
    m = _ffi.PgProblemDesc()
    C.memmove(C.addressof(m), C.addressof(p.desc), C.sizeof(m))
    for f in ("npat_ptr", "npat_col", "npat_row", "npat_map", "n2e_ptr", "n2e"):
        setattr(m, f, None)
    for b in range(4):
        m.rp[b] = None
        m.ci[b] = None
    h = C.c_void_p()
    L = _ffi.lib()
    _ffi.check(L.pgh_build_pattern(C.byref(m), C.byref(h)), "pgh_build_pattern")
    try:
        q = _ffi.PgProblemDesc()
        _ffi.check(L.pgh_get_desc(h, C.byref(q)), "pgh_get_desc")
        assert [q.nnz[b] for b in range(4)] == p.nnz and q.n_pairs == p.n_pairs
        rows = [p.n_free, p.n_free, p.n_pres, p.n_pres]
        for b in range(4):
            assert np.array_equal(np.ctypeslib.as_array(q.rp[b], (rows[b] + 1,)), p.rp[b])
            if p.nnz[b]:
                assert np.array_equal(np.ctypeslib.as_array(q.ci[b], (p.nnz[b],)), p.ci[b])
        nall = p.n_nodes + p.n_enr
        assert np.array_equal(np.ctypeslib.as_array(q.npat_map, (p.npat_map.size,)), p.npat_map)
        assert np.array_equal(np.ctypeslib.as_array(q.npat_col, (p.n_pairs,)), p.npat_col)
        assert np.array_equal(np.ctypeslib.as_array(q.n2e_ptr, (nall + 1,)), p.n2e_ptr)
        assert np.array_equal(np.ctypeslib.as_array(q.n2e, (p.n2e.size,)), p.n2e)
        # a broken mesh is rejected, not read out of range
        bad = _ffi.PgProblemDesc()
        C.memmove(C.addressof(bad), C.addressof(m), C.sizeof(m))
        bad.n_free += 1
        h2 = C.c_void_p()
        assert L.pgh_build_pattern(C.byref(bad), C.byref(h2)) != 0
    finally:
        L.pgh_free(h)


def test_pattern_flags_and_stage_times():
    """Unknown flag bits and a dof numbered twice are rejected; the stage
    timer reports every stage of a build."""
    import ctypes as C
    from pigfem_amd import _ffi
    from pigfem_amd.problem import RVEConfig
    with pytest.raises(Exception, match="unknown flag"):
        RVEProblem(RVEConfig(dim=2, n=8, n_incl=1, r_min=0.1, r_max=0.1, place=64))
    p = RVEProblem(small_config("op", n=24, n_incl=3))
    st = p.stage_seconds()
    assert set(st) == set(RVEProblem.STAGES) and st["pattern"] > 0 and st["instances"] == 0
    L = _ffi.lib()
    m = _ffi.PgProblemDesc()
    C.memmove(C.addressof(m), C.addressof(p.desc), C.sizeof(m))
    h = C.c_void_p()
    assert L.pgh_build_pattern_flags(C.byref(m), 8, C.byref(h)) == _ffi.PG_E_ARG
    dup = np.array(p.node_dof, copy=True)
    dup[1, 0] = dup[0, 0]
    m.node_dof = dup.ctypes.data_as(C.POINTER(C.c_int32))
    assert L.pgh_build_pattern(C.byref(m), C.byref(h)) == _ffi.PG_E_ARG
    assert b"twice" in L.pgh_last_error()


@pytest.mark.parametrize("kind", ["multi", "multi-op"])
def test_multi_inclusion_cut_topology(kind):
    """Tri3 elements cut by two or three inclusions (src/tri3.cpp:370-444): the
    children are counter-clockwise, tile the parent exactly (areas sum to it),
    polygons of 5 and 6 nodes appear, an inclusion child's parent nodes detect
    its inclusion and the matrix child's none, every enrichment node lies on the
    circle of its own inclusion, and each cohesive edge joins two enrichment
    nodes of one inclusion to their mirrors."""
    from pigfem_amd import _ffi

    def cross(a, b):
        return a[0] * b[1] - a[1] * b[0]
    p = RVEProblem(small_config(kind))
    seen = set()
    multi = 0
    for c in range(p.n_cut):
        r = p.cut_record(c)
        e = int(p.cut_elem[c])
        local = [int(x) for x in p.conn[e]] + [int(r.enr[k]) for k in range(r.n_enr)]
        X = p.coords[local]
        parea = 0.5 * cross(X[1] - X[0], X[2] - X[0])
        tot = 0.0
        for k in range(r.n_child):
            cn = _ffi.child_nodes(r, 2, k)
            seen.add(len(cn))
            P = X[cn]
            a = 0.5 * sum(cross(P[i], P[(i + 1) % len(P)]) for i in range(len(P)))
            assert a > 0
            tot += a
            det = {int(p.node_detect[local[l]]) for l in cn if l < 3}
            if r.child_mat[k] == 0:
                assert det <= {-1}
            else:
                assert len(det) == 1 and det.pop() >= 0
        assert abs(tot - parea) <= 1e-12 * parea
        for k in range(r.n_enr):
            g = local[3 + k]
            inc = int(p.enr_incl[g - p.n_nodes])
            assert abs(np.linalg.norm(p.coords[g] - p.incl_centers[inc]) - p.incl_radii[inc]) <= 1e-6 * p.incl_radii[inc]
        if r.n_child > 2:
            multi += 1
        for k in range(r.n_coh):
            lo = [local[r.coh_nodes[k][n]] for n in range(2)]
            up = [local[r.coh_nodes[k][n + 2]] for n in range(2)]
            assert len({int(p.enr_incl[g - p.n_nodes]) for g in lo + up}) == 1
            assert all(np.array_equal(p.coords[a], p.coords[b]) for a, b in zip(lo, up))
    assert multi > 50 and {3, 4, 5, 6} <= seen
