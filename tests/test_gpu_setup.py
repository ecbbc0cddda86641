"""The set-up on the GPU must equal the host build array for array:
  * the pattern stage (PGH_PATTERN_DEVICE): node incidence, node-pair pattern,
    4-block AIJ pattern and scatter indices built by radix sorts of packed keys;
  * detection and enrichment (PGH_SETUP_DEVICE): node detection, element
    analysis with the detection tables and prism decomposition, the enrichment
    nodes (edge intersections, positions) and the cut records with their ISV
    offsets -- and the same error code and message where the host refuses a
    mesh;
on every workload at full size (the 100 x 100 and 500 x 500 RVEs, the
2000 x 2000 Ortiz-Pandolfi RVE, the 120^3 x 6 tet4 cube, the batch of 256
RVEs) and on the small cases: cohesive tets (CohTri3, CohQuad4), elements cut
by two or three inclusions (polygon children), node-centred inclusions and no
inclusion."""
import ctypes as C
import time

import numpy as np
import pytest

from pigfem_amd import RVEProblem, _ffi, batch_config, config
from pigfem_amd.problem import RVEConfig, small_config

pytestmark = pytest.mark.gpu

ARRAYS = ("npat_ptr", "npat_col", "npat_row", "npat_map", "n2e_ptr", "n2e")


def _same_pattern(a, b):
    assert a.nnz == b.nnz and a.n_pairs == b.n_pairs
    for b_ in range(4):
        assert np.array_equal(a.rp[b_], b.rp[b_]), ("rp", b_)
        assert np.array_equal(a.ci[b_], b.ci[b_]), ("ci", b_)
    for f in ARRAYS:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def _build(name, device):
    if name == "C5":
        return RVEProblem.batch(batch_config(256), device_pattern=device)
    if name.startswith("small-"):
        kind = name[6:]
        if kind == "empty":
            return RVEProblem(RVEConfig(dim=2, n=16, n_incl=0, name="empty"), device_pattern=device)
        return RVEProblem(small_config(kind), device_pattern=device)
    return RVEProblem(config(name), device_pattern=device)


@pytest.mark.parametrize("name", ["small-op", "small-tet", "small-tet-op", "small-empty",
                                  "C1", "C2", "C3", "C4", "C5"])
def test_device_pattern_equals_host(name):
    print(f"device pattern {name}", flush=True)
    t = time.time()
    h = _build(name, False)
    th = time.time() - t
    t = time.time()
    d = _build(name, True)
    td = time.time() - t
    sh, sd = h.stage_seconds(), d.stage_seconds()
    print(f"device pattern {name}: pairs {h.n_pairs} nnz {h.nnz_total}; pattern host {sh['pattern']:.3f}s "
          f"device {sd['pattern']:.3f}s; whole set-up {th:.2f}s / {td:.2f}s", flush=True)
    _same_pattern(h, d)
    h.close()
    d.close()


def test_device_pattern_from_external_mesh():
    """pgh_build_pattern_flags with PGH_PATTERN_DEVICE on a mesh set up
    elsewhere rebuilds the generator's pattern exactly."""
    p = RVEProblem(small_config("op", n=24, n_incl=3))
    m = _ffi.PgProblemDesc()
    C.memmove(C.addressof(m), C.addressof(p.desc), C.sizeof(m))
    for f in ARRAYS:
        setattr(m, f, None)
    for b in range(4):
        m.rp[b] = None
        m.ci[b] = None
    L = _ffi.lib()
    h = C.c_void_p()
    _ffi.check(L.pgh_build_pattern_flags(C.byref(m), _ffi.PGH_PATTERN_DEVICE, C.byref(h)), "pgh_build_pattern_flags")
    try:
        q = _ffi.PgProblemDesc()
        _ffi.check(L.pgh_get_desc(h, C.byref(q)), "pgh_get_desc")
        assert [q.nnz[b] for b in range(4)] == p.nnz and q.n_pairs == p.n_pairs
        assert np.array_equal(np.ctypeslib.as_array(q.npat_map, (p.npat_map.size,)), p.npat_map)
        nall = p.n_nodes + p.n_enr
        assert np.array_equal(np.ctypeslib.as_array(q.n2e_ptr, (nall + 1,)), p.n2e_ptr)
        for b in range(4):
            if p.nnz[b]:
                assert np.array_equal(np.ctypeslib.as_array(q.ci[b], (p.nnz[b],)), p.ci[b])
    finally:
        L.pgh_free(h)


SETUP = ("coords", "conn", "elem_mat", "cut_of_elem", "cut_elem", "node_detect", "enr_incl", "enr_edge", "enr_alpha",
         "node_dof", "bc_values")


def _build_setup(name, device):
    if name == "C5":
        return RVEProblem.batch(batch_config(256), device_setup=device)
    if name.startswith("small-"):
        kind, n, ni, seed = name[6:].split(":")
        if kind == "empty":
            return RVEProblem(RVEConfig(dim=2, n=16, n_incl=0, name="empty"), device_setup=device)
        return RVEProblem(small_config(kind, n=int(n), n_incl=int(ni), seed=int(seed)), device_setup=device)
    if name == "big-sphere":  # one node against three and two against two (CohTri3 and CohQuad4)
        return RVEProblem(RVEConfig(dim=3, n=8, n_incl=1, r_min=0.3, r_max=0.3, cohesive=_ffi.PGH_COH_OP),
                          device_setup=device)
    return RVEProblem(config(name), device_setup=device)


SMALL = ["small-c0:24:3:11", "small-bl:20:3:12", "small-op:24:4:13", "small-multi:24:3:11", "small-multi-op:30:4:12",
         "small-multi-bl:30:4:13", "small-tet:8:1:11", "small-tet-op:10:2:12", "small-tet-op-q:10:3:11",
         "small-tet-bl-q:10:3:13", "small-empty:0:0:0", "big-sphere"]


@pytest.mark.parametrize("name", SMALL + ["C1", "C2", "C3", "C4", "C5"])
def test_device_setup_equals_host(name):
    print(f"device set-up {name}", flush=True)
    t = time.time()
    h = _build_setup(name, False)
    th = time.time() - t
    t = time.time()
    d = _build_setup(name, True)
    td = time.time() - t
    sh, sd = h.stage_seconds(), d.stage_seconds()
    print(f"device set-up {name}: cut {h.n_cut} enr {h.n_enr}; detect host {sh['detect']:.3f}s device "
          f"{sd['detect']:.3f}s, enrich host {sh['enrich']:.3f}s device {sd['enrich']:.3f}s; whole {th:.2f}s / "
          f"{td:.2f}s", flush=True)
    assert (h.n_nodes, h.n_enr, h.n_elem, h.n_cut, h.n_cqp, h.n_free, h.n_pres) == \
           (d.n_nodes, d.n_enr, d.n_elem, d.n_cut, d.n_cqp, d.n_free, d.n_pres)
    for f in SETUP:
        a, b = getattr(h, f), getattr(d, f)
        assert (a is None) == (b is None), f
        if a is not None:
            assert np.array_equal(a, b), f
    rh = np.frombuffer(b"".join(bytes(h.cut_record(i)) for i in range(h.n_cut)), np.uint8)
    rd = np.frombuffer(b"".join(bytes(d.cut_record(i)) for i in range(d.n_cut)), np.uint8)
    assert np.array_equal(rh, rd), "cut records"
    _same_pattern(h, d)
    h.close()
    d.close()


def test_device_setup_refuses_as_host():
    """Spheres packed closer than a tet cut elements by two inclusions, which the
    host refuses (PG_E_UNSUPPORTED): the device set-up returns the same code and
    message."""
    cfg = RVEConfig(dim=3, n=4, n_incl=40, r_min=0.1, r_max=0.15, gap_h=0.01, place=_ffi.PGH_PLACE_BEST_EFFORT)
    msgs = []
    for dev in (False, True):
        with pytest.raises(_ffi.PgError) as e:
            RVEProblem(cfg, device_setup=dev)
        msgs.append(str(e.value))
    assert msgs[0] == msgs[1] and f"({_ffi.PG_E_UNSUPPORTED})" in msgs[0] and "tet4 element cut by" in msgs[0]
