"""The pattern stage on the GPU (PGH_PATTERN_DEVICE): node incidence,
node-pair pattern, 4-block AIJ pattern and scatter indices built by radix
sorts of packed keys must equal the host build array for array, on every
workload at full size (the 100 x 100 and 500 x 500 RVEs, the 2000 x 2000
Ortiz-Pandolfi RVE, the 120^3 x 6 tet4 cube, the batch of 256 RVEs) and on the
small cases with cohesive tets, node-centred inclusions and no inclusion."""
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
