"""pg_csr_pcg_noalloc, the PCG solve on the library's fixed device workspace:
through the C ABI, two calls in a row (the second on a smaller system) give
pg_csr_pcg's solution and iteration count bit for bit; a system beyond the
fixed workspace is rejected before any device work."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from pigfem_amd import DeviceBuffer, PgError
from pigfem_amd._ffi import PG_E_ARG, lib

pytestmark = pytest.mark.gpu


def _laplacian(m, shift):
    """SPD 2D 5-point Laplacian on an m x m grid plus shift * I."""
    e = np.ones(m)
    T = sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1])
    A = sp.kron(sp.identity(m), T) + sp.kron(T, sp.identity(m)) + shift * sp.identity(m * m)
    return sp.csr_matrix(A)


def _solve(fn, A, b):
    n = A.shape[0]
    rp = DeviceBuffer.from_numpy(A.indptr.astype(np.int32))
    ci = DeviceBuffer.from_numpy(A.indices.astype(np.int32))
    v = DeviceBuffer.from_numpy(A.data.astype(np.float64))
    db = DeviceBuffer.from_numpy(b)
    dx = DeviceBuffer.zeros(n)
    it, rel = C.c_int(0), C.c_double(0.0)
    rc = fn(n, rp.ptr, ci.ptr, v.ptr, db.ptr, dx.ptr, 1e-12, 10 * n, C.byref(it), C.byref(rel), None)
    if rc != 0:
        raise PgError(f"rc {rc}: {lib().pg_asm_last_error().decode()}")
    return dx.download(n), it.value, rel.value


def test_noalloc_pcg_matches_the_allocating_solve():
    L = lib()
    rng = np.random.default_rng(5)
    for m, shift in ((120, 0.01), (40, 0.5)):  # the second call on a smaller system
        A = _laplacian(m, shift)
        b = rng.standard_normal(A.shape[0])
        print(f"no-alloc pcg: n = {A.shape[0]}", flush=True)
        x0, it0, r0 = _solve(L.pg_csr_pcg, A, b)
        x1, it1, r1 = _solve(L.pg_csr_pcg_noalloc, A, b)
        assert it1 == it0 and r1 == r0 and np.array_equal(x1, x0)
        assert r1 <= 1e-12
        assert np.linalg.norm(b - A @ x1) <= 1e-10 * np.linalg.norm(b)


def test_noalloc_pcg_rejects_a_system_beyond_its_workspace():
    L = lib()
    it, rel = C.c_int(0), C.c_double(0.0)
    rc = L.pg_csr_pcg_noalloc((1 << 20) + 1, None, None, None, None, None, 1e-12, 10, C.byref(it), C.byref(rel), None)
    assert rc == PG_E_ARG
    assert b"workspace" in L.pg_asm_last_error()
