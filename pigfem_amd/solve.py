"""Linear solve of a Newton iteration on the GPU: Jacobi-preconditioned CG on
the free-free block of K (pg_csr_pcg), in place of the KSP solve of
SolverNonlinearKSP (src/SolverNonlinearKSP.cpp:94, src/Solver.cpp:108-160)."""

import ctypes as C

import numpy as np
import scipy.sparse as sp

from ._ffi import PgError, check, lib
from .device import DeviceBuffer


class DeviceCSR:
    """A CSR structure kept on the device; values are uploaded per solve."""

    def __init__(self, indptr, indices):
        rp = np.ascontiguousarray(indptr, dtype=np.int64)
        ci = np.ascontiguousarray(indices, dtype=np.int64)
        self.n = len(rp) - 1
        if self.n < 0 or rp[0] != 0 or np.any(np.diff(rp) < 0) or rp[-1] != len(ci) or len(ci) >= 2 ** 31:
            raise PgError("malformed CSR row pointers")
        if len(ci) and (ci.min() < 0 or ci.max() >= self.n):
            raise PgError("CSR column index out of range")
        self.nnz = len(ci)
        self.rp = DeviceBuffer.from_numpy(rp.astype(np.int32))
        self.ci = DeviceBuffer.from_numpy(ci.astype(np.int32) if len(ci) else np.zeros(1, np.int32))

    def solve(self, values, b, x0=None, rtol=1e-12, max_iter=None):
        """Returns (x, iterations, relative residual).  Raises PgError if the
        relative residual stays above rtol after max_iter (default 10 n)."""
        values = np.ascontiguousarray(values, dtype=np.float64)[: self.nnz]
        if len(values) != self.nnz or len(b) != self.n:
            raise PgError("values or right-hand side of the wrong length")
        dV = DeviceBuffer.from_numpy(values if self.nnz else np.zeros(1))
        dB = DeviceBuffer.from_numpy(np.asarray(b, np.float64))
        dX = DeviceBuffer.from_numpy(np.zeros(self.n) if x0 is None else np.asarray(x0, np.float64))
        it, rel = C.c_int(0), C.c_double(0.0)
        mi = 10 * self.n if max_iter is None else int(max_iter)
        check(lib().pg_csr_pcg(self.n, self.rp.ptr, self.ci.ptr, dV.ptr, dB.ptr, dX.ptr, float(rtol), mi,
                               C.byref(it), C.byref(rel), None), "pg_csr_pcg")
        if not rel.value <= rtol:
            raise PgError(f"PCG did not converge: relative residual {rel.value:.3e} after {it.value} iterations")
        return dX.download(self.n), it.value, rel.value


def pcg_solve(A, b, x0=None, rtol=1e-12, max_iter=None):
    """Solve A x = b on the GPU for a symmetric positive definite scipy matrix A."""
    A = sp.csr_matrix(A)
    return DeviceCSR(A.indptr, A.indices).solve(A.data, b, x0, rtol, max_iter)


class DeviceKSP:
    """Newton's linear solve on the device (pg_ksp_*): K_ff x = b for the K_ff
    block of a device K buffer laid out as pg_asm_assemble_device writes it,
    preconditioned by multicolour symmetric Gauss-Seidel.  method: "cg",
    "fgmres" or "bicgstab"."""

    METHODS = {"cg": 0, "fgmres": 1, "bicgstab": 2}

    def __init__(self, problem, method="fgmres"):
        self.n = problem.n_free
        h = C.c_void_p()
        check(lib().pg_ksp_create(C.byref(problem.desc), self.METHODS[method], C.byref(h)), "pg_ksp_create")
        self._h = h

    def solve(self, dK, db, dx, rtol=1e-10, max_iter=20000):
        """dK, db, dx: DeviceBuffers (dx: initial guess in, solution out).
        Returns (iterations, true relative residual); raises PgError if the
        solve breaks down or stops above rtol."""
        it, rel = C.c_int(0), C.c_double(0.0)
        check(lib().pg_ksp_solve(self._h, dK.ptr, db.ptr, dx.ptr, float(rtol), int(max_iter), C.byref(it),
                                 C.byref(rel)), "pg_ksp_solve")
        if not rel.value <= 10 * rtol:
            raise PgError(f"device solve stopped at relative residual {rel.value:.3e} after {it.value} iterations")
        return it.value, rel.value

    def close(self):
        if self._h:
            lib().pg_ksp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dev_dot(n, x, y):
    out = C.c_double(0.0)
    check(lib().pg_dev_dot(int(n), x.ptr, y.ptr, C.byref(out)), "pg_dev_dot")
    return out.value


def dev_axpy(n, a, x, y):
    check(lib().pg_dev_axpy(int(n), float(a), x.ptr, y.ptr), "pg_dev_axpy")
