"""Host Newton / load-step driver around the assembly (harness, not hot path).

Restates PIGFEM's load stepping and Newton-Raphson loop so that the Newton
history of the GPU assembly can be compared with the FP64 oracle's:
  SolverNonlinear::solve            src/SolverNonlinear.cpp:70-318
  SolverNonlinear::update_time_step src/SolverNonlinear.cpp:386-422
  SolverNonlinear::explicit_solve   src/SolverNonlinear.cpp:430-510 (the last
                                    resort after 20 failed repeats or dt < min)
  SolverNonlinearKSP::nonlinearStep src/SolverNonlinearKSP.cpp:55-155
  Assembler::assemble_new_load_step src/Assembler.cpp:394-511 (Dirichlet values
                                    scaled by the time fraction, no body loads)
The linear solve of K_ff (MUMPS through PETSc KSP in the reference,
src/Solver.cpp:108-160) is a sparse direct LU: MKL PARDISO when the MKL runtime
is present (nested-dissection ordering, multi-threaded), else SuperLU (scipy).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl


@dataclass
class StepRecord:
    step: int
    t: float
    dt: float
    repeats: int
    converged: bool
    iterations: int
    residuals: list = field(default_factory=list)
    explicit: bool = False


class NewtonError(RuntimeError):
    pass


def _kff(problem, values):
    n = problem.n_free
    nnz = problem.nnz[0]
    return sp.csr_matrix((values[:nnz], problem.ci[0], problem.rp[0]), shape=(n, n))


_MKL_PATHS = ("/opt/conda/lib/libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1", "libmkl_rt.so")


def _load_mkl():
    for path in _MKL_PATHS:
        try:
            return C.CDLL(path)
        except OSError:
            continue
    return None


class PardisoLU:
    """K_ff x = r with MKL PARDISO (real unsymmetric, zero-based CSR): the
    symbolic analysis is done once for the fixed AIJ pattern, the numerical
    factorization at every Newton iterate."""

    def __init__(self, mkl, n, rp, ci):
        self.mkl = mkl
        self.n = n
        self.ia = np.ascontiguousarray(rp, dtype=np.int32)
        self.ja = np.ascontiguousarray(ci, dtype=np.int32)
        self.pt = (C.c_void_p * 64)()
        self.iparm = (C.c_int32 * 64)()
        self.mtype = C.c_int32(11)
        mkl.pardisoinit(self.pt, C.byref(self.mtype), self.iparm)
        self.iparm[34] = 1            # zero-based indices
        self.iparm[10] = 0            # no scaling and no weighted matching: both would
        self.iparm[12] = 0            # make the one-time analysis depend on the values
        self.a = np.zeros(len(self.ja))
        self._call(11, self.a)

    def _call(self, phase, a, b=None, x=None):
        one, ph, nn = C.c_int32(1), C.c_int32(phase), C.c_int32(self.n)
        msg, err = C.c_int32(0), C.c_int32(0)
        bp = None if b is None else b.ctypes.data_as(C.c_void_p)
        xp = None if x is None else x.ctypes.data_as(C.c_void_p)
        self.mkl.pardiso(self.pt, C.byref(one), C.byref(one), C.byref(self.mtype), C.byref(ph), C.byref(nn),
                         a.ctypes.data_as(C.c_void_p), self.ia.ctypes.data_as(C.c_void_p),
                         self.ja.ctypes.data_as(C.c_void_p), None, C.byref(one), self.iparm, C.byref(msg),
                         bp, xp, C.byref(err))
        return err.value

    def solve(self, values, r):
        self.a = np.ascontiguousarray(values, dtype=np.float64)
        if self._call(22, self.a) != 0:
            raise RuntimeError("PARDISO factorization failed")
        b = np.ascontiguousarray(r, dtype=np.float64)
        x = np.zeros(self.n)
        if self._call(33, self.a, b, x) != 0:
            raise RuntimeError("PARDISO solve failed")
        return x

    def close(self):
        if self.pt is not None:
            self._call(-1, self.a)
            self.pt = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LinearSolver:
    def __init__(self, problem, prefer="pardiso"):
        self.p = problem
        self._gpu = None
        if prefer == "none":  # the caller solves by itself (DeviceNewtonDriver)
            self.kind, self._pardiso = "none", None
            return
        if prefer == "gpu-pcg":
            from .solve import DeviceCSR
            self.kind, self._pardiso = "gpu-pcg", None
            self._gpu = DeviceCSR(problem.rp[0], problem.ci[0])
            return
        mkl = _load_mkl() if prefer == "pardiso" else None
        self.kind = "pardiso" if mkl is not None else "superlu"
        self._pardiso = PardisoLU(mkl, problem.n_free, problem.rp[0], problem.ci[0]) if mkl is not None else None

    def solve(self, K, r):
        if self._gpu is not None:
            return self._gpu.solve(K[: self.p.nnz[0]], r, rtol=1e-12)[0]
        if self._pardiso is not None:
            return self._pardiso.solve(K[: self.p.nnz[0]], r)
        return spl.splu(_kff(self.p, K).tocsc()).solve(r)


class NewtonDriver:
    """assemble(U, isv_committed) -> (K values [ff|fp|pf|pp], P [Pf|Pp], isv_trial)."""

    def __init__(self, problem, assemble, T_final=1.0, max_time_step=0.1, min_time_step=1e-5,
                 rel_tol=1e-8, abs_tol=1e-8, max_iter=25, max_repeats=20, explicit_steps=100,
                 solver="pardiso", log=None, max_steps=None):
        self.p = problem
        self.assemble = assemble
        self.explicit_steps = explicit_steps
        self.lin = LinearSolver(problem, solver)
        self.T_final, self.max_dt, self.min_dt = T_final, max_time_step, min_time_step
        self.rel_tol, self.abs_tol, self.max_iter, self.max_repeats = rel_tol, abs_tol, max_iter, max_repeats
        self.log = log
        self.max_steps = max_steps  # stop after this many load steps (None: run to T_final)
        self.history: list[StepRecord] = []
        self.U = np.zeros(problem.n_dof)
        self.isv = np.zeros(problem.n_cqp)
        self.stopped = None

    # SolverNonlinearKSP::nonlinearStep
    def _newton(self, U):
        p = self.p
        K, P, it = self.assemble(U, self.isv)
        res = -P[: p.n_free]                       # F_ext.f - P_int.f, F_ext.f = 0
        r = float(np.linalg.norm(res))
        r0 = r
        hist = [r]
        nr, inc = 0, 0
        while True:
            try:
                du = self.lin.solve(K, res)
            except RuntimeError:                     # singular factor: treat as divergence
                return False, nr, hist, it, U
            U = U.copy()
            U[: p.n_free] += du
            K, P, it = self.assemble(U, self.isv)
            prev = r
            res = -P[: p.n_free]
            r = float(np.linalg.norm(res))
            nr += 1
            hist.append(r)
            if r < self.rel_tol * r0 or r < self.abs_tol:
                return True, nr, hist, it, U
            if not math.isfinite(r):
                return False, nr, hist, it, U
            if nr >= self.max_iter:
                return False, nr, hist, it, U
            inc = inc + 1 if r > prev else 0
            if inc >= 3:
                return False, nr, hist, it, U

    # SolverNonlinear::explicit_solve: n steps of dt = min_dt, one assembly and
    # one linear solve each, history committed after every step
    def _explicit(self, t, step):
        p = self.p
        dt = self.min_dt
        n = self.explicit_steps
        n_till_end = int((self.T_final - t) / self.min_dt + 0.5)
        cont = True
        if n_till_end < n:
            n, cont = n_till_end, False
        bc = np.asarray(p.bc_values)
        for k in range(n):
            U = self.U.copy()
            U[p.n_free:] = ((t + dt) / self.T_final) * bc
            K, P, it = self.assemble(U, self.isv)
            res = -P[: p.n_free]
            r = float(np.linalg.norm(res))
            if not math.isfinite(r):
                raise NewtonError("invalid residual norm during an explicit step")
            U[: p.n_free] += self.lin.solve(K, res)
            self.U = U
            self.isv = it.copy()
            t += dt
            self.history.append(StepRecord(step, t, dt, 0, True, 1, [r], explicit=True))
        if self.log:
            self.log(f"[newton] explicit: {n} steps of dt={dt:.1e} to t={t:.6f}")
        return t, cont

    def solve(self):
        p = self.p
        t, dt, step = 0.0, self.max_dt, 0
        bc = np.asarray(p.bc_values)
        while t < self.T_final and (self.max_steps is None or step < self.max_steps):
            repeats, converged, last_failed = 0, False, False
            while not converged:
                if repeats == self.max_repeats or dt < self.min_dt:
                    if last_failed:
                        self.stopped = f"load step at t={t} did not converge after the explicit steps"
                        raise NewtonError(self.stopped)
                    last_failed = True
                    t, cont = self._explicit(t, step + 1)
                    if not cont:
                        break
                    dt *= 4                                    # the step size before the failure
                    dt = min(dt, self.T_final - t)
                repeats += 1
                U = self.U.copy()
                U[p.n_free:] = ((t + dt) / self.T_final) * bc   # assemble_new_load_step
                converged, nr, hist, isv_trial, U = self._newton(U)
                rec = StepRecord(step + 1, t + dt if converged else t, dt, repeats, converged, nr, hist)
                self.history.append(rec)
                if self.log:
                    self.log(f"[newton] step {rec.step} t={rec.t:.5f} dt={dt:.3e} rpt={repeats} "
                             f"conv={converged} iters={nr} res={hist[-1]:.3e}")
                if converged:
                    t += dt
                    self.U = U
                    self.isv = isv_trial.copy()              # InternalVars::copy_data_in
                # update_time_step
                if not converged:
                    dt *= 0.25
                elif repeats == 1:
                    dt *= 1.5
                remaining = self.T_final - t
                if dt > self.max_dt:
                    dt = self.max_dt
                if dt > remaining:
                    if remaining >= self.min_dt:
                        dt = remaining
                    else:
                        t = self.T_final
            step += 1
        return self.history

    def iteration_counts(self):
        """Newton iterations of every attempted load step (repeats included;
        an explicit step counts as one iteration)."""
        return [(r.step, r.repeats, r.converged, r.iterations) for r in self.history]


class DeviceNewtonDriver(NewtonDriver):
    """The same load stepping and Newton loop with the assembly and the linear
    solve on the device: U, K, P and the trial ISVs stay in HBM; K_ff is solved
    in place by preconditioned Krylov (DeviceKSP, iterative: its cost grows
    with the mesh); the residual norm is a
    device dot product.  Only scalars, the prescribed values at the start of an
    attempt and U / ISV snapshots of converged steps cross PCIe.  `asm` is an
    AssemblerGPU (any mode).

    method: "cg", "fgmres", "bicgstab" or "auto".  "auto" solves with CG
    (at most cg_max_iter iterations) while the tangent lets it converge --
    a tangent without softening is symmetric positive definite -- and from the
    first CG failure on (an indefinite tangent after the onset of cohesive
    softening) with FGMRES; the failed attempt restarts from zero.  "cg" and
    "bicgstab" keep their method, but a solve of theirs that breaks down or
    stalls (at most cg_max_iter iterations) is done again from zero by FGMRES:
    the reference's direct solve has no Krylov breakdown, so one must not
    count as a failed Newton iteration."""

    def __init__(self, problem, asm, method="auto", lin_rtol=1e-10, lin_max_iter=20000, cg_max_iter=5000, **kw):
        from .device import DeviceBuffer
        from .solve import DeviceKSP
        kw["solver"] = "none"
        super().__init__(problem, None, **kw)
        self.lin.kind = f"gpu-{method}"
        self.asm = asm
        # sticky: after the first failure of the primary method, FGMRES only
        self.sticky = method == "auto"
        if method == "fgmres":
            self.ksp = DeviceKSP(problem, method)
            self.ksp_fallback = None
        else:
            self.ksp = DeviceKSP(problem, "cg" if method == "auto" else method)
            self.ksp_fallback = DeviceKSP(problem, "fgmres")
        self.cg_max_iter = cg_max_iter
        self.lin_rtol, self.lin_max_iter = lin_rtol, lin_max_iter
        p = problem
        self.dU = DeviceBuffer.zeros(p.n_dof)
        self.dK = DeviceBuffer.zeros(max(p.nnz_total, 1))
        self.dP = DeviceBuffer.zeros(p.n_dof)
        self.dIc = DeviceBuffer.zeros(max(p.n_cqp, 1))
        self.dIt = DeviceBuffer.zeros(max(p.n_cqp, 1))
        self.ddu = DeviceBuffer.zeros(p.n_free)
        self.dzero = DeviceBuffer.zeros(p.n_free)
        self.lin_iters = []  # Krylov iterations of every solve

    def _assemble(self):
        """K, P at the device U; returns |P_f| (the residual norm, F_ext = 0)."""
        from .solve import dev_dot
        p = self.p
        c = self.dIc if p.n_cqp else None
        t = self.dIt if p.n_cqp else None
        self.asm.assemble_device(self.dU, c, t, self.dK, self.dP)
        self.asm.check_error()
        return math.sqrt(max(dev_dot(p.n_free, self.dP, self.dP), 0.0))

    def _solve_update(self):
        """U_f -= du with K_ff du = P_f (the Newton update du = K^-1 (F_ext - P))."""
        from .solve import dev_axpy
        p = self.p
        from ._ffi import PgError
        self.ddu.copy_from(self.dzero)
        if self.ksp_fallback is not None and self.ksp is not None:
            try:
                it, _ = self.ksp.solve(self.dK, self.dP, self.ddu, self.lin_rtol, self.cg_max_iter)
            except PgError:
                if self.sticky:
                    self.ksp.close()
                    self.ksp = None
                self.ddu.copy_from(self.dzero)
                it, _ = self.ksp_fallback.solve(self.dK, self.dP, self.ddu, self.lin_rtol, self.lin_max_iter)
        else:
            ksp = self.ksp if self.ksp is not None else self.ksp_fallback
            it, _ = ksp.solve(self.dK, self.dP, self.ddu, self.lin_rtol, self.lin_max_iter)
        self.lin_iters.append(it)
        dev_axpy(p.n_free, -1.0, self.ddu, self.dU)

    def _begin(self, U):
        p = self.p
        self.dU.upload(np.ascontiguousarray(U, dtype=np.float64))
        if p.n_cqp:
            self.dIc.upload(np.ascontiguousarray(self.isv, dtype=np.float64))

    def _trial_isv(self):
        p = self.p
        return self.dIt.download(p.n_cqp) if p.n_cqp else np.zeros(0)

    # SolverNonlinearKSP::nonlinearStep, device-resident
    def _newton(self, U):
        from .device import synchronize
        self._begin(U)
        r = self._assemble()
        r0 = r
        hist = [r]
        nr, inc = 0, 0
        while True:
            try:
                self._solve_update()
            except Exception:                        # breakdown: treat as divergence
                synchronize()
                return False, nr, hist, self._trial_isv(), self.dU.download(self.p.n_dof)
            prev = r
            r = self._assemble()
            nr += 1
            hist.append(r)
            done = None
            if r < self.rel_tol * r0 or r < self.abs_tol:
                done = True
            elif not math.isfinite(r) or nr >= self.max_iter:
                done = False
            else:
                inc = inc + 1 if r > prev else 0
                if inc >= 3:
                    done = False
            if done is not None:
                return done, nr, hist, self._trial_isv(), self.dU.download(self.p.n_dof)

    # SolverNonlinear::explicit_solve, device-resident
    def _explicit(self, t, step):
        p = self.p
        dt = self.min_dt
        n = self.explicit_steps
        n_till_end = int((self.T_final - t) / self.min_dt + 0.5)
        cont = True
        if n_till_end < n:
            n, cont = n_till_end, False
        bc = np.asarray(p.bc_values)
        for k in range(n):
            U = self.U.copy()
            U[p.n_free:] = ((t + dt) / self.T_final) * bc
            self._begin(U)
            r = self._assemble()
            if not math.isfinite(r):
                raise NewtonError("invalid residual norm during an explicit step")
            self._solve_update()
            self.U = self.dU.download(p.n_dof)
            self.isv = self._trial_isv().copy()
            t += dt
            self.history.append(StepRecord(step, t, dt, 0, True, 1, [r], explicit=True))
        if self.log:
            self.log(f"[newton] explicit: {n} steps of dt={dt:.1e} to t={t:.6f}")
        return t, cont


def gpu_assemble_fn(asm):
    """Adapter: AssemblerGPU -> assemble(U, isv) for the driver."""
    from .assembler import PMatrix, PVector
    p = asm.problem
    K, P = PMatrix(p), PVector(p)

    def f(U, isv):
        asm.committed_isvs[:] = isv
        it = np.zeros(p.n_cqp)
        asm.assemble(K, P, 0.0, it, U, True, True)
        return K.values, P.values, it
    return f
