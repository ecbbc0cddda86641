"""Host Newton / load-step driver around the assembly (harness, not hot path).

Restates PIGFEM's load stepping and Newton-Raphson loop so that the Newton
history of the GPU assembly can be compared with the FP64 oracle's:
  SolverNonlinear::solve            src/SolverNonlinear.cpp:70-385
  SolverNonlinear::update_time_step src/SolverNonlinear.cpp:386-422
  SolverNonlinearKSP::nonlinearStep src/SolverNonlinearKSP.cpp:55-155
  Assembler::assemble_new_load_step src/Assembler.cpp:394-511 (Dirichlet values
                                    scaled by the time fraction, no body loads)
The linear solve of K_ff (MUMPS Cholesky through PETSc KSP in the reference,
src/Solver.cpp:108-160) is a sparse LU (SuperLU via scipy) here.  The explicit
fallback the reference enters after 20 failed repeats is not restated: the driver
raises instead.
"""
from __future__ import annotations

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


class NewtonError(RuntimeError):
    pass


def _kff(problem, values):
    n = problem.n_free
    nnz = problem.nnz[0]
    return sp.csr_matrix((values[:nnz], problem.ci[0], problem.rp[0]), shape=(n, n))


class NewtonDriver:
    """assemble(U, isv_committed) -> (K values [ff|fp|pf|pp], P [Pf|Pp], isv_trial)."""

    def __init__(self, problem, assemble, T_final=1.0, max_time_step=0.1, min_time_step=1e-5,
                 rel_tol=1e-8, abs_tol=1e-8, max_iter=25, max_repeats=20, log=None):
        self.p = problem
        self.assemble = assemble
        self.T_final, self.max_dt, self.min_dt = T_final, max_time_step, min_time_step
        self.rel_tol, self.abs_tol, self.max_iter, self.max_repeats = rel_tol, abs_tol, max_iter, max_repeats
        self.log = log
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
                du = spl.splu(_kff(p, K).tocsc()).solve(res)
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

    def solve(self):
        p = self.p
        t, dt, step = 0.0, self.max_dt, 0
        bc = np.asarray(p.bc_values)
        while t < self.T_final:
            repeats, converged = 0, False
            while not converged:
                if repeats == self.max_repeats or dt < self.min_dt:
                    # the reference switches to its explicit solver here
                    self.stopped = f"load step at t={t} needs the explicit fallback (dt={dt:.3e}, repeats={repeats})"
                    raise NewtonError(self.stopped)
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
        """Newton iterations of every attempted load step (repeats included)."""
        return [(r.step, r.repeats, r.converged, r.iterations) for r in self.history]


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
