"""Host-side mirror of PIGFEM's Assembler plug-in API over the MI355X C ABI.

`AssemblerGPU.assemble(K, P_int, delta_t, update_ISVs, solution, assembleFunc,
assembleJac)` keeps the reference signature of Assembler::assemble
(include/Assembler.h:45-47, src/Assembler.cpp:65-146):
  * K is a `PMatrix` (the four AIJ blocks ff, fp, pf, pp, include/common.h:32-73),
  * P_int a `PVector` (f, p; include/common.h:74-101),
  * update_ISVs the trial internal variables, reset from the committed ones at
    every call (src/Assembler.cpp:92) and updated by the cohesive laws,
  * solution the global-dof-ordered nodal solution (NodalData),
  * both flags False is a no-op; errors raise PgError (the reference returns a
    PetscErrorCode / calls err_message).
All arithmetic runs in the native library's HIP kernels; this module only moves
pointers.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import PG_FUNC, PG_JAC, PG_FP32, PG_BATCHED, PG_CUT64, check, lib


class PMatrix:
    """Values of the four K blocks, stored back to back as [ff | fp | pf | pp]."""

    def __init__(self, problem):
        self.problem = problem
        self.values = np.zeros(problem.nnz_total, dtype=np.float64)
        off = np.cumsum([0] + problem.nnz)
        self.offsets = off

    def block(self, b: int) -> np.ndarray:
        return self.values[self.offsets[b]:self.offsets[b + 1]]

    def scipy(self, b: int):
        import scipy.sparse as sp
        p = self.problem
        shape = [(p.n_free, p.n_free), (p.n_free, p.n_pres), (p.n_pres, p.n_free), (p.n_pres, p.n_pres)][b]
        return sp.csr_matrix((self.block(b), p.ci[b], p.rp[b]), shape=shape)

    @property
    def ff(self):
        return self.scipy(0)


class PVector:
    """Internal-force vector split into free and prescribed parts ([Pf | Pp])."""

    def __init__(self, problem):
        self.problem = problem
        self.values = np.zeros(problem.n_dof, dtype=np.float64)

    @property
    def f(self) -> np.ndarray:
        return self.values[: self.problem.n_free]

    @property
    def p(self) -> np.ndarray:
        return self.values[self.problem.n_free:]


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class AssemblerGPU:
    """Drop-in assembler for AssemblerSmallStrainStructural on one MI355X.

    mode: "fp64"     -- the plain FP64 drop-in: one thread per element (reference
                        arithmetic, no class binning), staged element matrices,
                        deterministic node-pair reduction;
          "binned64" -- class-binned schedule in FP64 (cut elements one wavefront
                        each, uncut elements recomputed in the node-pair kernel);
          "mixed"    -- class-binned with FP32 B^T.D.B / cohesive tangents and FP64
                        sliver Jacobians, damage update, CSR sums and residual;
          "mixed_cut64" -- class-binned with FP32 only for uncut elements: cut
                        elements (children, enrichment, cohesive tangents) in FP64.
    """

    def __init__(self, problem, mode: str = "fp64"):
        if mode not in ("fp64", "binned64", "mixed", "mixed_cut64"):
            raise ValueError(mode)
        self.problem = problem
        self.mode = mode
        L = lib()
        h = C.c_void_p()
        check(L.pg_asm_create(C.byref(problem.desc), C.byref(h)), "pg_asm_create")
        self._h = h
        self.committed_isvs = np.zeros(problem.n_cqp, dtype=np.float64)

    # -- reference-shaped API ------------------------------------------------
    def flags(self, func: bool, jac: bool) -> int:
        f = (PG_FUNC if func else 0) | (PG_JAC if jac else 0)
        if self.mode == "mixed":
            f |= PG_FP32 | PG_BATCHED
        elif self.mode == "mixed_cut64":
            f |= PG_FP32 | PG_BATCHED | PG_CUT64
        elif self.mode == "binned64":
            f |= PG_BATCHED
        return f

    def assemble(self, K: PMatrix, P_int: PVector, delta_t: float, update_ISVs: np.ndarray,
                 solution: np.ndarray, assembleFunc: bool = True, assembleJac: bool = True) -> int:
        if not assembleFunc and not assembleJac:
            return 0
        p = self.problem
        U = np.ascontiguousarray(solution, dtype=np.float64)
        if U.shape != (p.n_dof,):
            raise ValueError(f"solution must have {p.n_dof} entries")
        if update_ISVs.shape != (p.n_cqp,) or update_ISVs.dtype != np.float64:
            raise ValueError("update_ISVs must be a float64 array of n_cqp entries")
        Kb = [_ptr(K.block(b)) if assembleJac else None for b in range(4)]
        Pb = [_ptr(P_int.f), _ptr(P_int.p)] if assembleFunc else [None, None]
        rc = lib().pg_asm_assemble(
            self._h, _ptr(U), _ptr(self.committed_isvs) if p.n_cqp else None,
            _ptr(update_ISVs) if p.n_cqp else None, float(delta_t), self.flags(assembleFunc, assembleJac),
            *Kb, *Pb)
        check(rc, "pg_asm_assemble")
        return 0

    def commit(self, update_ISVs: np.ndarray) -> None:
        """InternalVars::copy_data_in after a converged load step (src/SolverNonlinear.cpp:301)."""
        self.committed_isvs[:] = update_ISVs

    # -- device-resident API ---------------------------------------------------
    def assemble_device(self, U, isv_c, isv_t, K, P, func=True, jac=True, stream=None, delta_t=0.0) -> None:
        """Arguments are device buffers: `DeviceBuffer`s, raw device addresses (int),
        or objects with `data_ptr()` (isv_* may be None for a perfect-bond problem).
        Launches asynchronously on `stream` (a hipStream_t address) or the handle's
        own stream."""
        def dp(t):
            if t is None:
                return None
            if hasattr(t, "ptr"):
                return t.ptr
            if hasattr(t, "data_ptr"):
                return C.c_void_p(t.data_ptr())
            return C.c_void_p(int(t))
        st = None if stream is None else C.c_void_p(int(stream))
        rc = lib().pg_asm_assemble_device(self._h, dp(U), dp(isv_c), dp(isv_t), float(delta_t),
                                          self.flags(func, jac), dp(K), dp(P), st)
        check(rc, "pg_asm_assemble_device")

    def set_temperature_change(self, temp_change: float) -> None:
        """Temperature change applied by the following assemblies (the thermal
        pre-load's input.temp_change; pg_asm_set_temperature_change)."""
        check(lib().pg_asm_set_temperature_change(self._h, float(temp_change)), "pg_asm_set_temperature_change")

    def check_error(self) -> None:
        """Waits for the last assembly and raises PgError (PG_E_MESH) if a cut
        element had a child of negative Jacobian (src/Mesh.cpp:1561)."""
        check(lib().pg_asm_check_error(self._h), "pg_asm_check_error")

    def cohesive_openings(self, U) -> tuple:
        """Cohesive openings of a 2D cohesive mesh (WriterCohesiveOpenings,
        src/WriterCohesiveOpenings.cpp:73-147) from the GPU: U is a host array in
        global-dof order.  Returns (inclusion, rows): per cohesive edge, the index
        of its inclusion and (angle about its centre, tangential opening / delta_c,
        normal opening / delta_c), unsorted."""
        from .device import DeviceBuffer, synchronize
        p = self.problem
        n_out = p.n_cqp // 2
        rows = np.zeros((n_out, 3))
        if n_out == 0:
            return np.zeros(0, np.int64), rows
        inc = p.edge_inclusions()
        centers = np.ascontiguousarray(p.incl_centers[inc], dtype=np.float64)
        dU, dC, dO = DeviceBuffer.from_numpy(np.asarray(U, np.float64)), DeviceBuffer.from_numpy(centers), \
            DeviceBuffer.zeros(n_out * 3)
        check(lib().pg_asm_cohesive_openings(self._h, dU.ptr, dC.ptr, dO.ptr, n_out, None), "pg_asm_cohesive_openings")
        synchronize()
        rows = dO.download(n_out * 3).reshape(n_out, 3)
        return inc, rows

    def cohesive_failure(self, U) -> np.ndarray:
        """Per-edge data of WriterCohesiveFailureLengths
        (src/WriterCohesiveFailureLengths.cpp:64-86,183-224) from the GPU for a 2D
        cohesive mesh: U is a host array in global-dof order.  Returns rows of
        (edge length, mean normalised failure over the edge's Gauss points), in
        cut-element order."""
        from .device import DeviceBuffer, synchronize
        n_out = self.problem.n_cqp // 2
        if n_out == 0:
            return np.zeros((0, 2))
        dU, dO = DeviceBuffer.from_numpy(np.asarray(U, np.float64)), DeviceBuffer.zeros(n_out * 2)
        check(lib().pg_asm_cohesive_failure(self._h, dU.ptr, dO.ptr, n_out, None), "pg_asm_cohesive_failure")
        synchronize()
        return dO.download(n_out * 2).reshape(n_out, 2)

    def plot_fields(self, U, enr_edge=None, enr_alpha=None):
        """The fields of the VTK writers from the GPU (pg_asm_plot_fields): U is a
        host array in global-dof order.  Returns (nodes, cells): per node its
        plotted displacement (n_nodes + n_enr rows of dim), per VTK cell in the
        writer's order [volume, strain integral (nstr), stress integral (nstr)].
        The enrichment edges default to the problem's (pgh_enrich_edges)."""
        from .device import DeviceBuffer, synchronize
        p = self.problem
        edge = p.enr_edge if enr_edge is None else np.asarray(enr_edge, np.int32)
        alpha = p.enr_alpha if enr_alpha is None else np.asarray(enr_alpha, np.float64)
        if p.n_enr and (edge is None or alpha is None):
            raise ValueError("the enrichment edges of this mesh are not known: pass enr_edge and enr_alpha")
        nc = C.c_int64()
        check(lib().pg_asm_plot_cells(self._h, C.byref(nc)), "pg_asm_plot_cells")
        nall = p.n_nodes + p.n_enr
        ns = 3 if p.dim == 2 else 6
        S = 1 + 2 * ns
        dU = DeviceBuffer.from_numpy(np.ascontiguousarray(U, np.float64))
        dE = DeviceBuffer.from_numpy(np.ascontiguousarray(edge, np.int32).ravel()) if p.n_enr else None
        dA = DeviceBuffer.from_numpy(np.ascontiguousarray(alpha, np.float64)) if p.n_enr else None
        dN, dC = DeviceBuffer.zeros(nall * p.dim), DeviceBuffer.zeros(nc.value * S)
        check(lib().pg_asm_plot_fields(self._h, dU.ptr, dE.ptr if dE else None, dA.ptr if dA else None, dN.ptr,
                                       dC.ptr, None), "pg_asm_plot_fields")
        synchronize()
        return dN.download(nall * p.dim).reshape(nall, p.dim), dC.download(nc.value * S).reshape(nc.value, S)

    def set_timing(self, on: bool) -> None:
        check(lib().pg_asm_set_timing(self._h, 1 if on else 0), "pg_asm_set_timing")

    def last_timing(self):
        """[total, permutation + cut / element kernels, node-pair kernels,
        overlapped node groups, remaining node groups] in ms."""
        ms = (C.c_double * 5)()
        check(lib().pg_asm_last_timing(self._h, ms), "pg_asm_last_timing")
        return list(ms)

    def set_tiles(self, group_nodes: int = 0, cut_variant: int = 0, group_variant: int = 0) -> None:
        """Schedule of the class-binned modes (see pg_asm_set_tiles)."""
        check(lib().pg_asm_set_tiles(self._h, int(group_nodes), int(cut_variant), int(group_variant)),
              "pg_asm_set_tiles")

    def algorithmic_bytes(self) -> int:
        return int(lib().pg_asm_algorithmic_bytes(self._h))

    def device_bytes(self) -> int:
        return int(lib().pg_asm_device_bytes(self._h))

    def grid_nodes(self) -> int:
        """Nodes assembled from the structured-grid formulas (pg_asm_grid_nodes)."""
        return int(lib().pg_asm_grid_nodes(self._h))

    def close(self):
        if self._h:
            lib().pg_asm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class AssemblerGPUThermalPreLoad(AssemblerGPU):
    """Drop-in for AssemblerSmallStrainStructural_ThermalPreLoad
    (src/AssemblerSmallStrainStructural_ThermalPreLoad.cpp): the structural
    assembler with a temperature change that strains every LEI material whose
    thermal expansion ratio is set (pg_material.thermal_exp_set).  The parameter
    interface and the load-step rule are the reference's (:213-267):
    set_parameter("DELTA_T" | "TEMP" | "DELTA_TEMP", v) sets the target change
    and resets the current one to 0; set_parameter("COMPLETE" | "CONVERGED", v)
    marks the relaxation complete (v != 0); assemble_new_load_step(frac) sets the
    current change to delta_T once relaxation is complete, else delta_T * frac
    (:197-211).  Assemblies use the current change."""

    def __init__(self, problem, mode: str = "fp64"):
        super().__init__(problem, mode)
        self._delta_T = 0.0
        self._curr_delta_T = 0.0
        self._temp_set = False
        self._relaxation_complete = False

    def set_parameter(self, name: str, val: float) -> None:
        name = name.upper()
        if name in ("TEMP", "DELTA_T", "DELTA_TEMP"):
            self._delta_T = float(val)
            self._curr_delta_T = 0.0
            self._temp_set = True
        elif name in ("COMPLETE", "CONVERGED"):
            self._relaxation_complete = val != 0.0
        else:
            raise ValueError(f"{name} is not a valid parameter name")
        self.set_temperature_change(self._curr_delta_T)

    def get_parameter(self, name: str) -> float:
        name = name.upper()
        if name in ("TEMP", "DELTA_T", "DELTA_TEMP"):
            if not self._temp_set:
                raise ValueError("Please set a temperature change value before attempting to get it.")
            return self._delta_T
        if name in ("COMPLETE", "CONVERGED"):
            return 1.0 if self._relaxation_complete else 0.0
        if name in ("CURRENT TEMP", "CURRENT_TEMP"):
            return self._curr_delta_T
        raise ValueError(f"{name} is not a valid parameter name")

    def assemble_new_load_step(self, current_time_frac: float) -> None:
        if not self._temp_set:
            raise ValueError("Please set a temperature change value before attemping to solve a residual strain problem.")
        self._curr_delta_T = self._delta_T if self._relaxation_complete else self._delta_T * current_time_frac
        self.set_temperature_change(self._curr_delta_T)

    def clear(self) -> None:
        self._delta_T = self._curr_delta_T = 0.0
        self._temp_set = self._relaxation_complete = False
        self.set_temperature_change(0.0)


class AssemblerGPUThermal:
    """Drop-in for AssemblerThermal (include/AssemblerThermal.h:16-45,
    src/AssemblerThermal.cpp): heat conduction on a scalar-field problem
    (RVEConfig(thermal=True), one dof per node) through pg_therm_*.  The
    Assembler::assemble signature is kept; there are no internal variables
    (update_ISVs is ignored, as the thermal kernels have none).  P_int is
    B^T (D grad T), i.e. K T for the linear law."""

    def __init__(self, problem):
        if problem.nndof != 1:
            raise ValueError("the thermal assembler needs a scalar-field problem (RVEConfig(thermal=True))")
        self.problem = problem
        h = C.c_void_p()
        check(lib().pg_therm_create(C.byref(problem.desc), C.byref(h)), "pg_therm_create")
        self._h = h

    def assemble(self, K: PMatrix, P_int: PVector, delta_t: float, update_ISVs, solution: np.ndarray,
                 assembleFunc: bool = True, assembleJac: bool = True) -> int:
        if not assembleFunc and not assembleJac:
            return 0
        p = self.problem
        T = np.ascontiguousarray(solution, dtype=np.float64)
        if T.shape != (p.n_dof,):
            raise ValueError(f"solution must have {p.n_dof} entries")
        Kb = [_ptr(K.block(b)) if assembleJac else None for b in range(4)]
        Pb = [_ptr(P_int.f), _ptr(P_int.p)] if assembleFunc else [None, None]
        flags = (PG_FUNC if assembleFunc else 0) | (PG_JAC if assembleJac else 0)
        check(lib().pg_therm_assemble(self._h, _ptr(T), flags, *Kb, *Pb), "pg_therm_assemble")
        return 0

    def assemble_device(self, T, K, P, func=True, jac=True, stream=None) -> None:
        """Device buffers (DeviceBuffer or raw addresses); asynchronous on stream."""
        def dp(t):
            if t is None:
                return None
            return t.ptr if hasattr(t, "ptr") else C.c_void_p(int(t))
        flags = (PG_FUNC if func else 0) | (PG_JAC if jac else 0)
        st = None if stream is None else C.c_void_p(int(stream))
        check(lib().pg_therm_assemble_device(self._h, dp(T), flags, dp(K), dp(P), st), "pg_therm_assemble_device")

    def close(self):
        if self._h:
            lib().pg_therm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cut_inclusions(problem) -> np.ndarray:
    """Inclusion that cuts each cut record: the detection value of its parent
    nodes inside an inclusion (single-inclusion cuts; -1 if none is inside)."""
    det = problem.node_detect
    parents = problem.conn[problem.cut_elem]
    return det[parents].max(axis=1).astype(np.int64)


def write_cohesive_openings(filename: str, inclusion, rows) -> None:
    """WriterCohesiveOpenings::write (src/WriterCohesiveOpenings.cpp:44-59,
    182-197): the rows sorted by (inclusion, angle, openings), one line
    "inclusion,angle,tangential,normal" each, 16 significant digits."""
    data = sorted((int(i), float(r[0]), float(r[1]), float(r[2])) for i, r in zip(inclusion, rows))
    with open(filename, "w") as f:
        for i, th, dt, dn in data:
            f.write(f"{i},{th:.16g},{dt:.16g},{dn:.16g}\n")


def write_cohesive_failure_lengths(filename: str, rows, step: int, individual: bool = False) -> None:
    """WriterCohesiveFailureLengths::writeConsecutive
    (src/WriterCohesiveFailureLengths.cpp:93-130,137-258) for one rank: step 0
    creates the file and first writes the edge lengths (each followed by a comma
    when individual, else their total); every step then appends either each
    edge's failure followed by a comma (individual) or, for j = 1..5, the total
    length of the edges whose failure exceeds j.  Values print as %.16g."""
    rows = np.asarray(rows, np.float64).reshape(-1, 2)
    g = lambda x: f"{float(x):.16g}"
    out = []
    if step == 0:
        if individual:
            out.append("".join(g(v) + "," for v in rows[:, 0]) + "\n")
        else:
            tot = 0.0
            for v in rows[:, 0]:
                tot += v
            out.append(g(tot) + "\n")
    if individual:
        out.append("".join(g(v) + "," for v in rows[:, 1]) + "\n")
    else:
        failed = [0.0] * 5
        for L, f in rows:
            for j in range(5):
                if f > j + 1:
                    failed[j] += L
        out.append(",".join(g(v) for v in failed) + "\n")
    with open(filename, "w" if step == 0 else "a") as fh:
        fh.write("".join(out))
