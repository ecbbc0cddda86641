"""ctypes declarations of the C ABI in include/pgfem_asm.h and include/pgfem_harness.h.

The native library (pigfem_amd/lib/libpgfem_asm.so) is built in-tree by
``pigfem_amd.build.build()``.  Loading fails loudly when it is missing: there is
no Python or CPU fallback for the assembly.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpgfem_asm.so")

PG_FUNC = 1
PG_JAC = 2
PG_FP32 = 4
PG_BATCHED = 8
PG_CUT64 = 16

PG_OK, PG_E_ARG, PG_E_HIP, PG_E_UNSUPPORTED, PG_E_MESH, PG_E_NOMEM = 0, -1, -2, -3, -4, -5

PG_MAT_LEI, PG_MAT_BL, PG_MAT_BL_NU, PG_MAT_OP, PG_MAT_OP_NU, PG_MAT_LT = range(6)
PGH_COH_NONE, PGH_COH_BL, PGH_COH_OP, PGH_COH_BL_NU, PGH_COH_OP_NU = range(5)
PGH_PLACE_BEST_EFFORT = 1
PGH_PLACE_AT_NODES = 2
PGH_PATTERN_DEVICE = 4
PGH_SETUP_DEVICE = 8
PGH_SCALAR_DOFS = 16
PGH_BATCH_COLS = 8
PGH_N_STAGES = 7
PG_CHILD_TRI3, PG_CHILD_QUAD4, PG_CHILD_TET4, PG_CHILD_POLY5, PG_CHILD_POLY6 = range(5)
PG_MAX_ENR, PG_MAX_CHILD, PG_MAX_COH = 12, 8, 3
PG_CHILD_NN = {PG_CHILD_TRI3: 3, PG_CHILD_QUAD4: 4, PG_CHILD_TET4: 4, PG_CHILD_POLY5: 5, PG_CHILD_POLY6: 6}
PG_CHILD_NQP = {PG_CHILD_TRI3: 4, PG_CHILD_QUAD4: 4, PG_CHILD_TET4: 5, PG_CHILD_POLY5: 20, PG_CHILD_POLY6: 24}


class PgMaterial(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("nstr", C.c_int32), ("D", C.c_double * 36),
        ("sigma_c", C.c_double), ("delta_c", C.c_double), ("delta_c1", C.c_double),
        ("delta_c2", C.c_double), ("beta2", C.c_double), ("alpha", C.c_double),
        ("thermal_exp", C.c_double), ("thermal_exp_set", C.c_int32), ("pad1", C.c_int32),
    ]


PG_SENS_LEI, PG_SENS_SIGMA_C, PG_SENS_DELTA_C = 0, 1, 2
PG_SENS_OWN_CHILD_B = 1


class PgSensParam(C.Structure):
    _fields_ = [("mat", C.c_int32), ("kind", C.c_int32), ("dD", C.c_double * 36)]


class PgCutRec(C.Structure):
    _fields_ = [
        ("enr", C.c_int32 * PG_MAX_ENR), ("isv_off", C.c_int32),
        ("n_enr", C.c_uint8), ("n_child", C.c_uint8), ("n_coh", C.c_uint8), ("n_vqp", C.c_uint8),
        ("child_type", C.c_uint8 * PG_MAX_CHILD), ("child_mat", C.c_uint8 * PG_MAX_CHILD),
        ("child_node", C.c_uint8 * 32),
        ("coh_nodes", (C.c_uint8 * 8) * PG_MAX_COH), ("coh_mat", C.c_uint8 * PG_MAX_COH),
        ("coh_nn", C.c_uint8), ("pad", C.c_uint8 * 4),
    ]


assert C.sizeof(PgCutRec) == 136


def child_nodes(rec, dim: int, k: int) -> list:
    """Element-local nodes of child k of a cut record (PG_CHILD_STRIDE layout)."""
    st = 8 if dim == 2 else 4
    return [int(rec.child_node[st * k + n]) for n in range(PG_CHILD_NN[int(rec.child_type[k])])]


class PgProblemDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("nen", C.c_int32), ("n_nodes", C.c_int32), ("n_enr", C.c_int32),
        ("n_elem", C.c_int32), ("n_cut", C.c_int32), ("n_mat", C.c_int32), ("n_cqp", C.c_int32),
        ("n_free", C.c_int32), ("n_pres", C.c_int32), ("cohesive", C.c_int32), ("nndof", C.c_int32),
        ("nnz", C.c_int64 * 4), ("n_pairs", C.c_int64),
        ("coords", C.POINTER(C.c_double)), ("conn", C.POINTER(C.c_int32)),
        ("elem_mat", C.POINTER(C.c_int32)), ("cut_of_elem", C.POINTER(C.c_int32)),
        ("cut_elem", C.POINTER(C.c_int32)), ("cut", C.POINTER(PgCutRec)),
        ("mats", C.POINTER(PgMaterial)), ("node_dof", C.POINTER(C.c_int32)),
        ("rp", C.POINTER(C.c_int32) * 4), ("ci", C.POINTER(C.c_int32) * 4),
        ("npat_ptr", C.POINTER(C.c_int32)), ("npat_col", C.POINTER(C.c_int32)),
        ("npat_row", C.POINTER(C.c_int32)), ("npat_map", C.POINTER(C.c_int32)),
        ("n2e_ptr", C.POINTER(C.c_int32)), ("n2e", C.POINTER(C.c_int32)),
    ]


class PghRveParams(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("n", C.c_int32), ("n_incl", C.c_int32), ("cohesive", C.c_int32),
        ("r_min", C.c_double), ("r_max", C.c_double), ("gap", C.c_double), ("seed", C.c_uint64),
        ("E_matrix", C.c_double), ("nu_matrix", C.c_double), ("E_fiber", C.c_double), ("nu_fiber", C.c_double),
        ("sigma_c", C.c_double), ("delta_c", C.c_double), ("delta_c1", C.c_double), ("delta_c2", C.c_double),
        ("beta", C.c_double), ("alpha", C.c_double), ("strain", C.c_double),
        ("max_tries", C.c_int32), ("flags", C.c_int32),
        ("kappa_matrix", C.c_double), ("kappa_fiber", C.c_double),
    ]


# exported symbols and their signatures
SIGNATURES = {
    # pgfem_asm.h
    "pg_asm_create": (C.c_int, [C.POINTER(PgProblemDesc), C.POINTER(C.c_void_p)]),
    "pg_asm_destroy": (C.c_int, [C.c_void_p]),
    "pg_asm_last_error": (C.c_char_p, []),
    "pg_asm_assemble": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_uint,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pg_asm_assemble_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_uint,
                                         C.c_void_p, C.c_void_p, C.c_void_p]),
    "pg_asm_sizes": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "pg_asm_last_timing": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "pg_asm_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "pg_asm_set_tiles": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "pg_asm_algorithmic_bytes": (C.c_int64, [C.c_void_p]),
    "pg_asm_device_bytes": (C.c_int64, [C.c_void_p]),
    "pg_asm_cohesive_openings": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "pg_asm_cohesive_failure": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "pg_asm_check_error": (C.c_int, [C.c_void_p]),
    "pg_asm_set_temperature_change": (C.c_int, [C.c_void_p, C.c_double]),
    "pg_asm_plot_cells": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "pg_asm_plot_fields": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    "pgh_enrich_edges": (C.c_int, [C.c_void_p, C.POINTER(C.POINTER(C.c_int32)), C.POINTER(C.POINTER(C.c_double))]),
    "pg_therm_create": (C.c_int, [C.POINTER(PgProblemDesc), C.POINTER(C.c_void_p)]),
    "pg_therm_assemble_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pg_therm_assemble": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p]),
    "pg_therm_destroy": (C.c_int, [C.c_void_p]),
    "pg_asm_sensitivity_dPdd": (C.c_int, [C.c_void_p, C.POINTER(PgSensParam), C.c_uint, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p]),
    "pg_asm_grid_nodes": (C.c_int64, [C.c_void_p]),
    "pg_csr_pcg": (C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                             C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_void_p]),
    "pg_csr_pcg_noalloc": (C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_void_p]),
    "pg_dev_alloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_int64]),
    "pg_dev_free": (C.c_int, [C.c_void_p]),
    "pg_dev_upload": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "pg_dev_download": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "pg_dev_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "pg_dev_synchronize": (C.c_int, []),
    "pg_dev_axpy": (C.c_int, [C.c_int64, C.c_double, C.c_void_p, C.c_void_p]),
    "pg_dev_dot": (C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]),
    "pg_ksp_create": (C.c_int, [C.POINTER(PgProblemDesc), C.c_int, C.POINTER(C.c_void_p)]),
    "pg_ksp_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int,
                               C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "pg_ksp_destroy": (C.c_int, [C.c_void_p]),
    "pg_ksp_last_error": (C.c_char_p, []),
    # pgfem_harness.h
    "pgh_build_rve": (C.c_int, [C.POINTER(PghRveParams), C.POINTER(C.c_void_p)]),
    "pgh_get_desc": (C.c_int, [C.c_void_p, C.POINTER(PgProblemDesc)]),
    "pgh_build_pattern": (C.c_int, [C.POINTER(PgProblemDesc), C.POINTER(C.c_void_p)]),
    "pgh_build_pattern_flags": (C.c_int, [C.POINTER(PgProblemDesc), C.c_int32, C.POINTER(C.c_void_p)]),
    "pgh_stage_seconds": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int32]),
    "pgh_build_batch": (C.c_int, [C.POINTER(PghRveParams), C.c_int32, C.POINTER(C.c_void_p)]),
    "pgh_batch_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.POINTER(C.c_int64))]),
    "pgh_free": (None, [C.c_void_p]),
    "pgh_last_error": (C.c_char_p, []),
    "pgh_inclusions": (C.c_int, [C.c_void_p, C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.POINTER(C.c_double)),
                                 C.POINTER(C.c_int32)]),
    "pgh_node_detect": (C.POINTER(C.c_int32), [C.c_void_p]),
    "pgh_enrich_inclusion": (C.POINTER(C.c_int32), [C.c_void_p]),
    "pgh_bc_values": (C.POINTER(C.c_double), [C.c_void_p]),
    "pgh_elem_qps": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
}

_lib = None


def lib() -> C.CDLL:
    """Load the in-tree native library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"native library {LIB_PATH} is missing: run `python -c 'import __graft_entry__ as g; g.build()'`"
            )
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class PgError(RuntimeError):
    pass


def check(rc: int, where: str) -> None:
    if rc != 0:
        L = lib()
        msg = (L.pgh_last_error() if where.startswith("pgh") else
               L.pg_ksp_last_error() if where.startswith("pg_ksp") else L.pg_asm_last_error()) or b""
        raise PgError(f"{where} failed ({rc}): {msg.decode(errors='replace')}")
