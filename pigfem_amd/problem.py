"""Synthetic IGFEM RVE problems (host set-up) and the benchmark configurations.

`RVEProblem` wraps the native harness (include/pgfem_harness.h), which restates
PIGFEM's Problem::init set-up on flat arrays: structured mesh
(src/Mesh.cpp:236-630), inclusion detection (src/Mesh.cpp:2883-2930), cut
classification and child topology (src/tri3.cpp:287-596, src/tet4.cpp:228-520),
enrichment nodes (src/Mesh.cpp:3104-3251), dofs (src/DofObject.cpp:241-538) and
the 4-block AIJ pattern.  Array views are zero-copy numpy views of the native
buffers and live as long as the problem object.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _ffi
from ._ffi import PgProblemDesc, PghRveParams, check, lib


@dataclass
class RVEConfig:
    """Generator parameters of one RVE (units: MPa, unit-length domain)."""

    dim: int = 2
    n: int = 100
    n_incl: int = 16
    r_min: float = 0.03
    r_max: float = 0.05
    gap_h: float = 2.0           # minimum surface gap in element sizes h = 1/n
    cohesive: int = _ffi.PGH_COH_NONE
    seed: int = 1
    E_matrix: float = 3.5e3
    nu_matrix: float = 0.35
    E_fiber: float = 7.0e4
    nu_fiber: float = 0.2
    sigma_c: float = 50.0
    delta_c: float = 1e-3
    K0: float = 1e6
    beta: float = 1.0
    alpha: float = 50.0
    strain: float = 0.02
    place: int = 0               # PGH_PLACE_* flags of the inclusion placement
    thermal: bool = False        # a scalar field (heat conduction): one dof per node
    kappa_matrix: float = 0.2    # conductivities of the thermal problem
    kappa_fiber: float = 1.0
    name: str = "rve"

    def params(self) -> PghRveParams:
        p = PghRveParams()
        p.dim, p.n, p.n_incl, p.cohesive = self.dim, self.n, self.n_incl, self.cohesive
        p.r_min, p.r_max = self.r_min, self.r_max
        h = 1.0 / self.n
        p.gap = self.gap_h * h * (math.sqrt(3.0) if self.dim == 3 else 1.0)
        p.seed = self.seed
        p.E_matrix, p.nu_matrix, p.E_fiber, p.nu_fiber = self.E_matrix, self.nu_matrix, self.E_fiber, self.nu_fiber
        p.sigma_c = self.sigma_c
        p.delta_c = self.delta_c
        # bilinear: peak opening sigma_c / K0, full separation delta_c
        p.delta_c1 = self.sigma_c / self.K0
        p.delta_c2 = self.delta_c
        p.beta, p.alpha, p.strain = self.beta, self.alpha, self.strain
        p.max_tries = 0
        p.flags = self.place | (_ffi.PGH_SCALAR_DOFS if self.thermal else 0)
        p.kappa_matrix, p.kappa_fiber = self.kappa_matrix, self.kappa_fiber
        return p


# The five workloads (sizes and laws of the reference RVE studies).
def config(name: str) -> RVEConfig:
    name = name.upper()
    if name == "C1":   # 100x100 tri3, 16 fibers, perfect bond
        return RVEConfig(dim=2, n=100, n_incl=16, r_min=0.03, r_max=0.05, seed=1, name="C1")
    if name == "C2":   # 500x500 tri3, 200 fibers, bilinear cohesive
        return RVEConfig(dim=2, n=500, n_incl=200, r_min=0.01, r_max=0.02, seed=2,
                         cohesive=_ffi.PGH_COH_BL, name="C2")
    if name == "C3":   # 2000x2000 tri3, 3000 fibers, Ortiz-Pandolfi cohesive
        return RVEConfig(dim=2, n=2000, n_incl=3000, r_min=0.004, r_max=0.008, seed=3,
                         cohesive=_ffi.PGH_COH_OP, name="C3")
    if name == "C4":   # 120^3 cubes x 6 tet4, 500 spheres, perfect bond
        return RVEConfig(dim=3, n=120, n_incl=500, r_min=0.02, r_max=0.05, seed=4, name="C4")
    raise KeyError(name)


def small_config(kind: str = "c0", n: int = 24, n_incl: int = 3, seed: int = 11) -> RVEConfig:
    """Small RVEs for parity tests (oracle finishes in milliseconds)."""
    if kind == "c0":
        return RVEConfig(dim=2, n=n, n_incl=n_incl, r_min=0.08, r_max=0.14, seed=seed, name=f"small-c0-{n}")
    if kind == "bl":
        return RVEConfig(dim=2, n=n, n_incl=n_incl, r_min=0.08, r_max=0.14, seed=seed,
                         cohesive=_ffi.PGH_COH_BL, name=f"small-bl-{n}")
    if kind == "op":
        return RVEConfig(dim=2, n=n, n_incl=n_incl, r_min=0.08, r_max=0.14, seed=seed,
                         cohesive=_ffi.PGH_COH_OP, name=f"small-op-{n}")
    if kind == "tet":
        return RVEConfig(dim=3, n=n, n_incl=n_incl, r_min=0.12, r_max=0.2, seed=seed, name=f"small-tet-{n}")
    if kind in ("multi", "multi-op", "multi-bl"):
        # 120 small circles packed with a gap of 0.05 h: tri3 elements cut by
        # two inclusions (a matrix pentagon, or a matrix band between two
        # children) and by three (three triangles around a matrix hexagon)
        coh = {"multi": _ffi.PGH_COH_NONE, "multi-op": _ffi.PGH_COH_OP, "multi-bl": _ffi.PGH_COH_BL}[kind]
        return RVEConfig(dim=2, n=n, n_incl=120 if n_incl == 3 else n_incl, r_min=0.02, r_max=0.035, gap_h=0.05,
                         seed=1 if seed == 11 else seed, cohesive=coh, place=_ffi.PGH_PLACE_BEST_EFFORT,
                         name=f"small-{kind}-{n}")
    if kind in ("tet-op-q", "tet-bl-q"):
        # spheres anywhere: the interfaces cut tets one node against three
        # (CohTri3) and two against two (CohQuad4)
        return RVEConfig(dim=3, n=n, n_incl=n_incl, r_min=0.12, r_max=0.2, seed=seed,
                         cohesive=_ffi.PGH_COH_OP if kind == "tet-op-q" else _ffi.PGH_COH_BL, name=f"small-{kind}-{n}")
    if kind in ("tet-op", "tet-bl"):
        # spheres centred at mesh nodes, r < h / 2: CohTri3 interfaces only
        h = 1.0 / n
        return RVEConfig(dim=3, n=n, n_incl=n_incl, r_min=0.25 * h, r_max=0.4 * h, gap_h=0.5, seed=seed,
                         cohesive=_ffi.PGH_COH_OP if kind == "tet-op" else _ffi.PGH_COH_BL,
                         place=_ffi.PGH_PLACE_AT_NODES, name=f"small-{kind}-{n}")
    raise KeyError(kind)


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(int(n),)).view(dtype)


def batch_config(n_inst: int = 256, seed: int = 5, n_min: int = 64, n_max: int = 512) -> list:
    """The batch of independent cohesive RVEs: mesh n x n with n log-uniform in
    [n_min, n_max], 5-150 fibers each (count capped at an expected area fraction
    of 0.45), radius U[4h, 10h], bilinear cohesive law of the 500x500 study,
    one seed per instance."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_inst):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        h = 1.0 / n
        r_min, r_max = 4 * h, 10 * h
        mean_area = math.pi * (r_min * r_min + r_min * r_max + r_max * r_max) / 3.0
        nf = int(rng.integers(5, 151))
        nf = max(1, min(nf, int(0.45 / mean_area)))
        out.append(RVEConfig(dim=2, n=n, n_incl=nf, r_min=r_min, r_max=r_max, seed=seed * 100003 + i,
                             cohesive=_ffi.PGH_COH_BL, name=f"C5-{i}"))
    return out


class RVEProblem:
    """A built RVE (or a batch of RVEs merged into one block-diagonal problem):
    the static problem description plus numpy views of it."""

    def __init__(self, cfg: RVEConfig, _cfgs=None, device_pattern: bool = False, device_setup: bool = False):
        """device_pattern: build the pattern stage (incidence, node-pair and
        AIJ patterns, scatter indices) on the GPU; device_setup: run node
        detection, element analysis and enrichment on the GPU.  Either result
        is identical to the host build."""
        self.cfg = cfg
        L = lib()
        h = C.c_void_p()
        dev = (_ffi.PGH_PATTERN_DEVICE if device_pattern else 0) | (_ffi.PGH_SETUP_DEVICE if device_setup else 0)
        if _cfgs is None:
            self._params = cfg.params()
            self._params.flags |= dev
            check(L.pgh_build_rve(C.byref(self._params), C.byref(h)), "pgh_build_rve")
        else:
            arr = (PghRveParams * len(_cfgs))()
            for i, c in enumerate(_cfgs):
                arr[i] = c.params()
                arr[i].flags = c.place | _ffi.PGH_PLACE_BEST_EFFORT | dev
                arr[i].max_tries = 2_000_000
            self._params = arr
            check(L.pgh_build_batch(arr, len(_cfgs), C.byref(h)), "pgh_build_batch")
        self._h = h
        self.desc = PgProblemDesc()
        check(L.pgh_get_desc(h, C.byref(self.desc)), "pgh_get_desc")
        d = self.desc
        self.dim, self.nen = d.dim, d.nen
        self.nndof = d.nndof if d.nndof else d.dim   # dofs per node
        self.n_nodes, self.n_enr, self.n_elem, self.n_cut = d.n_nodes, d.n_enr, d.n_elem, d.n_cut
        self.n_free, self.n_pres, self.n_cqp = d.n_free, d.n_pres, d.n_cqp
        self.n_dof = self.n_free + self.n_pres
        self.nnz = [int(d.nnz[b]) for b in range(4)]
        self.nnz_total = sum(self.nnz)
        nall = self.n_nodes + self.n_enr
        self.coords = _arr(d.coords, nall * self.dim, np.float64).reshape(nall, self.dim)
        self.conn = _arr(d.conn, self.n_elem * self.nen, np.int32).reshape(self.n_elem, self.nen)
        self.elem_mat = _arr(d.elem_mat, self.n_elem, np.int32)
        self.cut_of_elem = _arr(d.cut_of_elem, self.n_elem, np.int32)
        self.cut_elem = _arr(d.cut_elem, self.n_cut, np.int32)
        self.node_dof = _arr(d.node_dof, nall * self.nndof, np.int32).reshape(nall, self.nndof)
        rows = [self.n_free, self.n_free, self.n_pres, self.n_pres]
        self.rp = [_arr(d.rp[b], rows[b] + 1, np.int32) for b in range(4)]
        self.ci = [_arr(d.ci[b], self.nnz[b], np.int32) for b in range(4)]
        self.n_pairs = int(d.n_pairs)
        self.npat_ptr = _arr(d.npat_ptr, nall + 1, np.int32)
        self.npat_col = _arr(d.npat_col, self.n_pairs, np.int32)
        self.npat_row = _arr(d.npat_row, self.n_pairs, np.int32)
        self.npat_map = _arr(d.npat_map, self.n_pairs * self.nndof * self.nndof, np.int32)
        self.n2e_ptr = _arr(d.n2e_ptr, nall + 1, np.int32)
        self.n2e = _arr(d.n2e, int(self.n2e_ptr[-1]), np.int32)
        self.node_detect = _arr(L.pgh_node_detect(h), self.n_nodes, np.int32)
        ei = L.pgh_enrich_inclusion(h)
        self.enr_incl = _arr(ei, self.n_enr, np.int32) if ei else None
        ep, ap = C.POINTER(C.c_int32)(), C.POINTER(C.c_double)()
        if L.pgh_enrich_edges(h, C.byref(ep), C.byref(ap)) == 0 and self.n_enr:
            self.enr_edge = _arr(ep, 2 * self.n_enr, np.int32).reshape(self.n_enr, 2)
            self.enr_alpha = _arr(ap, self.n_enr, np.float64)
        else:
            self.enr_edge = np.zeros((0, 2), np.int32) if self.n_enr == 0 else None
            self.enr_alpha = np.zeros(0) if self.n_enr == 0 else None
        self.bc_values = _arr(L.pgh_bc_values(h), self.n_pres, np.float64)
        cptr, rptr, ni = C.POINTER(C.c_double)(), C.POINTER(C.c_double)(), C.c_int32()
        check(L.pgh_inclusions(h, C.byref(cptr), C.byref(rptr), C.byref(ni)), "pgh_inclusions")
        self.incl_centers = _arr(cptr, ni.value * self.dim, np.float64).reshape(ni.value, self.dim)
        self.incl_radii = _arr(rptr, ni.value, np.float64)
        self.cut = [d.cut[i] for i in range(self.n_cut)] if self.n_cut < 200000 else None
        ni, off = C.c_int32(), C.POINTER(C.c_int64)()
        check(L.pgh_batch_info(h, C.byref(ni), C.byref(off)), "pgh_batch_info")
        self.n_inst = ni.value
        self.inst_off = (_arr(off, (ni.value + 1) * _ffi.PGH_BATCH_COLS, np.int64)
                         .reshape(ni.value + 1, _ffi.PGH_BATCH_COLS) if ni.value else None)

    @classmethod
    def batch(cls, cfgs, device_pattern: bool = False, device_setup: bool = False):
        """Merge independent RVEs into one problem (pgh_build_batch)."""
        return cls(cfgs[0], _cfgs=list(cfgs), device_pattern=device_pattern, device_setup=device_setup)

    STAGES = ("mesh", "place", "detect", "enrich", "dofs", "pattern", "instances")

    def stage_seconds(self) -> dict:
        """Wall seconds of each set-up stage (pgh_stage_seconds); 'instances'
        is the parallel per-instance build of a batch."""
        out = (C.c_double * _ffi.PGH_N_STAGES)()
        check(lib().pgh_stage_seconds(self._h, out, _ffi.PGH_N_STAGES), "pgh_stage_seconds")
        return {k: float(out[i]) for i, k in enumerate(self.STAGES)}

    def instance_ranges(self, i: int):
        """Value ranges of instance i in the concatenated K [ff|fp|pf|pp] and P [f|p]."""
        o, q = self.inst_off[i], self.inst_off[i + 1]
        f0, f1, p0, p1 = int(o[3]), int(q[3]), int(o[4]), int(q[4])
        base = np.cumsum([0] + self.nnz)
        rows = [(f0, f1), (f0, f1), (p0, p1), (p0, p1)]
        k = [(int(base[b] + self.rp[b][r0]), int(base[b] + self.rp[b][r1])) for b, (r0, r1) in enumerate(rows)]
        p = [(f0, f1), (self.n_free + p0, self.n_free + p1)]
        return k, p

    def instance_elements(self, i: int) -> int:
        return int(self.inst_off[i + 1][0] - self.inst_off[i][0])

    def cut_record(self, i: int):
        return self.desc.cut[i]

    def edge_inclusions(self) -> np.ndarray:
        """Inclusion of each 2D cohesive edge, at row isv_off / 2 + k for edge k
        of a cut record: the inclusion of the enrichment node that is the edge's
        first lower node (Elem::get_inclusion_from_enrich, as
        src/WriterCohesiveOpenings.cpp:128 takes it)."""
        out = np.zeros(self.n_cqp // 2, np.int64)
        if self.enr_incl is None:
            raise ValueError("the inclusions of a mesh set up elsewhere are not known")
        for c in range(self.n_cut):
            r = self.desc.cut[c]
            for k in range(r.n_coh):
                g = int(r.enr[r.coh_nodes[k][0] - self.nen])
                out[r.isv_off // 2 + k] = int(self.enr_incl[g - self.n_nodes])
        return out

    def random_u(self, seed: int = 0, scale: float = 1e-3) -> np.ndarray:
        """U ~ U[-scale, scale] on every dof (free and prescribed, regular and enrichment)."""
        rng = np.random.default_rng(seed)
        return rng.uniform(-scale, scale, self.n_dof)

    def stats(self) -> dict:
        return {
            "elements": self.n_elem, "nodes": self.n_nodes, "enrichment_nodes": self.n_enr,
            "cut_elements": self.n_cut, "cohesive_qps": self.n_cqp, "dofs": self.n_dof,
            "free_dofs": self.n_free, "nnz": self.nnz_total, "node_pairs": self.n_pairs,
            "inclusions": int(len(self.incl_radii)),
        }

    def close(self):
        if self._h:
            lib().pgh_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
