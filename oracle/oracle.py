"""ctypes wrapper of the CPU oracle (test infrastructure only).

`Oracle(problem)` builds its own 4-block AIJ pattern by inserting every element
dof block (src/Assembler.cpp:659-679 semantics) and caches shape data
(src/Mesh.cpp:1475-1643, src/Assembler.cpp:602-621); `assemble` restates
Assembler::assemble (src/Assembler.cpp:65-368) in FP64.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "liboracle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)
    return SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        L = C.CDLL(SO)
        L.oracle_new.restype = C.c_void_p
        L.oracle_new.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.oracle_free.argtypes = [C.c_void_p]
        L.oracle_error.restype = C.c_char_p
        L.oracle_error.argtypes = [C.c_void_p]
        L.oracle_nnz.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        L.oracle_pattern.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.POINTER(C.c_int32)),
                                     C.POINTER(C.POINTER(C.c_int32))]
        L.oracle_assemble.restype = C.c_int
        L.oracle_assemble.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                      C.c_void_p, C.c_void_p]
        L.oracle_assemble_threads.restype = C.c_int
        L.oracle_assemble_threads.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                              C.c_void_p, C.c_void_p, C.c_int]
        _lib = L
    return _lib


class Oracle:
    def __init__(self, problem, e0: int = 0, e1: int = -1):
        self.problem = problem
        L = lib()
        self._h = L.oracle_new(C.addressof(problem.desc), int(e0), int(e1))
        if not self._h:
            raise MemoryError("oracle_new failed")
        err = L.oracle_error(self._h)
        if err:
            raise RuntimeError("oracle set-up: " + err.decode())
        nnz = (C.c_int64 * 4)()
        L.oracle_nnz(self._h, nnz)
        self.nnz = [int(x) for x in nnz]
        rows = [problem.n_free, problem.n_free, problem.n_pres, problem.n_pres]
        self.rp, self.ci = [], []
        for b in range(4):
            rp, ci = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)()
            L.oracle_pattern(self._h, b, C.byref(rp), C.byref(ci))
            self.rp.append(np.ctypeslib.as_array(rp, shape=(rows[b] + 1,)).copy())
            self.ci.append(np.ctypeslib.as_array(ci, shape=(max(self.nnz[b], 1),))[: self.nnz[b]].copy())

    def assemble(self, U, isv_committed=None, func=True, jac=True, threads=0, temp_change=0.0):
        """threads = 0: the element loop in element order (the checker);
        threads > 0: colour by colour on that many OpenMP threads (the threaded
        CPU baseline; equal to the serial result up to rounding).
        temp_change: the thermal pre-load's current temperature change."""
        p = self.problem
        _set_temp(self._h, temp_change)
        U = np.ascontiguousarray(U, dtype=np.float64)
        K = np.zeros(sum(self.nnz), dtype=np.float64)
        P = np.zeros(p.n_dof, dtype=np.float64)
        ic = np.ascontiguousarray(isv_committed if isv_committed is not None else np.zeros(p.n_cqp), dtype=np.float64)
        it = np.zeros(max(p.n_cqp, 1), dtype=np.float64)
        args = (self._h, U.ctypes.data, ic.ctypes.data if p.n_cqp else None, it.ctypes.data, int(func), int(jac),
                K.ctypes.data, P.ctypes.data)
        rc = lib().oracle_assemble_threads(*args, int(threads)) if threads > 0 else lib().oracle_assemble(*args)
        if rc != 0:
            raise RuntimeError("oracle_assemble: " + lib().oracle_error(self._h).decode())
        return K, P, it[: p.n_cqp]

    def sens_dPdd(self, prm, U, isv=None, isv_sens=None, flags=0):
        """dP/dd of one material parameter (prm: a PgSensParam), the oracle's
        restatement of SensitivityParameter::assemble_dPdd.  Raises
        NotImplementedError where the reference stops with an error."""
        p = self.problem
        L = lib()
        L.oracle_sens_dPdd.restype = C.c_int
        L.oracle_sens_dPdd.argtypes = [C.c_void_p, C.c_void_p, C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p]
        U = np.ascontiguousarray(U, dtype=np.float64)
        ic = np.ascontiguousarray(isv if isv is not None else np.zeros(max(p.n_cqp, 1)), dtype=np.float64)
        iS = np.ascontiguousarray(isv_sens if isv_sens is not None else np.zeros(max(p.n_cqp, 1)), dtype=np.float64)
        dP = np.zeros(p.n_dof)
        rc = L.oracle_sens_dPdd(self._h, C.addressof(prm), int(flags), U.ctypes.data, ic.ctypes.data,
                                iS.ctypes.data, dP.ctypes.data)
        if rc == -2:
            raise NotImplementedError(L.oracle_error(self._h).decode())
        if rc != 0:
            raise RuntimeError("oracle_sens_dPdd: " + L.oracle_error(self._h).decode())
        return dP

    def close(self):
        if self._h:
            lib().oracle_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _set_temp(h, temp_change):
    L = lib()
    L.oracle_set_temperature_change.restype = None
    L.oracle_set_temperature_change.argtypes = [C.c_void_p, C.c_double]
    L.oracle_set_temperature_change(h, float(temp_change))


def assemble_in_ranges(problem, U, isv_committed=None, chunk=2_000_000, threads=0, temp_change=0.0):
    """The oracle's assembly of the whole mesh into the harness's own 4-block
    pattern (problem.rp / problem.ci), element range by element range, so that
    only one range's shape cache is held at a time.  threads = 0 gives the
    serial element loop's result bit for bit; threads > 0 runs each range
    colour by colour.  Returns (K, P, trial ISVs)."""
    L = lib()
    L.oracle_new_desc_pattern.restype = C.c_void_p
    L.oracle_new_desc_pattern.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.oracle_assemble_range.restype = C.c_int
    L.oracle_assemble_range.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                        C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    p = problem
    U = np.ascontiguousarray(U, dtype=np.float64)
    K = np.zeros(max(p.nnz_total, 1), dtype=np.float64)
    P = np.zeros(p.n_dof, dtype=np.float64)
    ic = np.ascontiguousarray(isv_committed if isv_committed is not None else np.zeros(p.n_cqp), dtype=np.float64)
    it = np.zeros(max(p.n_cqp, 1), dtype=np.float64)
    for e0 in range(0, p.n_elem, chunk):
        h = L.oracle_new_desc_pattern(C.addressof(p.desc), e0, min(p.n_elem, e0 + chunk))
        if not h:
            raise MemoryError("oracle_new_desc_pattern failed")
        try:
            err = L.oracle_error(h)
            if err:
                raise RuntimeError("oracle set-up: " + err.decode())
            _set_temp(h, temp_change)
            rc = L.oracle_assemble_range(h, U.ctypes.data, ic.ctypes.data if p.n_cqp else None, it.ctypes.data, 1, 1,
                                         K.ctypes.data, P.ctypes.data, int(e0 == 0), int(threads))
            if rc != 0:
                raise RuntimeError("oracle_assemble_range: " + L.oracle_error(h).decode())
        finally:
            L.oracle_free(h)
    return K[: p.nnz_total], P, it[: p.n_cqp]


def cohesive_law(mat, dim, delta, isv=0.0):
    """The oracle's cohesive constitutive update: returns (delta_used, isv_out, t, D)."""
    L = lib()
    L.oracle_cohesive_law.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    d = np.ascontiguousarray(np.array(delta, dtype=np.float64))
    h = np.array([isv], dtype=np.float64)
    t = np.zeros(dim)
    D = np.zeros(dim * dim)
    L.oracle_cohesive_law(C.addressof(mat), dim, d.ctypes.data, h.ctypes.data, t.ctypes.data, D.ctypes.data)
    return d, float(h[0]), t, D.reshape(dim, dim)


def child_qp(dim, ctype, X, q):
    """Reference point, weight, shape values, physical gradients and Jacobian
    of qp q of a child of type ctype with node coordinates X (nn x dim)."""
    L = lib()
    L.oracle_child_qp.restype = C.c_int
    L.oracle_child_qp.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int] + [C.c_void_p] * 5
    nn = len(X)
    cx = np.zeros((nn, 3))
    cx[:, :dim] = X
    xi, w, N, g, J = np.zeros(3), np.zeros(1), np.zeros(nn), np.zeros((nn, 3)), np.zeros(1)
    rc = L.oracle_child_qp(dim, ctype, cx.ctypes.data, q, xi.ctypes.data, w.ctypes.data, N.ctypes.data,
                           g.ctypes.data, J.ctypes.data)
    if rc:
        raise ValueError("qp out of range")
    return xi[:dim], float(w[0]), N, g[:, :dim], float(J[0])


def polygon_shape(nn, xi):
    """Polygon shape values and reference gradients at reference point xi."""
    L = lib()
    L.oracle_polygon_shape.restype = None
    L.oracle_polygon_shape.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    x = np.ascontiguousarray(xi, dtype=np.float64)
    N, dN = np.zeros(nn), np.zeros((nn, 2))
    L.oracle_polygon_shape(nn, x.ctypes.data, N.ctypes.data, dN.ctypes.data)
    return N, dN


def elem_cache(orc, e):
    L = lib()
    L.oracle_elem_cache.restype = C.c_int64
    L.oracle_elem_cache.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.POINTER(C.c_double))]
    ptr = C.POINTER(C.c_double)()
    n = L.oracle_elem_cache(orc._h, int(e), C.byref(ptr))
    if n < 0:
        raise IndexError(e)
    return np.ctypeslib.as_array(ptr, shape=(int(n),)).copy()


def cohesive_openings(problem, U):
    """Cohesive openings of a 2D cohesive mesh, restated from
    WriterCohesiveOpenings::writeFromStream (src/WriterCohesiveOpenings.cpp:73-147):
    for cohesive edge k of cut record c, row isv_off / 2 + k holds the angle of the
    midpoint of its lower nodes about the centre of its inclusion, then the opening
    R . sum_n N_n (U_upper - U_lower) at N = (1/2, 1/2) (cohesiveDeltaFast,
    src/Utilities.cpp:528-545) in the edge frame of CohEdge2
    (src/CohesiveElem.cpp:145-238: t = dx/dr over |t|, n = (-t_y, t_x) over |n|),
    tangential then normal, each over delta_c of the Ortiz-Pandolfi law."""
    Un = np.asarray(U, np.float64)[problem.node_dof]
    rows = np.zeros((problem.n_cqp // 2, 3))
    inc = np.zeros(problem.n_cqp // 2, np.int64)
    for c in range(problem.n_cut):
        r = problem.cut_record(c)
        e = int(problem.cut_elem[c])
        gl = [int(x) for x in problem.conn[e]] + [int(r.enr[k]) for k in range(r.n_enr)]
        for k in range(r.n_coh):
            # the inclusion of the edge's first lower (enrichment) node's cut-edge
            # slot, el->get_inclusion_from_enrich (src/WriterCohesiveOpenings.cpp:128)
            ic = int(problem.enr_incl[gl[r.coh_nodes[k][0]] - problem.n_nodes])
            dc = problem.desc.mats[r.coh_mat[k]].delta_c
            lo = [gl[r.coh_nodes[k][n]] for n in range(2)]
            up = [gl[r.coh_nodes[k][n + 2]] for n in range(2)]
            x0, x1 = problem.coords[lo[0]], problem.coords[lo[1]]
            t = 0.5 * (x1 - x0)
            nrm = np.array([-t[1], t[0]])
            R = np.array([t / np.linalg.norm(t), nrm / np.linalg.norm(nrm)])
            xm = 0.5 * x0 + 0.5 * x1
            dl = 0.5 * (Un[up[0]] - Un[lo[0]]) + 0.5 * (Un[up[1]] - Un[lo[1]])
            d = R @ dl
            cen = problem.incl_centers[ic]
            row = r.isv_off // 2 + k
            rows[row] = (np.arctan2(xm[1] - cen[1], xm[0] - cen[0]), d[0] / dc, d[1] / dc)
            inc[row] = ic
    return inc, rows


def cohesive_failure(problem, U):
    """Per-edge data of WriterCohesiveFailureLengths, restated from store_mesh and
    writeFromStream (src/WriterCohesiveFailureLengths.cpp:64-86,183-224): for
    cohesive edge k of cut record c, row isv_off / 2 + k holds the distance between
    its two lower nodes and the mean over the two Gauss points (xi = -+1/sqrt(3),
    src/CohEdge2.cpp:70-78) of getNormalizedCohesiveFailure of the opening in the
    edge frame: delta_eff / delta_c for OPCohesiveNU
    (src/Material_OPCohesiveNU.cpp:43-60,217-221), delta_eff / delta_c1 for
    BilinearCohesiveNU (src/Material_BilinearCohesiveNU.cpp:48-65,250-254), 0 for
    the other laws (include/material.h:123); delta_eff = sqrt(d_n^2 + beta2 d_t^2)
    after an opening of squared norm below 1e-40 is replaced by 1e-20 in each
    component."""
    Un = np.asarray(U, np.float64)[problem.node_dof]
    rows = np.zeros((problem.n_cqp // 2, 2))
    g = 1.0 / np.sqrt(3.0)
    for c in range(problem.n_cut):
        r = problem.cut_record(c)
        e = int(problem.cut_elem[c])
        gl = [int(x) for x in problem.conn[e]] + [int(r.enr[k]) for k in range(r.n_enr)]
        for k in range(r.n_coh):
            m = problem.desc.mats[r.coh_mat[k]]
            lo = [gl[r.coh_nodes[k][n]] for n in range(2)]
            up = [gl[r.coh_nodes[k][n + 2]] for n in range(2)]
            x0, x1 = problem.coords[lo[0]], problem.coords[lo[1]]
            t = 0.5 * (x1 - x0)
            R = np.array([t / np.linalg.norm(t), np.array([-t[1], t[0]]) / np.linalg.norm(t)])
            f = 0.0
            for xi in (-g, g):
                N = (0.5 * (1 - xi), 0.5 * (1 + xi))
                d = R @ (N[0] * (Un[up[0]] - Un[lo[0]]) + N[1] * (Un[up[1]] - Un[lo[1]]))
                if d @ d < 1e-40:
                    d = np.full(2, 1e-20)
                de = np.sqrt(d[1] ** 2 + m.beta2 * d[0] ** 2)
                f += de / m.delta_c if m.type == 4 else de / m.delta_c1 if m.type == 2 else 0.0
            rows[r.isv_off // 2 + k] = (np.linalg.norm(x1 - x0), f / 2)
    return rows


def plot_fields(problem, U):
    """The fields of the VTK writers, restated in numpy (the checker of
    pg_asm_plot_fields): Writer_VTK::gather_nodal_data with
    Problem::interpolate_nodal_enrichment (src/Writer_VTK.cpp:642-727,
    src/Problem.cpp:538-570) for the nodes, and per VTK cell (each uncut element,
    each integration child) the sums over its qps of J w, J w strain and
    J w stress (Integrator::integrateElem, include/Integrator.h:151-228) from the
    oracle's cached B matrices; the stress is the LEI law at that strain."""
    p = problem
    dim, nstr = p.dim, (3 if p.dim == 2 else 6)
    Un = np.asarray(U, np.float64)[p.node_dof]
    nodes = Un.copy()
    for k in range(p.n_enr):
        n1, n2 = (int(x) for x in p.enr_edge[k])
        a = float(p.enr_alpha[k])
        nodes[p.n_nodes + k] = Un[n1] * (1.0 - a) + Un[n2] * a + Un[p.n_nodes + k] * 1.0
    o = Oracle(p)
    rows = []
    for e in range(p.n_elem):
        c = int(p.cut_of_elem[e])
        if c < 0:
            local = [int(x) for x in p.conn[e]]
            groups = [(int(p.elem_mat[e]), list(range(4 if dim == 2 else 5)))]
        else:
            r = p.cut_record(c)
            local = [int(x) for x in p.conn[e]] + [int(r.enr[k]) for k in range(r.n_enr)]
            groups, q0 = [], 0
            for k in range(r.n_child):
                nq = {0: 4, 1: 4, 2: 5, 3: 20, 4: 24}[int(r.child_type[k])]
                groups.append((int(r.child_mat[k]), list(range(q0, q0 + nq))))
                q0 += nq
        Ue = Un[local].ravel()
        nd = len(Ue)
        cache = elem_cache(o, e)
        stride = 2 + nstr * nd
        for m, qps in groups:
            D = np.array(p.desc.mats[m].D[: nstr * nstr]).reshape(nstr, nstr)
            vol, ie, is_ = 0.0, np.zeros(nstr), np.zeros(nstr)
            for q in qps:
                W, J = cache[q * stride], cache[q * stride + 1]
                B = cache[q * stride + 2: (q + 1) * stride].reshape(nstr, nd)
                eps = B @ Ue
                wgt = J * W
                vol += wgt * 1.0
                ie += wgt * eps
                is_ += wgt * (D @ eps)
            rows.append(np.concatenate([[vol], ie, is_]))
    o.close()
    return nodes, np.array(rows)
