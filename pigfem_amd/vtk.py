"""Legacy VTK output from device data: PIGFEM's Writer_ASCII_VTK and
Writer_Binary_VTK (src/Writer_ASCII_VTK.cpp, src/Writer_Binary_VTK.cpp) over the
fields that Writer_VTK gathers (src/Writer_VTK.cpp:164-1000), computed on the GPU
by pg_asm_plot_fields.

The file layout, the float conversions and the text formats are the reference's:
points and displacements as float (regular nodes, then enrichment nodes), one
cell per uncut element and per integration child of a cut element, the cell's
material index, its average strain and stress (float(integral) scaled by
float(1 / float(volume))), and the von Mises stress of the float average.  ASCII
numbers print as C++ streams print them (6 significant digits); binary data is
big-endian.  Cohesive cells (the writer's "COHESIVE" option) and the partition
field are not written.
"""
from __future__ import annotations

import math

import numpy as np

from . import _ffi

# Writer_VTK::create_map (include/Writer_VTK.h:74-91)
_VTK_TYPE = {_ffi.PG_CHILD_TRI3: 5, _ffi.PG_CHILD_QUAD4: 9, _ffi.PG_CHILD_TET4: 10, _ffi.PG_CHILD_POLY5: 7,
             _ffi.PG_CHILD_POLY6: 7}
_VOIGT = {3: ["11", "22", "12"], 6: ["11", "22", "33", "23", "13", "12"]}


def vtk_cells(problem):
    """Writer_VTK::store_elements: per cell its node list (VTK point ids), VTK
    type and material index, elements in order, a cut element's integration
    children in child order."""
    p = problem
    conn, types, mats = [], [], []
    parent_type = 5 if p.dim == 2 else 10
    for e in range(p.n_elem):
        c = int(p.cut_of_elem[e])
        if c < 0:
            conn.append([int(x) for x in p.conn[e]])
            types.append(parent_type)
            mats.append(int(p.elem_mat[e]))
            continue
        r = p.cut_record(c)
        local = [int(x) for x in p.conn[e]] + [int(r.enr[k]) for k in range(r.n_enr)]
        for k in range(r.n_child):
            conn.append([local[l] for l in _ffi.child_nodes(r, p.dim, k)])
            types.append(_VTK_TYPE[int(r.child_type[k])])
            mats.append(int(r.child_mat[k]))
    return conn, types, mats


def cell_averages(cells, nstr):
    """The writer's averages: float(integral) * float(1. / float(volume))
    (Writer_VTK::gather_element_data with Utilities::dtos and scale)."""
    vol = cells[:, 0].astype(np.float32)
    inv = (1.0 / vol.astype(np.float64)).astype(np.float32)
    strain = cells[:, 1:1 + nstr].astype(np.float32) * inv[:, None]
    stress = cells[:, 1 + nstr:1 + 2 * nstr].astype(np.float32) * inv[:, None]
    return strain.astype(np.float32), stress.astype(np.float32)


def von_mises(s):
    """Utilities::von_mises of a float stress (src/Utilities.cpp:857-867): the
    float products and sums as C++ evaluates them, the 3.0 / 0.5 / 6.0 terms and
    pow in double, the result as float."""
    f = np.float32
    if len(s) == 3:
        a = f(f(s[0] * s[0]) - f(s[0] * s[1]))
        a = f(a + f(s[1] * s[1]))
        return f(math.sqrt(float(a) + 3.0 * float(s[2]) * float(s[2])))
    d01, d12, d20 = float(f(s[0] - s[1])), float(f(s[1] - s[2])), float(f(s[2] - s[0]))
    return f(math.sqrt(0.5 * (d01 ** 2 + d12 ** 2 + d20 ** 2 +
                              6.0 * (float(s[3]) ** 2 + float(s[4]) ** 2 + float(s[5]) ** 2))))


def _g(x) -> str:
    """A float or double printed by a default C++ ostream (%g, 6 digits)."""
    return "%g" % float(x)


def write_vtk(filename: str, problem, nodes, cells, curr_t: float = 0.0, binary: bool = False,
              igfem: bool = True) -> None:
    """Write the legacy VTK file of Writer_ASCII_VTK (binary=False) or
    Writer_Binary_VTK (binary=True) for one rank.  nodes, cells: the output of
    AssemblerGPU.plot_fields."""
    p = problem
    nall = p.n_nodes + p.n_enr
    nstr = 3 if p.dim == 2 else 6
    conn, types, mats = vtk_cells(p)
    n_cells = len(conn)
    if cells.shape[0] != n_cells:
        raise ValueError("cell rows do not match the mesh")
    n_points = sum(len(c) for c in conn)
    xyz = np.zeros((nall, 3), np.float32)
    xyz[:, :p.dim] = p.coords.astype(np.float32)
    disp = np.zeros((nall, 3), np.float32)
    disp[:, :p.dim] = np.asarray(nodes, np.float64).astype(np.float32)
    strain, stress = cell_averages(np.asarray(cells, np.float64), nstr)
    vm = np.array([von_mises(s) for s in stress], np.float32)
    title = f"{'IGFEM' if igfem else 'FEM'} {p.dim}D simulation: t={_g(curr_t)}"
    if not binary:
        out = ["# vtk DataFile Version 3.0\n", title + "\n", "ASCII\n\n", "DATASET UNSTRUCTURED_GRID\n",
               f"POINTS {nall} float\n"]
        out += [f"{_g(a)} {_g(b)} {_g(c)}\n" for a, b, c in xyz]
        out.append(f"\nCELLS {n_cells} {n_cells + n_points}\n")
        out += [f"{len(c)} " + " ".join(str(i) for i in c) + "\n" for c in conn]
        out.append(f"\nCELL_TYPES {n_cells}\n")
        out += [f"{t}\n" for t in types]
        out.append(f"\nPOINT_DATA {nall}\nVECTORS displacement float \n")
        out += [f"{_g(a)} {_g(b)} {_g(c)}\n" for a, b, c in disp]
        out.append(f"\nCELL_DATA {n_cells}\nSCALARS materials int\nLOOKUP_TABLE default\n")
        out += [f"{m}\n" for m in mats]
        for name, arr in [("e" + v, strain[:, i]) for i, v in enumerate(_VOIGT[nstr])] + \
                         [("s" + v, stress[:, i]) for i, v in enumerate(_VOIGT[nstr])] + [("von_mises", vm)]:
            out.append(f"\nSCALARS {name} float\nLOOKUP_TABLE default\n")
            out += [f"{_g(x)}\n" for x in arr]
        with open(filename, "w") as f:
            f.write("".join(out))
        return
    be_f, be_i, be_u = np.dtype(">f4"), np.dtype(">i4"), np.dtype(">u4")
    parts = [("# vtk DataFile Version 3.0\n" + title + "\nBINARY\nDATASET UNSTRUCTURED_GRID\n").encode(),
             f"POINTS {nall} float\n".encode(), xyz.astype(be_f).tobytes(),
             f"\nCELLS {n_cells} {n_cells + n_points}\n".encode()]
    parts += [np.array([len(c)] + c, dtype=be_u).tobytes() for c in conn]
    parts += [f"\nCELL_TYPES {n_cells}\n".encode(), np.array(types, dtype=be_i).tobytes(),
              f"\nPOINT_DATA {nall}\nVECTORS displacement float\n".encode(), disp.astype(be_f).tobytes(),
              f"\nCELL_DATA {n_cells}\nSCALARS materials int\nLOOKUP_TABLE default\n".encode(),
              np.array(mats, dtype=be_u).tobytes()]
    for name, arr in [("e" + v, strain[:, i]) for i, v in enumerate(_VOIGT[nstr])] + \
                     [("s" + v, stress[:, i]) for i, v in enumerate(_VOIGT[nstr])] + [("von_mises", vm)]:
        parts += [f"\nSCALARS {name} float\nLOOKUP_TABLE default\n".encode(), np.asarray(arr).astype(be_f).tobytes()]
    with open(filename, "wb") as f:
        f.write(b"".join(parts))
