// Set-up arithmetic shared by the host harness (harness.cpp, g++) and the
// device set-up (setup_dev.hip, hipcc): node detection, the element detection
// tables with the prism decomposition, and the edge intersection.  One source
// for both builds, on fixed-size arrays, so the device set-up reproduces the
// host set-up bit for bit (both compile to plain IEEE double operations: g++
// emits no fused multiply-add on x86-64 without -mfma, hipcc runs with
// -ffp-contract=off).  Citations are to brandyb2/PIGFEM.
#pragma once
#include <math.h>
#include <stdint.h>

#include "pgfem_asm.h"

#if defined(__HIPCC__)
#define PGH_HD __host__ __device__
#else
#define PGH_HD
#endif

namespace pgh {

// Error codes of the set-up arithmetic (the host turns them into PG_E_* codes
// and messages: cut_error_code, cut_error_text).
enum CutErr : int {
    CUT_OK = 0,
    CUT_E_TET_MULTI,     // tet4 element cut by more than one interface
    CUT_E_TET_TWO_INCL,  // tet4 element cut by two inclusions
    CUT_E_TET_1SEP,      // tet4: one node separated in an unknown manner
    CUT_E_TET_2SEP,      // tet4: two nodes separated in an unknown manner
    CUT_E_PRISM,         // impossible case for prism decomposition
    CUT_E_CHILDREN,      // too many integration children
    CUT_E_NOROOT,        // no valid intersection on a cut edge
    CUT_E_SURFACE,       // a node lies on an inclusion surface
};

inline int cut_error_code(int e) {
    return (e == CUT_E_TET_MULTI || e == CUT_E_TET_TWO_INCL || e == CUT_E_CHILDREN) ? PG_E_UNSUPPORTED : PG_E_MESH;
}

inline const char* cut_error_text(int e) {
    switch (e) {
        case CUT_E_TET_MULTI: return "tet4 element cut by more than one interface";
        case CUT_E_TET_TWO_INCL: return "tet4 element cut by two inclusions is not supported";
        case CUT_E_TET_1SEP: return "tet4 element has 1 node separated in an unknown manner";
        case CUT_E_TET_2SEP: return "tet4 element has 2 nodes separated in an unknown manner";
        case CUT_E_PRISM: return "impossible case for prism decomposition";
        case CUT_E_CHILDREN: return "too many integration children";
        case CUT_E_NOROOT: return "no valid intersection on a cut edge";
        case CUT_E_SURFACE: return "lies on an inclusion surface";
        default: return "set-up error";
    }
}

// Local edges: Tri3 src/tri3.cpp:22-27, Tet4 src/tet4.cpp:22-30.
PGH_HD inline int tri_edge(int e, int end) {
    return end == 0 ? e : (e == 2 ? 0 : e + 1);  // (0,1) (1,2) (2,0)
}
PGH_HD inline int tet_edge(int e, int end) {
    // (0,1) (1,2) (0,2) (0,3) (1,3) (2,3)
    const int a = (e == 1 || e == 4) ? 1 : (e == 5 ? 2 : 0);
    const int b = e == 0 ? 1 : (e == 1 || e == 2) ? 2 : 3;
    return end == 0 ? a : b;
}

// --------------------------------------------------------------------------
// Node detection, Mesh::detect_node (src/Mesh.cpp:2883-2930) for one
// inclusion: 0 outside, 1 on the surface, 2 inside.
// Circle_Inclusion::get_domain_lims / detect_node, src/InclusionCircle.cpp:141-163;
// Sphere_Inclusion, src/InclusionSphere.cpp:163-183 (the z upper bound uses center[1]).
// --------------------------------------------------------------------------
PGH_HD inline int detect_one(int dim, const double* c, double r, const double* x) {
    if (dim == 2) {
        if (!(c[0] - r <= x[0] && c[1] - r <= x[1] && x[0] <= c[0] + r && x[1] <= c[1] + r)) return 0;
        double dx = x[0] - c[0], dy = x[1] - c[1];
        double val = (sqrt(dx * dx + dy * dy) - r) / r;
        if (val > 1e-12) return 0;
        if (val < -1e-12) return 2;
        return 1;
    }
    if (!(c[0] - r <= x[0] && c[1] - r <= x[1] && c[2] - r <= x[2] && x[0] <= c[0] + r && x[1] <= c[1] + r &&
          x[2] <= c[1] + r))
        return 0;
    double gx = (x[0] - c[0]) / r, gy = (x[1] - c[1]) / r, gz = (x[2] - c[2]) / r;
    double val = gx * gx + gy * gy + gz * gz - 1.0;
    if (val > 1e-12) return 0;
    if (val < -1e-12) return 2;
    return 1;
}

// --------------------------------------------------------------------------
// Element detection tables, output in the cohesive (C-1) numbering:
//   Tri3::detection, polygon-decomposition branch (src/tri3.cpp:283-452): one
//   inclusion against the matrix (two children), two inclusions with or
//   without matrix nodes, three inclusions (three triangles around a hexagon
//   of matrix); children of 5 or 6 nodes are polygons.
//   Tet4::detection, prism-decomposition branch: src/tet4.cpp:493-558, then
//   Tet4::decomposePrisms / partitionPrism: src/tet4.cpp:673-870.
// child_val: the detection value (-1 matrix, k >= 0 inclusion k) whose
// material the child takes.
// --------------------------------------------------------------------------
struct CutDetect {
    int ncut;
    uint8_t cut_edge[PG_MAX_ENR];   // local edge index per slot
    int32_t slot_incl[PG_MAX_ENR];  // inclusion per slot
    int nchild;
    uint8_t child_nn[PG_MAX_CHILD];
    uint8_t child_nodes[PG_MAX_CHILD][6];
    int32_t child_val[PG_MAX_CHILD];
    int ncoh;
    uint8_t coh_nn;
    uint8_t coh_nodes[PG_MAX_COH][8];
};

PGH_HD inline void cd_child(CutDetect& D, int val, int n, int a, int b, int c, int d = 0, int e = 0, int f = 0) {
    const int k = D.nchild++;
    const int v[6] = {a, b, c, d, e, f};
    D.child_nn[k] = uint8_t(n);
    for (int j = 0; j < 6; ++j) D.child_nodes[k][j] = uint8_t(j < n ? v[j] : 0);
    D.child_val[k] = val;
}

PGH_HD inline void cd_coh(CutDetect& D, int n, int a, int b, int c, int d, int e = 0, int f = 0, int g = 0,
                          int h = 0) {
    const int k = D.ncoh++;
    const int v[8] = {a, b, c, d, e, f, g, h};
    D.coh_nn = uint8_t(n);
    for (int j = 0; j < 8; ++j) D.coh_nodes[k][j] = uint8_t(j < n ? v[j] : 0);
}

PGH_HD inline void cd_slots(CutDetect& D, int n, const int* edges, const int* incl) {
    D.ncut = n;
    for (int s = 0; s < n; ++s) {
        D.cut_edge[s] = uint8_t(edges[s]);
        D.slot_incl[s] = incl[s];
    }
}

PGH_HD inline void cd_clear(CutDetect& D) {
    D.ncut = 0;
    D.nchild = 0;
    D.ncoh = 0;
    D.coh_nn = 0;
}

// one inclusion against the matrix, src/tri3.cpp:336-367
PGH_HD inline void tri3_one(CutDetect& D, int ce0, int ce1, int sep, int other, int a0, int a1, int a2,
                            bool sep_lower) {
    const int m = sep > other ? sep : other;
    const int e[2] = {ce0, ce1}, in[2] = {m, m};
    cd_slots(D, 2, e, in);
    if (sep_lower) {
        cd_child(D, sep, 3, a0, 4, 3);
        cd_child(D, other, 4, a1, a2, 5, 6);
        cd_coh(D, 4, 3, 4, 5, 6);
    } else {
        cd_child(D, sep, 3, a0, 6, 5);
        cd_child(D, other, 4, a1, a2, 3, 4);
        cd_coh(D, 4, 4, 3, 6, 5);
    }
}

// Tri3 detection from the three nodes' detection values; false when uncut.
PGH_HD inline bool detect_tri3(const int nd[3], CutDetect& D) {
    cd_clear(D);
    const int n0 = nd[0], n1 = nd[1], n2 = nd[2];
    if (n0 == n1 && n1 == n2) return false;
    const int vmin = n0 < n1 ? (n0 < n2 ? n0 : n2) : (n1 < n2 ? n1 : n2);
    const int uniq = 1 + (n0 != n1 ? 1 : 0) + ((n2 != n0 && n2 != n1) ? 1 : 0);
    if (uniq == 2 && vmin < 0) {  // one inclusion against the matrix
        if (n1 == n2) tri3_one(D, 2, 0, n0, n1, 0, 1, 2, n0 < n1);         // node 0 separate
        else if (n0 == n2) tri3_one(D, 0, 1, n1, n0, 1, 2, 0, n1 < n0);    // node 1 separate
        else tri3_one(D, 1, 2, n2, n0, 2, 0, 1, n2 < n0);                  // node 2 separate
    } else if (uniq == 2) {  // two inclusions, no matrix node, src/tri3.cpp:370-398
        if (n1 == n2) {
            const int e[4] = {2, 0, 2, 0}, in[4] = {n0, n0, n1, n1};
            cd_slots(D, 4, e, in);
            cd_child(D, n0, 3, 0, 8, 7);
            cd_child(D, -1, 4, 4, 6, 5, 3);
            cd_child(D, n1, 4, 1, 2, 9, 10);
        } else if (n0 == n2) {
            const int e[4] = {0, 1, 0, 1}, in[4] = {n1, n1, n2, n2};
            cd_slots(D, 4, e, in);
            cd_child(D, n1, 3, 1, 8, 7);
            cd_child(D, -1, 4, 4, 6, 5, 3);
            cd_child(D, n2, 4, 2, 0, 9, 10);
        } else {
            const int e[4] = {1, 2, 1, 2}, in[4] = {n2, n2, n0, n0};
            cd_slots(D, 4, e, in);
            cd_child(D, n2, 3, 2, 8, 7);
            cd_child(D, -1, 4, 4, 6, 5, 3);
            cd_child(D, n0, 4, 0, 1, 9, 10);
        }
        cd_coh(D, 4, 4, 3, 8, 7);
        cd_coh(D, 4, 5, 6, 9, 10);
    } else if (vmin < 0) {  // two inclusions and a matrix node: a pentagon, src/tri3.cpp:407-433
        if (n0 < 0) {
            const int e[4] = {1, 0, 2, 1}, in[4] = {n1, n1, n2, n2};
            cd_slots(D, 4, e, in);
            cd_child(D, n1, 3, 1, 7, 8);
            cd_child(D, -1, 5, 0, 4, 3, 6, 5);
            cd_child(D, n2, 3, 2, 9, 10);
        } else if (n1 < 0) {
            const int e[4] = {2, 1, 0, 2}, in[4] = {n2, n2, n0, n0};
            cd_slots(D, 4, e, in);
            cd_child(D, n2, 3, 2, 7, 8);
            cd_child(D, -1, 5, 1, 4, 3, 6, 5);
            cd_child(D, n0, 3, 0, 9, 10);
        } else {
            const int e[4] = {0, 2, 1, 0}, in[4] = {n0, n0, n1, n1};
            cd_slots(D, 4, e, in);
            cd_child(D, n0, 3, 0, 7, 8);
            cd_child(D, -1, 5, 2, 4, 3, 6, 5);
            cd_child(D, n1, 3, 1, 9, 10);
        }
        cd_coh(D, 4, 3, 4, 7, 8);
        cd_coh(D, 4, 5, 6, 9, 10);
    } else {  // three inclusions: three triangles and a hexagon, src/tri3.cpp:437-444
        const int e[6] = {0, 2, 2, 1, 1, 0}, in[6] = {n0, n0, n2, n2, n1, n1};
        cd_slots(D, 6, e, in);
        cd_child(D, n0, 3, 0, 9, 10);
        cd_child(D, n2, 3, 2, 11, 12);
        cd_child(D, n1, 3, 1, 13, 14);
        cd_child(D, -1, 6, 8, 7, 6, 5, 4, 3);
        cd_coh(D, 4, 3, 4, 9, 10);
        cd_coh(D, 4, 5, 6, 11, 12);
        cd_coh(D, 4, 7, 8, 13, 14);
    }
    return true;
}

// Tet4::partitionPrism, src/tet4.cpp:771-870 (including its comparison of
// g[3] with g[5] in the second branch): three tets from prism l with ids g.
PGH_HD inline int partition_prism(const int l[6], const int64_t g[6], bool bottom_known, int r[3][4]) {
    auto set = [&](int a0, int a1, int a2, int a3, int b0, int b1, int b2, int b3, int c0, int c1, int c2, int c3) {
        const int v[12] = {a0, a1, a2, a3, b0, b1, b2, b3, c0, c1, c2, c3};
        for (int t = 0; t < 3; ++t)
            for (int k = 0; k < 4; ++k) r[t][k] = l[v[t * 4 + k]];
    };
    if (bottom_known) {
        if (g[0] < g[1] && g[0] < g[2]) {
            if (g[1] < g[2]) set(0, 1, 2, 5, 0, 1, 5, 4, 0, 4, 5, 3);
            else set(0, 1, 2, 4, 0, 4, 2, 5, 0, 4, 5, 3);
        } else if (g[1] < g[0] && g[1] < g[2]) {
            if (g[0] < g[2]) set(0, 1, 2, 5, 0, 1, 5, 3, 1, 5, 3, 4);
            else set(0, 1, 2, 3, 1, 2, 3, 5, 1, 5, 3, 4);
        } else if (g[2] < g[0] && g[2] < g[1]) {
            if (g[0] < g[1]) set(0, 1, 2, 4, 0, 4, 2, 3, 2, 3, 4, 5);
            else set(0, 1, 2, 3, 1, 2, 3, 4, 2, 3, 4, 5);
        } else {
            return CUT_E_PRISM;
        }
    } else {
        int min_enrich = 0;
        if (g[1] < g[2] && g[1] < g[4] && g[1] < g[5]) min_enrich = 1;
        else if (g[2] < g[1] && g[2] < g[4] && g[3] < g[5]) min_enrich = 2;
        else if (g[4] < g[1] && g[4] < g[2] && g[4] < g[5]) min_enrich = 4;
        else if (g[5] < g[1] && g[5] < g[2] && g[5] < g[4]) min_enrich = 5;
        const bool a15 = (min_enrich == 1 || min_enrich == 5);
        if (g[0] < g[3]) {
            if (a15) set(0, 1, 2, 5, 0, 1, 5, 4, 0, 4, 5, 3);
            else set(0, 1, 2, 4, 0, 4, 2, 5, 0, 4, 5, 3);
        } else if (g[3] < g[0]) {
            if (a15) set(0, 1, 2, 3, 1, 2, 3, 5, 1, 5, 3, 4);
            else set(0, 1, 2, 3, 1, 2, 3, 4, 2, 3, 4, 5);
        } else {
            return CUT_E_PRISM;
        }
    }
    return CUT_OK;
}

// a child of the tet4 tables: a tet, or a prism decomposed into three tets
// (decomposePrisms, src/tet4.cpp:673-735)
PGH_HD inline int tet4_child(CutDetect& D, const int32_t gid[4], int val, int n, const int* l) {
    if (n == 4) {
        cd_child(D, val, 4, l[0], l[1], l[2], l[3]);
        return CUT_OK;
    }
    int64_t ids[6];
    int known = 0;
    for (int k = 0; k < 6; ++k) {
        if (l[k] < 4) { ids[k] = gid[l[k]]; ++known; }
        else ids[k] = l[k];
    }
    int t[3][4];
    const int rc = partition_prism(l, ids, known != 2, t);
    if (rc) return rc;
    for (int k = 0; k < 3; ++k) cd_child(D, val, 4, t[k][0], t[k][1], t[k][2], t[k][3]);
    return CUT_OK;
}

// Tet4 detection (one inclusion against the matrix): returns CUT_OK with
// *cut set, or an error code.
PGH_HD inline int detect_tet4(const int nd[4], const int32_t gid[4], CutDetect& D, bool* cut) {
    cd_clear(D);
    *cut = false;
    const int n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
    const bool n0e1 = n0 == n1, n0e2 = n0 == n2, n0e3 = n0 == n3, n1e2 = n1 == n2, n1e3 = n1 == n3, n2e3 = n2 == n3;
    // the distinct values, ascending
    int v[4] = {n0, n1, n2, n3};
    for (int i = 1; i < 4; ++i)
        for (int j = i; j > 0 && v[j - 1] > v[j]; --j) {
            const int t = v[j];
            v[j] = v[j - 1];
            v[j - 1] = t;
        }
    int nu = 1;
    for (int i = 1; i < 4; ++i)
        if (v[i] != v[nu - 1]) v[nu++] = v[i];
    if (nu == 1) return CUT_OK;
    if (nu > 2) return CUT_E_TET_MULTI;
    if (v[0] >= 0) return CUT_E_TET_TWO_INCL;
    const int count0 = (n0 == v[0]) + (n1 == v[0]) + (n2 == v[0]) + (n3 == v[0]);
    int ints[2][6], nint[2], mats[2];
    bool lo;  // the tables' "_lo" cohesive ordering
    int edges[4];
    if (count0 == 1 || count0 == 3) {
        nint[0] = 4;
        nint[1] = 6;
        int sep, c0, c1, c2;  // the separated node and the three others in table order
        if (n1e2 && n1e3) {
            edges[0] = 0; edges[1] = 3; edges[2] = 2; mats[0] = n0; mats[1] = n1;
            sep = 0; c0 = 1; c1 = 3; c2 = 2; lo = n0 < n1;
        } else if (n0e2 && n0e3) {
            edges[0] = 0; edges[1] = 1; edges[2] = 4; mats[0] = n1; mats[1] = n0;
            sep = 1; c0 = 0; c1 = 2; c2 = 3; lo = n1 < n0;
        } else if (n0e1 && n0e3) {
            edges[0] = 1; edges[1] = 2; edges[2] = 5; mats[0] = n2; mats[1] = n1;
            sep = 2; c0 = 1; c1 = 0; c2 = 3; lo = n2 < n0;
        } else if (n0e1 && n1e2) {
            edges[0] = 3; edges[1] = 4; edges[2] = 5; mats[0] = n3; mats[1] = n1;
            sep = 3; c0 = 0; c1 = 1; c2 = 2; lo = n3 < n1;
        } else {
            return CUT_E_TET_1SEP;
        }
        const int e[3] = {edges[0], edges[1], edges[2]}, in[3] = {v[1], v[1], v[1]};
        cd_slots(D, 3, e, in);
        const int a = lo ? 4 : 7, b = lo ? 7 : 4;
        const int t0[4] = {a, a + 1, a + 2, sep};
        const int t1[6] = {c0, c1, c2, b, b + 1, b + 2};
        for (int k = 0; k < 4; ++k) ints[0][k] = t0[k];
        for (int k = 0; k < 6; ++k) ints[1][k] = t1[k];
        if (lo) cd_coh(D, 6, 4, 6, 5, 7, 9, 8);
        else cd_coh(D, 6, 4, 5, 6, 7, 8, 9);
    } else {
        nint[0] = 6;
        nint[1] = 6;
        int p0, p1, q0, q1;  // the first pair of nodes (p) and the second (q) as the tables list them
        if (n0e1 && n2e3 && !n0e2) {
            edges[0] = 1; edges[1] = 2; edges[2] = 3; edges[3] = 4; mats[0] = n0; mats[1] = n2;
            p0 = 0; p1 = 1; q0 = 2; q1 = 3; lo = n0 < n2;
        } else if (n1e2 && n0e3 && !n0e1) {
            edges[0] = 2; edges[1] = 0; edges[2] = 4; edges[3] = 5; mats[0] = n1; mats[1] = n0;
            p0 = 1; p1 = 2; q0 = 0; q1 = 3; lo = n1 < n3;
        } else if (n0e2 && n1e3 && !n0e1) {
            edges[0] = 0; edges[1] = 1; edges[2] = 5; edges[3] = 3; mats[0] = n2; mats[1] = n1;
            p0 = 2; p1 = 0; q0 = 1; q1 = 3; lo = n0 < n1;
        } else {
            return CUT_E_TET_2SEP;
        }
        const int e[4] = {edges[0], edges[1], edges[2], edges[3]}, in[4] = {v[1], v[1], v[1], v[1]};
        cd_slots(D, 4, e, in);
        if (lo) {
            const int t0[6] = {p0, 5, 6, p1, 4, 7}, t1[6] = {q0, 9, 8, q1, 10, 11};
            for (int k = 0; k < 6; ++k) { ints[0][k] = t0[k]; ints[1][k] = t1[k]; }
            cd_coh(D, 8, 4, 5, 6, 7, 8, 9, 10, 11);
        } else {
            const int t0[6] = {p0, 9, 10, p1, 8, 11}, t1[6] = {q0, 5, 4, q1, 6, 7};
            for (int k = 0; k < 6; ++k) { ints[0][k] = t0[k]; ints[1][k] = t1[k]; }
            cd_coh(D, 8, 7, 6, 5, 4, 11, 10, 9, 8);
        }
    }
    for (int o = 0; o < 2; ++o) {
        const int rc = tet4_child(D, gid, mats[o], nint[o], ints[o]);
        if (rc) return rc;
    }
    *cut = true;
    return CUT_OK;
}

// The cut record's child and cohesive fields from a detection, in the
// regular numbering for perfect bond (transformDetectionCohToReg,
// src/elem.cpp:547-559; a cohesive element keeps the mirrors); the
// enrichment nodes and the ISV offset are attached later.  Child material 0
// is the matrix, 1 the fibre; cohesive children take material 2.  Returns the
// element's cohesive qps (CohEdge2 2-point Gauss, CohTri3 / CohQuad4 4
// points), or -CUT_E_CHILDREN.
PGH_HD inline int cut_record(const CutDetect& D, int dim, int nen, bool coh, pg_cut_rec& R) {
    uint32_t* w = reinterpret_cast<uint32_t*>(&R);
    for (unsigned k = 0; k < sizeof(pg_cut_rec) / 4; ++k) w[k] = 0u;
    if (D.nchild > (dim == 2 ? 4 : PG_MAX_CHILD)) return -CUT_E_CHILDREN;
    R.n_enr = uint8_t(coh ? 2 * D.ncut : D.ncut);
    R.n_child = uint8_t(D.nchild);
    int nvqp = 0;
    for (int k = 0; k < D.nchild; ++k) {
        const int nn = D.child_nn[k];
        const int t = dim == 3 ? PG_CHILD_TET4
                               : (nn == 3 ? PG_CHILD_TRI3 : nn == 4 ? PG_CHILD_QUAD4 : nn == 5 ? PG_CHILD_POLY5 : PG_CHILD_POLY6);
        R.child_type[k] = uint8_t(t);
        R.child_mat[k] = uint8_t(D.child_val[k] < 0 ? 0 : 1);
        for (int j = 0; j < nn; ++j) {
            int v = D.child_nodes[k][j];
            if (!coh && v >= nen + D.ncut) v -= D.ncut;
            R.child_node[PG_CHILD_STRIDE(dim) * k + j] = uint8_t(v);
        }
        nvqp += PG_CHILD_NQP(t);
    }
    R.n_vqp = uint8_t(nvqp);
    int ncq = 0;
    if (coh) {
        R.n_coh = uint8_t(D.ncoh);
        R.coh_nn = D.coh_nn;
        for (int k = 0; k < D.ncoh; ++k) {
            for (int j = 0; j < D.coh_nn; ++j) R.coh_nodes[k][j] = D.coh_nodes[k][j];
            R.coh_mat[k] = 2;
            ncq += (D.coh_nn == 4) ? 2 : 4;
        }
    }
    return ncq;
}

// --------------------------------------------------------------------------
// Edge intersection, Inclusion::find_intersection (src/Inclusion.cpp:32-120):
// GSL Steffensen iteration from alpha = 0.5 (GSL 2.1 roots/steffenson.c, the
// published algorithm: Newton steps, Aitken delta^2 acceleration from the 3rd
// iterate on), stopping when |alpha - alpha_prev| < 1e-7, 25 iterations max.
// The reference then falls back to GSL Brent on [0,1]; here the fallback is
// bisection to the same 1e-7 bracket (a path the generated RVEs do not take).
// Circle_Inclusion::int_f / int_df, src/InclusionCircle.cpp:248-287 (the
// sphere's are the same with a third term).
// --------------------------------------------------------------------------
PGH_HD inline void edge_fdf(int dim, const double* x0, const double* x1, const double* c, double r, double a,
                            double& f, double& df) {
    double g[3], dd[3];
    for (int d = 0; d < dim; ++d) {
        g[d] = x0[d] + a * (x1[d] - x0[d]) - c[d];
        dd[d] = x1[d] - x0[d];
    }
    if (dim == 2) {
        f = (g[0] / r) * (g[0] / r) + (g[1] / r) * (g[1] / r) - 1;
        df = (1 / (r * r)) * 2.0 * (g[0] * dd[0] + g[1] * dd[1]);
    } else {
        f = (g[0] / r) * (g[0] / r) + (g[1] / r) * (g[1] / r) + (g[2] / r) * (g[2] / r) - 1.0;
        df = (1 / (r * r)) * 2.0 * (g[0] * dd[0] + g[1] * dd[1] + g[2] * dd[2]);
    }
}

PGH_HD inline double edge_f(int dim, const double* x0, const double* x1, const double* c, double r, double a) {
    double f, df;
    edge_fdf(dim, x0, x1, c, r, a, f, df);
    return f;
}

// alpha of the intersection on edge x0 -> x1, or CUT_E_NOROOT in *err
PGH_HD inline double intersect_edge(int dim, const double* x0, const double* x1, const double* c, double r,
                                    int* err) {
    *err = CUT_OK;
    if (edge_f(dim, x0, x1, c, r, 0.0) * edge_f(dim, x0, x1, c, r, 1.0) >= 0) {
        *err = CUT_E_NOROOT;
        return 0.0;
    }
    double x = 0.5, x_1 = 0.0, fx, dfx;
    edge_fdf(dim, x0, x1, c, r, x, fx, dfx);
    int count = 1, iter = 0;
    double root = 0.5, alpha = 0.5, alpha0;
    bool done = false;
    do {
        ++iter;
        if (dfx == 0.0) break;
        double x_new = x - fx / dfx, f_new, df_new;
        edge_fdf(dim, x0, x1, c, r, x_new, f_new, df_new);
        double x1_old = x_1;
        x_1 = x;
        double xo = x;
        x = x_new;
        fx = f_new;
        dfx = df_new;
        if (count < 3) {
            root = x_new;
            ++count;
        } else {
            double u = xo - x1_old, v = x_new - 2 * xo + x1_old;
            root = (v == 0) ? x_new : x1_old - u * u / v;
        }
        alpha0 = alpha;
        alpha = root;
        done = fabs(alpha - alpha0) < 1e-7;
    } while (!done && iter < 25);
    if (iter == 25 || alpha < 0.0 || alpha > 1.0 || !done) {
        double lo = 0.0, hi = 1.0, flo = edge_f(dim, x0, x1, c, r, lo);
        for (int it = 0; it < 100 && hi - lo >= 1e-7; ++it) {
            double mid = 0.5 * (lo + hi), fm = edge_f(dim, x0, x1, c, r, mid);
            if ((fm < 0) == (flo < 0)) { lo = mid; flo = fm; } else hi = mid;
        }
        alpha = 0.5 * (lo + hi);
    }
    return alpha;
}

}  // namespace pgh
