/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of PIGFEM's K / R assembly for AssemblerSmallStrainStructural.
 * It is the checker for the GPU library: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  Nothing in the product links it.
 *
 * What it restates (brandyb2/PIGFEM, file:line):
 *   Mesh::store_shape_functions                     src/Mesh.cpp:1475-1643
 *   Assembler::storeBmats / Assemble_Small_Strain_B_Mat
 *                                                   src/Assembler.cpp:602-621,
 *                                                   src/Utilities.cpp:68-127
 *   Elem::ShapeFunctions / transform_gradients,
 *   Utilities::det / inverse                        src/elem.cpp:115-142,
 *                                                   src/Utilities.cpp:1017-1070
 *   Tri3 / Quad4 / Tet4 / Edge2 quadrature + shapes src/tri3.cpp:93-257,
 *                                                   src/quad4.cpp:65-150,
 *                                                   src/tet4.cpp:87-160, src/edge2.cpp:101-126
 *   CohesiveElem::ShapeFunctions / rotation         src/CohesiveElem.cpp:145-238,
 *                                                   src/CohEdge2.cpp:55-209
 *   Assembler::assemble / assemble_elem             src/Assembler.cpp:65-368
 *   KernelVolumetric (B variant) / KernelCohesive   src/AssemblerSmallStrainStructural.cpp:23-103
 *   B_U, Bt_D_B, Bt_F (CBLAS dgemv/dgemm order)     src/Utilities.cpp:491-513
 *   cohesiveDeltaFast / KFast / InternalForceFast   src/Utilities.cpp:528-594
 *   LinearElasticIsotropicMaterial::Constitutive    src/material_lei.cpp:232-299
 *   Bilinear / OP cohesive Constitutive             src/Material_BilinearCohesive.cpp:38-125,
 *                                                   src/Material_BilinearCohesiveNU.cpp:48-138,
 *                                                   src/Material_OPCohesive.cpp:38-98,
 *                                                   src/Material_OPCohesiveNU.cpp:43-111
 *   set_matrix / set_vec (MatSetValue ADD_VALUES,
 *     modelled as a binary-search add into CSR)     src/Assembler.cpp:659-701
 *   AIJ pattern = union of full element dof blocks (explicit zeros inserted)
 *
 * Tet4 parents: the reference inverse-maps child qps with GSL multiroots
 * (src/elem.cpp:600-765); for an affine tet the closed-form inverse is used here.
 * The parent gradients of a linear parent do not depend on that point.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/pgfem_asm.h"

#define MAXND 48

typedef struct {
    int32_t n;
    int32_t cap;
    int32_t* c;
} row_t;

typedef struct oracle {
    const pg_problem_desc* d;
    int e0, e1;
    int nstr;
    int nndof; /* dofs per node: dim (structural) or 1 (heat conduction) */
    /* own AIJ pattern per block, built by insertion (PETSc MatSetValue analogue) */
    int32_t nrows[4];
    row_t* rows[4];
    int32_t* rp[4];
    int32_t* ci[4];
    int64_t nnz[4];
    int64_t base[4];
    /* cached shape data per element (store_shape_functions + storeBmats) */
    int64_t* off;     /* per element offset into cache */
    double* cache;
    /* element colouring for the threaded baseline (built on first use):
     * elements of one colour share no node, so they write disjoint K rows,
     * P entries and ISVs */
    int ncolour;
    int32_t* colour_ptr;  /* [ncolour + 1] */
    int32_t* colour_elem; /* elements grouped by colour, ascending within one */
    int own_pattern;      /* 0: rp / ci are the description's (harness) pattern */
    /* thermal pre-load: input.temp_change of every volumetric qp
     * (AssemblerSmallStrainStructural_ThermalPreLoad::KernelVolumetric,
     * src/AssemblerSmallStrainStructural_ThermalPreLoad.cpp:58-59) */
    double temp_change;
    char err[256];
} oracle;

/* ------------------------------------------------------------------------ */
static const double TRI_X[4] = {0.333333333333333333333333333333333, 0.200000000000000000000000000000000,
                                0.200000000000000000000000000000000, 0.600000000000000000000000000000000};
static const double TRI_Y[4] = {0.333333333333333333333333333333333, 0.200000000000000000000000000000000,
                                0.600000000000000000000000000000000, 0.200000000000000000000000000000000};
static const double TRI_W[4] = {-0.281250000000000000000000000000000, 0.260416666666666666666666666666665,
                                0.260416666666666666666666666666665, 0.260416666666666666666666666666665};
static const double G2 = 0.57735026918962576450914878050196;
static const double TET_X[5][3] = {{0.25, 0.25, 0.25},
                                   {0.5, 0.166666666666667, 0.166666666666667},
                                   {0.166666666666667, 0.166666666666667, 0.166666666666667},
                                   {0.166666666666667, 0.166666666666667, 0.5},
                                   {0.166666666666667, 0.5, 0.166666666666667}};
static const double TET_W[5] = {-0.8, 0.45, 0.45, 0.45, 0.45};

/* q_point + compute_shape_grad for the element types on the path */
static int nqp_of(int type) { return PG_CHILD_NQP(type); }

/* PI of include/common.h:215 */
static const double PG_PI = 3.14159265358979323846264338327950;

/* Polygon::PolyTriangulate (src/Polygon.cpp:51-63): vertex n of the regular
 * reference polygon at angle 2 pi (n + 1) / nn; sub-triangle t is (centre,
 * vertex t, vertex t + 1), the last one closing on vertex 0. */
static void poly_vertex(int nn, int n, double* p) {
    if (n == nn) n = 0;
    p[0] = cos(2.0 * PG_PI * (n + 1) / nn);
    p[1] = sin(2.0 * PG_PI * (n + 1) / nn);
}

/* Polygon::q_point (src/Polygon.cpp:84-131) with the tri3 rule of order 3
 * (the Elem* default): qp q is point q % 4 of sub-triangle q / 4 around the
 * origin; the weight is the tri3 weight times the sub-triangle's Jacobian. */
static void poly_qpoint(int nn, int q, double* xi, double* w) {
    const int t = q / 4, qq = q % 4;
    const double tN[3] = {1.0 - TRI_X[qq] - TRI_Y[qq], TRI_X[qq], TRI_Y[qq]};
    const double tdN[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
    double pT[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    poly_vertex(nn, t, pT[1]);
    poly_vertex(nn, t + 1, pT[2]);
    double J0[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int n = 0; n < 3; ++n) J0[i][j] += pT[n][i] * tdN[n][j];
    const double det_J0 = J0[0][0] * J0[1][1] - J0[0][1] * J0[1][0];
    *w = det_J0 * TRI_W[qq];
    xi[0] = 0.0;
    xi[1] = 0.0;
    for (int n = 0; n < 3; ++n) {
        xi[0] += tN[n] * pT[n][0];
        xi[1] += tN[n] * pT[n][1];
    }
}

/* Polygon::compute_shape / compute_shape_grad (src/Polygon.cpp:133-222):
 * with A_n the signed area of (x, vertex n - 1, vertex n) (A_0 = A_nn),
 * alpha_n = 1 / (A_n A_{n+1}), N_n = alpha_n / sum alpha and its gradient
 * from dA / dx. */
static void poly_shape(int nn, const double* xi, double* N, double dN[][3]) {
    double A[7], dA[7][2], alpha[6], dalpha[6][2];
    for (int n = 0; n < nn; ++n) {
        double p1[2], p2[2];
        poly_vertex(nn, n, p1);
        poly_vertex(nn, n + 1, p2);
        A[n + 1] = 0.5 * ((p1[0] * p2[1] - p1[1] * p2[0]) - (xi[0] * p2[1] - xi[1] * p2[0]) +
                          (xi[0] * p1[1] - xi[1] * p1[0]));
        dA[n + 1][0] = 0.5 * (p2[1] - p1[1]);
        dA[n + 1][1] = 0.5 * (p1[0] - p2[0]);
    }
    A[0] = A[nn];
    dA[0][0] = dA[nn][0];
    dA[0][1] = dA[nn][1];
    double sum_alpha = 0.0, sum_dalpha[2] = {0.0, 0.0};
    for (int n = 0; n < nn; ++n) {
        alpha[n] = 1.0 / (A[n] * A[n + 1]);
        dalpha[n][0] = -1.0 * alpha[n] * (dA[n][0] / A[n] + dA[n + 1][0] / A[n + 1]);
        dalpha[n][1] = -1.0 * alpha[n] * (dA[n][1] / A[n] + dA[n + 1][1] / A[n + 1]);
        sum_alpha += alpha[n];
        sum_dalpha[0] += dalpha[n][0];
        sum_dalpha[1] += dalpha[n][1];
    }
    for (int n = 0; n < nn; ++n) {
        const double Nn = alpha[n] / sum_alpha;
        N[n] = Nn;
        dN[n][0] = (dalpha[n][0] - Nn * sum_dalpha[0]) / sum_alpha;
        dN[n][1] = (dalpha[n][1] - Nn * sum_dalpha[1]) / sum_alpha;
    }
}

static void qpoint(int type, int q, double* xi, double* w) {
    if (type == PG_CHILD_POLY5 || type == PG_CHILD_POLY6) poly_qpoint(PG_CHILD_NN(type), q, xi, w);
    else if (type == PG_CHILD_TRI3) { xi[0] = TRI_X[q]; xi[1] = TRI_Y[q]; *w = TRI_W[q]; }
    else if (type == PG_CHILD_QUAD4) {
        /* Quad4::q_point: edge qp (q % 2) then (q / 2), w = w0 * w1 */
        xi[0] = (q % 2 == 0) ? -G2 : G2;
        xi[1] = (q / 2 == 0) ? -G2 : G2;
        *w = 1.0 * 1.0;
    } else { xi[0] = TET_X[q][0]; xi[1] = TET_X[q][1]; xi[2] = TET_X[q][2]; *w = TET_W[q]; }
}

static void shape(int type, const double* xi, double* N, double dN[][3]) {
    if (type == PG_CHILD_POLY5 || type == PG_CHILD_POLY6) {
        poly_shape(PG_CHILD_NN(type), xi, N, dN);
    } else if (type == PG_CHILD_TRI3) {
        double r = xi[0], s = xi[1];
        N[0] = 1.0 - r - s; N[1] = r; N[2] = s;
        dN[0][0] = -1.0; dN[0][1] = -1.0; dN[1][0] = 1.0; dN[1][1] = 0.0; dN[2][0] = 0.0; dN[2][1] = 1.0;
    } else if (type == PG_CHILD_QUAD4) {
        double r = xi[0], s = xi[1];
        N[0] = 0.25 * (1.0 - r) * (1.0 - s); N[1] = 0.25 * (1.0 + r) * (1.0 - s);
        N[2] = 0.25 * (1.0 + r) * (1.0 + s); N[3] = 0.25 * (1.0 - r) * (1.0 + s);
        dN[0][0] = -0.25 * (1.0 - s); dN[0][1] = -0.25 * (1.0 - r);
        dN[1][0] = 0.25 * (1.0 - s);  dN[1][1] = -0.25 * (1.0 + r);
        dN[2][0] = 0.25 * (1.0 + s);  dN[2][1] = 0.25 * (1.0 + r);
        dN[3][0] = -0.25 * (1.0 + s); dN[3][1] = 0.25 * (1.0 - r);
    } else {
        double r = xi[0], s = xi[1], t = xi[2];
        N[0] = 1.0 - r - s - t; N[1] = r; N[2] = s; N[3] = t;
        for (int n = 0; n < 4; ++n)
            for (int a = 0; a < 3; ++a) dN[n][a] = (n == 0) ? -1.0 : (n - 1 == a ? 1.0 : 0.0);
    }
}

static int nn_of(int type) { return PG_CHILD_NN(type); }

/* Elem::transform_gradients + Utilities::det / inverse */
static double transform(int dim, int nn, double dNdr[][3], double x[][3], double dNdx[][3]) {
    double Jac[3][3] = {{0}}, inv[3][3], J;
    for (int a = 0; a < dim; ++a)
        for (int i = 0; i < dim; ++i)
            for (int n = 0; n < nn; ++n) Jac[a][i] += dNdr[n][a] * x[n][i];
    if (dim == 2) {
        J = Jac[0][0] * Jac[1][1] - Jac[0][1] * Jac[1][0];
        inv[0][0] = Jac[1][1]; inv[1][1] = Jac[0][0];
        inv[0][1] = -1.0 * Jac[0][1]; inv[1][0] = -1.0 * Jac[1][0];
    } else {
        J = Jac[0][0] * (Jac[1][1] * Jac[2][2] - Jac[1][2] * Jac[2][1]) -
            Jac[0][1] * (Jac[1][0] * Jac[2][2] - Jac[1][2] * Jac[2][0]) +
            Jac[0][2] * (Jac[1][0] * Jac[2][1] - Jac[1][1] * Jac[2][0]);
        inv[0][0] = Jac[1][1] * Jac[2][2] - Jac[2][1] * Jac[1][2];
        inv[0][1] = -1.0 * (Jac[0][1] * Jac[2][2] - Jac[2][1] * Jac[0][2]);
        inv[0][2] = Jac[0][1] * Jac[1][2] - Jac[1][1] * Jac[0][2];
        inv[1][0] = -1.0 * (Jac[1][0] * Jac[2][2] - Jac[2][0] * Jac[1][2]);
        inv[1][1] = Jac[0][0] * Jac[2][2] - Jac[2][0] * Jac[0][2];
        inv[1][2] = -1.0 * (Jac[0][0] * Jac[1][2] - Jac[1][0] * Jac[0][2]);
        inv[2][0] = Jac[1][0] * Jac[2][1] - Jac[2][0] * Jac[1][1];
        inv[2][1] = -1.0 * (Jac[0][0] * Jac[2][1] - Jac[2][0] * Jac[0][1]);
        inv[2][2] = Jac[0][0] * Jac[1][1] - Jac[1][0] * Jac[0][1];
    }
    double rd = 1.0 / J;
    for (int a = 0; a < dim; ++a)
        for (int b = 0; b < dim; ++b) inv[a][b] *= rd;
    for (int n = 0; n < nn; ++n)
        for (int a = 0; a < dim; ++a) {
            double s = 0.0;
            for (int b = 0; b < dim; ++b) s += inv[a][b] * dNdr[n][b];
            dNdx[n][a] = s;
        }
    return J;
}

/* Assemble_Thermal_B_Mat (src/Utilities.cpp:147-160): B (dim x nn), B(j, n) =
 * the j-th gradient component of node n */
static void fill_B_thermal(int dim, int nn, double g[][3], double* B) {
    for (int j = 0; j < dim; ++j)
        for (int n = 0; n < nn; ++n) B[j * nn + n] = g[n][j];
}

/* Assemble_Small_Strain_B_Mat: B (nstr x dim*nn), row-major, zeros included */
static void fill_B(int dim, int nn, double g[][3], double* B) {
    int nstr = dim == 2 ? 3 : 6, nc = dim * nn;
    memset(B, 0, sizeof(double) * nstr * nc);
    for (int n = 0; n < nn; ++n)
        for (int i = 0; i < dim; ++i) B[i * nc + dim * n + i] = g[n][i];
    if (dim == 2) {
        for (int n = 0; n < nn; ++n) { B[2 * nc + 2 * n] = g[n][1]; B[2 * nc + 2 * n + 1] = g[n][0]; }
    } else {
        for (int n = 0; n < nn; ++n) {
            B[3 * nc + dim * n + 1] = g[n][2]; B[3 * nc + dim * n + 2] = g[n][1];
            B[4 * nc + dim * n] = g[n][2];     B[4 * nc + dim * n + 2] = g[n][0];
            B[5 * nc + dim * n] = g[n][1];     B[5 * nc + dim * n + 1] = g[n][0];
        }
    }
}

/* ------------------------------------------------------------------------ */
static int local_list(const pg_problem_desc* d, int e, int32_t* out) {
    int n = 0;
    for (int k = 0; k < d->nen; ++k) out[n++] = d->conn[(int64_t)e * d->nen + k];
    int c = d->cut_of_elem[e];
    if (c >= 0)
        for (int k = 0; k < d->cut[c].n_enr; ++k) out[n++] = d->cut[c].enr[k];
    return n;
}

static void pt(const pg_problem_desc* d, int node, double* x) {
    for (int i = 0; i < 3; ++i) x[i] = i < d->dim ? d->coords[(int64_t)node * d->dim + i] : 0.0;
}

/* per-element cache size (doubles) */
static int64_t cache_size(const oracle* o, int e) {
    const pg_problem_desc* d = o->d;
    int dim = d->dim, nstr = o->nstr;
    int c = d->cut_of_elem[e];
    if (c < 0) {
        int nq = dim == 2 ? 4 : 5;
        return (int64_t)nq * (2 + nstr * d->nen * o->nndof);
    }
    const pg_cut_rec* r = &d->cut[c];
    int nd = (d->nen + r->n_enr) * o->nndof;
    int ncq = r->n_coh * (r->coh_nn == 4 ? 2 : 4);
    return (int64_t)r->n_vqp * (2 + nstr * nd) + (int64_t)ncq * (2 + 4 + dim * dim);
}

/* Mesh::store_shape_functions + storeBmats for element e: per volumetric qp
 * [W, J, B...], per cohesive qp [W, J, N(4), R(dim*dim)]. */
static int store_shapes(oracle* o, int e, double* out) {
    const pg_problem_desc* d = o->d;
    int dim = d->dim, nstr = o->nstr, nen = d->nen;
    int c = d->cut_of_elem[e];
    double x[16][3], dN[16][3], g[16][3], N[16], xi[3], w;
    int32_t ln[16];
    int nl = local_list(d, e, ln);
    for (int n = 0; n < nl; ++n) pt(d, ln[n], x[n]);
    int ptype = dim == 2 ? PG_CHILD_TRI3 : PG_CHILD_TET4;
    if (c < 0) {
        int nq = nqp_of(ptype);
        for (int q = 0; q < nq; ++q) {
            qpoint(ptype, q, xi, &w);
            shape(ptype, xi, N, dN);
            double J = transform(dim, nen, dN, x, g);
            if (J < -1e-16) { strcpy(o->err, "negative element Jacobian"); return -1; }
            out[0] = w; out[1] = J;
            if (o->nndof == 1) fill_B_thermal(dim, nen, g, out + 2);
            else fill_B(dim, nen, g, out + 2);
            out += 2 + nstr * nen * o->nndof;
        }
        return 0;
    }
    const pg_cut_rec* r = &d->cut[c];
    int nd = nl * o->nndof;
    for (int k = 0; k < r->n_child; ++k) {
        int ct = r->child_type[k], cnn = nn_of(ct);
        const uint8_t* cl = &r->child_node[PG_CHILD_STRIDE(dim) * k];
        double cx[6][3], cg[6][3], cN[6];
        for (int n = 0; n < cnn; ++n)
            for (int i = 0; i < 3; ++i) cx[n][i] = x[cl[n]][i];
        for (int q = 0; q < nqp_of(ct); ++q) {
            qpoint(ct, q, xi, &w);
            shape(ct, xi, cN, dN);
            double Jc = transform(dim, cnn, dN, cx, cg);
            if (Jc < 0.0) { strcpy(o->err, "negative child Jacobian"); return -1; }
            /* global point, inverse map into the parent, parent shapes there */
            double gx[3] = {0, 0, 0};
            for (int i = 0; i < dim; ++i)
                for (int n = 0; n < cnn; ++n) gx[i] += cx[n][i] * cN[n];
            double rc[3] = {0, 0, 0};
            if (dim == 2) {
                /* Tri3::inverse_map closed form, src/tri3.cpp:242-257 */
                double A = 0.5 * (x[0][0] * (x[1][1] - x[2][1]) + x[1][0] * (x[2][1] - x[0][1]) + x[2][0] * (x[0][1] - x[1][1]));
                if (A < 0.0) A = -1.0 * A;
                rc[0] = (1.0 / (2.0 * A)) * ((x[2][1] - x[0][1]) * (gx[0] - x[0][0]) - (x[2][0] - x[0][0]) * (gx[1] - x[0][1]));
                rc[1] = (1.0 / (2.0 * A)) * (-1.0 * (x[1][1] - x[0][1]) * (gx[0] - x[0][0]) + (x[1][0] - x[0][0]) * (gx[1] - x[0][1]));
            } else {
                double Jm[3][3], in[3][3];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) Jm[i][j] = x[j + 1][i] - x[0][i];
                double det = Jm[0][0] * (Jm[1][1] * Jm[2][2] - Jm[1][2] * Jm[2][1]) -
                             Jm[0][1] * (Jm[1][0] * Jm[2][2] - Jm[1][2] * Jm[2][0]) +
                             Jm[0][2] * (Jm[1][0] * Jm[2][1] - Jm[1][1] * Jm[2][0]);
                in[0][0] = Jm[1][1] * Jm[2][2] - Jm[1][2] * Jm[2][1];
                in[0][1] = Jm[0][2] * Jm[2][1] - Jm[0][1] * Jm[2][2];
                in[0][2] = Jm[0][1] * Jm[1][2] - Jm[0][2] * Jm[1][1];
                in[1][0] = Jm[1][2] * Jm[2][0] - Jm[1][0] * Jm[2][2];
                in[1][1] = Jm[0][0] * Jm[2][2] - Jm[0][2] * Jm[2][0];
                in[1][2] = Jm[0][2] * Jm[1][0] - Jm[0][0] * Jm[1][2];
                in[2][0] = Jm[1][0] * Jm[2][1] - Jm[1][1] * Jm[2][0];
                in[2][1] = Jm[0][1] * Jm[2][0] - Jm[0][0] * Jm[2][1];
                in[2][2] = Jm[0][0] * Jm[1][1] - Jm[0][1] * Jm[1][0];
                for (int i = 0; i < 3; ++i) {
                    double s = 0;
                    for (int j = 0; j < 3; ++j) s += in[i][j] * (gx[j] - x[0][j]);
                    rc[i] = s / det;
                }
            }
            double tol = 1e-12; /* coords_inside tolerance order; the round trip below is the real check */
            double sum = 0;
            for (int i = 0; i < dim; ++i) { if (rc[i] < -1e-6 || rc[i] > 1.0 + 1e-6) tol = -1; sum += rc[i]; }
            if (tol < 0 || sum > 1.0 + 1e-6) { strcpy(o->err, "bad inverse map"); return -1; }
            double pN[4], pdN[4][3], pg_[4][3];
            shape(ptype, rc, pN, pdN);
            double pJ = transform(dim, nen, pdN, x, pg_);
            (void)pJ;
            double g2[3] = {0, 0, 0}, diff = 0;
            for (int i = 0; i < dim; ++i) {
                for (int n = 0; n < nen; ++n) g2[i] += x[n][i] * pN[n];
                diff += (g2[i] - gx[i]) * (g2[i] - gx[i]);
            }
            if (sqrt(diff) > 1e-9) { strcpy(o->err, "bad inverse map round trip"); return -1; }
            /* shape gradients over all local nodes: parent, then enrichment from the child */
            for (int n = 0; n < nl; ++n)
                for (int a = 0; a < 3; ++a) g[n][a] = n < nen ? pg_[n][a] : 0.0;
            for (int en = 0; en < r->n_enr; ++en)
                for (int n = 0; n < cnn; ++n)
                    if (cl[n] == nen + en) {
                        for (int a = 0; a < dim; ++a) g[nen + en][a] = 1.0 * cg[n][a];
                        break;
                    }
            out[0] = w; out[1] = Jc;
            if (o->nndof == 1) fill_B_thermal(dim, nl, g, out + 2);
            else fill_B(dim, nl, g, out + 2);
            out += 2 + nstr * nd;
        }
    }
    /* cohesive qps */
    int cnn = r->coh_nn / 2, cnq = r->coh_nn == 4 ? 2 : 4;
    for (int k = 0; k < r->n_coh; ++k)
        for (int q = 0; q < cnq; ++q) {
            double cN[4] = {0, 0, 0, 0}, dNr[4][2], R[3][3], dA;
            double Jc[2][3] = {{0, 0, 0}, {0, 0, 0}};
            if (r->coh_nn == 4) {
                double s = q == 0 ? -G2 : G2;
                w = 1.0;
                cN[0] = 0.5 * (1.0 - s); cN[1] = 0.5 * (1 + s);
                dNr[0][0] = -0.5; dNr[1][0] = 0.5;
            } else if (r->coh_nn == 8) {
                /* CohQuad4::q_point / compute_shape / compute_shape_grad
                 * (src/CohQuad4.cpp:58-140): edge qps (q % 2, q / 2), w = w0 w1 */
                double e0 = (q % 2) == 0 ? -G2 : G2, e1 = (q / 2) == 0 ? -G2 : G2;
                w = 1.0 * 1.0;
                double rr = e0, ss = e1;
                cN[0] = 0.25 * (1.0 - rr) * (1.0 - ss); cN[1] = 0.25 * (1.0 + rr) * (1.0 - ss);
                cN[2] = 0.25 * (1.0 + rr) * (1.0 + ss); cN[3] = 0.25 * (1.0 - rr) * (1.0 + ss);
                dNr[0][0] = -0.25 * (1.0 - ss); dNr[0][1] = -0.25 * (1.0 - rr);
                dNr[1][0] = 0.25 * (1.0 - ss);  dNr[1][1] = -0.25 * (1.0 + rr);
                dNr[2][0] = 0.25 * (1.0 + ss);  dNr[2][1] = 0.25 * (1.0 + rr);
                dNr[3][0] = -0.25 * (1.0 + ss); dNr[3][1] = 0.25 * (1.0 - rr);
            } else {
                w = TRI_W[q];
                cN[0] = 1.0 - TRI_X[q] - TRI_Y[q]; cN[1] = TRI_X[q]; cN[2] = TRI_Y[q];
                dNr[0][0] = -1.0; dNr[0][1] = -1.0; dNr[1][0] = 1.0; dNr[1][1] = 0.0; dNr[2][0] = 0.0; dNr[2][1] = 1.0;
            }
            for (int n = 0; n < cnn; ++n)
                for (int a = 0; a < dim - 1; ++a)
                    for (int i = 0; i < dim; ++i) Jc[a][i] += dNr[n][a] * x[r->coh_nodes[k][n]][i];
            double nr[3], mag[2];
            if (dim == 2) { nr[0] = -Jc[0][1]; nr[1] = Jc[0][0]; }
            else {
                nr[0] = Jc[0][1] * Jc[1][2] - Jc[0][2] * Jc[1][1];
                nr[1] = Jc[0][2] * Jc[1][0] - Jc[0][0] * Jc[1][2];
                nr[2] = Jc[0][0] * Jc[1][1] - Jc[0][1] * Jc[1][0];
            }
            for (int a = 0; a < dim - 1; ++a) {
                double s = 0;
                for (int i = 0; i < dim; ++i) s += Jc[a][i] * Jc[a][i];
                mag[a] = sqrt(s);
            }
            double s = 0;
            for (int i = 0; i < dim; ++i) s += nr[i] * nr[i];
            dA = sqrt(s);
            for (int i = 0; i < dim; ++i) R[dim - 1][i] = nr[i] / dA;
            for (int i = 0; i < dim; ++i)
                for (int a = 0; a < dim - 1; ++a) R[a][i] = Jc[a][i] / mag[a];
            out[0] = w; out[1] = dA;
            for (int n = 0; n < 4; ++n) out[2 + n] = n < cnn ? cN[n] * 1.0 : 0.0;
            for (int i = 0; i < dim; ++i)
                for (int j = 0; j < dim; ++j) out[6 + i * dim + j] = R[i][j];
            out += 2 + 4 + dim * dim;
        }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* AIJ row insertion (MatSetValue into a preallocated row, keeping it sorted) */
static int row_find(const row_t* r, int32_t c) {
    int lo = 0, hi = r->n;
    while (lo < hi) {
        int m = (lo + hi) >> 1;
        if (r->c[m] < c) lo = m + 1; else hi = m;
    }
    return lo;
}

static int row_insert(row_t* r, int32_t c) {
    int p = row_find(r, c);
    if (p < r->n && r->c[p] == c) return 0;
    if (r->n == r->cap) {
        int nc = r->cap ? 2 * r->cap : 16;
        int32_t* nn = (int32_t*)realloc(r->c, sizeof(int32_t) * nc);
        if (!nn) return -1;
        r->c = nn; r->cap = nc;
    }
    memmove(r->c + p + 1, r->c + p, sizeof(int32_t) * (r->n - p));
    r->c[p] = c;
    r->n++;
    return 0;
}

static int elem_dofs(const pg_problem_desc* d, int e, int32_t* dofs) {
    int32_t ln[16];
    int nl = local_list(d, e, ln);
    const int nn = d->nndof ? d->nndof : d->dim;
    for (int n = 0; n < nl; ++n)
        for (int j = 0; j < nn; ++j) dofs[n * nn + j] = d->node_dof[(int64_t)ln[n] * nn + j];
    return nl * nn;
}

static void oracle_free_rows(oracle* o) {
    for (int b = 0; b < 4; ++b) {
        if (o->rows[b]) {
            for (int r = 0; r < o->nrows[b]; ++r) free(o->rows[b][r].c);
            free(o->rows[b]);
            o->rows[b] = NULL;
        }
    }
}

void oracle_free(oracle* o) {
    if (!o) return;
    oracle_free_rows(o);
    if (o->own_pattern)
        for (int b = 0; b < 4; ++b) { free(o->rp[b]); free(o->ci[b]); }
    free(o->off);
    free(o->cache);
    free(o->colour_ptr);
    free(o->colour_elem);
    free(o);
}

static oracle* oracle_create(const pg_problem_desc* d, int e0, int e1, int own_pattern);

/* Build the pattern over elements [e0, e1) and cache their shape data. */
oracle* oracle_new(const pg_problem_desc* d, int e0, int e1) { return oracle_create(d, e0, e1, 1); }

/* The same over the description's own 4-block pattern (the harness's): K
 * values land in the layout of d->rp / d->ci, and an element entry missing
 * from it is an error.  Lets a large mesh be checked in element ranges with
 * bounded memory (oracle_assemble_range). */
oracle* oracle_new_desc_pattern(const pg_problem_desc* d, int e0, int e1) { return oracle_create(d, e0, e1, 0); }

static oracle* oracle_create(const pg_problem_desc* d, int e0, int e1, int own_pattern) {
    oracle* o = (oracle*)calloc(1, sizeof(oracle));
    if (!o) return NULL;
    o->d = d;
    o->e0 = e0 < 0 ? 0 : e0;
    o->e1 = (e1 < 0 || e1 > d->n_elem) ? d->n_elem : e1;
    o->nndof = d->nndof ? d->nndof : d->dim;
    o->nstr = o->nndof == 1 ? d->dim : (d->dim == 2 ? 3 : 6);
    o->own_pattern = own_pattern;
    int nf = d->n_free, np = d->n_pres;
    o->nrows[0] = nf; o->nrows[1] = nf; o->nrows[2] = np; o->nrows[3] = np;
    if (!own_pattern) {
        int64_t acc = 0;
        for (int b = 0; b < 4; ++b) {
            o->rp[b] = (int32_t*)d->rp[b];
            o->ci[b] = (int32_t*)d->ci[b];
            o->nnz[b] = d->nnz[b];
            o->base[b] = acc;
            acc += d->nnz[b];
        }
        goto cache;
    }
    for (int b = 0; b < 4; ++b) {
        o->rows[b] = (row_t*)calloc((size_t)(o->nrows[b] ? o->nrows[b] : 1), sizeof(row_t));
        if (!o->rows[b]) { oracle_free(o); return NULL; }
    }
    int32_t dofs[MAXND];
    for (int e = o->e0; e < o->e1; ++e) {
        int nd = elem_dofs(d, e, dofs);
        for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j) {
                int32_t r = dofs[i], c = dofs[j];
                int b = (r < nf ? 0 : 2) + (c < nf ? 0 : 1);
                if (row_insert(&o->rows[b][r < nf ? r : r - nf], c < nf ? c : c - nf)) { oracle_free(o); return NULL; }
            }
    }
    int64_t acc = 0;
    for (int b = 0; b < 4; ++b) {
        o->rp[b] = (int32_t*)malloc(sizeof(int32_t) * ((size_t)o->nrows[b] + 1));
        int64_t n = 0;
        for (int r = 0; r < o->nrows[b]; ++r) n += o->rows[b][r].n;
        o->ci[b] = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
        if (!o->rp[b] || !o->ci[b]) { oracle_free(o); return NULL; }
        o->rp[b][0] = 0;
        int64_t k = 0;
        for (int r = 0; r < o->nrows[b]; ++r) {
            memcpy(o->ci[b] + k, o->rows[b][r].c, sizeof(int32_t) * o->rows[b][r].n);
            k += o->rows[b][r].n;
            o->rp[b][r + 1] = (int32_t)k;
        }
        o->nnz[b] = n;
        o->base[b] = acc;
        acc += n;
    }
    oracle_free_rows(o);
cache:
    /* shape cache */
    o->off = (int64_t*)malloc(sizeof(int64_t) * ((size_t)(o->e1 - o->e0) + 1));
    if (!o->off) { oracle_free(o); return NULL; }
    o->off[0] = 0;
    for (int e = o->e0; e < o->e1; ++e) o->off[e - o->e0 + 1] = o->off[e - o->e0] + cache_size(o, e);
    o->cache = (double*)malloc(sizeof(double) * (size_t)(o->off[o->e1 - o->e0] + 1));
    if (!o->cache) { oracle_free(o); return NULL; }
    /* elements fill disjoint cache slots, so they run in parallel; the first
     * element (in element order) that fails is then run again alone to keep
     * its message */
    int bad = o->e1;
#pragma omp parallel for schedule(static, 4096) reduction(min : bad)
    for (int e = o->e0; e < o->e1; ++e) {
        oracle tmp = *o;
        if (store_shapes(&tmp, e, o->cache + o->off[e - o->e0]) && e < bad) bad = e;
    }
    o->err[0] = 0;
    if (bad < o->e1) store_shapes(o, bad, o->cache + o->off[bad - o->e0]); /* error kept in o->err */
    return o;
}

const char* oracle_error(const oracle* o) { return o ? o->err : "null oracle"; }

void oracle_nnz(const oracle* o, int64_t* nnz) { for (int b = 0; b < 4; ++b) nnz[b] = o->nnz[b]; }

void oracle_pattern(const oracle* o, int b, const int32_t** rp, const int32_t** ci) {
    *rp = o->rp[b];
    *ci = o->ci[b];
}

/* ------------------------------------------------------------------------ */
/* Material::Constitutive for the laws on the path */
static void lei(const pg_material* m, int dim, const double* e, double* s) {
    const double* D = m->D;
    if (dim == 2) {
        s[0] = D[0] * e[0] + D[1] * e[1];
        s[1] = D[3] * e[0] + D[4] * e[1];
        s[2] = D[8] * e[2];
    } else {
        s[0] = D[0] * e[0] + D[1] * e[1] + D[2] * e[2];
        s[1] = D[6] * e[0] + D[7] * e[1] + D[8] * e[2];
        s[2] = D[12] * e[0] + D[13] * e[1] + D[14] * e[2];
        s[3] = D[21] * e[3];
        s[4] = D[28] * e[4];
        s[5] = D[35] * e[5];
    }
}

static double kron(int i, int j) { return i == j ? 1.0 : 0.0; }

/* OPCohesiveNUMaterial::getDeltaEff (src/Material_OPCohesiveNU.cpp:43-61): an
 * opening of squared norm below 1e-40 becomes 1e-20 in every component. */
static double op_delta_eff(double* delta, int dim, double b2) {
    double norm = 0;
    for (int i = 0; i < dim; ++i) norm += delta[i] * delta[i];
    if (norm < 1e-40)
        for (int i = 0; i < dim; ++i) delta[i] = 1e-20;
    double de = pow(delta[dim - 1], 2);
    for (int i = 0; i < dim - 1; i++) de += b2 * pow(delta[i], 2);
    return sqrt(de);
}

static void cohesive(const pg_material* m, int dim, double* delta, double* dmax_io, double* t, double Dm[3][3]) {
    double norm = 0;
    for (int i = 0; i < dim; ++i) norm += delta[i] * delta[i];
    if (norm < 1e-40)
        for (int i = 0; i < dim; ++i) delta[i] = 1e-20;
    double b2 = m->beta2;
    double de = pow(delta[dim - 1], 2);
    for (int i = 0; i < dim - 1; i++) de += b2 * pow(delta[i], 2);
    de = sqrt(de);
    double tr = 0, dtr = 0, dmax = *dmax_io;
    int bl = (m->type == PG_MAT_BL || m->type == PG_MAT_BL_NU);
    if (bl) {
        double sc = m->sigma_c, d1 = m->delta_c1, d2 = m->delta_c2;
        if (m->type == PG_MAT_BL_NU || de >= dmax || de < 0.0) {
            if (de >= 0 && de < d1) { tr = de * (sc / d1); dtr = (sc / d1); }
            else if (de >= d1 && de < d2) { tr = sc * (1 - (de - d1) / (d2 - d1)); dtr = -1.0 * (sc / (d2 - d1)); }
            else { tr = 0; dtr = 0; }
        } else {
            double tm;
            if (dmax >= 0 && dmax < d1) tm = dmax * (sc / d1);
            else if (dmax >= d1 && dmax < d2) tm = sc * (1 - (dmax - d1) / (d2 - d1));
            else tm = 0;
            tr = tm * (de / dmax);
            dtr = tm / dmax;
        }
    } else {
        double sc = m->sigma_c, dc = m->delta_c;
        if (m->type == PG_MAT_OP_NU || de >= dmax || de < 0.0) {
            tr = sc * (de / dc) * exp(1.0 - de / dc);
            dtr = (sc / dc) * (1.0 - de / dc) * exp(1.0 - de / dc);
            if (m->type == PG_MAT_OP) *dmax_io = de;
        } else {
            double tm = sc * (dmax / dc) * exp(1.0 - dmax / dc);
            tr = tm * (de / dmax);
            dtr = tm / dmax;
        }
    }
    double br[3] = {0, 0, 0}, dd[3] = {0, 0, 0};
    for (int i = 0; i < dim; ++i) br[i] = b2 * delta[i];
    br[dim - 1] += (1.0 - b2) * delta[dim - 1];
    for (int i = 0; i < dim; ++i) t[i] = (tr / de) * br[i];
    dd[dim - 1] = delta[dim - 1] / de;
    for (int i = 0; i < dim - 1; i++) dd[i] = b2 * delta[i] / de;
    for (int i = 0; i < dim; ++i)
        for (int k = 0; k < dim; ++k)
            Dm[i][k] = br[i] * (dtr * de - tr) * (dd[k] / pow(de, 2)) +
                       (tr / de) * (b2 * kron(i, k) + (1.0 - b2) * kron(i, dim - 1) * kron(k, dim - 1));
    if (bl && delta[dim - 1] < 0) {
        t[dim - 1] = (m->sigma_c / m->alpha) * sinh(m->alpha * delta[dim - 1] / m->delta_c1);
        for (int i = 0; i < dim - 1; ++i) Dm[dim - 1][i] = 0.0;
        Dm[dim - 1][dim - 1] = (m->sigma_c / m->delta_c1) * cosh(m->alpha * delta[dim - 1] / m->delta_c1);
    }
}

/* ------------------------------------------------------------------------ */
/* MatSetValue(ADD_VALUES) into the block CSR: binary search of the column */
static int add_value(oracle* o, double* K, int32_t r, int32_t c, double v) {
    int nf = o->d->n_free;
    int b = (r < nf ? 0 : 2) + (c < nf ? 0 : 1);
    int32_t row = r < nf ? r : r - nf, col = c < nf ? c : c - nf;
    const int32_t* beg = o->ci[b] + o->rp[b][row];
    int lo = 0, hi = o->rp[b][row + 1] - o->rp[b][row];
    while (lo < hi) {
        int m = (lo + hi) >> 1;
        if (beg[m] < col) lo = m + 1; else hi = m;
    }
    if (lo >= o->rp[b][row + 1] - o->rp[b][row] || beg[lo] != col) return -1;
    K[o->base[b] + o->rp[b][row] + lo] += v;
    return 0;
}

static int assemble_elem(oracle* o, int e, const double* U, double* isv_t, int func, int jac, double* K, double* P);

/* Assembler::assemble over elements [e0, e1).  K: concatenated [ff|fp|pf|pp]
 * values in this oracle's pattern; P: [Pf|Pp]; isv_t reset from isv_c first. */
int oracle_assemble(oracle* o, const double* U, const double* isv_c, double* isv_t, int func, int jac, double* K,
                    double* P) {
    const pg_problem_desc* d = o->d;
    if (!func && !jac) return 0;
    if (o->err[0]) return -1;
    if (jac) memset(K, 0, sizeof(double) * (size_t)(o->nnz[0] + o->nnz[1] + o->nnz[2] + o->nnz[3]));
    if (func) memset(P, 0, sizeof(double) * (size_t)(d->n_free + d->n_pres));
    if (d->n_cqp > 0) memcpy(isv_t, isv_c, sizeof(double) * (size_t)d->n_cqp); /* InternalVars::copy_data */
    for (int e = o->e0; e < o->e1; ++e)
        if (assemble_elem(o, e, U, isv_t, func, jac, K, P)) return -1;
    return 0;
}

/* assemble_elem: one element of the loop of Assembler::assemble
 * (src/Assembler.cpp:100-140): assemble_elem + set_matrix / set_vec. */
static int assemble_elem(oracle* o, int e, const double* U, double* isv_t, int func, int jac, double* K, double* P) {
    const pg_problem_desc* d = o->d;
    const int dim = d->dim, nstr = o->nstr, nen = d->nen;
    int32_t dofs[MAXND];
    double Ue[MAXND], Ke[MAXND * MAXND], Pe[MAXND], Kt[MAXND * MAXND], Pt[MAXND], BtD[MAXND * 6];
    {
        const double* cc = o->cache + o->off[e - o->e0];
        int nd = elem_dofs(d, e, dofs);
        for (int i = 0; i < nd; ++i) Ue[i] = U[dofs[i]]; /* formElementalVector */
        memset(Ke, 0, sizeof(double) * nd * nd);
        memset(Pe, 0, sizeof(double) * nd);
        int c = d->cut_of_elem[e];
        int nvq = c < 0 ? (dim == 2 ? 4 : 5) : d->cut[c].n_vqp;
        for (int q = 0; q < nvq; ++q) {
            double W = cc[0], J = cc[1];
            const double* B = cc + 2;
            const pg_material* m;
            if (c < 0) m = &d->mats[d->elem_mat[e]];
            else {
                int k = 0, qq = q;
                while (qq >= nqp_of(d->cut[c].child_type[k])) { qq -= nqp_of(d->cut[c].child_type[k]); ++k; }
                m = &d->mats[d->cut[c].child_mat[k]];
            }
            /* KernelVolumetric: strain = B U (dgemv), stress, K = (B^T D) B (dgemm), P = B^T sigma */
            double eps[6], sig[6];
            for (int s = 0; s < nstr; ++s) {
                double t = 0;
                for (int j = 0; j < nd; ++j) t += U[dofs[j]] * B[s * nd + j];
                eps[s] = 0.0 + 1.0 * t;
            }
            if (m->thermal_exp_set) {
                /* LinearElasticIsotropicMaterial::Constitutive with the thermal
                 * expansion set (src/material_lei.cpp:272-291): the normal
                 * strains take the thermal strain, strain = 1.0 * thermal + strain */
                double strain_val = -1.0 * o->temp_change * m->thermal_exp;
                for (int i = 0; i < dim; ++i) eps[i] = 1.0 * strain_val + eps[i];
            }
            if (o->nndof == 1) {
                /* heat conduction: the flux q = D . grad T of the linear law (the
                 * reference's LinearThermalMaterial leaves its output flux unset,
                 * src/material_lt.cpp:51-60) */
                for (int k = 0; k < nstr; ++k) {
                    double t = 0;
                    for (int l = 0; l < nstr; ++l) t += m->D[k * nstr + l] * eps[l];
                    sig[k] = t;
                }
            } else {
                lei(m, dim, eps, sig);
            }
            if (jac) {
                for (int i = 0; i < nd; ++i)
                    for (int l = 0; l < nstr; ++l) {
                        double t = 0;
                        for (int k = 0; k < nstr; ++k) t += B[k * nd + i] * m->D[k * nstr + l];
                        BtD[i * nstr + l] = t;
                    }
                for (int i = 0; i < nd; ++i)
                    for (int j = 0; j < nd; ++j) {
                        double t = 0;
                        for (int l = 0; l < nstr; ++l) t += BtD[i * nstr + l] * B[l * nd + j];
                        Kt[i * nd + j] = t;
                    }
            }
            if (func)
                for (int i = 0; i < nd; ++i) {
                    double t = 0;
                    for (int k = 0; k < nstr; ++k) t += B[k * nd + i] * sig[k];
                    Pt[i] = t;
                }
            double wgt = J * W;
            if (jac)
                for (int i = 0; i < nd * nd; ++i) Ke[i] += wgt * Kt[i];
            if (func)
                for (int i = 0; i < nd; ++i) Pe[i] += wgt * Pt[i];
            cc += 2 + nstr * nd;
        }
        if (c >= 0 && d->cohesive) {
            const pg_cut_rec* r = &d->cut[c];
            int cnn = r->coh_nn / 2, cnq = r->coh_nn == 4 ? 2 : 4, ncd = r->coh_nn * dim;
            int cq = 0;
            for (int k = 0; k < r->n_coh; ++k) {
                double Kc[24 * 24], Pc[24], Kq[24 * 24], Pq[24], Uc[24];
                memset(Kc, 0, sizeof(double) * ncd * ncd);
                memset(Pc, 0, sizeof(double) * ncd);
                for (int n = 0; n < r->coh_nn; ++n)
                    for (int j = 0; j < dim; ++j) Uc[n * dim + j] = Ue[r->coh_nodes[k][n] * dim + j];
                for (int q = 0; q < cnq; ++q, ++cq) {
                    double W = cc[0], J = cc[1];
                    const double* N = cc + 2;
                    double R[3][3];
                    for (int i = 0; i < dim; ++i)
                        for (int j = 0; j < dim; ++j) R[i][j] = cc[6 + i * dim + j];
                    /* cohesiveDeltaFast */
                    double dl[3] = {0, 0, 0}, del[3];
                    for (int i = 0; i < dim; ++i)
                        for (int n = 0; n < cnn; ++n) dl[i] += N[n] * (Uc[(n + cnn) * dim + i] - Uc[n * dim + i]);
                    for (int i = 0; i < dim; ++i) {
                        double t = 0;
                        for (int j = 0; j < dim; ++j) t += R[i][j] * dl[j];
                        del[i] = t;
                    }
                    double tr[3], Dm[3][3];
                    int iq = r->isv_off + cq;
                    cohesive(&d->mats[r->coh_mat[k]], dim, del, &isv_t[iq], tr, Dm);
                    /* cohesiveKFast: RDR = (R^T D) R */
                    double RtD[3][3], RDR[3][3], tR[3];
                    for (int i = 0; i < dim; ++i)
                        for (int j = 0; j < dim; ++j) {
                            double t = 0;
                            for (int l = 0; l < dim; ++l) t += R[l][i] * Dm[l][j];
                            RtD[i][j] = t;
                        }
                    for (int i = 0; i < dim; ++i)
                        for (int j = 0; j < dim; ++j) {
                            double t = 0;
                            for (int l = 0; l < dim; ++l) t += RtD[i][l] * R[l][j];
                            RDR[i][j] = t;
                        }
                    for (int i = 0; i < dim; ++i) {
                        double t = 0;
                        for (int l = 0; l < dim; ++l) t += R[l][i] * tr[l];
                        tR[i] = t;
                    }
                    for (int n1 = 0; n1 < cnn; ++n1)
                        for (int n2 = 0; n2 < cnn; ++n2)
                            for (int d1 = 0; d1 < dim; ++d1)
                                for (int d2 = 0; d2 < dim; ++d2) {
                                    double v = N[n1] * N[n2] * RDR[d1][d2];
                                    Kq[(n1 * dim + d1) * ncd + n2 * dim + d2] = v;
                                    Kq[((n1 + cnn) * dim + d1) * ncd + n2 * dim + d2] = -v;
                                    Kq[(n1 * dim + d1) * ncd + (n2 + cnn) * dim + d2] = -v;
                                    Kq[((n1 + cnn) * dim + d1) * ncd + (n2 + cnn) * dim + d2] = v;
                                }
                    for (int n = 0; n < cnn; ++n)
                        for (int dd = 0; dd < dim; ++dd) {
                            double v = N[n] * tR[dd];
                            Pq[n * dim + dd] = -v;
                            Pq[(n + cnn) * dim + dd] = v;
                        }
                    double wgt = J * W;
                    if (jac)
                        for (int i = 0; i < ncd * ncd; ++i) Kc[i] += wgt * Kq[i];
                    if (func)
                        for (int i = 0; i < ncd; ++i) Pc[i] += wgt * Pq[i];
                    cc += 2 + 4 + dim * dim;
                }
                /* fold into the element matrix */
                for (int n1 = 0; n1 < r->coh_nn; ++n1) {
                    int rn = r->coh_nodes[k][n1];
                    for (int d1 = 0; d1 < dim; ++d1) {
                        int rd = rn * dim + d1, rl = n1 * dim + d1;
                        if (jac)
                            for (int n2 = 0; n2 < r->coh_nn; ++n2) {
                                int cn = r->coh_nodes[k][n2];
                                for (int d2 = 0; d2 < dim; ++d2)
                                    Ke[rd * nd + cn * dim + d2] += Kc[rl * ncd + n2 * dim + d2];
                            }
                        if (func) Pe[rd] += Pc[rl];
                    }
                }
            }
        }
        /* set_matrix / set_vec */
        if (jac)
            for (int i = 0; i < nd; ++i)
                for (int j = 0; j < nd; ++j)
                    if (add_value(o, K, dofs[i], dofs[j], Ke[i * nd + j])) { strcpy(o->err, "entry missing from pattern"); return -1; }
        if (func)
            for (int i = 0; i < nd; ++i)
                if (Pe[i] != 0.0) P[dofs[i]] += Pe[i];
    }
    (void)nen;
    return 0;
}

/* Greedy element colouring through the node -> element incidence: an element
 * takes the smallest colour none of its already coloured neighbours (elements
 * sharing a node, enrichment nodes included) holds. */
static int build_colours(oracle* o) {
    const pg_problem_desc* d = o->d;
    const int ne = o->e1 - o->e0;
    int32_t* col = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ne ? ne : 1));
    if (!col) return -1;
    int nc = 0;
    unsigned char used[256];
    int32_t ln[4 + PG_MAX_ENR];
    for (int e = o->e0; e < o->e1; ++e) {
        memset(used, 0, sizeof used);
        int nl = local_list(d, e, ln);
        for (int n = 0; n < nl; ++n)
            for (int32_t t = d->n2e_ptr[ln[n]]; t < d->n2e_ptr[ln[n] + 1]; ++t) {
                int f = d->n2e[t];
                if (f >= o->e0 && f < e) used[col[f - o->e0]] = 1;
            }
        int c = 0;
        while (c < 256 && used[c]) ++c;
        if (c == 256) { free(col); strcpy(o->err, "element colouring needs more than 256 colours"); return -1; }
        col[e - o->e0] = c;
        if (c + 1 > nc) nc = c + 1;
    }
    o->colour_ptr = (int32_t*)calloc((size_t)nc + 1, sizeof(int32_t));
    o->colour_elem = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ne ? ne : 1));
    if (!o->colour_ptr || !o->colour_elem) { free(col); return -1; }
    for (int k = 0; k < ne; ++k) o->colour_ptr[col[k] + 1]++;
    for (int c = 0; c < nc; ++c) o->colour_ptr[c + 1] += o->colour_ptr[c];
    int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nc ? nc : 1));
    if (!fill) { free(col); return -1; }
    memcpy(fill, o->colour_ptr, sizeof(int32_t) * (size_t)nc);
    for (int k = 0; k < ne; ++k) o->colour_elem[fill[col[k]]++] = o->e0 + k;
    free(fill);
    free(col);
    o->ncolour = nc;
    return 0;
}

/* The same assembly on `nthreads` OpenMP threads, colour by colour (the
 * threaded CPU baseline; within a colour no two elements write the same K
 * row, P entry or ISV).  The sums then run in colour order, not element
 * order: the results agree with oracle_assemble to rounding, not bit for bit. */
int oracle_assemble_threads(oracle* o, const double* U, const double* isv_c, double* isv_t, int func, int jac,
                            double* K, double* P, int nthreads) {
    const pg_problem_desc* d = o->d;
    if (!func && !jac) return 0;
    if (o->err[0]) return -1;
    if (!o->colour_ptr && build_colours(o)) return -1;
    if (jac) memset(K, 0, sizeof(double) * (size_t)(o->nnz[0] + o->nnz[1] + o->nnz[2] + o->nnz[3]));
    if (func) memset(P, 0, sizeof(double) * (size_t)(d->n_free + d->n_pres));
    if (d->n_cqp > 0) memcpy(isv_t, isv_c, sizeof(double) * (size_t)d->n_cqp);
    int bad = 0;
    for (int c = 0; c < o->ncolour; ++c) {
        const int k0 = o->colour_ptr[c], k1 = o->colour_ptr[c + 1];
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(| : bad)
        for (int k = k0; k < k1; ++k)
            bad |= assemble_elem(o, o->colour_elem[k], U, isv_t, func, jac, K, P) != 0;
    }
    return bad ? -1 : 0;
}

/* Assembly of this oracle's element range into K / P that already hold the
 * ranges before it: first != 0 clears K and P and resets the trial ISVs from
 * the committed ones (as oracle_assemble does), first == 0 adds to them.  Over
 * consecutive ranges in element order, nthreads == 0 gives oracle_assemble's
 * result over their union bit for bit (every entry receives its adds in
 * element order); nthreads > 0 runs each range colour by colour. */
int oracle_assemble_range(oracle* o, const double* U, const double* isv_c, double* isv_t, int func, int jac,
                          double* K, double* P, int first, int nthreads) {
    const pg_problem_desc* d = o->d;
    if (!func && !jac) return 0;
    if (o->err[0]) return -1;
    if (first) {
        if (jac) memset(K, 0, sizeof(double) * (size_t)(o->nnz[0] + o->nnz[1] + o->nnz[2] + o->nnz[3]));
        if (func) memset(P, 0, sizeof(double) * (size_t)(d->n_free + d->n_pres));
        if (d->n_cqp > 0) memcpy(isv_t, isv_c, sizeof(double) * (size_t)d->n_cqp);
    }
    if (nthreads <= 0) {
        for (int e = o->e0; e < o->e1; ++e)
            if (assemble_elem(o, e, U, isv_t, func, jac, K, P)) return -1;
        return 0;
    }
    if (!o->colour_ptr && build_colours(o)) return -1;
    int bad = 0;
    for (int c = 0; c < o->ncolour; ++c) {
        const int k0 = o->colour_ptr[c], k1 = o->colour_ptr[c + 1];
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(| : bad)
        for (int k = k0; k < k1; ++k)
            bad |= assemble_elem(o, o->colour_elem[k], U, isv_t, func, jac, K, P) != 0;
    }
    return bad ? -1 : 0;
}

void oracle_set_temperature_change(oracle* o, double temp_change) { o->temp_change = temp_change; }

/* ------------------------------------------------------------------------ */
/* Material-parameter sensitivity dP/dd */

/* OPCohesiveMaterial::SensitivityConstitutive (src/Material_OPCohesive.cpp:104-180)
 * and OPCohesiveNUMaterial::SensitivityConstitutive (src/Material_OPCohesiveNU.cpp:117-170):
 * the partial derivative of the traction, dt, at opening delta (ttn frame, modified
 * in place by getDeltaEff as the reference does); right: this law is the one
 * differentiated; dmax / ddmax: delta_max and its sensitivity (OP only). */
static void cohesive_sens(const pg_material* m, int dim, double* delta, int right, int kind, double dmax,
                          double ddmax, double* dt) {
    for (int i = 0; i < dim; ++i) dt[i] = 0.0;
    const double sc = m->sigma_c, dc = m->delta_c, b2 = m->beta2;
    double br[3];
    if (m->type == PG_MAT_OP) {
        double de = op_delta_eff(delta, dim, b2);
        for (int i = 0; i < dim; ++i) br[i] = delta[dim - 1];
        for (int i = 0; i < dim - 1; ++i) br[i] = b2 * delta[i];
        int loading = de >= dmax;
        if (right) {
            double dte = 0.0;
            if (kind == PG_SENS_SIGMA_C)
                dte = loading ? de / dc * exp(1 - de / dc) : de / dc * exp(1 - dmax / dc);
            else if (kind == PG_SENS_DELTA_C)
                dte = loading ? (sc * de / (dc * dc) * exp(1 - de / dc)) * (de / dc - 1)
                              : (sc * de / (dc * dc) * exp(1 - dmax / dc)) * (dmax / dc - 1);
            for (int i = 0; i < dim; ++i) dt[i] = (dte / de) * br[i];
        }
        if (!loading) {
            double dtm = (sc * de / dc) * exp(1 - dmax / dc) * (-1.0 / dc);
            for (int i = 0; i < dim; ++i) dt[i] += ddmax * (dtm / de) * br[i];
        }
    } else if (m->type == PG_MAT_OP_NU && right) {
        double de = op_delta_eff(delta, dim, b2);
        for (int i = 0; i < dim; ++i) br[i] = delta[dim - 1];
        for (int i = 0; i < dim - 1; ++i) br[i] = b2 * delta[i];
        double dte = 0.0;
        if (kind == PG_SENS_SIGMA_C) dte = de / dc * exp(1 - de / dc);
        else if (kind == PG_SENS_DELTA_C) dte = (sc * de / dc / dc * exp(1 - de / dc)) * (de / dc - 1);
        for (int i = 0; i < dim; ++i) dt[i] = (dte / de) * br[i];
    }
}

/* SensitivityParameter::assemble_dPdd (src/SensitivityParameter.cpp:108-151) with
 * SensitivityMaterialParameter::assemble_elem_dPdd (src/SensitivityMaterialParameter.cpp:54-264)
 * and the B-matrix kernels of SensitivityMaterialParameterStructural
 * (src/SensitivityMaterialParameterStructural.cpp:40-121): dP = [dP_f | dP_p],
 * overwritten.  Returns -2 where the reference stops with an error. */
int oracle_sens_dPdd(oracle* o, const pg_sens_param* prm, unsigned flags, const double* U, const double* isv,
                     const double* isv_sens, double* dP) {
    const pg_problem_desc* d = o->d;
    const int dim = d->dim, nstr = o->nstr, nen = d->nen;
    const int tgt = prm->mat;
    if (tgt < 0 || tgt >= d->n_mat) { strcpy(o->err, "bad material"); return -1; }
    const pg_material* tm = &d->mats[tgt];
    if (prm->kind == PG_SENS_LEI ? tm->type != PG_MAT_LEI
                                 : !(tm->type == PG_MAT_OP || tm->type == PG_MAT_OP_NU)) {
        strcpy(o->err, "parameter kind does not fit the material");
        return -2;
    }
    memset(dP, 0, sizeof(double) * (size_t)(d->n_free + d->n_pres));
    for (int e = o->e0; e < o->e1; ++e) {
        int c = d->cut_of_elem[e];
        const pg_cut_rec* r = c >= 0 ? &d->cut[c] : NULL;
        /* _assemble_elem (set_material_element_assemblies) || elemHasISVs */
        int visit = 0;
        if (!r) visit = d->elem_mat[e] == tgt;
        else {
            for (int k = 0; k < r->n_child; ++k) visit |= r->child_mat[k] == tgt;
            if (d->cohesive)
                for (int k = 0; k < r->n_coh; ++k) {
                    int t = d->mats[r->coh_mat[k]].type;
                    visit |= r->coh_mat[k] == tgt || t == PG_MAT_OP || t == PG_MAT_BL;
                }
        }
        if (!visit) continue;
        int32_t dofs[MAXND];
        double Ue[MAXND], dPe[MAXND], tmp[MAXND];
        int nd = elem_dofs(d, e, dofs);
        for (int i = 0; i < nd; ++i) Ue[i] = U[dofs[i]];
        memset(dPe, 0, sizeof(double) * nd);
        const double* cc = o->cache + o->off[e - o->e0];
        const int vstride = 2 + nstr * nd;
        /* one volumetric qp: strain = B U, dstress = dD strain, temp = B^T dstress */
        #define SENS_VQP(Bq, wgt)                                                        \
            do {                                                                         \
                double eps_[6], ds_[6];                                                  \
                for (int s_ = 0; s_ < nstr; ++s_) {                                      \
                    double t_ = 0;                                                       \
                    for (int j_ = 0; j_ < nd; ++j_) t_ += Ue[j_] * (Bq)[s_ * nd + j_];   \
                    eps_[s_] = 0.0 + 1.0 * t_;                                           \
                }                                                                        \
                pg_material sm_ = *tm;                                                   \
                memcpy(sm_.D, prm->dD, sizeof(sm_.D));                                   \
                lei(&sm_, dim, eps_, ds_);                                               \
                for (int i_ = 0; i_ < nd; ++i_) {                                        \
                    double t_ = 0;                                                       \
                    for (int k_ = 0; k_ < nstr; ++k_) t_ += (Bq)[k_ * nd + i_] * ds_[k_]; \
                    tmp[i_] = t_;                                                        \
                }                                                                        \
                for (int i_ = 0; i_ < nd; ++i_) dPe[i_] += (wgt) * tmp[i_];              \
            } while (0)
        if (!r) {
            if (d->elem_mat[e] == tgt)
                for (int q = 0; q < (dim == 2 ? 4 : 5); ++q) {
                    const double* q0 = cc + q * vstride;
                    SENS_VQP(q0 + 2, q0[1] * q0[0]);
                }
        } else {
            int cq = 0; /* curr_qp */
            for (int k = 0; k < r->n_child; ++k) {
                int nq = nqp_of(r->child_type[k]);
                if (r->child_mat[k] != tgt) { cq += nq; continue; }
                for (int q = 0; q < nq; ++q, ++cq) {
                    /* getBmat(l_elem, qp): the child-local index (the reference's
                     * call); J and W at curr_qp */
                    const double* Bq = cc + (size_t)((flags & PG_SENS_OWN_CHILD_B) ? cq : q) * vstride + 2;
                    const double* jw = cc + (size_t)cq * vstride;
                    SENS_VQP(Bq, jw[1] * jw[0]);
                }
            }
            if (d->cohesive) {
                const double* ccq = cc + (size_t)r->n_vqp * vstride;
                int cnn = r->coh_nn / 2, cnq = r->coh_nn == 4 ? 2 : 4, ncd = r->coh_nn * dim;
                for (int k = 0; k < r->n_coh; ++k) {
                    const pg_material* cm = &d->mats[r->coh_mat[k]];
                    int right = r->coh_mat[k] == tgt;
                    int has_isv = cm->type == PG_MAT_OP || cm->type == PG_MAT_BL;
                    if (!(has_isv || right)) continue;
                    if (cm->type == PG_MAT_BL || cm->type == PG_MAT_BL_NU) {
                        strcpy(o->err, "Sensitivity is not implemented for the bilinear cohesive model");
                        return -2;
                    }
                    double dPc[24], Uc[24];
                    memset(dPc, 0, sizeof(double) * ncd);
                    for (int n = 0; n < r->coh_nn; ++n)
                        for (int j = 0; j < dim; ++j) Uc[n * dim + j] = Ue[r->coh_nodes[k][n] * dim + j];
                    for (int q = 0; q < cnq; ++q) {
                        const double* c0 = ccq + (size_t)(k * cnq + q) * (2 + 4 + dim * dim);
                        const double W = c0[0], J = c0[1];
                        const double* N = c0 + 2;
                        double R[3][3];
                        for (int i = 0; i < dim; ++i)
                            for (int j = 0; j < dim; ++j) R[i][j] = c0[6 + i * dim + j];
                        double dl[3] = {0, 0, 0}, del[3], dt[3], tR[3];
                        for (int i = 0; i < dim; ++i)
                            for (int n = 0; n < cnn; ++n) dl[i] += N[n] * (Uc[(n + cnn) * dim + i] - Uc[n * dim + i]);
                        for (int i = 0; i < dim; ++i) {
                            double t = 0;
                            for (int j = 0; j < dim; ++j) t += R[i][j] * dl[j];
                            del[i] = t;
                        }
                        int iq = r->isv_off + k * cnq + q;
                        cohesive_sens(cm, dim, del, right, prm->kind, isv ? isv[iq] : 0.0,
                                      isv_sens ? isv_sens[iq] : 0.0, dt);
                        for (int i = 0; i < dim; ++i) {
                            double t = 0;
                            for (int l = 0; l < dim; ++l) t += R[l][i] * dt[l];
                            tR[i] = t;
                        }
                        for (int n = 0; n < cnn; ++n)
                            for (int dd = 0; dd < dim; ++dd) {
                                double v = N[n] * tR[dd];
                                tmp[n * dim + dd] = -v;
                                tmp[(n + cnn) * dim + dd] = v;
                            }
                        for (int i = 0; i < ncd; ++i) dPc[i] += (J * W) * tmp[i];
                    }
                    for (int n1 = 0; n1 < r->coh_nn; ++n1)
                        for (int d1 = 0; d1 < dim; ++d1) dPe[r->coh_nodes[k][n1] * dim + d1] += dPc[n1 * dim + d1];
                }
            }
        }
        #undef SENS_VQP
        /* set_dP: nonzero entries added */
        for (int i = 0; i < nd; ++i)
            if (dPe[i] != 0.0) dP[dofs[i]] += dPe[i];
    }
    (void)nen;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Test hooks: the oracle's constitutive routines, exposed so that tests can pin
 * them against closed forms and finite differences. */
void oracle_cohesive_law(const pg_material* m, int dim, double* delta, double* isv, double* t, double* D) {
    double Dm[3][3];
    cohesive(m, dim, delta, isv, t, Dm);
    for (int i = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j) D[i * dim + j] = Dm[i][j];
}

void oracle_lei_stress(const pg_material* m, int dim, const double* e, double* s) { lei(m, dim, e, s); }

/* Element-level shape data of element e: per volumetric qp the weight W, the
 * Jacobian J and the B matrix (nstr x ndof); used by tests of the quadrature
 * and of partition of unity on cut elements. Returns the number of doubles. */
int64_t oracle_elem_cache(const oracle* o, int e, const double** out) {
    if (e < o->e0 || e >= o->e1) return -1;
    *out = o->cache + o->off[e - o->e0];
    return o->off[e - o->e0 + 1] - o->off[e - o->e0];
}

/* Test access to the child shape data: for child type `type` with physical
 * node coordinates cx (nn x 3, row-major), qp q gives the reference point xi,
 * the weight w, the shape values N, the physical gradients grad (nn x 3) and
 * the Jacobian J (Elem::q_point + Elem::ShapeFunctions). */
int oracle_child_qp(int dim, int type, const double* cx, int q, double* xi, double* w, double* N, double* grad,
                    double* J) {
    const int nn = nn_of(type);
    if (q < 0 || q >= nqp_of(type)) return -1;
    double x[6][3], dN[6][3] = {{0}}, g[6][3];
    for (int n = 0; n < nn; ++n)
        for (int i = 0; i < 3; ++i) x[n][i] = cx[n * 3 + i];
    qpoint(type, q, xi, w);
    shape(type, xi, N, dN);
    *J = transform(dim, nn, dN, x, g);
    for (int n = 0; n < nn; ++n)
        for (int i = 0; i < 3; ++i) grad[n * 3 + i] = i < dim ? g[n][i] : 0.0;
    return 0;
}

/* Polygon shape values and reference gradients (nn x 2) at reference point xi. */
void oracle_polygon_shape(int nn, const double* xi, double* N, double* dN) {
    double d[6][3];
    poly_shape(nn, xi, N, d);
    for (int n = 0; n < nn; ++n) {
        dN[n * 2] = d[n][0];
        dN[n * 2 + 1] = d[n][1];
    }
}
