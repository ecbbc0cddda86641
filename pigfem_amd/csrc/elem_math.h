// Device element math shared by the assembly kernels: quadrature tables, shape
// functions, Jacobian transform, Voigt B columns, linear-elastic stress and the
// cohesive traction-separation laws.  Each routine names the reference routine
// it restates (brandyb2/PIGFEM, file:line).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "pgfem_asm.h"

namespace pg {

// Tri3 order-3 rule, src/tri3.cpp:118-126 (weights include the negative centroid
// weight; coordinates are the reference's decimal literals).
__constant__ const double TRI_QX[4] = {0.333333333333333333333333333333333, 0.200000000000000000000000000000000,
                                       0.200000000000000000000000000000000, 0.600000000000000000000000000000000};
__constant__ const double TRI_QY[4] = {0.333333333333333333333333333333333, 0.200000000000000000000000000000000,
                                       0.600000000000000000000000000000000, 0.200000000000000000000000000000000};
__constant__ const double TRI_QW[4] = {-0.281250000000000000000000000000000, 0.260416666666666666666666666666665,
                                       0.260416666666666666666666666666665, 0.260416666666666666666666666666665};
// Edge2 2-point Gauss (src/edge2.cpp:118-126); Quad4 tensor rule qp -> (qp%2, qp/2)
// (src/quad4.cpp:109-126); CohEdge2 uses the same points (src/CohEdge2.cpp:70-78).
#define PG_GAUSS2 0.57735026918962576450914878050196
// Tet4 order-3 rule, src/tet4.cpp:113-121 (weights sum to 1.0).
__constant__ const double TET_Q[5][3] = {{0.25, 0.25, 0.25},
                                         {0.5, 0.166666666666667, 0.166666666666667},
                                         {0.166666666666667, 0.166666666666667, 0.166666666666667},
                                         {0.166666666666667, 0.166666666666667, 0.5},
                                         {0.166666666666667, 0.5, 0.166666666666667}};
__constant__ const double TET_QW[5] = {-0.8, 0.45, 0.45, 0.45, 0.45};

// Entry q of a rule table for a q that varies per lane, as a chain of selects
// between immediates (an indexed load would put a memory round trip in the
// middle of the qp arithmetic).
template <int N>
__device__ inline double pick(const double (&t)[N], int q) {
    double v = t[N - 1];
#pragma unroll
    for (int k = N - 2; k >= 0; --k) {
        double c = t[k];
        asm("" : "+v"(c));  // keeps the chain from being folded back into a table load
        v = q == k ? c : v;
    }
    return v;
}

// Reference-coordinate gradients of the linear simplex shape functions
// (Tri3::compute_shape_grad src/tri3.cpp:219-225, Tet4 src/tet4.cpp:147-154).
__device__ inline double simplex_dNdr(int n, int a) {
    return n == 0 ? -1.0 : (n - 1 == a ? 1.0 : 0.0);
}

// Quad4::compute_shape_grad, src/quad4.cpp:138-148.
__device__ inline void quad4_dNdr(double r, double s, double dN[4][2]) {
    dN[0][0] = -0.25 * (1.0 - s); dN[0][1] = -0.25 * (1.0 - r);
    dN[1][0] = 0.25 * (1.0 - s);  dN[1][1] = -0.25 * (1.0 + r);
    dN[2][0] = 0.25 * (1.0 + s);  dN[2][1] = 0.25 * (1.0 + r);
    dN[3][0] = -0.25 * (1.0 + s); dN[3][1] = 0.25 * (1.0 - r);
}

// Elem::transform_gradients (src/elem.cpp:125-142) with Utilities::det /
// Utilities::inverse (src/Utilities.cpp:1017-1070): Jac(a,i) = sum_n dNdr[n][a] x_n[i],
// J = det(Jac), dNdx[n] = Jac^-1 dNdr[n] (adjugate scaled by 1/det).
template <int D>
__device__ inline double transform(int nn, const double (*dNdr)[D], const double (*x)[3], double (*dNdx)[3]) {
    double Jac[D][D];
    for (int a = 0; a < D; ++a)
        for (int i = 0; i < D; ++i) {
            double s = 0.0;
            for (int n = 0; n < nn; ++n) s += dNdr[n][a] * x[n][i];
            Jac[a][i] = s;
        }
    double J, inv[D][D];
    if constexpr (D == 2) {
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
    const double rd = 1.0 / J;
    for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b) inv[a][b] *= rd;
    for (int n = 0; n < nn; ++n)
        for (int a = 0; a < D; ++a) {
            double s = 0.0;
            for (int b = 0; b < D; ++b) s += inv[a][b] * dNdr[n][b];
            dNdx[n][a] = s;
        }
    return J;
}

// Nonzero entries of Voigt B column (node gradient g, component c):
// Assemble_Small_Strain_B_Mat, src/Utilities.cpp:68-127.  2D rows [xx, yy, xy];
// 3D rows [xx, yy, zz, yz, xz, xy].  Writes the full nstr-vector.
template <int D, typename T>
__device__ inline void bcol(const T* g, int c, T* b) {
    if constexpr (D == 2) {
        b[0] = c == 0 ? g[0] : T(0);
        b[1] = c == 1 ? g[1] : T(0);
        b[2] = c == 0 ? g[1] : g[0];
    } else {
        b[0] = c == 0 ? g[0] : T(0);
        b[1] = c == 1 ? g[1] : T(0);
        b[2] = c == 2 ? g[2] : T(0);
        b[3] = c == 0 ? T(0) : (c == 1 ? g[2] : g[1]);
        b[4] = c == 1 ? T(0) : (c == 0 ? g[2] : g[0]);
        b[5] = c == 2 ? T(0) : (c == 0 ? g[1] : g[0]);
    }
}

// The same column as (row, value) pairs of its D nonzeros, rows ascending: the
// products a dense column would add as exact zeros are skipped, which leaves
// every sum unchanged.
template <int D, typename T>
__device__ inline void bnz(const T* g, int c, int* r, T* v) {
    if constexpr (D == 2) {
        if (c == 0) { r[0] = 0; v[0] = g[0]; r[1] = 2; v[1] = g[1]; }
        else        { r[0] = 1; v[0] = g[1]; r[1] = 2; v[1] = g[0]; }
    } else {
        if (c == 0)      { r[0] = 0; v[0] = g[0]; r[1] = 4; v[1] = g[2]; r[2] = 5; v[2] = g[1]; }
        else if (c == 1) { r[0] = 1; v[0] = g[1]; r[1] = 3; v[1] = g[2]; r[2] = 5; v[2] = g[0]; }
        else             { r[0] = 2; v[0] = g[2]; r[1] = 3; v[1] = g[1]; r[2] = 4; v[2] = g[0]; }
    }
}

// The d x d block w B_a^T D B_b for the isotropic D of an LEI material
// (D = lambda 1 1^T + mu diag(2, .., 2, 1, .., 1) in Voigt order, the D that
// LinearElasticIsotropicMaterial builds, src/material_lei.cpp:52-97):
//   out[i][j] = w lambda ga_i gb_j + w mu ga_j gb_i + w mu (ga . gb) delta_ij,
// with wl = w lambda and wm = w mu.
// Written with explicit fused multiply-adds so that every kernel that inlines
// it rounds the same way (bit-identical results across the schedules).
template <typename T>
__device__ inline T fused(T a, T b, T c) {
    if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
    else return __builtin_fma(a, b, c);
}
template <int D, typename T>
__device__ inline void iso_block(const T* ga, const T* gb, T wl, T wm, T (*out)[D]) {
    T dot = ga[0] * gb[0];
    for (int k = 1; k < D; ++k) dot = fused(ga[k], gb[k], dot);
    const T wd = wm * dot;
    for (int i = 0; i < D; ++i) {
        const T t = wl * ga[i];
        for (int j = 0; j < D; ++j) {
            const T u = wm * ga[j];
            out[i][j] = fused(t, gb[j], fused(u, gb[i], i == j ? wd : T(0)));
        }
    }
}

// The sum over qps of iso_block(ga, gb_q, wl_q, wm_q) for a constant ga (a
// parent node's gradient): out[i][j] = ga_i vl_j + ga_j vm_i + delta_ij ga . vm
// with vl = sum_q wl_q gb_q and vm = sum_q wm_q gb_q.
template <int D, typename T>
__device__ inline void iso_block_v(const T* ga, const T* vl, const T* vm, T (*out)[D]) {
    T dot = ga[0] * vm[0];
    for (int k = 1; k < D; ++k) dot = fused(ga[k], vm[k], dot);
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) out[i][j] = fused(ga[i], vl[j], fused(ga[j], vm[i], i == j ? dot : T(0)));
}

// Thermal strain of LinearElasticIsotropicMaterial::Constitutive
// (src/material_lei.cpp:272-291): with its thermal expansion set, the material
// adds th = -temp_change * thermal_exp to the normal strains (the first D Voigt
// components) before computing the stress, strain = 1.0 * thermal + strain.
// th = 0 leaves the strain as it is.
template <int D>
__device__ inline void add_thermal(double* e, double th) {
    e[0] = th + e[0];
    e[1] = th + e[1];
    if constexpr (D == 3) e[2] = th + e[2];
}

// LinearElasticIsotropicMaterial::computeStress, src/material_lei.cpp:232-261.
template <int D>
__device__ inline void lei_stress(const double* Dm, const double* e, double* s) {
    if constexpr (D == 2) {
        s[0] = Dm[0] * e[0] + Dm[1] * e[1];
        s[1] = Dm[3] * e[0] + Dm[4] * e[1];
        s[2] = Dm[8] * e[2];
    } else {
        s[0] = Dm[0] * e[0] + Dm[1] * e[1] + Dm[2] * e[2];
        s[1] = Dm[6] * e[0] + Dm[7] * e[1] + Dm[8] * e[2];
        s[2] = Dm[12] * e[0] + Dm[13] * e[1] + Dm[14] * e[2];
        s[3] = Dm[21] * e[3];
        s[4] = Dm[28] * e[4];
        s[5] = Dm[35] * e[5];
    }
}

// The isotropic LEI stress from three entries of D: c0 = D[0][0] (lambda +
// 2 mu), c1 = D[0][1] (lambda) and c2 = D[last][last] (mu).  For the D that
// LinearElasticIsotropicMaterial builds (src/material_lei.cpp:52-97), whose
// equal entries are equal to the bit, this is lei_stress's arithmetic with its
// exact-zero products skipped; the node-group kernels load these three values
// instead of D.  The fourth coefficient is the material's thermal strain
// (add_thermal), 0 without a temperature change.
template <int D> struct LeiCoef {
    static constexpr int N = 4;
    __host__ __device__ static constexpr int index(int k) {
        return k == 0 ? 0 : (k == 1 ? 1 : (D == 2 ? 8 : 35));
    }
};
template <int D>
__device__ inline void lei_stress_c(const double* c, const double* e_in, double* s) {
    double e[D == 2 ? 3 : 6];
    for (int k = 0; k < (D == 2 ? 3 : 6); ++k) e[k] = e_in[k];
    add_thermal<D>(e, c[3]);
    if constexpr (D == 2) {
        s[0] = c[0] * e[0] + c[1] * e[1];
        s[1] = c[1] * e[0] + c[0] * e[1];
        s[2] = c[2] * e[2];
    } else {
        s[0] = c[0] * e[0] + c[1] * e[1] + c[1] * e[2];
        s[1] = c[1] * e[0] + c[0] * e[1] + c[1] * e[2];
        s[2] = c[1] * e[0] + c[1] * e[1] + c[0] * e[2];
        s[3] = c[2] * e[3];
        s[4] = c[2] * e[4];
        s[5] = c[2] * e[5];
    }
}

// Cohesive constitutive update.  getDeltaEff (src/Material_BilinearCohesiveNU.cpp:48-66,
// src/Material_OPCohesiveNU.cpp:43-61) overwrites a near-zero opening with 1e-20;
// traction and tangent follow BilinearCohesiveMaterial::Constitutive
// (src/Material_BilinearCohesive.cpp:38-125, delta_max read but never written),
// BilinearCohesiveNU (src/Material_BilinearCohesiveNU.cpp:69-138),
// OPCohesiveMaterial::Constitutive (src/Material_OPCohesive.cpp:38-98, delta_max
// updated on loading) and OPCohesiveNU (src/Material_OPCohesiveNU.cpp:64-111).
// delta: in/out (ttn frame, normal last); isv: trial delta_max (in/out).
// M: pg_material, or any type with its cohesive fields (type, sigma_c,
// delta_c, delta_c1, delta_c2, beta2, alpha), such as CohMat.
struct CohMat {
    int32_t type;
    double sigma_c, delta_c, delta_c1, delta_c2, beta2, alpha;
};
// the six doubles from sigma_c to alpha are contiguous in both (kernels copy them
// as an array)
static_assert(offsetof(pg_material, alpha) == offsetof(pg_material, sigma_c) + 5 * sizeof(double), "layout");
static_assert(offsetof(CohMat, alpha) == offsetof(CohMat, sigma_c) + 5 * sizeof(double), "layout");
template <int D, typename M>
__device__ inline void cohesive_law(const M& m, double* delta, double& isv, double* t, double Dm[D][D]) {
    double norm = 0;
    for (int i = 0; i < D; ++i) norm += delta[i] * delta[i];
    if (norm < 1e-40)
        for (int i = 0; i < D; ++i) delta[i] = 1e-20;
    const double b2 = m.beta2;
    double de = delta[D - 1] * delta[D - 1];
    for (int i = 0; i < D - 1; ++i) de += b2 * (delta[i] * delta[i]);
    de = sqrt(de);

    double tr = 0, dtr = 0;
    const bool bl = (m.type == PG_MAT_BL || m.type == PG_MAT_BL_NU);
    if (bl) {
        const double sc = m.sigma_c, d1 = m.delta_c1, d2 = m.delta_c2;
        const double dmax = isv;
        bool loading = (m.type == PG_MAT_BL_NU) || (de >= dmax || de < 0.0);
        if (loading) {
            if (de >= 0 && de < d1) { tr = de * (sc / d1); dtr = (sc / d1); }
            else if (de >= d1 && de < d2) { tr = sc * (1 - (de - d1) / (d2 - d1)); dtr = -1.0 * (sc / (d2 - d1)); }
            else { tr = 0; dtr = 0; }
        } else {
            double tmax;
            if (dmax >= 0 && dmax < d1) tmax = dmax * (sc / d1);
            else if (dmax >= d1 && dmax < d2) tmax = sc * (1 - (dmax - d1) / (d2 - d1));
            else tmax = 0;
            tr = tmax * (de / dmax);
            dtr = tmax / dmax;
        }
    } else {
        const double sc = m.sigma_c, dc = m.delta_c;
        const double dmax = isv;
        bool loading = (m.type == PG_MAT_OP_NU) || (de >= dmax || de < 0.0);
        if (loading) {
            tr = sc * (de / dc) * exp(1.0 - de / dc);
            dtr = (sc / dc) * (1.0 - de / dc) * exp(1.0 - de / dc);
            if (m.type == PG_MAT_OP) isv = de;
        } else {
            double tmax = sc * (dmax / dc) * exp(1.0 - dmax / dc);
            tr = tmax * (de / dmax);
            dtr = tmax / dmax;
        }
    }
    double br[D], dd[D];
    for (int i = 0; i < D; ++i) br[i] = b2 * delta[i];
    br[D - 1] += (1.0 - b2) * delta[D - 1];
    for (int i = 0; i < D; ++i) t[i] = (tr / de) * br[i];
    dd[D - 1] = delta[D - 1] / de;
    for (int i = 0; i < D - 1; ++i) dd[i] = b2 * delta[i] / de;
    const double de2 = de * de;
    for (int i = 0; i < D; ++i)
        for (int k = 0; k < D; ++k) {
            double kr = (i == k ? 1.0 : 0.0);
            double kn = (i == D - 1 && k == D - 1) ? 1.0 : 0.0;
            Dm[i][k] = br[i] * (dtr * de - tr) * (dd[k] / de2) + (tr / de) * (b2 * kr + (1.0 - b2) * kn);
        }
    if (bl && delta[D - 1] < 0) {  // contact branch, src/Material_BilinearCohesive.cpp:114-121
        const double a = m.alpha, sc = m.sigma_c, d1 = m.delta_c1;
        t[D - 1] = (sc / a) * sinh(a * delta[D - 1] / d1);
        for (int i = 0; i < D - 1; ++i) Dm[D - 1][i] = 0.0;
        Dm[D - 1][D - 1] = (sc / d1) * cosh(a * delta[D - 1] / d1);
    }
}

// Traction sensitivity of the Ortiz-Pandolfi laws: the partial derivative dt of
// the traction with respect to sigma_c or delta_c (kind 1 / 2) at opening delta
// (ttn frame; getDeltaEff's 1e-20 replacement applied in place).
// OPCohesiveMaterial::SensitivityConstitutive (src/Material_OPCohesive.cpp:104-180):
// the parameter's term where this law is the one differentiated (right), plus,
// on unloading, the history term ddmax * d t / d delta_max for every OP law;
// OPCohesiveNUMaterial::SensitivityConstitutive (src/Material_OPCohesiveNU.cpp:117-170):
// the parameter's term only.
template <int D, typename M>
__device__ inline void cohesive_sens(const M& m, double* delta, bool right, int kind, double dmax, double ddmax,
                                     double* dt) {
    for (int i = 0; i < D; ++i) dt[i] = 0.0;
    const bool op = m.type == PG_MAT_OP;
    if (!op && !(m.type == PG_MAT_OP_NU && right)) return;
    double norm = 0;
    for (int i = 0; i < D; ++i) norm += delta[i] * delta[i];
    if (norm < 1e-40)
        for (int i = 0; i < D; ++i) delta[i] = 1e-20;
    const double b2 = m.beta2, sc = m.sigma_c, dc = m.delta_c;
    double de = delta[D - 1] * delta[D - 1];
    for (int i = 0; i < D - 1; ++i) de += b2 * (delta[i] * delta[i]);
    de = sqrt(de);
    double br[D];
    for (int i = 0; i < D; ++i) br[i] = i < D - 1 ? b2 * delta[i] : delta[D - 1];
    const bool loading = !op || de >= dmax;  // the NU law has no unloading branch here
    if (right) {
        double dte = 0.0;
        if (kind == 1)
            dte = loading ? de / dc * exp(1 - de / dc) : de / dc * exp(1 - dmax / dc);
        else if (kind == 2)
            dte = op ? (loading ? (sc * de / (dc * dc) * exp(1 - de / dc)) * (de / dc - 1)
                                : (sc * de / (dc * dc) * exp(1 - dmax / dc)) * (dmax / dc - 1))
                     : (sc * de / dc / dc * exp(1 - de / dc)) * (de / dc - 1);
        for (int i = 0; i < D; ++i) dt[i] = (dte / de) * br[i];
    }
    if (!loading) {
        const double dtm = (sc * de / dc) * exp(1 - dmax / dc) * (-1.0 / dc);
        for (int i = 0; i < D; ++i) dt[i] += ddmax * (dtm / de) * br[i];
    }
}

}  // namespace pg
