// MI355X (gfx950) element assembly of K and R for IGFEM / cohesive RVEs.
//
// Two schedules behind one C ABI (pgfem_asm.h):
//   plain  (FP64, drop-in): one thread per parent element computes the element
//          matrix exactly as the reference does (per-qp B^T.D.B accumulation,
//          src/Assembler.cpp:148-368) and stages it in HBM; cut elements are
//          integrated by one wavefront each; a node-pair kernel then sums the
//          staged blocks into the CSR values in element order (deterministic,
//          no atomics) through the precomputed scatter indices.
//   mixed  (PG_FP32, class-binned): cut elements stage FP32 tangents (FP64
//          sliver Jacobians, damage update and residual); uncut elements are
//          never staged -- the node-pair kernel recomputes their 2x2 / 3x3
//          blocks on the fly from coordinates and accumulates in FP64.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <array>
#include <cmath>
#include <type_traits>
#include <string>
#include <mutex>
#include <vector>

#include "elem_math.h"
#include "pgfem_asm.h"

using namespace pg;

namespace {

thread_local std::string g_err;

#define HIP_OK(x)                                                                   \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            g_err = std::string(#x) + ": " + hipGetErrorString(e_);                \
            return PG_E_HIP;                                                        \
        }                                                                           \
    } while (0)

struct DevProb {
    int dim, nen, n_nodes, n_enr, n_elem, n_cut, n_free, n_pres, cohesive, n_cqp;
    int64_t n_pairs;
    const double* coords;
    const int32_t* conn;
    const int32_t* elem_mat;
    const int32_t* cut_of_elem;
    const int32_t* cut_elem;
    const pg_cut_rec* cut;
    const pg_material* mats;
    const int32_t* node_dof;
    const int32_t* npat_ptr;
    const int32_t* npat_row;
    const int32_t* npat_col;
    const int32_t* npat_map;
    const int32_t* n2e_ptr;
    const int32_t* n2e;
    const int64_t* cut_koff;   // per cut element: offset of its ndof^2 block
    const int64_t* cut_poff;   // per cut element: offset of its ndof vector
    // per group size, NG node entries per group in group-list order (groups
    // without cut incidences first): node, first incidence, first pair,
    // incidences | pairs << 16; entries past the last node count nothing
    const int32_t* gtab;
    // node-group kernels: per group size, each group's incidence records (d + 2
    // words) and slot words in a block of 7 NG (tri3) or 24 NG (tet4) slots
    // (group-list order, then the incidence order of the group's nodes), so
    // that a lane loads its record without the node table.  Uncut: the
    // element's nodes, owner first (owner_first order), and a word
    // 1 << 31 | material | lo << 3 | rank << 8; cut: block index of the
    // owner's staged row, residual offset, la | nloc << 8 | lo << 16 |
    // rank << 21, cut index (lo: the owner's position in the group; rank: the
    // incidence's position among the owner's; la: the owner's local index)
    const int32_t* incg;
    const uint64_t* slotg;
    // cut-kernel inputs in class order (small class, then general): the cut
    // records, each element's local node list (parent nodes, then enrichment
    // nodes; -1 past the last; CUT_NODES words) and the offsets of its staged
    // tangent and residual, so that a lane finds everything from its position
    const pg_cut_rec* ccut;
    const int32_t* cnodes;
    const int64_t* ckoff;
    const int64_t* cpoff;
    int n_mat;
    // thermal strain of each material for the handle's current temperature
    // change: -temp_change * thermal_exp where the material's expansion is
    // set, else 0 (src/material_lei.cpp:272-291); indexed by material & 7
    double eth[8];
};

// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8).
// Renumbering so that each XCD walks one contiguous range of the work keeps
// neighbouring node rows -- which share coordinates, U and staged tangents --
// in the same L2.  A bijection of [0, nb) for any grid size.
__device__ inline int xcd_order(int b, int nb) {
    constexpr int X = 8;
    const int per = nb / X, rem = nb % X, x = b % X;
    return x * per + min(x, rem) + b / X;
}

// Clears a flag word (a kernel rather than hipMemsetAsync, so that a caller
// may capture the launch sequence in a HIP graph).
__global__ void k_fill_int(int* __restrict__ p, int v) {
    if (threadIdx.x == 0) *p = v;
}

__global__ void k_fill_double(int64_t n, double* __restrict__ p, double v) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Row i of `in` (width w) to row pos[i] of `out`.
__global__ void k_scatter_rows(int64_t n, int w, const int64_t* __restrict__ pos, const double* __restrict__ in,
                               double* __restrict__ out) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n * w) return;
    const int64_t i = k / w, j = k % w;
    out[pos[i] * w + j] = in[k];
}

// --------------------------------------------------------------------------
// U in global-dof order -> node-major (node*dim + comp).  The reference gathers
// U per element through NodalData (formElementalVector, src/Utilities.cpp:599-617).
// --------------------------------------------------------------------------
__global__ void k_permute_u(int64_t n, const int32_t* __restrict__ node_dof, const double* __restrict__ U,
                            double* __restrict__ Un) {
    int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) Un[k] = U[node_dof[k]];
}

constexpr int MAXN2 = 3 + PG_MAX_ENR;    // 2D local nodes (tri3 + 6 cut-edge slots + mirrors)
constexpr int CUT_NODES = 4 + PG_MAX_ENR;  // words per local node list of a cut element
// internal launch flag: every node group takes the global-memory path (the
// fallback of groups that overflow the LDS tables; group variant bit 3)
constexpr unsigned PG_GLOBAL_PATH = 1u << 31;
constexpr int MAXN3 = 4 + PG_MAX_ENR;    // 3D local nodes
constexpr int MAXQ2 = 36;                // 2D volumetric qps per cut element (3 tri3 + hexagon)
constexpr int MAXQ3 = 32;                // 3D volumetric qps per cut element
constexpr int CUT_MAXC = PG_MAX_COH * 4; // cohesive qps per cut element
constexpr int MAX_LDS_MATS = 8;          // materials staged in LDS
typedef double dpair __attribute__((ext_vector_type(2), aligned(8)));  // two adjacent CSR values
constexpr int64_t KC_PAD = 8 * 9;        // staged cut tangents: padding past the last row

// Local node at position n of an uncut element listed from its local node la:
// cyclic for tri3, la xor n for tet4 -- both even permutations of the node
// order, so the orientation (the sign of the Jacobian) is kept.  The node-group
// kernels compute an incidence's element in this order, owner first.
__host__ __device__ inline int owner_first(int nen, int la, int n) {
    return nen == 3 ? (la + n) % 3 : (la ^ n);
}

// Coefficient k of material m's LEI stress (LeiCoef): three entries of D, then
// the thermal strain of the handle's current temperature change.
template <int D>
__device__ inline double lei_coef(const DevProb& p, int m, int k) {
    return k < 3 ? p.mats[m].D[LeiCoef<D>::index(k)] : p.eth[m & 7];
}

template <int D> struct CutLim;
template <> struct CutLim<2> { static constexpr int N = MAXN2, Q = MAXQ2; };
template <> struct CutLim<3> { static constexpr int N = MAXN3, Q = MAXQ3; };

// Regular reference polygon of Polygon::PolyTriangulate (src/Polygon.cpp:51-63):
// vertex n at angle 2 pi (n + 1) / nn, as cos / sin of (2.0 * PI * (n + 1)) / nn
// give them on the host (PI of include/common.h:215); pentagon, then hexagon.
__constant__ const double POLY_VX[11] = {
    0x1.3c6ef372fe950p-2, -0x1.9e3779b97f4a7p-1, -0x1.9e3779b97f4a9p-1, 0x1.3c6ef372fe94cp-2, 0x1.0000000000000p+0,
    0x1.0000000000001p-1, -0x1.ffffffffffffcp-2, -0x1.0000000000000p+0, -0x1.0000000000004p-1, 0x1.0000000000001p-1,
    0x1.0000000000000p+0};
__constant__ const double POLY_VY[11] = {
    0x1.e6f0e134454ffp-1, 0x1.2cf2304755a5fp-1, -0x1.2cf2304755a5dp-1, -0x1.e6f0e13445500p-1, -0x1.1a62633145c07p-52,
    0x1.bb67ae8584caap-1, 0x1.bb67ae8584cabp-1, 0x1.1a62633145c07p-53, -0x1.bb67ae8584ca8p-1, -0x1.bb67ae8584caap-1,
    -0x1.1a62633145c07p-52};

// Polygon child qp q (Polygon::q_point + compute_shape_grad, src/Polygon.cpp:84-222):
// point q % 4 of the tri3 order-3 rule on sub-triangle t = q / 4 (origin,
// vertex t, vertex t + 1), weight = tri3 weight x sub-triangle Jacobian;
// Wachspress gradients at that point with A_n the signed area of (x,
// vertex n - 1, vertex n), alpha_n = 1 / (A_n A_{n+1}), N_n = alpha_n / sum.
// As in the reference, dA is minus the derivative of A (src/Polygon.cpp:199-200),
// so these are -grad N; in 2D the Jacobian built from them flips too, and J and
// the physical gradients come out right.  NN is a template parameter so that
// every array index is a constant.
template <int NN>
__device__ inline double poly_qp(int q, double (&dN)[NN][2]) {
    constexpr int V0 = NN == 5 ? 0 : 5;
    const int t = q / 4, qq = q % 4;
    const double qx = pick(TRI_QX, qq), qy = pick(TRI_QY, qq);
    double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
#pragma unroll
    for (int n = 0; n < NN; ++n) {
        if (n == t) { ax = POLY_VX[V0 + n]; ay = POLY_VY[V0 + n]; }
        if (n == (t + 1 == NN ? 0 : t + 1)) { bx = POLY_VX[V0 + n]; by = POLY_VY[V0 + n]; }
    }
    // sub-triangle Jacobian (columns: vertex t, vertex t + 1) and the point
    const double w = (ax * by - bx * ay) * pick(TRI_QW, qq);
    const double n0 = 1.0 - qx - qy;
    const double xi0 = ((0.0 + n0 * 0.0) + qx * ax) + qy * bx;
    const double xi1 = ((0.0 + n0 * 0.0) + qx * ay) + qy * by;
    double A[NN + 1], dA0[NN + 1], dA1[NN + 1];
#pragma unroll
    for (int n = 0; n < NN; ++n) {
        const int m = n + 1 == NN ? 0 : n + 1;
        const double p1x = POLY_VX[V0 + n], p1y = POLY_VY[V0 + n], p2x = POLY_VX[V0 + m], p2y = POLY_VY[V0 + m];
        A[n + 1] = 0.5 * (((p1x * p2y - p1y * p2x) - (xi0 * p2y - xi1 * p2x)) + (xi0 * p1y - xi1 * p1x));
        dA0[n + 1] = 0.5 * (p2y - p1y);
        dA1[n + 1] = 0.5 * (p1x - p2x);
    }
    A[0] = A[NN];
    dA0[0] = dA0[NN];
    dA1[0] = dA1[NN];
    double alpha[NN], da0[NN], da1[NN], sa = 0.0, sd0 = 0.0, sd1 = 0.0;
#pragma unroll
    for (int n = 0; n < NN; ++n) {
        alpha[n] = 1.0 / (A[n] * A[n + 1]);
        da0[n] = (-1.0 * alpha[n]) * (dA0[n] / A[n] + dA0[n + 1] / A[n + 1]);
        da1[n] = (-1.0 * alpha[n]) * (dA1[n] / A[n] + dA1[n + 1] / A[n + 1]);
        sa += alpha[n];
        sd0 += da0[n];
        sd1 += da1[n];
    }
#pragma unroll
    for (int n = 0; n < NN; ++n) {
        const double N = alpha[n] / sa;
        dN[n][0] = (da0[n] - N * sd0) / sa;
        dN[n][1] = (da1[n] - N * sd1) / sa;
    }
    return w;
}

// A polygon child's qp: reference gradients, then the physical ones.
template <int NN>
__device__ inline double poly_child_qp(int q, const double (*cx)[3], double (*cg)[3], double& w) {
    double dN[NN][2];
    w = poly_qp<NN>(q, dN);
    return transform<2>(NN, dN, cx, cg);
}

// Geometry of one volumetric qp of a child: child gradients, child J and weight
// (Elem::q_point + Elem::ShapeFunctions on the child, src/Mesh.cpp:1535-1562).
// POLY: polygon children may occur (2D only; the kernels that never meet one
// leave the polygon code out).
template <int D, bool POLY = false>
__device__ inline double child_qp(int ct, int q, int cnn, const double (*cx)[3], double (*cg)[3], double& w) {
    if constexpr (POLY && D == 2) {
        if (ct == PG_CHILD_POLY5) return poly_child_qp<5>(q, cx, cg, w);
        if (ct == PG_CHILD_POLY6) return poly_child_qp<6>(q, cx, cg, w);
    }
    double cdN[4][D];
    if (ct == PG_CHILD_QUAD4) {
        const double r = (q % 2 == 0) ? -PG_GAUSS2 : PG_GAUSS2;
        const double s = (q / 2 == 0) ? -PG_GAUSS2 : PG_GAUSS2;
        double t4[4][2];
        quad4_dNdr(r, s, t4);
        for (int n = 0; n < 4; ++n)
            for (int a = 0; a < D; ++a) cdN[n][a] = t4[n][a];
        w = 1.0;
    } else {
        for (int n = 0; n < cnn; ++n)
            for (int a = 0; a < D; ++a) cdN[n][a] = simplex_dNdr(n, a);
        w = (D == 2) ? pick(TRI_QW, q) : pick(TET_QW, q);
    }
    return transform<D>(cnn, cdN, cx, cg);
}

// Volumetric units of a child in the binned cut kernel: one for a linear simplex
// (constant gradients), one per qp otherwise.
__host__ __device__ inline int cut_units(int ct) {
    return (ct == PG_CHILD_TRI3 || ct == PG_CHILD_TET4) ? 1 : PG_CHILD_NQP(ct);
}

// Lower nodes of a cohesive child: CohEdge2 in 2D, CohTri3 or CohQuad4 in 3D.
template <int D> struct CohLim { static constexpr int N = D == 2 ? 2 : 4; };

// Cohesive qp geometry: N, rotation R and dA on the lower surface
// (CohesiveElem::ShapeFunctions / compute_rotation_general, src/CohesiveElem.cpp:145-238;
// CohEdge2 2-pt Gauss src/CohEdge2.cpp:70-78,186-209; CohTri3 tri3 4-pt rule;
// CohQuad4 2 x 2 Gauss).
template <int D>
__device__ inline double cohesive_qp(int coh_nn, int q, const double (*lx)[3], double* N, double R[D][D], double& w) {
    double dNr[CohLim<D>::N][2];
    const int cnn = coh_nn / 2;
    if constexpr (D == 2) {
        const double xi = (q == 0) ? -PG_GAUSS2 : PG_GAUSS2;
        N[0] = 0.5 * (1.0 - xi); N[1] = 0.5 * (1 + xi);
        dNr[0][0] = -0.5; dNr[1][0] = 0.5;
        w = 1.0;
    } else if (coh_nn == 8) {
      if constexpr (D == 3) {
        // CohQuad4 (src/CohQuad4.cpp:58-140): the tensor product of the edge's
        // 2-pt Gauss rule, qp q at (xi_{q % 2}, xi_{q / 2}), w = 1 * 1; quad4 shapes
        const double r = (q & 1) ? PG_GAUSS2 : -PG_GAUSS2, s = (q & 2) ? PG_GAUSS2 : -PG_GAUSS2;
        N[0] = 0.25 * (1.0 - r) * (1.0 - s); N[1] = 0.25 * (1.0 + r) * (1.0 - s);
        N[2] = 0.25 * (1.0 + r) * (1.0 + s); N[3] = 0.25 * (1.0 - r) * (1.0 + s);
        dNr[0][0] = -0.25 * (1.0 - s); dNr[0][1] = -0.25 * (1.0 - r);
        dNr[1][0] = 0.25 * (1.0 - s);  dNr[1][1] = -0.25 * (1.0 + r);
        dNr[2][0] = 0.25 * (1.0 + s);  dNr[2][1] = 0.25 * (1.0 + r);
        dNr[3][0] = -0.25 * (1.0 + s); dNr[3][1] = 0.25 * (1.0 - r);
        w = 1.0 * 1.0;
      }
    } else {
      if constexpr (D == 3) {
        const double qx = pick(TRI_QX, q), qy = pick(TRI_QY, q);
        N[0] = 1.0 - qx - qy; N[1] = qx; N[2] = qy;
        dNr[0][0] = -1.0; dNr[0][1] = -1.0; dNr[1][0] = 1.0; dNr[1][1] = 0.0; dNr[2][0] = 0.0; dNr[2][1] = 1.0;
        w = pick(TRI_QW, q);
      }
    }
    double Jc[2][3] = {{0, 0, 0}, {0, 0, 0}};
    for (int n = 0; n < cnn; ++n)
        for (int a = 0; a < D - 1; ++a)
            for (int i = 0; i < D; ++i) Jc[a][i] += dNr[n][a] * lx[n][i];
    double nrm[3];
    if (D == 2) { nrm[0] = -Jc[0][1]; nrm[1] = Jc[0][0]; }
    else {
        nrm[0] = Jc[0][1] * Jc[1][2] - Jc[0][2] * Jc[1][1];
        nrm[1] = Jc[0][2] * Jc[1][0] - Jc[0][0] * Jc[1][2];
        nrm[2] = Jc[0][0] * Jc[1][1] - Jc[0][1] * Jc[1][0];
    }
    double mag[2];
    for (int a = 0; a < D - 1; ++a) {
        double s = 0;
        for (int i = 0; i < D; ++i) s += Jc[a][i] * Jc[a][i];
        mag[a] = sqrt(s);
    }
    double dA = 0;
    for (int i = 0; i < D; ++i) dA += nrm[i] * nrm[i];
    dA = sqrt(dA);
    for (int i = 0; i < D; ++i) R[D - 1][i] = nrm[i] / dA;
    for (int i = 0; i < D; ++i)
        for (int a = 0; a < D - 1; ++a) R[a][i] = Jc[a][i] / mag[a];
    return dA;
}

// One cohesive qp: opening, law (trial history in/out), R^T D R and R^T t
// (KernelCohesive, src/AssemblerSmallStrainStructural.cpp:86-103 with
// cohesiveDeltaFast / KFast / InternalForceFast, src/Utilities.cpp:528-594).
template <int D, typename M>
__device__ inline void cohesive_point(const M& m, int cnn, const double* N, const double R[D][D],
                                      const double* Ulo, const double* Uup, double& hist, double RDR[D][D],
                                      double tR[D]) {
    double dl[D], del[D];
    for (int i = 0; i < D; ++i) {
        double s = 0;
        for (int n = 0; n < cnn; ++n) s += N[n] * (Uup[n * D + i] - Ulo[n * D + i]);
        dl[i] = s;
    }
    for (int i = 0; i < D; ++i) {
        double s = 0;
        for (int j = 0; j < D; ++j) s += R[i][j] * dl[j];
        del[i] = s;
    }
    double t[D], Dm[D][D];
    cohesive_law<D, M>(m, del, hist, t, Dm);
    double RD[D][D];
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            double s = 0;
            for (int l = 0; l < D; ++l) s += R[l][i] * Dm[l][j];
            RD[i][j] = s;
        }
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            double s = 0;
            for (int l = 0; l < D; ++l) s += RD[i][l] * R[l][j];
            RDR[i][j] = s;
        }
    for (int i = 0; i < D; ++i) {
        double s = 0;
        for (int l = 0; l < D; ++l) s += R[l][i] * t[l];
        tR[i] = s;
    }
}

// --------------------------------------------------------------------------
// Plain FP64 schedule (the drop-in, no class binning): element kernels that
// follow Assembler::assemble_elem (src/Assembler.cpp:148-368) with
// KernelVolumetric's B-matrix variant -- per qp K_tmp = (B^T D) B,
// P_tmp = B^T sigma, K_el += (J w) K_tmp, P_el += (J w) P_tmp -- stage every
// element matrix in HBM, and a node-pair kernel sums the staged blocks into the
// CSR values in element order.
//
// Staged layout of uncut elements: blocks of 64 consecutive elements; inside a
// block, for each local block (la, lb) of the element matrix the 64 elements'
// d x d blocks one after another (each d x d block contiguous), then for each
// local node la the 64 elements' d residual entries.  A thread's stores then
// fill contiguous stretches together with its neighbours', and a node-pair
// lane reads one d x d block contiguously.  Cut elements stage row-major
// nd x nd matrices (the reference's K_el) at cut_koff and nd vectors at cut_poff.
// --------------------------------------------------------------------------
template <int D> struct PlainStage {
    static constexpr int NEN = D + 1;
    static constexpr int64_t KB = int64_t(NEN) * NEN * D * D * 64;  // Ke part of a block of 64 elements
    static constexpr int64_t W = KB + int64_t(NEN) * D * 64;        // doubles per block of 64 elements
    __host__ __device__ static int64_t kblk(int64_t e, int la, int lb) {
        return (e >> 6) * W + (int64_t(la * NEN + lb) * 64 + (e & 63)) * (D * D);
    }
    __host__ __device__ static int64_t pvec(int64_t e, int la) {
        return (e >> 6) * W + KB + (int64_t(la) * 64 + (e & 63)) * D;
    }
};

// Stages the D matrices of the materials in LDS (zero past the last one).
template <int NS>
__device__ inline void stage_dmats(const DevProb& p, double (*sD)[NS * NS]) {
    for (int k = threadIdx.x; k < MAX_LDS_MATS * NS * NS; k += blockDim.x) {
        const int m = k / (NS * NS), j = k % (NS * NS);
        const double v = p.mats[m < p.n_mat ? m : p.n_mat - 1].D[j];
        sD[m][j] = m < p.n_mat ? v : 0.0;
    }
}

// Uncut elements, one thread each.  A linear simplex has the same B, D and
// sigma at every qp, so each entry accumulates (J w_q) K_tmp over the qps in
// the reference order from one K_tmp value.
template <int D>
__global__ __launch_bounds__(256) void k_elem_plain(DevProb p, const double* __restrict__ Un,
                                                    double* __restrict__ stage) {
    constexpr int NEN = D + 1, NS = (D == 2 ? 3 : 6), NQ = (D == 2 ? 4 : 5), ND = NEN * D;
    __shared__ double sD[MAX_LDS_MATS][NS * NS];
    stage_dmats<NS>(p, sD);
    __syncthreads();
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= p.n_elem || p.cut_of_elem[e] >= 0) return;
    double px[NEN][3], pdN[NEN][D], pg[NEN][3], u[ND];
#pragma unroll
    for (int n = 0; n < NEN; ++n) {
        const int node = p.conn[int64_t(e) * NEN + n];
        for (int i = 0; i < 3; ++i) px[n][i] = i < D ? p.coords[int64_t(node) * D + i] : 0.0;
        for (int i = 0; i < D; ++i) u[n * D + i] = Un[int64_t(node) * D + i];
        for (int a = 0; a < D; ++a) pdN[n][a] = simplex_dNdr(n, a);
    }
    const double J = transform<D>(NEN, pdN, px, pg);
    const double* Dm = sD[p.elem_mat[e]];
    double wq[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) wq[q] = J * (D == 2 ? TRI_QW[q] : TET_QW[q]);
    // strain = B U (dgemv order), stress
    double eps[NS], sig[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) eps[k] = 0.0;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
        double b[NS];
        bcol<D, double>(pg[i / D], i % D, b);
#pragma unroll
        for (int k = 0; k < NS; ++k) eps[k] += b[k] * u[i];
    }
    add_thermal<D>(eps, p.eth[p.elem_mat[e] & 7]);
    lei_stress<D>(Dm, eps, sig);
#pragma unroll
    for (int i = 0; i < ND; ++i) {
        double bi[NS], btd[NS];
        bcol<D, double>(pg[i / D], i % D, bi);
#pragma unroll
        for (int l = 0; l < NS; ++l) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < NS; ++k) t += bi[k] * Dm[k * NS + l];
            btd[l] = t;
        }
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            double bj[NS];
            bcol<D, double>(pg[j / D], j % D, bj);
            double kt = 0.0;
#pragma unroll
            for (int l = 0; l < NS; ++l) kt += btd[l] * bj[l];
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc += wq[q] * kt;
            stage[PlainStage<D>::kblk(e, i / D, j / D) + (i % D) * D + j % D] = acc;
        }
        double pt = 0.0;
#pragma unroll
        for (int k = 0; k < NS; ++k) pt += bi[k] * sig[k];
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc += wq[q] * pt;
        stage[PlainStage<D>::pvec(e, i / D) + i % D] = acc;
    }
}

// Cut elements, one wavefront each (store_shape_functions for its children,
// src/Mesh.cpp:1521-1641, then assemble_elem's cut and cohesive branches,
// src/Assembler.cpp:223-365).  Phase 1: one lane per child qp computes the
// child Jacobian, the gradient of every local node (parent gradient, the child
// gradient of the child node that is the enrichment node, else 0;
// src/Mesh.cpp:1596-1614), J w and the stress; one lane per cohesive qp runs
// the law with the history update.  Phase 2: one lane per entry of K_el sums
// its volumetric qps in qp order, then adds each cohesive element's K_c (its
// qps summed first, as the reference folds K_c into K_el); likewise P_el.
//
// POLY: the launch over the cut elements with polygon children (list: their
// cut indices); the other launch (list null, one block per cut element) leaves
// those elements to it, so that its kernel carries no polygon code.
template <int D, bool POLY>
__global__ __launch_bounds__(64) void k_cut_plain(DevProb p, const double* __restrict__ Un,
                                                  const double* __restrict__ isv_c, double* __restrict__ isv_t,
                                                  double* __restrict__ Ks, double* __restrict__ Ps,
                                                  int* __restrict__ err, const int32_t* __restrict__ list) {
    constexpr int NEN = D + 1, NS = (D == 2 ? 3 : 6);
    constexpr int MAXN = CutLim<D>::N, MAXQ = CutLim<D>::Q, MAXC = PG_MAX_COH * 4;
    __shared__ double sD[MAX_LDS_MATS][NS * NS];
    __shared__ double lx[MAXN][3];
    __shared__ double ue[MAXN * D];
    __shared__ double gq[MAXQ][MAXN][D];  // gradient of each local node at each volumetric qp
    __shared__ double wq[MAXQ];            // J w of each volumetric qp
    __shared__ double sq[MAXQ][NS];        // stress at each volumetric qp
    __shared__ int mq[MAXQ];               // material of each volumetric qp (its child's)
    constexpr int CNN = CohLim<D>::N;
    __shared__ double cN[MAXC][CNN], cw[MAXC], cRDR[MAXC][D][D], ctR[MAXC][D];
    __shared__ pg_cut_rec rec;
    const int c = list ? list[blockIdx.x] : int(blockIdx.x), lane = threadIdx.x;
    stage_dmats<NS>(p, sD);
    constexpr int RW = int(sizeof(pg_cut_rec) / 4);
    if (lane < RW) reinterpret_cast<int32_t*>(&rec)[lane] = reinterpret_cast<const int32_t*>(&p.cut[c])[lane];
    const int e = p.cut_elem[c];
    __syncthreads();
    if constexpr (!POLY && D == 2) {
        bool poly = false;
        for (int k = 0; k < rec.n_child; ++k)
            poly = poly || rec.child_type[k] == PG_CHILD_POLY5 || rec.child_type[k] == PG_CHILD_POLY6;
        if (poly) return;  // the polygon launch takes it (the whole block returns)
    }
    const int nloc = NEN + rec.n_enr, nd = nloc * D;
    if (lane < nloc) {
        const int node = lane < NEN ? p.conn[int64_t(e) * NEN + lane] : rec.enr[lane - NEN];
        for (int i = 0; i < 3; ++i) lx[lane][i] = i < D ? p.coords[int64_t(node) * D + i] : 0.0;
        for (int i = 0; i < D; ++i) ue[lane * D + i] = Un[int64_t(node) * D + i];
    }
    __syncthreads();
    double px[NEN][3], pdN[NEN][D], pgr[NEN][3];
    for (int n = 0; n < NEN; ++n) {
        for (int i = 0; i < 3; ++i) px[n][i] = lx[n][i];
        for (int a = 0; a < D; ++a) pdN[n][a] = simplex_dNdr(n, a);
    }
    transform<D>(NEN, pdN, px, pgr);
    const int nvq = rec.n_vqp;
    for (int t = lane; t < nvq; t += 64) {
        int k = 0, q = t;
        for (; k < rec.n_child; ++k) {
            const int nq = PG_CHILD_NQP(rec.child_type[k]);
            if (q < nq) break;
            q -= nq;
        }
        constexpr int CN = (D == 2 && POLY) ? 6 : 4;  // nodes of the largest child
        const int ct = rec.child_type[k], ccn = PG_CHILD_NN(ct);
        double cx[CN][3], cg[CN][3], w;
        int cl[CN];
        for (int n = 0; n < CN; ++n) {
            cl[n] = n < ccn ? rec.child_node[PG_CHILD_STRIDE(D) * k + n] : -1;
            for (int i = 0; i < 3; ++i) cx[n][i] = (n < ccn && i < D) ? lx[cl[n]][i] : 0.0;
        }
        const double Jc = child_qp<D, POLY>(ct, q, ccn, cx, cg, w);
        if (Jc < 0.0) atomicOr(err, 1);  // store_shape_functions aborts (src/Mesh.cpp:1561)
        double eps[NS];
        for (int s2 = 0; s2 < NS; ++s2) eps[s2] = 0.0;
        for (int l = 0; l < nloc; ++l) {
            double gr[3] = {0.0, 0.0, 0.0};
            if (l < NEN) {
                for (int a = 0; a < D; ++a) gr[a] = pgr[l][a];
            } else {
                for (int n = 0; n < CN; ++n)
                    if (cl[n] == l)
                        for (int a = 0; a < D; ++a) gr[a] = cg[n][a];
            }
            for (int a = 0; a < D; ++a) gq[t][l][a] = gr[a];
            for (int cc = 0; cc < D; ++cc) {
                double b[NS];
                bcol<D, double>(gr, cc, b);
                for (int s2 = 0; s2 < NS; ++s2) eps[s2] += b[s2] * ue[l * D + cc];
            }
        }
        double sig[NS];
        add_thermal<D>(eps, p.eth[rec.child_mat[k] & 7]);
        lei_stress<D>(sD[rec.child_mat[k]], eps, sig);
        wq[t] = Jc * w;
        mq[t] = rec.child_mat[k];
        for (int s2 = 0; s2 < NS; ++s2) sq[t][s2] = sig[s2];
    }
    const int cnn = rec.coh_nn / 2, cnq = rec.coh_nn == 4 ? 2 : 4;
    const int ncq = rec.n_coh * cnq;
    for (int cq = lane; cq < ncq; cq += 64) {
        const int k = cq / cnq, q = cq % cnq;
        double lo[CNN][3], ulo[CNN * D], uup[CNN * D], N[CNN], R[D][D], w, RDR[D][D], tR[D];
        for (int n = 0; n < CNN; ++n)
            for (int i = 0; i < D; ++i) {
                const bool in = n < cnn;
                lo[n][i] = in ? lx[rec.coh_nodes[k][n]][i] : 0.0;
                ulo[n * D + i] = in ? ue[rec.coh_nodes[k][n] * D + i] : 0.0;
                uup[n * D + i] = in ? ue[rec.coh_nodes[k][n + cnn] * D + i] : 0.0;
            }
        const double dA = cohesive_qp<D>(rec.coh_nn, q, lo, N, R, w);
        const int iq = rec.isv_off + cq;
        double hist = isv_c ? isv_c[iq] : 0.0;
        cohesive_point<D>(p.mats[rec.coh_mat[k]], cnn, N, R, ulo, uup, hist, RDR, tR);
        if (isv_t) isv_t[iq] = hist;
        for (int i = 0; i < D; ++i) {
            for (int j = 0; j < D; ++j) cRDR[cq][i][j] = RDR[i][j];
            ctR[cq][i] = tR[i];
        }
        for (int n = 0; n < CNN; ++n) cN[cq][n] = n < cnn ? N[n] : 0.0;
        cw[cq] = dA * w;
    }
    __syncthreads();
    double* Ke = Ks + p.cut_koff[c];
    for (int idx = lane; idx < nd * nd; idx += 64) {
        const int i = idx / nd, j = idx % nd;
        const int li = i / D, di = i % D, lj = j / D, dj = j % D;
        double acc = 0.0;
        for (int q = 0; q < nvq; ++q) {
            const double* Dm = sD[mq[q]];
            double bi[NS], bj[NS], btd[NS];
            bcol<D, double>(gq[q][li], di, bi);
            bcol<D, double>(gq[q][lj], dj, bj);
            for (int l = 0; l < NS; ++l) {
                double t = 0.0;
                for (int k = 0; k < NS; ++k) t += bi[k] * Dm[k * NS + l];
                btd[l] = t;
            }
            double kt = 0.0;
            for (int l = 0; l < NS; ++l) kt += btd[l] * bj[l];
            acc += wq[q] * kt;
        }
        for (int k = 0; k < rec.n_coh; ++k) {
            int m1 = -1, m2 = -1;
            for (int m = 0; m < rec.coh_nn; ++m) {
                if (rec.coh_nodes[k][m] == li) m1 = m;
                if (rec.coh_nodes[k][m] == lj) m2 = m;
            }
            if (m1 < 0 || m2 < 0) continue;
            const int n1 = m1 % cnn, n2 = m2 % cnn;
            const double sg = ((m1 < cnn) == (m2 < cnn)) ? 1.0 : -1.0;
            double kc = 0.0;
            for (int q = 0; q < cnq; ++q) {
                const int cq = k * cnq + q;
                kc += cw[cq] * (sg * (cN[cq][n1] * cN[cq][n2] * cRDR[cq][di][dj]));
            }
            acc += kc;
        }
        Ke[idx] = acc;
    }
    double* Pe = Ps + p.cut_poff[c];
    for (int i = lane; i < nd; i += 64) {
        const int l = i / D, di = i % D;
        double acc = 0.0;
        for (int q = 0; q < nvq; ++q) {
            double bi[NS];
            bcol<D, double>(gq[q][l], di, bi);
            double pt = 0.0;
            for (int k = 0; k < NS; ++k) pt += bi[k] * sq[q][k];
            acc += wq[q] * pt;
        }
        for (int k = 0; k < rec.n_coh; ++k)
            for (int m = 0; m < rec.coh_nn; ++m)
                if (rec.coh_nodes[k][m] == l) {
                    const int n = m % cnn;
                    const double sgn = m < cnn ? -1.0 : 1.0;
                    double pc = 0.0;
                    for (int q = 0; q < cnq; ++q) {
                        const int cq = k * cnq + q;
                        pc += cw[cq] * (sgn * (cN[cq][n] * ctR[cq][di]));
                    }
                    acc += pc;
                }
        Pe[i] = acc;
    }
}

// --------------------------------------------------------------------------
// Material-parameter sensitivity dP/dd (SensitivityParameter::assemble_dPdd,
// src/SensitivityParameter.cpp:108-151; SensitivityMaterialParameter::
// assemble_elem_dPdd, src/SensitivityMaterialParameter.cpp:54-264).  The element
// vectors go to the FP64 drop-in's staging buffers (uncut: PlainStage residual
// entries; cut: cut_poff), and k_pair_staged's residual path sums them per node
// in element order.
// --------------------------------------------------------------------------
struct SensDev {
    int mat;       // material differentiated against
    int kind;      // PG_SENS_*
    int own_b;     // PG_SENS_OWN_CHILD_B
    double dD[36]; // PG_SENS_LEI: dD/dd
};

// Uncut elements, one thread each: (J w_q) B^T (dD . B U) summed over the qps
// for elements of the LEI material differentiated against
// (SensitivityMaterialParameterStructural::KernelVolumetric, B variant,
// src/SensitivityMaterialParameterStructural.cpp:40-60); 0 for the others,
// which the reference does not visit.
template <int D>
__global__ __launch_bounds__(256) void k_sens_elem(DevProb p, const double* __restrict__ Un, SensDev s,
                                                   double* __restrict__ stage) {
    constexpr int NEN = D + 1, NS = (D == 2 ? 3 : 6), NQ = (D == 2 ? 4 : 5), ND = NEN * D;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= p.n_elem || p.cut_of_elem[e] >= 0) return;
    if (s.kind != PG_SENS_LEI || p.elem_mat[e] != s.mat) {
        for (int i = 0; i < ND; ++i) stage[PlainStage<D>::pvec(e, i / D) + i % D] = 0.0;
        return;
    }
    double px[NEN][3], pdN[NEN][D], pg[NEN][3], u[ND];
#pragma unroll
    for (int n = 0; n < NEN; ++n) {
        const int node = p.conn[int64_t(e) * NEN + n];
        for (int i = 0; i < 3; ++i) px[n][i] = i < D ? p.coords[int64_t(node) * D + i] : 0.0;
        for (int i = 0; i < D; ++i) u[n * D + i] = Un[int64_t(node) * D + i];
        for (int a = 0; a < D; ++a) pdN[n][a] = simplex_dNdr(n, a);
    }
    const double J = transform<D>(NEN, pdN, px, pg);
    double wq[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) wq[q] = J * (D == 2 ? TRI_QW[q] : TET_QW[q]);
    double eps[NS], ds[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) eps[k] = 0.0;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
        double b[NS];
        bcol<D, double>(pg[i / D], i % D, b);
#pragma unroll
        for (int k = 0; k < NS; ++k) eps[k] += b[k] * u[i];
    }
    lei_stress<D>(s.dD, eps, ds);  // computeStress with the D sensitivity
#pragma unroll
    for (int i = 0; i < ND; ++i) {
        double bi[NS];
        bcol<D, double>(pg[i / D], i % D, bi);
        double pt = 0.0;
#pragma unroll
        for (int k = 0; k < NS; ++k) pt += bi[k] * ds[k];
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc += wq[q] * pt;
        stage[PlainStage<D>::pvec(e, i / D) + i % D] = acc;
    }
}

// Cut elements, one wavefront each.  Phase 1: one lane per child qp computes
// the gradients of the local nodes and J w (as k_cut_plain does).  Phase 2: the
// qp's strain and dstress = dD . B U where its child is of the material
// differentiated against, with the B of the element's qp numbered by the
// child-local index (getBmat(l_elem, qp), src/SensitivityMaterialParameter.cpp:
// 138-150) unless own_b; one lane per cohesive qp of a visited cohesive child
// (a law with internal variables, or the one differentiated) computes R^T dt.
// Phase 3: one lane per element dof sums (J w) B^T dstress over the qps in
// order, then folds each cohesive child's sum (:165-258).
template <int D, bool POLY>
__global__ __launch_bounds__(64) void k_sens_cut(DevProb p, const double* __restrict__ Un, SensDev s,
                                                 const double* __restrict__ isv, const double* __restrict__ isvs,
                                                 double* __restrict__ Ps, const int32_t* __restrict__ list) {
    constexpr int NEN = D + 1, NS = (D == 2 ? 3 : 6);
    constexpr int MAXN = CutLim<D>::N, MAXQ = CutLim<D>::Q, MAXC = PG_MAX_COH * 4;
    __shared__ double lx[MAXN][3];
    __shared__ double ue[MAXN * D];
    __shared__ double gq[MAXQ][MAXN][D];  // gradient of each local node at each volumetric qp
    __shared__ double wq[MAXQ];            // J w of each volumetric qp
    __shared__ double sq[MAXQ][NS];        // dstress at each volumetric qp (0 where not visited)
    __shared__ int bq[MAXQ];               // the qp whose B each qp uses
    constexpr int CNN = CohLim<D>::N;
    __shared__ double cN[MAXC][CNN], cw[MAXC], ctR[MAXC][D];
    __shared__ pg_cut_rec rec;
    const int c = list ? list[blockIdx.x] : int(blockIdx.x), lane = threadIdx.x;
    constexpr int RW = int(sizeof(pg_cut_rec) / 4);
    if (lane < RW) reinterpret_cast<int32_t*>(&rec)[lane] = reinterpret_cast<const int32_t*>(&p.cut[c])[lane];
    const int e = p.cut_elem[c];
    __syncthreads();
    if constexpr (!POLY && D == 2) {
        bool poly = false;
        for (int k = 0; k < rec.n_child; ++k)
            poly = poly || rec.child_type[k] == PG_CHILD_POLY5 || rec.child_type[k] == PG_CHILD_POLY6;
        if (poly) return;  // the polygon launch takes it (the whole block returns)
    }
    const int nloc = NEN + rec.n_enr, nd = nloc * D;
    if (lane < nloc) {
        const int node = lane < NEN ? p.conn[int64_t(e) * NEN + lane] : rec.enr[lane - NEN];
        for (int i = 0; i < 3; ++i) lx[lane][i] = i < D ? p.coords[int64_t(node) * D + i] : 0.0;
        for (int i = 0; i < D; ++i) ue[lane * D + i] = Un[int64_t(node) * D + i];
    }
    __syncthreads();
    double px[NEN][3], pdN[NEN][D], pgr[NEN][3];
    for (int n = 0; n < NEN; ++n) {
        for (int i = 0; i < 3; ++i) px[n][i] = lx[n][i];
        for (int a = 0; a < D; ++a) pdN[n][a] = simplex_dNdr(n, a);
    }
    transform<D>(NEN, pdN, px, pgr);
    const int nvq = rec.n_vqp;
    for (int t = lane; t < nvq; t += 64) {
        int k = 0, q = t;
        for (; k < rec.n_child; ++k) {
            const int nq = PG_CHILD_NQP(rec.child_type[k]);
            if (q < nq) break;
            q -= nq;
        }
        constexpr int CN = (D == 2 && POLY) ? 6 : 4;
        const int ct = rec.child_type[k], ccn = PG_CHILD_NN(ct);
        double cx[CN][3], cg[CN][3], w;
        int cl[CN];
        for (int n = 0; n < CN; ++n) {
            cl[n] = n < ccn ? rec.child_node[PG_CHILD_STRIDE(D) * k + n] : -1;
            for (int i = 0; i < 3; ++i) cx[n][i] = (n < ccn && i < D) ? lx[cl[n]][i] : 0.0;
        }
        const double Jc = child_qp<D, POLY>(ct, q, ccn, cx, cg, w);
        for (int l = 0; l < nloc; ++l) {
            double gr[3] = {0.0, 0.0, 0.0};
            if (l < NEN) {
                for (int a = 0; a < D; ++a) gr[a] = pgr[l][a];
            } else {
                for (int n = 0; n < CN; ++n)
                    if (cl[n] == l)
                        for (int a = 0; a < D; ++a) gr[a] = cg[n][a];
            }
            for (int a = 0; a < D; ++a) gq[t][l][a] = gr[a];
        }
        wq[t] = Jc * w;
        bq[t] = s.own_b ? t : q;
    }
    __syncthreads();
    for (int t = lane; t < nvq; t += 64) {
        int k = 0, q = t;
        for (; k < rec.n_child; ++k) {
            const int nq = PG_CHILD_NQP(rec.child_type[k]);
            if (q < nq) break;
            q -= nq;
        }
        double ds[NS];
        for (int s2 = 0; s2 < NS; ++s2) ds[s2] = 0.0;
        if (s.kind == PG_SENS_LEI && rec.child_mat[k] == s.mat) {
            const int tb = bq[t];
            double eps[NS];
            for (int s2 = 0; s2 < NS; ++s2) eps[s2] = 0.0;
            for (int l = 0; l < nloc; ++l)
                for (int cc = 0; cc < D; ++cc) {
                    double b[NS];
                    bcol<D, double>(gq[tb][l], cc, b);
                    for (int s2 = 0; s2 < NS; ++s2) eps[s2] += b[s2] * ue[l * D + cc];
                }
            lei_stress<D>(s.dD, eps, ds);
        }
        for (int s2 = 0; s2 < NS; ++s2) sq[t][s2] = ds[s2];
    }
    const int cnn = rec.coh_nn / 2, cnq = rec.coh_nn == 4 ? 2 : 4;
    const int ncq = p.cohesive ? rec.n_coh * cnq : 0;
    for (int cq = lane; cq < ncq; cq += 64) {
        const int k = cq / cnq, q = cq % cnq;
        const pg_material& m = p.mats[rec.coh_mat[k]];
        const bool right = rec.coh_mat[k] == s.mat;
        const bool visit = right || m.type == PG_MAT_OP || m.type == PG_MAT_BL;
        double lo[CNN][3], N[CNN], R[D][D], w, tR[D];
        for (int n = 0; n < CNN; ++n)
            for (int i = 0; i < D; ++i) lo[n][i] = n < cnn ? lx[rec.coh_nodes[k][n]][i] : 0.0;
        const double dA = cohesive_qp<D>(rec.coh_nn, q, lo, N, R, w);
        for (int i = 0; i < D; ++i) tR[i] = 0.0;
        if (visit) {
            // cohesiveDeltaFast (src/Utilities.cpp:528-545)
            double dl[D], del[D], dt[D];
            for (int i = 0; i < D; ++i) {
                double sm = 0;
                for (int n = 0; n < cnn; ++n)
                    sm += N[n] * (ue[rec.coh_nodes[k][n + cnn] * D + i] - ue[rec.coh_nodes[k][n] * D + i]);
                dl[i] = sm;
            }
            for (int i = 0; i < D; ++i) {
                double sm = 0;
                for (int j = 0; j < D; ++j) sm += R[i][j] * dl[j];
                del[i] = sm;
            }
            const int iq = rec.isv_off + cq;
            cohesive_sens<D>(m, del, right, s.kind, isv ? isv[iq] : 0.0, isvs ? isvs[iq] : 0.0, dt);
            for (int i = 0; i < D; ++i) {
                double sm = 0;
                for (int l = 0; l < D; ++l) sm += R[l][i] * dt[l];
                tR[i] = sm;
            }
        }
        for (int i = 0; i < D; ++i) ctR[cq][i] = tR[i];
        for (int n = 0; n < CNN; ++n) cN[cq][n] = n < cnn ? N[n] : 0.0;
        cw[cq] = dA * w;
    }
    __syncthreads();
    double* Pe = Ps + p.cut_poff[c];
    for (int i = lane; i < nd; i += 64) {
        const int l = i / D, di = i % D;
        double acc = 0.0;
        for (int t = 0; t < nvq; ++t) {
            double bi[NS];
            bcol<D, double>(gq[bq[t]][l], di, bi);
            double pt = 0.0;
            for (int k = 0; k < NS; ++k) pt += bi[k] * sq[t][k];
            acc += wq[t] * pt;
        }
        for (int k = 0; k < (p.cohesive ? int(rec.n_coh) : 0); ++k)
            for (int m = 0; m < rec.coh_nn; ++m)
                if (rec.coh_nodes[k][m] == l) {
                    const int n = m % cnn;
                    const double sgn = m < cnn ? -1.0 : 1.0;
                    double pc = 0.0;
                    for (int q = 0; q < cnq; ++q) {
                        const int cq = k * cnq + q;
                        pc += cw[cq] * (sgn * (cN[cq][n] * ctR[cq][di]));
                    }
                    acc += pc;
                }
        Pe[i] = acc;
    }
}

// --------------------------------------------------------------------------
// Cut-element kernel of the class-binned schedule (src/Mesh.cpp:1521-1641 +
// src/Assembler.cpp:223-365).  EPW intersected elements share one wavefront, a
// group of 64/EPW lanes each:
//   phase 1: one lane per volumetric unit computes the child Jacobian in FP64
//            (sliver children), the gradients of all local nodes, wJ and the
//            FP64 stress; one lane per cohesive qp runs the FP64 law and the
//            history update;
//   phase 2: every lane accumulates its K entries b_i . (wJ D) b_j over the units
//            in TK, adds the cohesive fold and stores; P (FP64) likewise.
// A linear simplex child (tri3, tet4) has the same gradients, Jacobian and so
// strain and stress at each of its qps: it is one unit, weighted by J times the
// sum of its rule's weights (as the node-group kernels weight an uncut element).
// A quad4 or polygon child contributes one unit per qp.  MAXQ counts units.
// --------------------------------------------------------------------------
template <int D, typename TK, int EPW, int MAXN, int MAXQ, int MAXC, int OCC>
__global__ void __attribute__((amdgpu_flat_work_group_size(64, 64), amdgpu_waves_per_eu(OCC))) k_cut(DevProb p, int cbase, int n_list,
                                            const double* __restrict__ Un,
                                            const double* __restrict__ isv_c, double* __restrict__ isv_t,
                                            TK* __restrict__ Kst, double* __restrict__ Pst, int* __restrict__ err) {
    constexpr int NS = (D == 2 ? 3 : 6), NEN = D + 1, GS = 64 / EPW;
    // MAXC == 0: a mesh without cohesive children (no cohesive qp phase, fold or
    // LDS arrays)
    constexpr bool COH = MAXC > 0;
    constexpr int CUT_MAXC = COH ? MAXC : 1;
    __shared__ double xs[EPW][MAXN][3];
    __shared__ double us[EPW][MAXN * D];
    // gradients: a parent node's is the same at every qp (linear simplex
    // parent), an enrichment node's is its child's at that qp (or zero)
    __shared__ double pg[EPW][NEN][D];
    __shared__ double qg[EPW][MAXQ][MAXN - NEN][D];
    __shared__ double qw[EPW][MAXQ];
    __shared__ TK qwl[EPW][MAXQ][2];        // w lambda, w mu of each qp (its child's material)
    // qps at which a local node's gradient is not zero (one bit per qp)
    using QMask = typename std::conditional<(MAXQ > 32), unsigned long long, uint32_t>::type;
    static_assert(MAXQ <= 64, "one mask bit per qp");
    __shared__ QMask qmask[EPW][MAXN];
    // nodes of the largest child
    constexpr int CN = (D == 2 && MAXQ > 8) ? 6 : 4;
    __shared__ double qs[EPW][MAXQ][NS];
    __shared__ double cN[EPW][CUT_MAXC][CohLim<D>::N];
    __shared__ double cw[EPW][CUT_MAXC];
    __shared__ double cRDR[EPW][CUT_MAXC][D][D];
    __shared__ double ctR[EPW][CUT_MAXC][D];
    __shared__ pg_cut_rec recs[EPW];
    // 3D: sums over the qps of w lambda and w mu, and of w lambda g_l and
    // w mu g_l for each enrichment node l (the blocks of a parent node, whose
    // gradient is the same at every qp, follow from them)
    __shared__ TK sS[EPW][2];
    __shared__ TK sV[EPW][D == 3 ? MAXN - NEN : 1][2][D];
    __shared__ TK sLM[MAX_LDS_MATS][2];  // lambda, mu of each LEI material
    __shared__ double sLC[MAX_LDS_MATS][LeiCoef<D>::N];  // LEI stress coefficients
    __shared__ CohMat sCM[MAX_LDS_MATS];                  // cohesive law parameters

    const int g = threadIdx.x / GS, gl = threadIdx.x % GS;
    const int li_ = xcd_order(blockIdx.x, gridDim.x) * EPW + g;
    const bool live = li_ < n_list;
    const int64_t pos = int64_t(cbase) + (live ? li_ : 0);  // class position
    // every set-up load goes out before the first wait: the materials' lambda,
    // mu and stress coefficients, the record and the local node list (both
    // indexed by the class position); the node data follows the node list.
    // Loads use clamped indices so that none sits behind a branch.
    const int mm = threadIdx.x >> 2, mj = threadIdx.x & 3;  // 64 lanes: 16 x 4 >= 8 materials x 4
    const double mv0 = lei_coef<D>(p, mm < p.n_mat ? mm : p.n_mat - 1, mj);
    // cohesive parameters: lane t < 48 loads field t % 6 (sigma_c .. alpha,
    // contiguous) of material t / 6, and every lane the type of material t % 8
    const int cm = threadIdx.x < 48 ? threadIdx.x / 6 : 0, cf = threadIdx.x < 48 ? threadIdx.x % 6 : 0;
    const int ct8 = threadIdx.x & 7;
    const double cv0 = (&p.mats[cm < p.n_mat ? cm : p.n_mat - 1].sigma_c)[cf];
    const int32_t cty0 = p.mats[ct8 < p.n_mat ? ct8 : p.n_mat - 1].type;
    constexpr int RW = int(sizeof(pg_cut_rec) / 4), RWL = (RW + GS - 1) / GS;
    static_assert(MAXN <= GS, "one local node per lane");
    int32_t rw[RWL];
#pragma unroll
    for (int k = 0; k < RWL; ++k) {
        const int kk = gl + k * GS;
        rw[k] = reinterpret_cast<const int32_t*>(&p.ccut[pos])[kk < RW ? kk : RW - 1];
    }
    const int nodeid = p.cnodes[pos * CUT_NODES + (gl < MAXN ? gl : MAXN - 1)];
    const int nid = nodeid < 0 ? 0 : nodeid;
    double nx[D], nu[D];
    for (int i = 0; i < D; ++i) {
        nx[i] = p.coords[int64_t(nid) * D + i];
        nu[i] = Un[int64_t(nid) * D + i];
    }
    {
        const double v = mm < p.n_mat ? mv0 : 0.0;
        if (mm < MAX_LDS_MATS) sLC[mm][mj] = v;
        if (mm < MAX_LDS_MATS && mj == 1) sLM[mm][0] = TK(v);  // D[0][1] = lambda
        if (mm < MAX_LDS_MATS && mj == 2) sLM[mm][1] = TK(v);  // D[last][last] = mu
        if (COH && threadIdx.x < 48) (&sCM[cm].sigma_c)[cf] = cv0;
        if (COH && threadIdx.x < MAX_LDS_MATS) sCM[ct8].type = cty0;
    }
#pragma unroll
    for (int k = 0; k < RWL; ++k) {
        const int kk = gl + k * GS;
        if (live && kk < RW) reinterpret_cast<int32_t*>(&recs[g])[kk] = rw[k];
    }
    if (gl < MAXN) {
        qmask[g][gl] = QMask(0);
        if (live && nodeid >= 0)
            for (int i = 0; i < D; ++i) {
                xs[g][gl][i] = nx[i];
                us[g][gl * D + i] = nu[i];
            }
    }
    __syncthreads();
    const pg_cut_rec& rec = recs[g];
    const int nloc = live ? NEN + rec.n_enr : 0;
    const int nd = nloc * D;

    int nvq = 0;  // volumetric units
    if (live)
        for (int k = 0; k < rec.n_child; ++k) nvq += cut_units(rec.child_type[k]);
    const int cnn = rec.coh_nn / 2;
    const int cnq = (rec.coh_nn == 4) ? 2 : 4;
    const int ncq = (COH && live) ? rec.n_coh * cnq : 0;
    // the committed history of this lane's first cohesive qp goes out before
    // the volumetric work (any valid index when the element has none)
    double hist0 = 0.0;
    if (COH && isv_c && p.n_cqp > 0) hist0 = isv_c[ncq > 0 ? rec.isv_off + (gl < ncq ? gl : ncq - 1) : 0];
    for (int t = gl; t < nvq; t += GS) {
        {
            // ---- volumetric unit: child k (simplex) or child k's qp q ----
            int k = 0, q = t;
            if constexpr (D == 3) {
                k = t;
                q = 0;
            } else {
                for (; k < rec.n_child; ++k) {
                    const int nu = cut_units(rec.child_type[k]);
                    if (q < nu) break;
                    q -= nu;
                }
            }
            const int ct = rec.child_type[k];
            const int ccn = PG_CHILD_NN(ct);
            double cx[CN][3], cg[CN][3], w;
            int cl[CN];
            for (int n = 0; n < CN; ++n) {
                cl[n] = n < ccn ? rec.child_node[PG_CHILD_STRIDE(D) * k + n] : -1;
                for (int i = 0; i < 3; ++i) cx[n][i] = (n < ccn && i < D) ? xs[g][cl[n]][i] : 0.0;
            }
            const double Jc = child_qp<D, CN == 6>(ct, q, ccn, cx, cg, w);
            if (Jc < 0.0) atomicOr(err, 1);  // store_shape_functions aborts (src/Mesh.cpp:1561)
            if (ct == PG_CHILD_TRI3 || ct == PG_CHILD_TET4) {
                w = 0.0;
                if constexpr (D == 2)
                    for (int r = 0; r < 4; ++r) w += TRI_QW[r];
                else
                    for (int r = 0; r < 5; ++r) w += TET_QW[r];
            }
            double px[NEN][3], pdN[NEN][D], pgr[NEN][3];
            for (int n = 0; n < NEN; ++n) {
                for (int i = 0; i < D; ++i) px[n][i] = xs[g][n][i];
                for (int a = 0; a < D; ++a) pdN[n][a] = simplex_dNdr(n, a);
            }
            transform<D>(NEN, pdN, px, pgr);
            double eps[NS];
            for (int s = 0; s < NS; ++s) eps[s] = 0.0;
            for (int l = 0; l < nloc; ++l) {
                double gr[D];
                if (l < NEN) {
                    for (int a = 0; a < D; ++a) gr[a] = pgr[l][a];
                } else {
                    for (int a = 0; a < D; ++a) gr[a] = 0.0;
#pragma unroll
                    for (int n = 0; n < CN; ++n)
                        if (cl[n] == l)
                            for (int a = 0; a < D; ++a) gr[a] = cg[n][a];
                }
                for (int a = 0; a < D; ++a) {
                    if (l < NEN) pg[g][l][a] = gr[a];  // (every qp lane writes the same value)
                    else qg[g][t][l - NEN][a] = gr[a];
                }
                bool nz = false;
                for (int a = 0; a < D; ++a) nz = nz || gr[a] != 0.0;
                if (nz) atomicOr(&qmask[g][l], QMask(1) << t);
                for (int c = 0; c < D; ++c) {
                    double b[NS];
                    bcol<D, double>(gr, c, b);
                    for (int s = 0; s < NS; ++s) eps[s] += b[s] * us[g][l * D + c];
                }
            }
            qw[g][t] = Jc * w;
            {
                const TK wt = TK(Jc * w);
                qwl[g][t][0] = wt * sLM[rec.child_mat[k]][0];
                qwl[g][t][1] = wt * sLM[rec.child_mat[k]][1];
            }
            double sig[NS];
            lei_stress_c<D>(sLC[rec.child_mat[k]], eps, sig);
            for (int s = 0; s < NS; ++s) qs[g][t][s] = sig[s];
        }
    }
    if constexpr (COH)
    for (int cq = gl; cq < ncq; cq += GS) {
        {
            // ---- cohesive qp ----
            const int k = cq / cnq, q = cq % cnq;
            constexpr int CNN = CohLim<D>::N;
            double lo[CNN][3], ulo[CNN * D], uup[CNN * D], N[CNN], R[D][D], w, RDR[D][D], tR[D];
            for (int n = 0; n < CNN; ++n)
                for (int i = 0; i < D; ++i) {
                    const bool in = n < cnn;
                    lo[n][i] = in ? xs[g][rec.coh_nodes[k][n]][i] : 0.0;
                    ulo[n * D + i] = in ? us[g][rec.coh_nodes[k][n] * D + i] : 0.0;
                    uup[n * D + i] = in ? us[g][rec.coh_nodes[k][n + cnn] * D + i] : 0.0;
                }
            const double dA = cohesive_qp<D>(rec.coh_nn, q, lo, N, R, w);
            const int iq = rec.isv_off + cq;
            double hist = cq == gl ? hist0 : (isv_c ? isv_c[iq] : 0.0);
            cohesive_point<D>(sCM[rec.coh_mat[k]], cnn, N, R, ulo, uup, hist, RDR, tR);
            if (isv_t) isv_t[iq] = hist;
            for (int i = 0; i < D; ++i) {
                for (int j = 0; j < D; ++j) cRDR[g][cq][i][j] = RDR[i][j];
                ctR[g][cq][i] = tR[i];
            }
            for (int n = 0; n < CNN; ++n) cN[g][cq][n] = n < cnn ? N[n] : 0.0;
            cw[g][cq] = dA * w;
        }
    }
    __syncthreads();
    if constexpr (D == 3) {
        // one lane per sum: S (2), then V (2 x D per enrichment node), each over
        // its qps in ascending order in FP64, rounded once to TK
        const int ne = nloc > NEN ? nloc - NEN : 0;
        for (int task = gl; task < 2 + 2 * D * ne; task += GS) {
            if (task < 2) {
                double a = 0.0;
                for (int q = 0; q < nvq; ++q) a += double(qwl[g][q][task]);
                sS[g][task] = TK(a);
            } else {
                const int k = task - 2, l = k / (2 * D), kind = (k / D) & 1, c = k % D;
                double a = 0.0;
                for (QMask qm = qmask[g][NEN + l]; qm; qm &= qm - QMask(1)) {
                    const int q = sizeof(QMask) == 8 ? __builtin_ctzll(qm) : __builtin_ctz(uint32_t(qm));
                    a += double(qwl[g][q][kind]) * qg[g][q][l][c];
                }
                sV[g][l][kind][c] = TK(a);
            }
        }
        __syncthreads();
    }
    if (!live) return;

    // gradient of local node l at qp q (parent gradient, the child gradient of the
    // child node that is enrichment node l, or 0; src/Mesh.cpp:1596-1614)
    auto grad = [&](int q, int l, double* gr) {
        for (int a = 0; a < D; ++a) gr[a] = l < NEN ? pg[g][l][a] : qg[g][q][l - NEN][a];
    };
    // local-node-pair blocks: K(li, lj) = V(li, lj) + C(li, lj) with the
    // volumetric part V(li, lj) = sum_q B_li^T (wJ D) B_lj (isotropic closed
    // form) and the cohesive
    // fold C.  D is symmetric, so V(lj, li) = V(li, lj)^T, and C(lj, li) =
    // C(li, lj) (the fold scales R^T D R by N_li N_lj): one lane per unordered
    // pair li <= lj writes both blocks.  Staged block-major: the d x d block of
    // (li, lj) is contiguous at (li * nloc + lj) * d * d.
    TK* Ke = Kst + p.ckoff[pos];
    const int nup = nloc * (nloc + 1) / 2;
    for (int pb = gl; pb < nup; pb += GS) {
        int li = 0, rem = pb;
        while (rem >= nloc - li) { rem -= nloc - li; ++li; }
        const int lj = li + rem;
        TK vol[D][D];
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) vol[i][j] = TK(0);
        bool summed = false;
        if constexpr (D == 3) {
            if (li < NEN) {
                // a parent row: its gradient is the same at every qp, so the sum
                // over the qps collapses onto the sums S (parent column) or V
                TK gi[D];
                for (int a = 0; a < D; ++a) gi[a] = TK(pg[g][li][a]);
                if (lj < NEN) {
                    TK gj[D];
                    for (int a = 0; a < D; ++a) gj[a] = TK(pg[g][lj][a]);
                    iso_block<D, TK>(gi, gj, sS[g][0], sS[g][1], vol);
                } else {
                    iso_block_v<D, TK>(gi, sV[g][lj - NEN][0], sV[g][lj - NEN][1], vol);
                }
                summed = true;
            }
        }
        // the qps at which both gradients are not zero, ascending (the others
        // contribute nothing)
        if (!summed)
        for (QMask qm = qmask[g][li] & qmask[g][lj]; qm; qm &= qm - QMask(1)) {
            const int q = sizeof(QMask) == 8 ? __builtin_ctzll(qm) : __builtin_ctz(uint32_t(qm));
            double gid[D], gjd[D];
            grad(q, li, gid);
            grad(q, lj, gjd);
            TK gi[D], gj[D];
            for (int a = 0; a < D; ++a) { gi[a] = TK(gid[a]); gj[a] = TK(gjd[a]); }

            TK blk[D][D];
            iso_block<D, TK>(gi, gj, qwl[g][q][0], qwl[g][q][1], blk);
            for (int i = 0; i < D; ++i)
                for (int j = 0; j < D; ++j) vol[i][j] += blk[i][j];
        }
        TK coh[D][D];
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) coh[i][j] = TK(0);
        if constexpr (COH)
        for (int k = 0; k < rec.n_coh; ++k) {
            int m1 = -1, m2 = -1;
            for (int m = 0; m < rec.coh_nn; ++m) {
                if (rec.coh_nodes[k][m] == li) m1 = m;
                if (rec.coh_nodes[k][m] == lj) m2 = m;
            }
            if (m1 < 0 || m2 < 0) continue;
            const int n1 = m1 % cnn, n2 = m2 % cnn;
            const TK sgn = ((m1 < cnn) == (m2 < cnn)) ? TK(1) : TK(-1);
            for (int q = 0; q < cnq; ++q) {
                const int cq = k * cnq + q;
                const TK f = TK(cw[g][cq]) * sgn * TK(cN[g][cq][n1]) * TK(cN[g][cq][n2]);
                for (int i = 0; i < D; ++i)
                    for (int j = 0; j < D; ++j) coh[i][j] += f * TK(cRDR[g][cq][i][j]);
            }
        }
        TK* B1 = Ke + int64_t(li * nloc + lj) * D * D;
        TK* B2 = Ke + int64_t(lj * nloc + li) * D * D;
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) B1[i * D + j] = vol[i][j] + coh[i][j];
        if (li != lj)
            for (int i = 0; i < D; ++i)
                for (int j = 0; j < D; ++j) B2[i * D + j] = vol[j][i] + coh[i][j];
    }
    double* Pe = Pst + p.cpoff[pos];
    for (int i = gl; i < nd; i += GS) {
        const int l = i / D, c = i % D;
        double a = 0.0;
        for (QMask qm = qmask[g][l]; qm; qm &= qm - QMask(1)) {
            const int q = sizeof(QMask) == 8 ? __builtin_ctzll(qm) : __builtin_ctz(uint32_t(qm));
            double gr[D], v[D];
            int r[D];
            grad(q, l, gr);
            bnz<D, double>(gr, c, r, v);
            double pt = 0.0;
            for (int n = 0; n < D; ++n) pt += v[n] * qs[g][q][r[n]];
            a += qw[g][q] * pt;
        }
        double pc = 0.0;
        if constexpr (COH)
        for (int k = 0; k < rec.n_coh; ++k)
            for (int m = 0; m < rec.coh_nn; ++m)
                if (rec.coh_nodes[k][m] == l) {
                    const int n = m % cnn;
                    const double sgn = m < cnn ? -1.0 : 1.0;
                    double s = 0.0;
                    for (int q = 0; q < cnq; ++q) {
                        const int cq = k * cnq + q;
                        s += cw[g][cq] * (sgn * (cN[g][cq][n] * ctR[g][cq][c]));
                    }
                    pc += s;
                }
        Pe[i] = a + pc;
    }
}

// Local index of `node` in element e's node list (parents, then enrichment).
__device__ inline int find_local(const DevProb& p, int e, int ci, int node, int nen) {
    for (int n = 0; n < nen; ++n)
        if (p.conn[int64_t(e) * nen + n] == node) return n;
    if (ci >= 0) {
        const pg_cut_rec& r = p.cut[ci];
        for (int k = 0; k < r.n_enr; ++k)
            if (r.enr[k] == node) return nen + k;
    }
    return -1;
}

// --------------------------------------------------------------------------
// Node-pair kernel of the plain schedule: thread per node pair (a, b); the d x d
// block K[a-dofs, b-dofs] sums, over the elements incident to a (ascending: the
// reference's element loop order, src/Assembler.cpp:100) that also contain b,
// the staged element matrices, and is written once through the scatter indices
// (set_matrix, src/Assembler.cpp:659-679; no atomics).  Diagonal pairs also sum
// P (set_vec, src/Assembler.cpp:681-701).
// --------------------------------------------------------------------------
// Cohesive openings of a 2D cohesive mesh (WriterCohesiveOpenings::
// writeFromStream, src/WriterCohesiveOpenings.cpp:73-147): one thread per cut
// record; for each cohesive edge, N = (1/2, 1/2) at its midpoint, the angle of
// the midpoint of its lower nodes about the inclusion centre, and the opening
// R * sum_n N_n (U_upper - U_lower) (cohesiveDeltaFast, src/Utilities.cpp:528-545,
// with the edge's rotation, constant along a straight CohEdge2) divided by
// delta_c.  err bit 2: a law other than Ortiz-Pandolfi.
__global__ void k_coh_openings(DevProb p, const double* __restrict__ Un, const double* __restrict__ centers,
                               double* __restrict__ out, int* err) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= p.n_cut) return;
    const pg_cut_rec& rec = p.cut[c];
    if (rec.n_coh == 0 || rec.coh_nn != 4) return;
    const int e = p.cut_elem[c];
    auto gnode = [&](int l) { return l < 3 ? p.conn[int64_t(e) * 3 + l] : rec.enr[l - 3]; };
    for (int k = 0; k < rec.n_coh; ++k) {
        const pg_material& m = p.mats[rec.coh_mat[k]];
        if (m.type != PG_MAT_OP && m.type != PG_MAT_OP_NU) { atomicOr(err, 4); return; }
        const double N[3] = {0.5, 0.5, 0.0};
        double lo[3][3], ulo[6], uup[6];
        for (int n = 0; n < 2; ++n) {
            const int a = gnode(rec.coh_nodes[k][n]), b = gnode(rec.coh_nodes[k][n + 2]);
            for (int i = 0; i < 3; ++i) lo[n][i] = i < 2 ? p.coords[int64_t(a) * 2 + i] : 0.0;
            for (int i = 0; i < 2; ++i) {
                ulo[n * 2 + i] = Un[int64_t(a) * 2 + i];
                uup[n * 2 + i] = Un[int64_t(b) * 2 + i];
            }
        }
        double Nq[3], R[2][2], w;
        cohesive_qp<2>(4, 0, lo, Nq, R, w);
        double xm[2] = {0.0, 0.0}, dl[2], del[2];
        for (int d = 0; d < 2; ++d)
            for (int n = 0; n < 2; ++n) xm[d] += N[n] * lo[n][d];
        for (int i = 0; i < 2; ++i) {
            double s = 0;
            for (int n = 0; n < 2; ++n) s += N[n] * (uup[n * 2 + i] - ulo[n * 2 + i]);
            dl[i] = s;
        }
        for (int i = 0; i < 2; ++i) {
            double s = 0;
            for (int j = 0; j < 2; ++j) s += R[i][j] * dl[j];
            del[i] = s;
        }
        double* o = out + (int64_t(rec.isv_off) / 2 + k) * 3;
        const int64_t row = int64_t(rec.isv_off) / 2 + k;
        o[0] = atan2(xm[1] - centers[row * 2 + 1], xm[0] - centers[row * 2]);
        o[1] = del[0] / m.delta_c;
        o[2] = del[1] / m.delta_c;
    }
}

// Cohesive failure lengths of a 2D cohesive mesh (WriterCohesiveFailureLengths,
// src/WriterCohesiveFailureLengths.cpp:64-86 and 183-224): one thread per cut
// record; for each cohesive edge, the distance between its two lower nodes and
// the mean over its two Gauss points of the law's normalised failure
// (getNormalizedCohesiveFailure): delta_eff / delta_c for OPCohesiveNU
// (src/Material_OPCohesiveNU.cpp:43-60,217-221), delta_eff / delta_c1 for
// BilinearCohesiveNU (src/Material_BilinearCohesiveNU.cpp:48-65,250-254), and 0
// for the laws that keep the base definition (include/material.h:123).  An
// opening of squared norm below 1e-40 counts as 1e-20 in each component.
__global__ void k_coh_failure(DevProb p, const double* __restrict__ Un, double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= p.n_cut) return;
    const pg_cut_rec& rec = p.cut[c];
    if (rec.n_coh == 0 || rec.coh_nn != 4) return;
    const int e = p.cut_elem[c];
    auto gnode = [&](int l) { return l < 3 ? p.conn[int64_t(e) * 3 + l] : rec.enr[l - 3]; };
    for (int k = 0; k < rec.n_coh; ++k) {
        const pg_material& m = p.mats[rec.coh_mat[k]];
        double lo[3][3], ulo[6], uup[6];
        for (int n = 0; n < 2; ++n) {
            const int a = gnode(rec.coh_nodes[k][n]), b = gnode(rec.coh_nodes[k][n + 2]);
            for (int i = 0; i < 3; ++i) lo[n][i] = i < 2 ? p.coords[int64_t(a) * 2 + i] : 0.0;
            for (int i = 0; i < 2; ++i) {
                ulo[n * 2 + i] = Un[int64_t(a) * 2 + i];
                uup[n * 2 + i] = Un[int64_t(b) * 2 + i];
            }
        }
        double fsum = 0.0;
        for (int q = 0; q < 2; ++q) {
            double Nq[3], R[2][2], w;
            cohesive_qp<2>(4, q, lo, Nq, R, w);
            double dl[2], del[2];
            for (int i = 0; i < 2; ++i) {
                double s = 0;
                for (int n = 0; n < 2; ++n) s += Nq[n] * (uup[n * 2 + i] - ulo[n * 2 + i]);
                dl[i] = s;
            }
            for (int i = 0; i < 2; ++i) {
                double s = 0;
                for (int j = 0; j < 2; ++j) s += R[i][j] * dl[j];
                del[i] = s;
            }
            if (del[0] * del[0] + del[1] * del[1] < 1e-40) del[0] = del[1] = 1e-20;
            const double de = sqrt(del[1] * del[1] + m.beta2 * (del[0] * del[0]));
            fsum += m.type == PG_MAT_OP_NU ? de / m.delta_c : m.type == PG_MAT_BL_NU ? de / m.delta_c1 : 0.0;
        }
        const double dx = lo[1][0] - lo[0][0], dy = lo[1][1] - lo[0][1];
        double* o = out + (int64_t(rec.isv_off) / 2 + k) * 2;
        o[0] = sqrt(dx * dx + dy * dy);
        o[1] = fsum / 2.0;
    }
}

// --------------------------------------------------------------------------
// Jacobi-preconditioned conjugate gradients on a CSR matrix (the linear solve
// of each Newton iteration, SolverNonlinearKSP / Solver::solve,
// src/SolverNonlinearKSP.cpp:94, src/Solver.cpp:108-160).  One thread per row
// for the product; dot products reduce over a fixed grid and then one block in
// a fixed order, so results are bitwise repeatable.  The CG scalars stay on
// the device in sc[]: sc[0], sc[1] hold r.z of alternate iterations, sc[2]
// p.Ap, sc[3] r.r, sc[4] a second copy a fused dot does not need.
constexpr int PCG_T = 256, PCG_B = 512;

__global__ void k_csr_diag_inv(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                               const double* __restrict__ v, double* __restrict__ dinv) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
        if (ci[k] == i) d = v[k];
    dinv[i] = d != 0.0 ? 1.0 / d : 1.0;  // a zero diagonal keeps the identity, as PCJACOBI does
}

__global__ void k_csr_spmv(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                           const double* __restrict__ v, const double* __restrict__ x, double* __restrict__ y) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s = fma(v[k], x[ci[k]], s);
    y[i] = s;
}

// part[2 * block + {0, 1}] = partial sums of a.b and c.d
__global__ void __launch_bounds__(PCG_T) k_dot2(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                                const double* __restrict__ c, const double* __restrict__ d,
                                                double* __restrict__ part) {
    __shared__ double l0[PCG_T], l1[PCG_T];
    double s0 = 0.0, s1 = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * PCG_T + threadIdx.x; i < n; i += int64_t(gridDim.x) * PCG_T) {
        s0 = fma(a[i], b[i], s0);
        s1 = fma(c[i], d[i], s1);
    }
    l0[threadIdx.x] = s0;
    l1[threadIdx.x] = s1;
    __syncthreads();
    for (int w = PCG_T / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) {
            l0[threadIdx.x] += l0[threadIdx.x + w];
            l1[threadIdx.x] += l1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = l0[0];
        part[2 * blockIdx.x + 1] = l1[0];
    }
}

__global__ void __launch_bounds__(PCG_T) k_dot_fin(const double* __restrict__ part, int nb, double* __restrict__ sc,
                                                   int o0, int o1) {
    __shared__ double l0[PCG_T], l1[PCG_T];
    double s0 = 0.0, s1 = 0.0;
    for (int i = threadIdx.x; i < nb; i += PCG_T) {
        s0 += part[2 * i];
        s1 += part[2 * i + 1];
    }
    l0[threadIdx.x] = s0;
    l1[threadIdx.x] = s1;
    __syncthreads();
    for (int w = PCG_T / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) {
            l0[threadIdx.x] += l0[threadIdx.x + w];
            l1[threadIdx.x] += l1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sc[o0] = l0[0];
        sc[o1] = l1[0];
    }
}

// r = b - q, z = dinv r, p = z
__global__ void k_pcg_init(int64_t n, const double* __restrict__ b, const double* __restrict__ q,
                           const double* __restrict__ dinv, double* __restrict__ r, double* __restrict__ z,
                           double* __restrict__ pv) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ri = b[i] - q[i];
    r[i] = ri;
    z[i] = dinv[i] * ri;
    pv[i] = dinv[i] * ri;
}

// alpha = sc[irz] / sc[2]; x += alpha p; r -= alpha q; z = dinv r
__global__ void k_pcg_xr(int64_t n, double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
                         const double* __restrict__ pv, const double* __restrict__ q, const double* __restrict__ dinv,
                         const double* __restrict__ sc, int irz) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double alpha = sc[irz] / sc[2];
    x[i] = fma(alpha, pv[i], x[i]);
    const double ri = fma(-alpha, q[i], r[i]);
    r[i] = ri;
    z[i] = dinv[i] * ri;
}

// y += a x
__global__ void k_axpy(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = fma(a, x[i], y[i]);
}

// beta = sc[inew] / sc[iold]; p = z + beta p
__global__ void k_pcg_p(int64_t n, double* __restrict__ pv, const double* __restrict__ z,
                        const double* __restrict__ sc, int inew, int iold) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pv[i] = fma(sc[inew] / sc[iold], pv[i], z[i]);
}

template <int D>
__global__ void k_pair_staged(DevProb p, const double* __restrict__ stage, const double* __restrict__ Ks,
                              const double* __restrict__ Ps, double* __restrict__ K, double* __restrict__ P,
                              unsigned flags) {
    constexpr int NEN = D + 1;
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= p.n_pairs) return;
    const int a = p.npat_row[q], b = p.npat_col[q];
    double acc[D][D];
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) acc[i][j] = 0.0;
    double pa[D];
    for (int i = 0; i < D; ++i) pa[i] = 0.0;
    const bool diag = (a == b) && (flags & PG_FUNC);
    const bool jac = (flags & PG_JAC) != 0;
    if (!jac && !diag) return;  // a residual-only call needs the diagonal pairs only
    const int e0 = p.n2e_ptr[a], e1 = p.n2e_ptr[a + 1];
    for (int t = e0; t < e1; ++t) {
        const int e = p.n2e[t];
        const int ci = p.cut_of_elem[e];
        if (ci < 0) {
            int la = -1, lb = -1;
            for (int n = 0; n < NEN; ++n) {
                const int node = p.conn[int64_t(e) * NEN + n];
                if (node == a) la = n;
                if (node == b) lb = n;
            }
            if (la < 0 || lb < 0) continue;
            if (jac) {
                const double* Kb = stage + PlainStage<D>::kblk(e, la, lb);
                for (int i = 0; i < D; ++i)
                    for (int j = 0; j < D; ++j) acc[i][j] += Kb[i * D + j];
            }
            if (diag)
                for (int i = 0; i < D; ++i) pa[i] += stage[PlainStage<D>::pvec(e, la) + i];
            continue;
        }
        const int la = find_local(p, e, ci, a, NEN);
        const int lb = (a == b) ? la : find_local(p, e, ci, b, NEN);
        if (la < 0 || lb < 0) continue;
        if (jac) {
            const int nd = (NEN + p.cut[ci].n_enr) * D;
            const double* Ke = Ks + p.cut_koff[ci];
            for (int i = 0; i < D; ++i)
                for (int j = 0; j < D; ++j) acc[i][j] += Ke[int64_t(la * D + i) * nd + lb * D + j];
        }
        if (diag)
            for (int i = 0; i < D; ++i) pa[i] += Ps[p.cut_poff[ci] + la * D + i];
    }
    if (jac) {
        const int32_t* mp = p.npat_map + q * D * D;
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) K[mp[i * D + j]] = acc[i][j];
    }
    if (diag)
        for (int i = 0; i < D; ++i) P[p.node_dof[int64_t(a) * D + i]] = pa[i];
}

// --------------------------------------------------------------------------
// Heat conduction (AssemblerThermal, src/AssemblerThermal.cpp:24-85): one dof
// per node, B = the shape gradients (Assemble_Thermal_B_Mat,
// src/Utilities.cpp:147-160), per qp K_tmp = (B^T D) B and P_tmp = B^T q with
// the flux q = D . B T, K_el += (J w) K_tmp, P_el += (J w) P_tmp (the
// reference's order, src/Assembler.cpp:148-368), staged per element and summed
// per node pair in element order as the FP64 drop-in does.
// Staged layout of uncut elements: blocks of 64 elements; per local entry
// (la, lb) the 64 elements' values, then per la the 64 residuals.
// --------------------------------------------------------------------------
template <int NEN> struct ThermStage {
    static constexpr int64_t KB = int64_t(NEN) * NEN * 64;
    static constexpr int64_t W = KB + int64_t(NEN) * 64;
    __host__ __device__ static int64_t kval(int64_t e, int la, int lb) {
        return (e >> 6) * W + int64_t(la * NEN + lb) * 64 + (e & 63);
    }
    __host__ __device__ static int64_t pval(int64_t e, int la) { return (e >> 6) * W + KB + int64_t(la) * 64 + (e & 63); }
};

// Uncut elements, one thread each.  Tn: T per node (node order).
template <int D>
__global__ __launch_bounds__(256) void k_therm_elem(DevProb p, const double* __restrict__ Tn,
                                                    double* __restrict__ stage) {
    constexpr int NEN = D + 1, NQ = (D == 2 ? 4 : 5);
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= p.n_elem || p.cut_of_elem[e] >= 0) return;
    double px[NEN][3], pdN[NEN][D], pg[NEN][3], t[NEN];
#pragma unroll
    for (int n = 0; n < NEN; ++n) {
        const int node = p.conn[int64_t(e) * NEN + n];
        for (int i = 0; i < 3; ++i) px[n][i] = i < D ? p.coords[int64_t(node) * D + i] : 0.0;
        t[n] = Tn[node];
        for (int a = 0; a < D; ++a) pdN[n][a] = simplex_dNdr(n, a);
    }
    const double J = transform<D>(NEN, pdN, px, pg);
    const double* Dm = p.mats[p.elem_mat[e]].D;  // D x D, row-major
    double Dl[D][D];
#pragma unroll
    for (int k = 0; k < D; ++k)
        for (int l = 0; l < D; ++l) Dl[k][l] = Dm[k * D + l];
    double wq[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) wq[q] = J * (D == 2 ? TRI_QW[q] : TET_QW[q]);
    // grad T = B T (dgemv order: nodes ascending), flux q = D grad T
    double gT[D], fl[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        double v = 0.0;
#pragma unroll
        for (int a = 0; a < NEN; ++a) v += pg[a][k] * t[a];
        gT[k] = v;
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
        double v = 0.0;
#pragma unroll
        for (int l = 0; l < D; ++l) v += Dl[k][l] * gT[l];
        fl[k] = v;
    }
#pragma unroll
    for (int a = 0; a < NEN; ++a) {
        double btd[D];
#pragma unroll
        for (int l = 0; l < D; ++l) {
            double v = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) v += pg[a][k] * Dl[k][l];
            btd[l] = v;
        }
#pragma unroll
        for (int b = 0; b < NEN; ++b) {
            double kt = 0.0;
#pragma unroll
            for (int l = 0; l < D; ++l) kt += btd[l] * pg[b][l];
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc += wq[q] * kt;
            stage[ThermStage<NEN>::kval(e, a, b)] = acc;
        }
        double pt = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) pt += pg[a][k] * fl[k];
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc += wq[q] * pt;
        stage[ThermStage<NEN>::pval(e, a)] = acc;
    }
}

// Cut elements, one wavefront each: one lane per child qp computes the
// gradients of the local nodes (parent, and a child's own gradient for its
// enrichment nodes, src/Mesh.cpp:1596-1614), J w and the flux; then one lane
// per entry of K_el (and of P_el) sums its qps in order.  Ks / Ps: row-major
// nloc x nloc matrices and nloc vectors at cut_koff / cut_poff.  POLY as in
// k_cut_plain.
template <int D, bool POLY>
__global__ __launch_bounds__(64) void k_therm_cut(DevProb p, const double* __restrict__ Tn, double* __restrict__ Ks,
                                                  double* __restrict__ Ps, int* __restrict__ err,
                                                  const int32_t* __restrict__ list) {
    constexpr int NEN = D + 1;
    constexpr int MAXN = CutLim<D>::N, MAXQ = CutLim<D>::Q;
    __shared__ double lx[MAXN][3];
    __shared__ double te[MAXN];
    __shared__ double gq[MAXQ][MAXN][D];
    __shared__ double wq[MAXQ];
    __shared__ double fq[MAXQ][D];   // flux at each qp
    __shared__ int mq[MAXQ];
    __shared__ pg_cut_rec rec;
    const int c = list ? list[blockIdx.x] : int(blockIdx.x), lane = threadIdx.x;
    constexpr int RW = int(sizeof(pg_cut_rec) / 4);
    if (lane < RW) reinterpret_cast<int32_t*>(&rec)[lane] = reinterpret_cast<const int32_t*>(&p.cut[c])[lane];
    const int e = p.cut_elem[c];
    __syncthreads();
    if constexpr (!POLY && D == 2) {
        bool poly = false;
        for (int k = 0; k < rec.n_child; ++k)
            poly = poly || rec.child_type[k] == PG_CHILD_POLY5 || rec.child_type[k] == PG_CHILD_POLY6;
        if (poly) return;  // the polygon launch takes it (the whole block returns)
    }
    const int nloc = NEN + rec.n_enr;
    if (lane < nloc) {
        const int node = lane < NEN ? p.conn[int64_t(e) * NEN + lane] : rec.enr[lane - NEN];
        for (int i = 0; i < 3; ++i) lx[lane][i] = i < D ? p.coords[int64_t(node) * D + i] : 0.0;
        te[lane] = Tn[node];
    }
    __syncthreads();
    double px[NEN][3], pdN[NEN][D], pgr[NEN][3];
    for (int n = 0; n < NEN; ++n) {
        for (int i = 0; i < 3; ++i) px[n][i] = lx[n][i];
        for (int a = 0; a < D; ++a) pdN[n][a] = simplex_dNdr(n, a);
    }
    transform<D>(NEN, pdN, px, pgr);
    const int nvq = rec.n_vqp;
    for (int t = lane; t < nvq; t += 64) {
        int k = 0, q = t;
        for (; k < rec.n_child; ++k) {
            const int nq = PG_CHILD_NQP(rec.child_type[k]);
            if (q < nq) break;
            q -= nq;
        }
        constexpr int CN = (D == 2 && POLY) ? 6 : 4;
        const int ct = rec.child_type[k], ccn = PG_CHILD_NN(ct);
        double cx[CN][3], cg[CN][3], w;
        int cl[CN];
        for (int n = 0; n < CN; ++n) {
            cl[n] = n < ccn ? rec.child_node[PG_CHILD_STRIDE(D) * k + n] : -1;
            for (int i = 0; i < 3; ++i) cx[n][i] = (n < ccn && i < D) ? lx[cl[n]][i] : 0.0;
        }
        const double Jc = child_qp<D, POLY>(ct, q, ccn, cx, cg, w);
        if (Jc < 0.0) atomicOr(err, 1);  // store_shape_functions aborts (src/Mesh.cpp:1561)
        double gT[D];
        for (int a = 0; a < D; ++a) gT[a] = 0.0;
        for (int l = 0; l < nloc; ++l) {
            double gr[3] = {0.0, 0.0, 0.0};
            if (l < NEN) {
                for (int a = 0; a < D; ++a) gr[a] = pgr[l][a];
            } else {
                for (int n = 0; n < CN; ++n)
                    if (cl[n] == l)
                        for (int a = 0; a < D; ++a) gr[a] = cg[n][a];
            }
            for (int a = 0; a < D; ++a) {
                gq[t][l][a] = gr[a];
                gT[a] += gr[a] * te[l];
            }
        }
        const double* Dm = p.mats[rec.child_mat[k]].D;
        for (int a = 0; a < D; ++a) {
            double v = 0.0;
            for (int b = 0; b < D; ++b) v += Dm[a * D + b] * gT[b];
            fq[t][a] = v;
        }
        wq[t] = Jc * w;
        mq[t] = rec.child_mat[k];
    }
    __syncthreads();
    double* Ke = Ks + p.cut_koff[c];
    for (int idx = lane; idx < nloc * nloc; idx += 64) {
        const int i = idx / nloc, j = idx % nloc;
        double acc = 0.0;
        for (int q = 0; q < nvq; ++q) {
            const double* Dm = p.mats[mq[q]].D;
            double kt = 0.0;
            for (int l = 0; l < D; ++l) {
                double btd = 0.0;
                for (int k = 0; k < D; ++k) btd += gq[q][i][k] * Dm[k * D + l];
                kt += btd * gq[q][j][l];
            }
            acc += wq[q] * kt;
        }
        Ke[idx] = acc;
    }
    double* Pe = Ps + p.cut_poff[c];
    for (int i = lane; i < nloc; i += 64) {
        double acc = 0.0;
        for (int q = 0; q < nvq; ++q) {
            double pt = 0.0;
            for (int k = 0; k < D; ++k) pt += gq[q][i][k] * fq[q][k];
            acc += wq[q] * pt;
        }
        Pe[i] = acc;
    }
}

// Node pair (a, b) of the scalar pattern: the staged entries of every element
// incident to a that contains b, summed in element order, written once through
// the pair's scatter index; the diagonal pair also sums P (set_matrix /
// set_vec, src/Assembler.cpp:659-701).
template <int D>
__global__ void k_therm_pair(DevProb p, const double* __restrict__ stage, const double* __restrict__ Ks,
                             const double* __restrict__ Ps, double* __restrict__ K, double* __restrict__ P,
                             unsigned flags) {
    constexpr int NEN = D + 1;
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= p.n_pairs) return;
    const int a = p.npat_row[q], b = p.npat_col[q];
    const bool diag = (a == b) && (flags & PG_FUNC);
    const bool jac = (flags & PG_JAC) != 0;
    if (!jac && !diag) return;
    double acc = 0.0, pa = 0.0;
    for (int t = p.n2e_ptr[a]; t < p.n2e_ptr[a + 1]; ++t) {
        const int e = p.n2e[t];
        const int ci = p.cut_of_elem[e];
        int la = -1, lb = -1;
        if (ci < 0) {
            for (int n = 0; n < NEN; ++n) {
                const int node = p.conn[int64_t(e) * NEN + n];
                if (node == a) la = n;
                if (node == b) lb = n;
            }
            if (la < 0 || lb < 0) continue;
            if (jac) acc += stage[ThermStage<NEN>::kval(e, la, lb)];
            if (diag) pa += stage[ThermStage<NEN>::pval(e, la)];
            continue;
        }
        la = find_local(p, e, ci, a, NEN);
        lb = (a == b) ? la : find_local(p, e, ci, b, NEN);
        if (la < 0 || lb < 0) continue;
        const int nl = NEN + p.cut[ci].n_enr;
        if (jac) acc += Ks[p.cut_koff[ci] + int64_t(la) * nl + lb];
        if (diag) pa += Ps[p.cut_poff[ci] + la];
    }
    if (jac) K[p.npat_map[q]] = acc;
    if (diag) P[p.node_dof[a]] = pa;
}

// --------------------------------------------------------------------------
// Fields of the VTK writers from device data (Writer_VTK::gather_nodal_data and
// gather_element_data, src/Writer_VTK.cpp:642-1000).
// --------------------------------------------------------------------------
// Nodal displacement for plotting: a regular node's own U; an enrichment node
// U(n1) (1 - alpha) + U(n2) alpha + U_enr * 1 (Problem::interpolate_nodal_enrichment,
// src/Problem.cpp:538-570, scale factor 1): the field's value at the node.
__global__ void k_plot_nodes(DevProb p, const double* __restrict__ Un, const int32_t* __restrict__ edge,
                             const double* __restrict__ alpha, double* __restrict__ out) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int D = p.dim;
    if (k >= int64_t(p.n_nodes) + p.n_enr) return;
    if (k < p.n_nodes) {
        for (int i = 0; i < D; ++i) out[k * D + i] = Un[k * D + i];
        return;
    }
    const int64_t j = k - p.n_nodes;
    const int n1 = edge[2 * j], n2 = edge[2 * j + 1];
    const double a = alpha[j];
    for (int i = 0; i < D; ++i)
        out[k * D + i] = Un[int64_t(n1) * D + i] * (1.0 - a) + Un[int64_t(n2) * D + i] * a + Un[k * D + i] * 1.0;
}

// SmallStrainFast (src/Utilities.cpp:300-362): each strain component summed
// over the nodes in turn; grad(l, a) and u(l, i) give node l's gradient and U.
template <int D, typename G, typename UF>
__device__ inline void small_strain_fast(int nl, G grad, UF u, double* e) {
    for (int k = 0; k < (D == 2 ? 3 : 6); ++k) e[k] = 0.0;
    if constexpr (D == 2) {
        for (int n = 0; n < nl; ++n) e[0] += grad(n, 0) * u(n, 0);
        for (int n = 0; n < nl; ++n) e[1] += grad(n, 1) * u(n, 1);
        for (int n = 0; n < nl; ++n) e[2] += grad(n, 1) * u(n, 0) + grad(n, 0) * u(n, 1);
    } else {
        for (int n = 0; n < nl; ++n) e[0] += grad(n, 0) * u(n, 0);
        for (int n = 0; n < nl; ++n) e[1] += grad(n, 1) * u(n, 1);
        for (int n = 0; n < nl; ++n) e[2] += grad(n, 2) * u(n, 2);
        for (int n = 0; n < nl; ++n) e[3] += grad(n, 2) * u(n, 1) + grad(n, 1) * u(n, 2);
        for (int n = 0; n < nl; ++n) e[4] += grad(n, 2) * u(n, 0) + grad(n, 0) * u(n, 2);
        for (int n = 0; n < nl; ++n) e[5] += grad(n, 1) * u(n, 0) + grad(n, 0) * u(n, 1);
    }
}

// One thread per VTK cell (an uncut element, or one integration child of a cut
// element: Writer_VTK::store_elements, src/Writer_VTK.cpp:221-312): the
// integrals over the cell's qps of 1, the strain and the stress, each qp
// adding (J w) times its value (Integrator::integrateElem with the strain and
// stress functors and the volume functor, include/Integrator.h:151-228,
// include/AssemblerSmallStrainStructural.h:23-79).  out: per cell
// [volume, strain (nstr), stress (nstr)].  cells[c] = (element, child or -1).
template <int D, bool POLY>
__global__ __launch_bounds__(128) void k_plot_cells(DevProb p, const double* __restrict__ Un,
                                                    const int2* __restrict__ cells, int n,
                                                    double* __restrict__ out) {
    constexpr int NEN = D + 1, NS = (D == 2 ? 3 : 6), NQ = (D == 2 ? 4 : 5), S = 1 + 2 * NS;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int e = cells[c].x, kc = cells[c].y;
    double px[NEN][3], pdN[NEN][D], pg[NEN][3];
    for (int m = 0; m < NEN; ++m) {
        const int node = p.conn[int64_t(e) * NEN + m];
        for (int i = 0; i < 3; ++i) px[m][i] = i < D ? p.coords[int64_t(node) * D + i] : 0.0;
        for (int a = 0; a < D; ++a) pdN[m][a] = simplex_dNdr(m, a);
    }
    const double J = transform<D>(NEN, pdN, px, pg);
    double vol = 0.0, ie[NS], is[NS];
    for (int k = 0; k < NS; ++k) ie[k] = is[k] = 0.0;
    if (kc < 0) {
        double eps[NS], sig[NS];
        small_strain_fast<D>(
            NEN, [&](int m, int a) { return pg[m][a]; },
            [&](int m, int i) { return Un[int64_t(p.conn[int64_t(e) * NEN + m]) * D + i]; }, eps);
        lei_stress<D>(p.mats[p.elem_mat[e]].D, eps, sig);
        for (int q = 0; q < NQ; ++q) {
            const double wgt = J * (D == 2 ? TRI_QW[q] : TET_QW[q]);
            vol += wgt * 1.0;
            for (int k = 0; k < NS; ++k) ie[k] += wgt * eps[k];
            for (int k = 0; k < NS; ++k) is[k] += wgt * sig[k];
        }
    } else {
        const int ci = p.cut_of_elem[e];
        const pg_cut_rec& rec = p.cut[ci];
        const int nloc = NEN + rec.n_enr;
        constexpr int CN = (D == 2 && POLY) ? 6 : 4;
        const int ct = rec.child_type[kc], ccn = PG_CHILD_NN(ct);
        double cx[CN][3], cg[CN][3], w;
        int cl[CN];
        for (int m = 0; m < CN; ++m) {
            cl[m] = m < ccn ? rec.child_node[PG_CHILD_STRIDE(D) * kc + m] : -1;
            const int l = cl[m] < 0 ? 0 : cl[m];
            const int node = l < NEN ? p.conn[int64_t(e) * NEN + l] : rec.enr[l - NEN];
            for (int i = 0; i < 3; ++i) cx[m][i] = (m < ccn && i < D) ? p.coords[int64_t(node) * D + i] : 0.0;
        }
        const double* Dm = p.mats[rec.child_mat[kc]].D;
        auto nodeof = [&](int l) { return l < NEN ? p.conn[int64_t(e) * NEN + l] : rec.enr[l - NEN]; };
        for (int q = 0; q < PG_CHILD_NQP(ct); ++q) {
            const double Jc = child_qp<D, POLY>(ct, q, ccn, cx, cg, w);
            auto grad = [&](int l, int a) {
                if (l < NEN) return pg[l][a];
                double g = 0.0;
                for (int m = 0; m < CN; ++m)
                    if (cl[m] == l) g = cg[m][a];
                return g;
            };
            double eps[NS], sig[NS];
            small_strain_fast<D>(nloc, grad, [&](int l, int i) { return Un[int64_t(nodeof(l)) * D + i]; }, eps);
            lei_stress<D>(Dm, eps, sig);
            const double wgt = Jc * w;
            vol += wgt * 1.0;
            for (int k = 0; k < NS; ++k) ie[k] += wgt * eps[k];
            for (int k = 0; k < NS; ++k) is[k] += wgt * sig[k];
        }
    }
    double* o = out + int64_t(c) * S;
    o[0] = vol;
    for (int k = 0; k < NS; ++k) o[1 + k] = ie[k];
    for (int k = 0; k < NS; ++k) o[1 + NS + k] = is[k];
}

// --------------------------------------------------------------------------
// Node-group kernels of the class-binned schedule.  A block owns NG nodes of a
// group list (rows of the node-pair pattern).  Phase A takes one lane per
// (node a, incident element e) incidence: for an uncut element it computes the
// simplex gradients from FP64 coordinate differences, the FP64 internal force
// w B_a^T sigma (w = J * sum of the reference qp weights) and the element's
// d x d blocks (a, m) with the isotropic closed form of B^T D B in TC; for a
// cut element it reads the row of a in the staged tangent and residual.  Each
// contribution lands at its static (slot, rank) -- rank = ascending element
// order among a's incidences that contain b -- so phase B, one lane per node
// pair (a, b), sums it in the reference's element order, in FP64, and writes
// the block once through the scatter indices.  A group whose incidences or
// pairs overflow the LDS tables takes the same arithmetic from global memory.
//
// The kernels are bound by the latency of their dependent loads, so each phase
// issues all of its loads before its first wait: the set-up (node table,
// incidence records, slot words, materials), then the pair metadata with the
// element data (coordinates, U and stress coefficients of an uncut element,
// listed owner first so that no lane selects its row at run time; the whole
// staged row of a cut element), then the stores.
// --------------------------------------------------------------------------
template <int D> struct GroupCfg;
// per-group LDS capacity (incidences, pairs) for NG nodes: 7 incidences and 16
// pairs per node in 2D (structured tri3: at most 8 and 6 on average per regular
// node, 2 per enrichment node), 24 and 36 in 3D (6-tet cubes: at most 24 and
// 27 + enrichment).  A group beyond these takes the global-memory path.
template <int D, int NG> struct GroupLim {
    static constexpr int MAXI = (D == 2 ? 7 : 24) * NG;
    static constexpr int MAXP = (D == 2 ? 16 : 36) * NG;
};

// Closed-form gradients of a linear simplex (adjugate of the affine Jacobian
// over its determinant) from FP64 coordinate differences; returns J.
template <int D, typename T>
__device__ inline T simplex_grads(const double (*x)[3], T (*g)[D]) {
    T d[D][D];  // d[k][i] = x_{k+1}[i] - x_0[i]  (row k: derivative along r_k)
    for (int k = 0; k < D; ++k)
        for (int i = 0; i < D; ++i) d[k][i] = T(x[k + 1][i] - x[0][i]);
    if constexpr (D == 2) {
        const T J = d[0][0] * d[1][1] - d[0][1] * d[1][0];
        const T rJ = T(1) / J;
        g[1][0] = d[1][1] * rJ; g[1][1] = -d[1][0] * rJ;
        g[2][0] = -d[0][1] * rJ; g[2][1] = d[0][0] * rJ;
        g[0][0] = -(g[1][0] + g[2][0]); g[0][1] = -(g[1][1] + g[2][1]);
        return J;
    } else {
        T c[3][3];
        c[0][0] = d[1][1] * d[2][2] - d[1][2] * d[2][1];
        c[0][1] = d[1][2] * d[2][0] - d[1][0] * d[2][2];
        c[0][2] = d[1][0] * d[2][1] - d[1][1] * d[2][0];
        c[1][0] = d[0][2] * d[2][1] - d[0][1] * d[2][2];
        c[1][1] = d[0][0] * d[2][2] - d[0][2] * d[2][0];
        c[1][2] = d[0][1] * d[2][0] - d[0][0] * d[2][1];
        c[2][0] = d[0][1] * d[1][2] - d[0][2] * d[1][1];
        c[2][1] = d[0][2] * d[1][0] - d[0][0] * d[1][2];
        c[2][2] = d[0][0] * d[1][1] - d[0][1] * d[1][0];
        const T J = d[0][0] * c[0][0] + d[0][1] * c[0][1] + d[0][2] * c[0][2];
        const T rJ = T(1) / J;
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < 3; ++i) g[k + 1][i] = c[k][i] * rJ;
        for (int i = 0; i < 3; ++i) g[0][i] = -(g[1][i] + g[2][i] + g[3][i]);
        return J;
    }
}

// Row la of a NEN x D register array (a select chain: la varies per lane).
template <int NEN, int D, typename T>
__device__ inline void sel_row(const T (*a)[D], int la, T* out) {
#pragma unroll
    for (int c = 0; c < D; ++c) {
        T v = a[NEN - 1][c];
#pragma unroll
        for (int n = NEN - 2; n >= 0; --n) v = la == n ? a[n][c] : v;
        out[c] = v;
    }
}

// FP64 internal-force contribution w B_la^T sigma of an uncut element, from
// the element's nodal U and the LEI coefficients of its material (LeiCoef).
// The sums run over the nonzero entries of the Voigt B only, in the order of
// the dense products (B U by node and component, then B^T sigma by row): the
// skipped terms are exact zeros.
template <int D>
__device__ inline void plain_p(const double (*g)[D], const double (*u)[D], int la, double w, const double* lc,
                               double* out) {
    constexpr int NEN = D + 1, NS = (D == 2 ? 3 : 6);
    double e[NS], sig[NS], gl[D];
    if constexpr (D == 2) {
        // B column (n, x): rows xx (g_x), xy (g_y); column (n, y): rows yy (g_y), xy (g_x)
        e[0] = g[0][0] * u[0][0];
        e[2] = g[0][1] * u[0][0];
        e[1] = g[0][1] * u[0][1];
        e[2] = e[2] + g[0][0] * u[0][1];
        for (int n = 1; n < NEN; ++n) {
            e[0] = e[0] + g[n][0] * u[n][0];
            e[2] = e[2] + g[n][1] * u[n][0];
            e[1] = e[1] + g[n][1] * u[n][1];
            e[2] = e[2] + g[n][0] * u[n][1];
        }
    } else {
        // columns (n, x): rows xx, xz, xy; (n, y): yy, yz, xy; (n, z): zz, yz, xz
        e[0] = g[0][0] * u[0][0];
        e[4] = g[0][2] * u[0][0];
        e[5] = g[0][1] * u[0][0];
        e[1] = g[0][1] * u[0][1];
        e[3] = g[0][2] * u[0][1];
        e[5] = e[5] + g[0][0] * u[0][1];
        e[2] = g[0][2] * u[0][2];
        e[3] = e[3] + g[0][1] * u[0][2];
        e[4] = e[4] + g[0][0] * u[0][2];
        for (int n = 1; n < NEN; ++n) {
            e[0] = e[0] + g[n][0] * u[n][0];
            e[4] = e[4] + g[n][2] * u[n][0];
            e[5] = e[5] + g[n][1] * u[n][0];
            e[1] = e[1] + g[n][1] * u[n][1];
            e[3] = e[3] + g[n][2] * u[n][1];
            e[5] = e[5] + g[n][0] * u[n][1];
            e[2] = e[2] + g[n][2] * u[n][2];
            e[3] = e[3] + g[n][1] * u[n][2];
            e[4] = e[4] + g[n][0] * u[n][2];
        }
    }
    lei_stress_c<D>(lc, e, sig);
    sel_row<NEN, D, double>(g, la, gl);
    if constexpr (D == 2) {
        out[0] = w * (gl[0] * sig[0] + gl[1] * sig[2]);
        out[1] = w * (gl[1] * sig[1] + gl[0] * sig[2]);
    } else {
        out[0] = w * (gl[0] * sig[0] + gl[2] * sig[4] + gl[1] * sig[5]);
        out[1] = w * (gl[1] * sig[1] + gl[2] * sig[3] + gl[0] * sig[5]);
        out[2] = w * (gl[2] * sig[2] + gl[1] * sig[3] + gl[0] * sig[4]);
    }
}

// Stages lambda, mu of every material in LDS (zero past the last one).  The
// loads use a clamped index and a select, so that they go out together with
// the other set-up loads.
template <typename TC, int BS>
__device__ inline void stage_lm(const DevProb& p, int ns, TC (*lm)[2]) {
    for (int k = threadIdx.x; k < MAX_LDS_MATS * 2; k += BS) {
        const int m = k >> 1, j = k & 1;
        const double v0 = p.mats[m < p.n_mat ? m : p.n_mat - 1].D[j == 0 ? 1 : ns * ns - 1];
        lm[m][j] = m < p.n_mat ? TC(v0) : TC(0);
    }
}

// Threads per node group: TV = 0 gives about one lane per node pair (8 per
// node in 2D, 16 in 3D), TV = 1 twice that, so that a group's pairs fit in one
// pass even where nodes have more pairs; whole wavefronts, at most 512.
template <int D, int NG, int TV> struct GroupThreads {
    static constexpr int per = (D == 2 ? 8 : 16) * (TV ? 2 : 1) * NG;
    static constexpr int value = per < 64 ? 64 : (per > 512 ? 512 : (per + 63) / 64 * 64);
};

// Node-pair metadata of a group's local pair kq: owner slot lo, row node a,
// global pair index q, column node b, the row's dofs and (2D) the scatter
// indices of the d x d block.  It depends only on the group's node table, so
// the 2D kernels load it before phase A and its latency overlaps the incidence
// loads (the 3D list kernel loads it in phase B: at 128 VGPRs, holding it
// through phase A spills).
template <int D> struct PairMeta {
    int lo, a, q, b;
    int dof[D];
    int4 map;  // 2D scatter indices
};

template <int D, int NG>
__device__ inline PairMeta<D> pair_meta(const DevProb& p, const int* s_qptr, const int* s_qbeg, const int* s_node,
                                        int ng, int kq, bool jac) {
    PairMeta<D> m;
    int lo = 0, hi = ng;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_qptr[mid] <= kq) lo = mid; else hi = mid;
    }
    m.lo = lo;
    m.a = s_node[lo];
    m.q = s_qbeg[lo] + (kq - s_qptr[lo]);
    m.b = p.npat_col[m.q];
    for (int i = 0; i < D; ++i) m.dof[i] = p.node_dof[int64_t(m.a) * D + i];
    if constexpr (D == 2) {
        if (jac) m.map = reinterpret_cast<const int4*>(p.npat_map)[m.q];
        else m.map = make_int4(0, 0, 0, 0);
    } else {
        m.map = make_int4(0, 0, 0, 0);
    }
    return m;
}

// Diagonal block (la, la) of an uncut element.  With FP32 tangents it is minus
// the FP64 sum of the row's off-diagonal FP32 blocks: the gradients of a
// simplex sum to zero, so the exact block equals that sum, and taking it this
// way makes every element row sum to zero to FP64 rounding -- the assembled
// tangent keeps annihilating rigid translations, which FP32 rounding of each
// block on its own would spoil (loosely bonded fibres turn that spurious
// stiffness into slow Newton convergence).  FP64 tangents use the closed form.
// The off-diagonal blocks of row la come in ascending order of the other
// local node m = k < la ? k : k + 1 (k = 0 .. NEN - 2), which is also the
// order of the slot fields; diag_acc folds block k into the diagonal.
template <int D, typename TC>
__device__ inline void diag_acc(int k, const TC (*blk)[D], double (*out)[D]) {
    if constexpr (sizeof(TC) == 4)
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) out[i][j] = k == 0 ? -double(blk[i][j]) : out[i][j] - double(blk[i][j]);
}
template <int D, typename TC>
__device__ inline void diag_fin(const TC* ga, TC wl, TC wm, double (*out)[D]) {
    if constexpr (sizeof(TC) != 4) {
        TC blk[D][D];
        iso_block<D, TC>(ga, ga, wl, wm, blk);
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) out[i][j] = double(blk[i][j]);
    }
}
template <int D, typename TC>
__device__ inline void diag_block(const TC (*gt)[D], int la, TC wl, TC wm, double (*out)[D]) {
    constexpr int NEN = D + 1;
    TC ga[D];
    sel_row<NEN, D, TC>(gt, la, ga);
    if constexpr (sizeof(TC) == 4) {
        for (int k = 0; k < NEN - 1; ++k) {
            TC gb[D], blk[D][D];
            for (int c = 0; c < D; ++c) gb[c] = k < la ? gt[k][c] : gt[k + 1][c];
            iso_block<D, TC>(ga, gb, wl, wm, blk);
            diag_acc<D, TC>(k, blk, out);
        }
    } else {
        diag_fin<D, TC>(ga, wl, wm, out);
    }
}

// Diagonal block (la, la) of a staged cut-element tangent, from the element's
// row la (Ke points at block (la, 0)).  With FP32 tangents and la a parent
// node it is minus the FP64 sum of the row's other parent-node blocks: the
// parent gradients sum to zero at every child qp and the cohesive fold couples
// only enrichment nodes, so the exact block equals that sum, and taking it this
// way keeps rigid translations (parent dofs equal, enrichment dofs zero) in the
// null space of the assembled tangent, as for uncut elements.  Otherwise the
// staged block itself.
template <int D, typename TK>
__device__ inline void cut_diag(const TK* Ke, int la, double* out) {
    constexpr int NEN = D + 1;
    if constexpr (sizeof(TK) == 4) {
        if (la < NEN) {
            for (int k = 0; k < D * D; ++k) out[k] = 0.0;
            for (int m = 0; m < NEN; ++m) {
                if (m == la) continue;
                for (int k = 0; k < D * D; ++k) out[k] -= double(Ke[m * D * D + k]);
            }
            return;
        }
    }
    for (int k = 0; k < D * D; ++k) out[k] = double(Ke[la * D * D + k]);
}

// cut_diag from the row's blocks in registers (kb: blocks 0 .. M - 1 of the
// row, kd: block la), with the same arithmetic.
template <int D, typename TK, int M>
__device__ inline void cut_diag_rows(const TK (*kb)[D * D], const TK* kd, int la, double* out) {
    constexpr int NEN = D + 1;
    if constexpr (sizeof(TK) == 4) {
        if (la < NEN) {
            for (int k = 0; k < D * D; ++k) out[k] = 0.0;
#pragma unroll
            for (int m = 0; m < NEN; ++m) {
                if (m == la) continue;
                for (int k = 0; k < D * D; ++k) out[k] -= double(kb[m][k]);
            }
            return;
        }
    }
    for (int k = 0; k < D * D; ++k) out[k] = double(kd[k]);
}

// One node pair (a, b) from global memory: the incidences [t0, t1) of a in
// element order, with the arithmetic of the LDS-staged paths (used by groups
// whose incidences or pairs exceed the LDS tables).
template <int D, typename TC, typename TKC>
__device__ inline void pair_from_global(const DevProb& p, const double* __restrict__ Un, const TKC* __restrict__ Kc,
                                        const double* __restrict__ Pc, const TC (*sLM)[2], double wsum,
                                        int t0, int t1, int a, int b, bool jac, bool diag, double acc[D][D],
                                        double pa[D]) {
    constexpr int NEN = D + 1;
    for (int t = t0; t < t1; ++t) {
        const int e = p.n2e[t];
        const int ci = p.cut_of_elem[e];
        const int la = find_local(p, e, ci, a, NEN);
        const int lb = (a == b) ? la : find_local(p, e, ci, b, NEN);
        if (la < 0 || lb < 0) continue;
        if (ci < 0) {
            // the element owner first, as the node-group kernels list it
            const int lbr = NEN == 3 ? (lb - la + 3) % 3 : (lb ^ la);  // b's position in that order
            int nodes[NEN];
            double x[NEN][3], u[NEN][D];
            for (int n = 0; n < NEN; ++n) {
                nodes[n] = p.conn[int64_t(e) * NEN + owner_first(NEN, la, n)];
                for (int c = 0; c < D; ++c) {
                    x[n][c] = p.coords[int64_t(nodes[n]) * D + c];
                    u[n][c] = Un[int64_t(nodes[n]) * D + c];
                }
            }
            const int mi = p.elem_mat[e];
            double g[NEN][D];
            const double J = simplex_grads<D, double>(x, g);
            const double w = J * wsum;
            if (jac) {
                TC gt[NEN][D], blk[D][D];
                for (int n = 0; n < NEN; ++n)
                    for (int c = 0; c < D; ++c) gt[n][c] = TC(g[n][c]);
                const TC wl = TC(w) * sLM[mi][0], wm = TC(w) * sLM[mi][1];
                if (la == lb) {
                    double dd[D][D];
                    diag_block<D, TC>(gt, 0, wl, wm, dd);
                    for (int i = 0; i < D; ++i)
                        for (int j = 0; j < D; ++j) acc[i][j] += dd[i][j];
                } else {
                    iso_block<D, TC>(gt[0], gt[lbr], wl, wm, blk);
                    for (int i = 0; i < D; ++i)
                        for (int j = 0; j < D; ++j) acc[i][j] += double(blk[i][j]);
                }
            }
            if (diag) {
                double pp[D], lc[LeiCoef<D>::N];
                for (int k = 0; k < LeiCoef<D>::N; ++k) lc[k] = lei_coef<D>(p, mi, k);
                plain_p<D>(g, u, 0, w, lc, pp);
                for (int i = 0; i < D; ++i) pa[i] += pp[i];
            }
        } else {
            const int nl = NEN + p.cut[ci].n_enr;
            if (jac) {
                const TKC* Ke = Kc + p.cut_koff[ci] + int64_t(la * nl) * D * D;  // row la
                double blk[D * D];
                if (la == lb) {
                    cut_diag<D, TKC>(Ke, la, blk);
                } else {
                    for (int k = 0; k < D * D; ++k) blk[k] = double(Ke[lb * D * D + k]);
                }
                for (int i = 0; i < D; ++i)
                    for (int j = 0; j < D; ++j) acc[i][j] += blk[i * D + j];
            }
            if (diag)
                for (int i = 0; i < D; ++i) pa[i] += Pc[p.cut_poff[ci] + la * D + i];
        }
    }
}

// Node-group kernel (tri3 meshes).
template <int D, typename TC, int NG, int TV, int BS = GroupThreads<D, NG, TV>::value>
__global__ __launch_bounds__(BS) void k_pair_group(DevProb p, const double* __restrict__ Un,
                                                    const TC* __restrict__ Kc, const double* __restrict__ Pc,
                                                    double* __restrict__ K, double* __restrict__ P, unsigned flags,
                                                    int gbase, int64_t ibase) {
    constexpr int NEN = D + 1, NS = (D == 2 ? 3 : 6), NQ = (D == 2 ? 4 : 5);
    constexpr int MAXI = GroupLim<D, NG>::MAXI, MAXP = GroupLim<D, NG>::MAXP;
    constexpr int MAXLOC = (D == 2 ? 7 : 8);  // local nodes per incidence
    // contributions per off-diagonal pair: the elements sharing the edge (a, b),
    // 2 in a tri3 mesh (cut elements included: an enrichment node lies on one
    // cut edge) and up to 6 around a tet4 edge; the diagonal pair of a node takes
    // one contribution per incidence
    constexpr int CAPO = (D == 2 ? 2 : 8);
    constexpr int CAPD = (D == 2 ? 8 : 24);
    __shared__ int s_node[NG];               // the group's nodes (from the node order)
    __shared__ int s_ibeg[NG], s_qbeg[NG];   // their first incidence and first node pair
    __shared__ int s_iptr[NG + 1];           // local prefix of incidences per node
    __shared__ int s_qptr[NG + 1];           // local prefix of node pairs per node
    __shared__ TC s_blk[MAXP][CAPO][D * D];   // off-diagonal d x d blocks by pair slot and rank (0 when absent)
    __shared__ double s_dblk[NG][CAPD][D * D];  // diagonal blocks by node and incidence
    __shared__ double s_dP[NG][CAPD][D];      // internal force by node and incidence
    __shared__ TC sLM[MAX_LDS_MATS][2];       // lambda, mu of each LEI material
    __shared__ int s_bad;

    const bool func = (flags & PG_FUNC) != 0, jac = (flags & PG_JAC) != 0;
    const int tid = threadIdx.x;
    const int grp = xcd_order(blockIdx.x, gridDim.x);
    const int4* gt = reinterpret_cast<const int4*>(p.gtab) + gbase + int64_t(grp) * NG;
    constexpr int ng = NG;
    // every set-up load goes out before the first wait: the group's node table
    // (NG entries, every lane loads one), its incidence records and slot words
    // (their block does not depend on the node table; lanes past the block
    // reload its last slot) and lambda, mu of the materials
    const int4 r = gt[tid & (NG - 1)];
    const int64_t ib = ibase + int64_t(grp) * MAXI;
    const int is = tid < MAXI ? tid : MAXI - 1;
    const int4 rec0 = reinterpret_cast<const int4*>(p.incg)[ib + is];
    const uint64_t sw0 = p.slotg[ib + is];
    stage_lm<TC, BS>(p, NS, sLM);
    // the group's nodes and their incidence / pair counts, prefix-summed across
    // the wavefront (NG <= 32: the first wavefront holds the group)
    int ci = tid < ng ? (r.w & 0xffff) : 0, cq = tid < ng ? (r.w >> 16) : 0;
#pragma unroll
    for (int off = 1; off < NG; off <<= 1) {
        const int ti = __shfl_up(ci, off), tq = __shfl_up(cq, off);
        if ((tid & 63) >= off) { ci += ti; cq += tq; }
    }
    if (tid < ng) {
        s_node[tid] = r.x;
        s_ibeg[tid] = r.y;
        s_qbeg[tid] = r.z;
        s_iptr[tid + 1] = ci;
        s_qptr[tid + 1] = cq;
    }
    if (tid == 0) {
        s_iptr[0] = 0;
        s_qptr[0] = 0;
        s_bad = 0;
    }
    __syncthreads();
    const int ni = s_iptr[ng], nq = s_qptr[ng];
    const bool staged = ni <= MAXI && nq <= MAXP && !(flags & PG_GLOBAL_PATH);
    double wsum = 0.0;
    for (int k = 0; k < NQ; ++k) wsum += (D == 2 ? TRI_QW[k] : TET_QW[k]);

    PairMeta<D> pm0;
    if (int(threadIdx.x) < nq) pm0 = pair_meta<D, NG>(p, s_qptr, s_qbeg, s_node, ng, threadIdx.x, jac);

    // ---- phase A: every incidence computes its element's d x d blocks of the
    // owner's row and stores them at their (slot, rank) -- ascending element
    // order is the rank order, so phase B sums them in the reference order ----
    if (staged) {
        for (int k = threadIdx.x; k < nq * CAPO * D * D; k += blockDim.x)
            s_blk[k / (CAPO * D * D)][(k / (D * D)) % CAPO][k % (D * D)] = TC(0);
        __syncthreads();
        for (int i = threadIdx.x; i < ni; i += blockDim.x) {
            // staged: ni <= MAXI, so slot i lies in the group's block
            const int4 rec = i == int(threadIdx.x) ? rec0 : reinterpret_cast<const int4*>(p.incg)[ib + i];
            const uint64_t sw = i == int(threadIdx.x) ? sw0 : p.slotg[ib + i];
            const int tag = rec.w;
            const int lo = tag < 0 ? (tag >> 3) & 31 : (rec.z >> 16) & 31;  // owner's position in the group
            const int rd = tag < 0 ? (tag >> 8) & 31 : (rec.z >> 21) & 31;  // rank among the owner's incidences
            if (rd >= CAPD || (sw >> 63)) { s_bad = 1; continue; }
            const int ps0 = s_qptr[lo], np = s_qptr[lo + 1] - s_qptr[lo];
            if (tag < 0) {
                // the element's nodes, owner first (the record lists them so)
                const int ln[NEN] = {rec.x, rec.y, rec.z};
                double x[NEN][3], u[NEN][D];  // U goes out with the coordinates
                for (int n = 0; n < NEN; ++n)
                    for (int c = 0; c < D; ++c) {
                        x[n][c] = p.coords[int64_t(ln[n]) * D + c];
                        u[n][c] = Un[int64_t(ln[n]) * D + c];
                    }
                const int mi = tag & 7;
                double lc[LeiCoef<D>::N];
                for (int k = 0; k < LeiCoef<D>::N; ++k) lc[k] = lei_coef<D>(p, mi, k);
                // every load of the incidence is out before the arithmetic
                __builtin_amdgcn_sched_barrier(0);
                double g[NEN][D];
                const double J = simplex_grads<D, double>(x, g);
                const double w = J * wsum;
                // (also without PG_FUNC, where phase B ignores it: the U loads
                // then stay ahead of the gradient arithmetic)
                plain_p<D>(g, u, 0, w, lc, s_dP[lo][rd]);
                if (!jac) continue;
                TC gt[NEN][D];
                for (int n = 0; n < NEN; ++n)
                    for (int c = 0; c < D; ++c) gt[n][c] = TC(g[n][c]);
                const TC wl = TC(w) * sLM[mi][0], wm = TC(w) * sLM[mi][1];
                // each off-diagonal block (owner, n) once: to its slot, and
                // (FP32) into the row-sum diagonal
                double dd[D][D];
#pragma unroll
                for (int k = 0; k < NEN - 1; ++k) {
                    TC blk[D][D];
                    iso_block<D, TC>(gt[0], gt[k + 1], wl, wm, blk);
                    diag_acc<D, TC>(k, blk, dd);
                    const int f = int((sw >> (9 * k)) & 511u);
                    const int slot = f & 63, rank = f >> 6;
                    if (slot >= np || rank >= CAPO) { s_bad = 1; continue; }
                    for (int i = 0; i < D; ++i)
                        for (int j = 0; j < D; ++j) s_blk[ps0 + slot][rank][i * D + j] = blk[i][j];
                }
                diag_fin<D, TC>(gt[0], wl, wm, dd);
                for (int i = 0; i < D; ++i)
                    for (int j = 0; j < D; ++j) s_dblk[lo][rd][i * D + j] = dd[i][j];
            } else {
                // cut element: the record holds the staged row of a directly.
                // The row's MAXLOC blocks (the staged rows are padded), the
                // owner's diagonal block and its residual go out at once.
                const int la = rec.z & 255, nloc = (rec.z >> 8) & 255;
                const TC* Ke = Kc + int64_t(uint32_t(rec.x)) * (D * D);  // blocks (la, 0 .. nloc - 1)
                TC kb[MAXLOC][D * D], kd[D * D];
                double pc[D];
                for (int m = 0; m < MAXLOC; ++m)
                    for (int kk = 0; kk < D * D; ++kk) kb[m][kk] = Ke[m * D * D + kk];
                const int lc = la < MAXLOC ? la : 0;
                for (int kk = 0; kk < D * D; ++kk) kd[kk] = Ke[lc * D * D + kk];
                for (int c = 0; c < D; ++c) pc[c] = Pc[int64_t(rec.y) + c];
                __builtin_amdgcn_sched_barrier(0);
                if (nloc > MAXLOC || la >= nloc) { s_bad = 1; continue; }  // does not fit: global path
                for (int c = 0; c < D; ++c) s_dP[lo][rd][c] = pc[c];
                if (!jac) continue;
                cut_diag_rows<D, TC, MAXLOC>(kb, kd, la, s_dblk[lo][rd]);
#pragma unroll
                for (int m = 0; m < MAXLOC; ++m) {
                    if (m >= nloc || m == la) continue;
                    const int k = m < la ? m : m - 1;  // slot field of m
                    const int f = int((sw >> (9 * k)) & 511u);
                    const int slot = f & 63, rank = f >> 6;
                    if (slot >= np || rank >= CAPO) { s_bad = 1; continue; }
                    for (int kk = 0; kk < D * D; ++kk) s_blk[ps0 + slot][rank][kk] = kb[m][kk];
                }
            }
        }
    }
    __syncthreads();
    const bool use_lists = staged && !s_bad;

    // ---- phase B: node pairs ----
    for (int kq = threadIdx.x; kq < nq; kq += blockDim.x) {
        const PairMeta<D> pm = kq == int(threadIdx.x) ? pm0 : pair_meta<D, NG>(p, s_qptr, s_qbeg, s_node, ng, kq, jac);
        const int lo = pm.lo, a = pm.a, q = pm.q, b = pm.b;
        const bool diag = (a == b) && func;
        double acc[D][D];
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) acc[i][j] = 0.0;
        double pa[D];
        for (int i = 0; i < D; ++i) pa[i] = 0.0;
        if (use_lists && a == b) {
            // diagonal pair: every incidence of a, in ascending element order
            const int nd = s_iptr[lo + 1] - s_iptr[lo];
            for (int r = 0; r < nd; ++r) {
                if (func)
                    for (int i = 0; i < D; ++i) pa[i] += s_dP[lo][r][i];
                if (jac)
                    for (int i = 0; i < D; ++i)
                        for (int j = 0; j < D; ++j) acc[i][j] += s_dblk[lo][r][i * D + j];
            }
        } else if (use_lists) {
            // off-diagonal pair: its contributions by rank (absent ranks hold 0)
            if (jac) {
                const int s = kq;
#pragma unroll
                for (int r = 0; r < CAPO; ++r)
                    for (int i = 0; i < D; ++i)
                        for (int j = 0; j < D; ++j) acc[i][j] += double(s_blk[s][r][i * D + j]);
            }
        } else {
            pair_from_global<D, TC, TC>(p, Un, Kc, Pc, sLM, wsum, s_ibeg[lo], s_ibeg[lo] + (s_iptr[lo + 1] - s_iptr[lo]),
                                    a, b, jac, diag, acc, pa);
        }
        if (jac) {
            if (pm.map.y == pm.map.x + 1 && pm.map.w == pm.map.z + 1) {
                // both columns of b in one block, side by side: one 16-byte
                // store per row (8-byte aligned)
                *reinterpret_cast<dpair*>(K + pm.map.x) = dpair{acc[0][0], acc[0][1]};
                *reinterpret_cast<dpair*>(K + pm.map.z) = dpair{acc[1][0], acc[1][1]};
            } else {
                K[pm.map.x] = acc[0][0];
                K[pm.map.y] = acc[0][1];
                K[pm.map.z] = acc[1][0];
                K[pm.map.w] = acc[1][1];
            }
        }
        if (diag)
            for (int i = 0; i < D; ++i) P[pm.dof[i]] = pa[i];
    }
}

// Structured tri3 grid of the reference's generate_mesh (src/Mesh.cpp:418-441):
// cell (cx, cy) of an n x n grid (node (x, y) = x + y (n + 1), elements
// 2 (cx + cy n) and 2 (cx + cy n) + 1) is split along (x, y)-(x+1, y+1) when
// cx + cy is even and along (x+1, y)-(x, y+1) otherwise.  An interior node
// with x + y even has 8 incident elements and 9 node pairs, one with x + y odd
// 4 and 5.  GridTpl holds, for both parities, what the node-group kernel
// otherwise reads from its incidence records, slot words and pair tables: per
// incidence in ascending element id (the reference's loop order) its cell
// offset and half, its element's nodes owner first as grid offsets, and the
// (slot, rank) fields of the two other nodes; per node pair in ascending
// column node id the column's grid offset.  pg_asm_create accepts a node for
// this path only after checking its incidences, element node order, pattern
// columns, scatter indices and dofs against these formulas.
struct GridTpl {
    int8_t pdx[2][9], pdy[2][9];           // column node of slot s
    int8_t ecx[2][8], ecy[2][8], eh[2][8];  // cell offset and half of incidence r
    int8_t ndx[2][8][3], ndy[2][8][3];      // the element's nodes, owner first
    int16_t sr[2][8][2];                    // slot | rank << 6 of owner-first nodes 1 and 2
};
static_assert(sizeof(GridTpl) % 4 == 0, "the template is copied to LDS as words");
constexpr int GRID_NI[2] = {8, 4}, GRID_NP[2] = {9, 5}, GRID_DIAG[2] = {4, 2};
// the same by arithmetic for the kernels: a parity-indexed read of the arrays
// is a global memory load, whose wait also waits for every store before it
__device__ inline int grid_ni(int pr) { return 8 >> pr; }
__device__ inline int grid_np(int pr) { return 9 - 4 * pr; }
__device__ inline int grid_diag(int pr) { return 4 >> pr; }

// Node-group kernel for the nodes of the structured grid whose incidences are
// all uncut (tri3).  Each node comes with one record -- node, index of its first
// K value (row 0, first column; row 1 follows after the 2 np values of row 0),
// first dof, element id of its cell (x, y) -- and a word with the grid's row
// stride S = n + 1 and the parity of x + y.  Everything else is arithmetic, so
// the loads are the records, then the element data; the arithmetic, the
// summation order and the stores are those of k_pair_group, bit for bit.
template <typename TC, int NG, int TV, int BS = GroupThreads<2, NG, TV>::value>
__global__ __launch_bounds__(BS) void k_pair_struct(DevProb p, const int32_t* __restrict__ tpl,
                                                     const int4* __restrict__ srec, const int32_t* __restrict__ sinf,
                                                     const double* __restrict__ Un, double* __restrict__ K,
                                                     double* __restrict__ P, unsigned flags, int ngroups, int per) {
    constexpr int D = 2, NEN = 3, NQ = 4;
    constexpr int MAXP = 9 * NG;
    constexpr int CAPO = 2, CAPD = 8;
    static_assert(NG <= 32, "the first wavefront holds the group's nodes");
    __shared__ int s_node[NG], s_kb[NG], s_dof[NG], s_eb[NG], s_inf[NG];
    __shared__ int s_iptr[NG + 1], s_qptr[NG + 1];
    __shared__ TC s_blk[MAXP][CAPO][D * D];      // off-diagonal blocks by pair and rank
    __shared__ double s_dblk[NG][CAPD][D * D];   // diagonal blocks by node and incidence
    __shared__ double s_dP[NG][CAPD][D];         // internal force by node and incidence
    __shared__ TC sLM[MAX_LDS_MATS][2];          // lambda, mu of each LEI material
    __shared__ double sLC[MAX_LDS_MATS][LeiCoef<D>::N];  // LEI stress coefficients
    __shared__ GridTpl sT;

    const bool func = (flags & PG_FUNC) != 0, jac = (flags & PG_JAC) != 0;
    const int tid = threadIdx.x;
    // each block takes `per` consecutive groups (XCD-aware block order)
    const int g0 = xcd_order(blockIdx.x, gridDim.x) * per, g1 = min(ngroups, g0 + per);
    // set-up loads, all independent: the first group's records, the template
    // and the materials
    int4 r = make_int4(-1, 0, 0, 0);
    int inf = 0;
    if (g0 < g1) {
        const int64_t k0 = int64_t(g0) * NG + (tid & (NG - 1));
        r = srec[k0];
        inf = sinf[k0];
    }
    constexpr int TW = int(sizeof(GridTpl) / 4);
    const int tw = tpl[tid < TW ? tid : TW - 1];
    stage_lm<TC, BS>(p, 3, sLM);
    for (int k = tid; k < MAX_LDS_MATS * LeiCoef<D>::N; k += BS) {
        const int m = k / LeiCoef<D>::N, j = k % LeiCoef<D>::N;
        const double v = lei_coef<D>(p, m < p.n_mat ? m : p.n_mat - 1, j);
        sLC[m][j] = m < p.n_mat ? v : 0.0;
    }
    if (tid < TW) reinterpret_cast<int32_t*>(&sT)[tid] = tw;
    double wsum = 0.0;
    for (int k = 0; k < NQ; ++k) wsum += TRI_QW[k];
    for (int g = g0; g < g1; ++g) {
        const bool live = tid < NG && r.x >= 0;
        const int par = (inf >> 24) & 1;
        int ci = live ? grid_ni(par) : 0, cq = live ? grid_np(par) : 0;
#pragma unroll
        for (int off = 1; off < NG; off <<= 1) {
            const int ti = __shfl_up(ci, off), tq = __shfl_up(cq, off);
            if ((tid & 63) >= off) { ci += ti; cq += tq; }
        }
        if (tid < NG) {
            s_node[tid] = r.x;
            s_kb[tid] = r.y;
            s_dof[tid] = r.z;
            s_eb[tid] = r.w;
            s_inf[tid] = inf;
            s_iptr[tid + 1] = ci;
            s_qptr[tid + 1] = cq;
        }
        if (tid == 0) {
            s_iptr[0] = 0;
            s_qptr[0] = 0;
        }
        __syncthreads();
        // the next group's records go out now and land during this group's work
        if (g + 1 < g1) {
            const int64_t k1 = int64_t(g + 1) * NG + (tid & (NG - 1));
            r = srec[k1];
            inf = sinf[k1];
        }
        const int ni = s_iptr[NG], nq = s_qptr[NG];

        // ---- phase A: incidences (as k_pair_group, records from the template) ----
        for (int i = tid; i < ni; i += BS) {
            int lo = 0, hi = NG;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_iptr[mid] <= i) lo = mid; else hi = mid;
            }
            const int rr = i - s_iptr[lo];
            const int a = s_node[lo], inf0 = s_inf[lo], pr = (inf0 >> 24) & 1, S = inf0 & 0xffffff;
            const int e = s_eb[lo] + 2 * (sT.ecx[pr][rr] + sT.ecy[pr][rr] * (S - 1)) + sT.eh[pr][rr];
            int ln[NEN];
            for (int n = 0; n < NEN; ++n) ln[n] = a + sT.ndx[pr][rr][n] + sT.ndy[pr][rr][n] * S;
            double x[NEN][3], u[NEN][D];  // U goes out with the coordinates
            for (int n = 0; n < NEN; ++n)
                for (int c = 0; c < D; ++c) {
                    x[n][c] = p.coords[int64_t(ln[n]) * D + c];
                    u[n][c] = Un[int64_t(ln[n]) * D + c];
                }
            const int mi = p.elem_mat[e] & (MAX_LDS_MATS - 1);
            // every load of the incidence is out before the arithmetic
            __builtin_amdgcn_sched_barrier(0);
            double lc[LeiCoef<D>::N];
            for (int k = 0; k < LeiCoef<D>::N; ++k) lc[k] = sLC[mi][k];
            double g[NEN][D];
            const double J = simplex_grads<D, double>(x, g);
            const double w = J * wsum;
            plain_p<D>(g, u, 0, w, lc, s_dP[lo][rr]);
            if (!jac) continue;
            TC gt[NEN][D];
            for (int n = 0; n < NEN; ++n)
                for (int c = 0; c < D; ++c) gt[n][c] = TC(g[n][c]);
            const TC wl = TC(w) * sLM[mi][0], wm = TC(w) * sLM[mi][1];
            const int ps0 = s_qptr[lo];
            double dd[D][D];
#pragma unroll
            for (int k = 0; k < NEN - 1; ++k) {
                TC blk[D][D];
                iso_block<D, TC>(gt[0], gt[k + 1], wl, wm, blk);
                diag_acc<D, TC>(k, blk, dd);
                const int f = sT.sr[pr][rr][k];
                const int slot = f & 63, rank = f >> 6;
                for (int ii = 0; ii < D; ++ii)
                    for (int jj = 0; jj < D; ++jj) s_blk[ps0 + slot][rank][ii * D + jj] = blk[ii][jj];
            }
            diag_fin<D, TC>(gt[0], wl, wm, dd);
            for (int ii = 0; ii < D; ++ii)
                for (int jj = 0; jj < D; ++jj) s_dblk[lo][rr][ii * D + jj] = dd[ii][jj];
        }
        __syncthreads();

        // ---- phase B: node pairs ----
        for (int kq = tid; kq < nq; kq += BS) {
            int lo = 0, hi = NG;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_qptr[mid] <= kq) lo = mid; else hi = mid;
            }
            const int slot = kq - s_qptr[lo];
            const int pr = (s_inf[lo] >> 24) & 1, np = grid_np(pr);
            const bool isdiag = slot == grid_diag(pr);
            double acc[D][D];
            for (int i = 0; i < D; ++i)
                for (int j = 0; j < D; ++j) acc[i][j] = 0.0;
            double pa[D] = {0.0, 0.0};
            if (isdiag) {
                const int nd = grid_ni(pr);
                for (int rr = 0; rr < nd; ++rr) {
                    if (func)
                        for (int i = 0; i < D; ++i) pa[i] += s_dP[lo][rr][i];
                    if (jac)
                        for (int i = 0; i < D; ++i)
                            for (int j = 0; j < D; ++j) acc[i][j] += s_dblk[lo][rr][i * D + j];
                }
            } else if (jac) {
#pragma unroll
                for (int rk = 0; rk < CAPO; ++rk)
                    for (int i = 0; i < D; ++i)
                        for (int j = 0; j < D; ++j) acc[i][j] += double(s_blk[kq][rk][i * D + j]);
            }
            if (jac) {
                const int64_t k0v = int64_t(s_kb[lo]) + 2 * slot;
                *reinterpret_cast<dpair*>(K + k0v) = dpair{acc[0][0], acc[0][1]};
                *reinterpret_cast<dpair*>(K + k0v + 2 * np) = dpair{acc[1][0], acc[1][1]};
            }
            if (isdiag && func)
                for (int i = 0; i < D; ++i) P[s_dof[lo] + i] = pa[i];
        }
        __syncthreads();  // the next group rewrites the tables
    }
}

// Structured tet4 grid of generate_mesh (src/Mesh.cpp:443-630): cube
// (cx, cy, cz) of the n^3 grid (node (x, y, z) = x + y S + z S^2, S = n + 1)
// holds elements 6 (cx + cy n + cz n^2) + t, t = 0 .. 5, the same six tets in
// every cube around its diagonal from corner (0, 0, 0) to (1, 1, 1).  Every
// interior node has 24 incident elements and 15 node pairs; an off-diagonal
// pair gets 4 or 6 contributions.  GridTpl3 holds what k_pair_lists reads from
// its records for such a node: per incidence in ascending element id the cube
// offset, the tet, the owner's local index and the element's nodes owner first
// (la ^ n) as grid offsets; per node pair in ascending column id the column's
// offset and its contributions by rank (incidence << 2 | the column's
// owner-first position).  pg_asm_create accepts a node only after checking its
// incidences, element nodes, pattern columns, scatter indices and dofs.
constexpr int G3_NI = 24, G3_NP = 15, G3_CAP = 6;
struct GridTpl3 {
    int8_t ec[G3_NI][4];        // cube offset dx, dy, dz and tet t of incidence r
    int8_t la[G3_NI];           // the owner's local index in the element
    int8_t nd[G3_NI][4][3];     // the element's nodes, owner first, as grid offsets
    int8_t pd[G3_NP][3];        // column node of slot s
    uint8_t pc[G3_NP][G3_CAP];  // contributions to slot s by rank (0xff: none)
    int8_t diag;                // slot of the diagonal pair
};
static_assert(sizeof(GridTpl3) % 4 == 0, "the template is copied to LDS as words");

// Element tangents of the structured tet4 grid from a table.  On a cube grid
// of spacing h every tet has gradients g = G / h with G integer and J = Jt h^3
// (Jt = +-1), so the isotropic block (owner, owner-first node lb) of
// incidence r, J w (lambda g_a,i g_b,j + mu g_a,j g_b,i + delta_ij mu g_a.g_b)
// (w = the sum of the rule's weights), is h Jt w (lambda G_a,i G_b,j + ...):
// it depends on the material, the incidence and h, not on U.  GridTab3 holds
// these blocks in FP32, computed in FP64 from the integer G and rounded once,
// for materials 0 .. G3_MT - 1.  pg_asm_create uses it only when every
// structured node's 26 neighbours lie at the grid offsets times one spacing h
// to 1e-10 h and every element of those nodes has a material in the table, so
// a table block equals the element's exact FP32 tangent to within FP32
// rounding.
constexpr int G3_MT = 2;
struct GridTab3 {
    float tk[G3_MT][G3_NI][3][9];  // block (owner, owner-first node 1 + k), row-major 3 x 3
};
static_assert(sizeof(GridTab3) % 16 == 0, "the table is copied to LDS as 16-byte words");

// Node-group kernel for the structured tet4 nodes whose incidences are all
// uncut.  Each node comes with one record -- node, index of its first K value
// (row 0, first column; rows 1 and 2 follow after 3 np values each), first
// dof, element id of its cube's first tet -- and the row stride S.  The
// arithmetic, the summation order and the values are those of k_pair_lists,
// bit for bit; the loads are the records, then the element data.
template <typename TC, int NG, int TV, bool TAB, int BS = GroupThreads<3, NG, TV>::value>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(TAB && !TV ? 5 : 4))) void k_pair_struct3(
    DevProb p, const int32_t* __restrict__ tpl, const int4* __restrict__ tab, const int4* __restrict__ srec,
    const int32_t* __restrict__ sinf, const double* __restrict__ Un, double* __restrict__ K, double* __restrict__ P,
    unsigned flags, int ngroups, int per) {
    static_assert(!TAB || sizeof(TC) == 4, "the table holds FP32 tangents");
    constexpr int D = 3, NEN = 4, NQ = 5;
    constexpr int MAXI = G3_NI * NG;
    static_assert(NG <= 32, "the first wavefront holds the group's nodes");
    __shared__ int s_node[NG], s_kb[NG], s_dof[NG], s_eb[NG], s_S[NG];
    constexpr int MX = TAB ? 1 : MAXI;      // the table form keeps no per-incidence blocks
    __shared__ TC s_g[MX][NEN][D];
    __shared__ TC s_wlm[MX][2];             // w lambda, w mu of each incidence
    __shared__ double s_P[MAXI][D];
    __shared__ double s_dblk[MX][D * D];     // diagonal block (la, la) of each incidence
    __shared__ uint8_t s_mi[TAB ? MAXI : 1]; // material of each incidence (table form)
    __shared__ float sTk[TAB ? G3_MT * G3_NI * 3 * 9 : 4];  // GridTab3
    __shared__ TC sLM[MAX_LDS_MATS][2];      // lambda, mu of each LEI material
    __shared__ double sLC[MAX_LDS_MATS][LeiCoef<D>::N];  // LEI stress coefficients
    __shared__ GridTpl3 sT;

    const bool func = (flags & PG_FUNC) != 0, jac = (flags & PG_JAC) != 0;
    const int tid = threadIdx.x;
    // each block takes `per` consecutive groups (XCD-aware block order)
    const int g0 = xcd_order(blockIdx.x, gridDim.x) * per, g1 = min(ngroups, g0 + per);
    int4 r = make_int4(-1, 0, 0, 0);
    int inf = 0;
    if (g0 < g1) {
        const int64_t k0 = int64_t(g0) * NG + (tid & (NG - 1));
        r = srec[k0];
        inf = sinf[k0];
    }
    constexpr int TW = int(sizeof(GridTpl3) / 4);
    int tw[(TW + BS - 1) / BS];
#pragma unroll
    for (int k = 0; k < (TW + BS - 1) / BS; ++k) {
        const int w = tid + k * BS;
        tw[k] = tpl[w < TW ? w : TW - 1];
    }
    stage_lm<TC, BS>(p, 6, sLM);
    if constexpr (TAB)
        for (int k = tid; k < int(sizeof(GridTab3) / 16); k += BS) reinterpret_cast<int4*>(sTk)[k] = tab[k];
    for (int k = tid; k < MAX_LDS_MATS * LeiCoef<D>::N; k += BS) {
        const int m = k / LeiCoef<D>::N, j = k % LeiCoef<D>::N;
        const double v = lei_coef<D>(p, m < p.n_mat ? m : p.n_mat - 1, j);
        sLC[m][j] = m < p.n_mat ? v : 0.0;
    }
#pragma unroll
    for (int k = 0; k < (TW + BS - 1) / BS; ++k) {
        const int w = tid + k * BS;
        if (w < TW) reinterpret_cast<int32_t*>(&sT)[w] = tw[k];
    }
    double wsum = 0.0;
    for (int k = 0; k < NQ; ++k) wsum += TET_QW[k];
    for (int g = g0; g < g1; ++g) {
        if (tid < NG) {
            s_node[tid] = r.x;
            s_kb[tid] = r.y;
            s_dof[tid] = r.z;
            s_eb[tid] = r.w;
            s_S[tid] = inf;
        }
        __syncthreads();
        // the next group's records go out now and land during this group's work
        if (g + 1 < g1) {
            const int64_t k1 = int64_t(g + 1) * NG + (tid & (NG - 1));
            r = srec[k1];
            inf = sinf[k1];
        }

        // ---- phase A: incidences (as k_pair_lists, records from the template) ----
        for (int i = tid; i < MAXI; i += BS) {
            const int lo = i / G3_NI, rr = i - lo * G3_NI;
            const int a = s_node[lo];
            if (a < 0) continue;  // padding past the last node
            const int64_t S = s_S[lo], n = S - 1;
            const int64_t e = int64_t(s_eb[lo]) + 6 * (sT.ec[rr][0] + sT.ec[rr][1] * n + sT.ec[rr][2] * n * n) +
                              sT.ec[rr][3];
            int64_t ln[NEN];
            for (int k = 0; k < NEN; ++k) ln[k] = a + sT.nd[rr][k][0] + sT.nd[rr][k][1] * S + sT.nd[rr][k][2] * S * S;
            double x[NEN][3], u[NEN][D];  // U goes out with the coordinates
            for (int k = 0; k < NEN; ++k)
                for (int c = 0; c < D; ++c) {
                    x[k][c] = p.coords[ln[k] * D + c];
                    u[k][c] = Un[ln[k] * D + c];
                }
            const int mi = p.elem_mat[e] & (MAX_LDS_MATS - 1);
            // every load of the incidence is out before the arithmetic
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (TAB) {
                s_mi[i] = uint8_t(mi);
                if (!func) continue;
            }
            double lc[LeiCoef<D>::N];
            for (int k = 0; k < LeiCoef<D>::N; ++k) lc[k] = sLC[mi][k];
            double gr[NEN][D];
            const double J = simplex_grads<D, double>(x, gr);
            const double w = J * wsum;
            plain_p<D>(gr, u, 0, w, lc, s_P[i]);
            if constexpr (TAB) continue;
            TC gt[NEN][D];
            for (int k = 0; k < NEN; ++k)
                for (int c = 0; c < D; ++c) { gt[k][c] = TC(gr[k][c]); s_g[i][k][c] = gt[k][c]; }
            if (jac) {
                const TC wl = TC(w) * sLM[mi][0], wm = TC(w) * sLM[mi][1];
                s_wlm[i][0] = wl;
                s_wlm[i][1] = wm;
                diag_block<D, TC>(gt, 0, wl, wm, reinterpret_cast<double (*)[D]>(s_dblk[i]));
            }
        }
        __syncthreads();

        // ---- phase B: node pairs ----
        // lanes [0, NOFF): one off-diagonal pair each (node lo, its k-th
        // off-diagonal slot), the closed-form blocks of its contributions;
        // lanes [DB, DB + NDE): one entry each of a node's diagonal block
        // (c < 9) or internal force (c >= 9), summed over its incidences.  DB
        // starts a wavefront, so no wavefront runs both kinds of work.
        constexpr int NOFF = (G3_NP - 1) * NG, DB = (NOFF + 63) / 64 * 64, NDE = (D * D + D) * NG;
        for (int kq = tid; kq < DB + NDE; kq += BS) {
            if (kq < NOFF) {
                const int lo = kq / (G3_NP - 1), k = kq - lo * (G3_NP - 1);
                if (s_node[lo] < 0 || !jac) continue;
                const int slot = k < sT.diag ? k : k + 1;
                double acc[D][D];
                for (int i = 0; i < D; ++i)
                    for (int j = 0; j < D; ++j) acc[i][j] = 0.0;
#pragma unroll
                for (int r = 0; r < G3_CAP; ++r) {
                    const unsigned code = sT.pc[slot][r];
                    if (code == 0xffu) continue;
                    const int t = lo * G3_NI + int(code >> 2), lb = int(code & 3);
                    if constexpr (TAB) {
                        const float* tb = &sTk[((int(s_mi[t]) * G3_NI + int(code >> 2)) * 3 + lb - 1) * 9];
                        for (int i = 0; i < D; ++i)
                            for (int j = 0; j < D; ++j) acc[i][j] += double(tb[i * D + j]);
                    } else {
                        TC blk[D][D];
                        iso_block<D, TC>(s_g[t][0], s_g[t][lb], s_wlm[t][0], s_wlm[t][1], blk);
                        for (int i = 0; i < D; ++i)
                            for (int j = 0; j < D; ++j) acc[i][j] += double(blk[i][j]);
                    }
                }
                const int64_t k0v = int64_t(s_kb[lo]) + 3 * slot;
                for (int i = 0; i < D; ++i)
                    for (int j = 0; j < D; ++j) K[k0v + int64_t(i) * 3 * G3_NP + j] = acc[i][j];
            } else if (kq >= DB) {
                const int e = kq - DB, lo = e / (D * D + D), c = e - lo * (D * D + D);
                if (s_node[lo] < 0 || (c < D * D ? !jac : !func)) continue;
                // every incidence of the node, in ascending element order
                const int t0 = lo * G3_NI;
                double a = 0.0;
                if (c < D * D) {
                    if constexpr (TAB) {
                        // minus the FP64 sum of each incidence's three table
                        // blocks, in the order diag_acc forms it
#pragma unroll 8
                        for (int t = t0; t < t0 + G3_NI; ++t) {
                            const float* tb = &sTk[(int(s_mi[t]) * G3_NI + (t - t0)) * 27 + c];
                            double v = -double(tb[0]);
                            v = v - double(tb[9]);
                            v = v - double(tb[18]);
                            a += v;
                        }
                    } else {
#pragma unroll 8
                        for (int t = t0; t < t0 + G3_NI; ++t) a += s_dblk[t][c];
                    }
                    K[int64_t(s_kb[lo]) + 3 * sT.diag + int64_t(c / D) * 3 * G3_NP + c % D] = a;
                } else {
#pragma unroll 8
                    for (int t = t0; t < t0 + G3_NI; ++t) a += s_P[t][c - D * D];
                    P[s_dof[lo] + c - D * D] = a;
                }
            }
        }
        __syncthreads();  // the next group rewrites the tables
    }
}

// The same schedule for tet4 meshes, where the per-pair block tables would not
// fit in LDS (and a tri3 variant): phase A stores each incidence's TC
// gradients, w lambda and w mu, its diagonal block and internal force, and
// registers (incidence, local node) codes at their static (slot, rank); phase
// B evaluates the closed-form block of each contribution.
// (at most 128 VGPRs: four wavefronts per SIMD where LDS allows them)
template <int D, typename TC, typename TKC, int NG, int TV, int BS = GroupThreads<D, NG, TV>::value>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4))) void k_pair_lists(DevProb p, const double* __restrict__ Un,
                                                    const TKC* __restrict__ Kc, const double* __restrict__ Pc,
                                                    double* __restrict__ K, double* __restrict__ P, unsigned flags,
                                                    int gbase, int64_t ibase) {
    constexpr int NEN = D + 1, NS = (D == 2 ? 3 : 6), NQ = (D == 2 ? 4 : 5);
    constexpr int MAXI = GroupLim<D, NG>::MAXI, MAXP = GroupLim<D, NG>::MAXP;
    constexpr int MAXLOC = (D == 2 ? 7 : 8);  // local nodes per incidence
    // contributions per off-diagonal pair: the elements sharing the edge (a, b),
    // 2 in a tri3 mesh and up to 6 around a tet4 edge (cut elements add their
    // enrichment pairs); the diagonal pair reads the node's incidences directly
    constexpr int CAP = (D == 2 ? 4 : 8);
    static_assert(MAXI <= 4096 && MAXLOC <= 16, "contribution codes are (incidence << 4 | local node)");
    __shared__ int s_node[NG];               // the group's nodes (from the node order)
    __shared__ int s_ibeg[NG], s_qbeg[NG];   // their first incidence and first node pair
    __shared__ int s_iptr[NG + 1];           // local prefix of incidences per node
    __shared__ int s_qptr[NG + 1];           // local prefix of node pairs per node
    __shared__ unsigned short s_list[MAXP][CAP];  // contributions per pair, by rank (0xffff: none)
    __shared__ int s_info[MAXI];             // la | nloc << 8 | (cut ? 1 << 16 : 0); la = 255: absent
    __shared__ int64_t s_koff[MAXI];
    __shared__ TC s_g[MAXI][NEN][D];
    __shared__ TC s_wlm[MAXI][2];            // w lambda, w mu of each uncut incidence
    __shared__ double s_P[MAXI][D];
    __shared__ double s_dblk[MAXI][D * D];    // diagonal block (la, la) of each incidence
    __shared__ TC sLM[MAX_LDS_MATS][2];       // lambda, mu of each LEI material
    __shared__ int s_dq[NG];                  // local index of each node's diagonal pair (-1: none)
    __shared__ int s_bad;

    const bool func = (flags & PG_FUNC) != 0, jac = (flags & PG_JAC) != 0;
    const int tid = threadIdx.x;
    const int grp = xcd_order(blockIdx.x, gridDim.x);
    const int4* gt = reinterpret_cast<const int4*>(p.gtab) + gbase + int64_t(grp) * NG;
    constexpr int ng = NG;
    // every set-up load goes out before the first wait: the group's node table
    // (NG entries, every lane loads one), the first pass's incidence records and
    // slot words (their block does not depend on the node table; lanes past the
    // block reload its last slot) and lambda, mu of the materials
    const int4 r = gt[tid & (NG - 1)];
    const int64_t ib = ibase + int64_t(grp) * MAXI;
    const int is = tid < MAXI ? tid : MAXI - 1;
    int rec0[NEN + 1];
    for (int n = 0; n <= NEN; ++n) rec0[n] = p.incg[(ib + is) * (NEN + 1) + n];
    const uint64_t sw0 = p.slotg[ib + is];
    stage_lm<TC, BS>(p, NS, sLM);
    // the group's nodes and their incidence / pair counts, prefix-summed across
    // the wavefront (NG <= 32: the first wavefront holds the group)
    int ci = tid < ng ? (r.w & 0xffff) : 0, cq = tid < ng ? (r.w >> 16) : 0;
#pragma unroll
    for (int off = 1; off < NG; off <<= 1) {
        const int ti = __shfl_up(ci, off), tq = __shfl_up(cq, off);
        if ((tid & 63) >= off) { ci += ti; cq += tq; }
    }
    if (tid < ng) {
        s_node[tid] = r.x;
        s_ibeg[tid] = r.y;
        s_qbeg[tid] = r.z;
        s_iptr[tid + 1] = ci;
        s_qptr[tid + 1] = cq;
        s_dq[tid] = -1;
    }
    if (tid == 0) {
        s_iptr[0] = 0;
        s_qptr[0] = 0;
        s_bad = 0;
    }
    __syncthreads();
    const int ni = s_iptr[ng], nq = s_qptr[ng];
    const bool staged = ni <= MAXI && nq <= MAXP && !(flags & PG_GLOBAL_PATH);
    double wsum = 0.0;
    for (int k = 0; k < NQ; ++k) wsum += (D == 2 ? TRI_QW[k] : TET_QW[k]);

    PairMeta<D> pm0;
    if (D == 2 && int(threadIdx.x) < nq) pm0 = pair_meta<D, NG>(p, s_qptr, s_qbeg, s_node, ng, threadIdx.x, jac);

    // ---- phase A: incidences ----
    if (staged) {
        for (int k = threadIdx.x; k < nq * CAP; k += blockDim.x) s_list[k / CAP][k % CAP] = 0xffffu;
        __syncthreads();
        for (int i = threadIdx.x; i < ni; i += blockDim.x) {
            // staged: ni <= MAXI, so slot i lies in the group's block
            const bool first = i == int(threadIdx.x);
            int rec[NEN + 1];
            for (int n = 0; n <= NEN; ++n) rec[n] = first ? rec0[n] : p.incg[(ib + i) * (NEN + 1) + n];
            const uint64_t sw = first ? sw0 : p.slotg[ib + i];
            const int tag = rec[NEN];
            const int lo = tag < 0 ? (tag >> 3) & 31 : (rec[2] >> 16) & 31;  // owner's position in the group
            int la = -1;
            int nloc = NEN;
            if (tag < 0) {
                // the element's nodes, owner first (the record lists them so)
                int ln[NEN];
                for (int n = 0; n < NEN; ++n) ln[n] = rec[n];
                la = 0;
                double x[NEN][3], u[NEN][D];  // U goes out with the coordinates
                for (int n = 0; n < NEN; ++n)
                    for (int c = 0; c < D; ++c) {
                        x[n][c] = p.coords[int64_t(ln[n]) * D + c];
                        u[n][c] = Un[int64_t(ln[n]) * D + c];
                    }
                const int mi = tag & 7;
                double lc[LeiCoef<D>::N];
                for (int k = 0; k < LeiCoef<D>::N; ++k) lc[k] = lei_coef<D>(p, mi, k);
                // every load of the incidence is out before the arithmetic
                __builtin_amdgcn_sched_barrier(0);
                double g[NEN][D];
                const double J = simplex_grads<D, double>(x, g);
                const double w = J * wsum;
                TC gt[NEN][D];
                for (int n = 0; n < NEN; ++n)
                    for (int c = 0; c < D; ++c) { gt[n][c] = TC(g[n][c]); s_g[i][n][c] = gt[n][c]; }
                // (also without PG_FUNC, where phase B ignores it: the U loads
                // then stay ahead of the gradient arithmetic)
                plain_p<D>(g, u, 0, w, lc, s_P[i]);
                if (jac) {
                    const TC wl = TC(w) * sLM[mi][0], wm = TC(w) * sLM[mi][1];
                    s_wlm[i][0] = wl;
                    s_wlm[i][1] = wm;
                    diag_block<D, TC>(gt, 0, wl, wm, reinterpret_cast<double (*)[D]>(s_dblk[i]));
                }
            } else {
                // cut element: the record holds the staged row of a directly
                la = rec[2] & 255;
                nloc = (rec[2] >> 8) & 255;
                if (la >= nloc) {
                    // no row of a in this element: its diagonal entries add zero
                    la = -1;
                    for (int c = 0; c < D; ++c) s_P[i][c] = 0.0;
                    for (int k = 0; k < D * D; ++k) s_dblk[i][k] = 0.0;
                }
                if (nloc > MAXLOC) {
                    s_bad = 1;  // local list does not fit: this group takes the global path
                } else {
                    s_koff[i] = int64_t(uint32_t(rec[0])) * (D * D);  // blocks (la, 0 .. nloc - 1)
                    if (func && la >= 0)
                        for (int c = 0; c < D; ++c) s_P[i][c] = Pc[int64_t(rec[1]) + c];
                    if (jac && la >= 0) cut_diag<D, TKC>(Kc + s_koff[i], la, s_dblk[i]);
                }
            }
            s_info[i] = (la < 0 ? 255 : la) | (nloc << 8) | (tag >= 0 ? (1 << 16) : 0);
            if (la < 0 || nloc > MAXLOC) continue;
            // contribution of this incidence to each off-diagonal pair (a, ln[m]) of
            // a's row, at its precomputed slot and rank (ascending element order)
            if (sw >> 63) { s_bad = 1; continue; }
            const int ps0 = s_qptr[lo], np = s_qptr[lo + 1] - s_qptr[lo];
            int k = 0;
            for (int m = 0; m < nloc; ++m) {
                if (m == la) continue;
                const int f = int((sw >> (9 * k)) & 511u);
                ++k;
                const int slot = f & 63, rank = f >> 6;
                if (slot >= np || rank >= CAP) { s_bad = 1; continue; }
                s_list[ps0 + slot][rank] = (unsigned short)((i << 4) | m);
            }
        }
    }
    __syncthreads();
    const bool use_lists = staged && !s_bad;

    // ---- phase B: node pairs ----
    // Staged: the off-diagonal pairs first (a diagonal pair's lane only notes
    // its index), then one lane per entry of a node's diagonal block or internal
    // force, summed over the node's incidences -- so that no wavefront runs both
    // the contribution blocks and those sums.
    for (int kq = threadIdx.x; kq < nq; kq += blockDim.x) {
        const PairMeta<D> pm = (D == 2 && kq == int(threadIdx.x)) ? pm0 : pair_meta<D, NG>(p, s_qptr, s_qbeg, s_node, ng, kq, jac);
        const int lo = pm.lo, a = pm.a, q = pm.q, b = pm.b;
        if (use_lists && a == b) {
            s_dq[lo] = kq;
            continue;
        }
        const bool diag = (a == b) && func;
        double acc[D][D];
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) acc[i][j] = 0.0;
        double pa[D];
        for (int i = 0; i < D; ++i) pa[i] = 0.0;
        if (use_lists) {
            // the slot's contributions, stored by rank = ascending element order
            const int s = kq;
            if (jac) {
#pragma unroll
                for (int k = 0; k < CAP; ++k) {
                    const unsigned code = s_list[s][k];
                    if (code == 0xffffu) continue;
                    const int t = int(code >> 4), lb = int(code & 15);
                    const int info = s_info[t];
                    if (!(info & (1 << 16))) {
                        TC blk[D][D];  // uncut: the owner is local node 0
                        iso_block<D, TC>(s_g[t][0], s_g[t][lb], s_wlm[t][0], s_wlm[t][1], blk);
                        for (int i = 0; i < D; ++i)
                            for (int j = 0; j < D; ++j) acc[i][j] += double(blk[i][j]);
                    } else {
                        const TKC* Ke = Kc + s_koff[t] + int64_t(lb) * D * D;
                        for (int i = 0; i < D; ++i)
                            for (int j = 0; j < D; ++j) acc[i][j] += double(Ke[i * D + j]);
                    }
                }
            }
        } else {
            pair_from_global<D, TC, TKC>(p, Un, Kc, Pc, sLM, wsum, s_ibeg[lo], s_ibeg[lo] + (s_iptr[lo + 1] - s_iptr[lo]),
                                    a, b, jac, diag, acc, pa);
        }
        if (jac) {
            // 2D: a row's two columns of b side by side in one block go out as
            // one 16-byte store (8-byte aligned)
            if constexpr (D == 2) {
                if (pm.map.y == pm.map.x + 1 && pm.map.w == pm.map.z + 1) {
                    *reinterpret_cast<dpair*>(K + pm.map.x) = dpair{acc[0][0], acc[0][1]};
                    *reinterpret_cast<dpair*>(K + pm.map.z) = dpair{acc[1][0], acc[1][1]};
                } else {
                    K[pm.map.x] = acc[0][0];
                    K[pm.map.y] = acc[0][1];
                    K[pm.map.z] = acc[1][0];
                    K[pm.map.w] = acc[1][1];
                }
            } else {
                // (merging a 3D row's stores, 16 + 8 bytes, was 2 % slower on C4)
                const int32_t* mp = p.npat_map + int64_t(q) * D * D;
                for (int i = 0; i < D; ++i)
                    for (int j = 0; j < D; ++j) K[mp[i * D + j]] = acc[i][j];
            }
        }
        if (diag)
            for (int i = 0; i < D; ++i) P[pm.dof[i]] = pa[i];
    }
    if (!use_lists) return;
    __syncthreads();
    // diagonal entries: every incidence of the node in ascending element order
    // (an incidence without a row holds zeros: no test, so that the LDS reads of
    // later incidences go out early)
    constexpr int NDE = D * D + D;
    for (int e = threadIdx.x; e < ng * NDE; e += blockDim.x) {
        const int lo = e / NDE, c = e - lo * NDE;
        const int kq = s_dq[lo];
        if (kq < 0 || (c < D * D ? !jac : !func)) continue;
        const int t0 = s_iptr[lo], t1 = s_iptr[lo + 1];
        const int64_t q = int64_t(s_qbeg[lo]) + (kq - s_qptr[lo]);
        double a = 0.0;
        if (c < D * D) {
#pragma unroll 8
            for (int t = t0; t < t1; ++t) a += s_dblk[t][c];
            K[p.npat_map[q * D * D + c]] = a;
        } else {
#pragma unroll 8
            for (int t = t0; t < t1; ++t) a += s_P[t][c - D * D];
            P[p.node_dof[int64_t(s_node[lo]) * D + c - D * D]] = a;
        }
    }
}

}  // namespace

struct pg_asm {
    DevProb d{};
    int64_t n_dof = 0, nnz_total = 0;
    int64_t nnz[4] = {0, 0, 0, 0};
    int64_t cut_k_total = 0, cut_p_total = 0;
    std::vector<void*> allocs;
    double* Un = nullptr;        // node-major U
    double* stage = nullptr;     // plain schedule: uncut elements' Ke + Pe (FP64, PlainStage layout)
    double* stage_k = nullptr;   // plain schedule: cut elements' K_el, row-major, at cut_koff
    double* stage_p = nullptr;   // plain schedule: cut elements' P_el at cut_poff
    double* Kc64 = nullptr;      // cut tangents (FP64)
    float* Kc32 = nullptr;       // cut tangents (FP32)
    double* Pc = nullptr;        // cut residuals (FP64)
    int* err = nullptr;
    int32_t* class_list = nullptr;  // cut indices: small class, then general class
    int n_class[2] = {0, 0};
    // host-API buffers
    double *dU = nullptr, *dK = nullptr, *dP = nullptr, *dIc = nullptr, *dIt = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    hipStream_t side = nullptr;     // second stream: node groups without cut elements overlap the cut kernel
    hipEvent_t sync[2] = {nullptr, nullptr};
    // completion of the handle's last call that used its scratch (Un, staged
    // tangents, the error flag) and the stream it ran on: a call on another
    // stream waits for it first
    hipEvent_t done = nullptr;
    hipStream_t done_st = nullptr;
    bool done_valid = false;
    int gopt_ng[3] = {0, 0, 0};     // group sizes with a group list
    int goff[3] = {0, 0, 0};        // first gtab entry of each group list
    int64_t gioff[3] = {0, 0, 0};   // first incg / slotg slot of each group list
    int ngroups[3] = {0, 0, 0};     // groups of each list
    int nclean[3] = {0, 0, 0};      // groups without cut incidences (listed first)
    int nenrg[3] = {0, 0, 0};       // groups of enrichment nodes (listed last)
    int nstruct[3] = {0, 0, 0};     // groups of structured-grid nodes (listed first)
    int64_t n_struct = 0;           // structured-grid nodes (k_pair_struct records)
    int32_t* stpl = nullptr;        // GridTpl words
    int4* srec = nullptr;           // per structured node: node, first K value, first dof, cell element
    int32_t* sinf = nullptr;        // per structured node: row stride | parity << 24
    int4* stab3 = nullptr;          // GridTab3 (tet4 grid, FP32 tangents), or null: no table
    int32_t n_poly = 0;             // cut elements with polygon children (2D)
    int32_t* poly_list = nullptr;   // their cut indices (the FP64 drop-in's polygon launch)
    bool timing = false;
    double last_ms[5] = {0, 0, 0, 0, 0};
    int pair_block = 256;
    int group_nodes = 0;   // node-group size of the binned schedule (0: default)
    int cut_variant = 0;   // small-class cut kernel variant (elements per wave, occupancy)
    int group_variant = 0; // bit 0: two threads per node pair; bit 1: list form (2D); bit 2: no overlap;
                           // bit 3: global-memory path; bit 4: two threads per pair for enrichment-node groups;
                           // bits 5-6: 1 + option index of the group size of the groups after the cut kernels;
                           // bit 7: structured nodes through the generic kernel; bits 8-9: log2 groups per block
                           // of the structured kernel; bit 10: FP32 structured tet4 nodes without GridTab3
    int64_t alg_bytes = 0;
    int64_t dev_bytes = 0;  // device memory held by the handle
    // thermal pre-load: each material's expansion ratio (0 where not set) and
    // the temperature change of the following assemblies
    double thermal_exp[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool thermal_set[8] = {false, false, false, false, false, false, false, false};
    double temp_change = 0.0;
    // VTK cells (uncut elements and integration children), built at the first
    // pg_asm_plot_fields: (element, child or -1), polygon children listed apart
    int2* plot_cells = nullptr;
    int32_t n_plot[2] = {0, 0};  // cells without / with polygon children
    int64_t n_cells = 0;
    int64_t* plot_pos = nullptr;  // each listed cell's position in the writer's cell order
    double* plot_tmp = nullptr;   // the kernels' rows, in list order
    int mat_type[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // PG_MAT_* of each material
    unsigned coh_law_mask = 0;                   // bit PG_MAT_* set for each law of a cohesive child
};

namespace {

// GridTpl built from the union-jack rule, with each incidence's element nodes
// in their local (connectivity) order for the host check.
struct GridTplHost {
    GridTpl t;
    int8_t ldx[2][8][3], ldy[2][8][3];
};

// The two triangles of cell (cx, cy), nodes in the reference's local order
// (src/Mesh.cpp:418-441).
void grid_tri(int cx, int cy, int half, int v[3][2]) {
    static const int E[2][2][3][2] = {{{{0, 0}, {1, 0}, {1, 1}}, {{0, 0}, {1, 1}, {0, 1}}},
                                      {{{0, 0}, {1, 0}, {0, 1}}, {{1, 0}, {1, 1}, {0, 1}}}};
    const int k = ((cx + cy) % 2 + 2) % 2;
    for (int m = 0; m < 3; ++m) {
        v[m][0] = cx + E[k][half][m][0];
        v[m][1] = cy + E[k][half][m][1];
    }
}

bool build_grid_tpl(GridTplHost& H) {
    memset(&H, 0, sizeof H);
    for (int par = 0; par < 2; ++par) {
        const int x = 2 + par, y = 2;
        int r = 0;
        std::vector<std::pair<int, int>> nb;  // (dy, dx): ascending node id for any row stride > 2
        for (int dcy = -1; dcy <= 0; ++dcy)
            for (int dcx = -1; dcx <= 0; ++dcx)
                for (int hh = 0; hh < 2; ++hh) {
                    int v[3][2];
                    grid_tri(x + dcx, y + dcy, hh, v);
                    int la = -1;
                    for (int m = 0; m < 3; ++m)
                        if (v[m][0] == x && v[m][1] == y) la = m;
                    if (la < 0) continue;
                    if (r >= GRID_NI[par]) return false;
                    H.t.ecx[par][r] = int8_t(dcx);
                    H.t.ecy[par][r] = int8_t(dcy);
                    H.t.eh[par][r] = int8_t(hh);
                    for (int m = 0; m < 3; ++m) {
                        H.ldx[par][r][m] = int8_t(v[m][0] - x);
                        H.ldy[par][r][m] = int8_t(v[m][1] - y);
                        const int mo = owner_first(3, la, m);
                        H.t.ndx[par][r][m] = int8_t(v[mo][0] - x);
                        H.t.ndy[par][r][m] = int8_t(v[mo][1] - y);
                        nb.push_back({v[m][1] - y, v[m][0] - x});
                    }
                    ++r;
                }
        if (r != GRID_NI[par]) return false;
        std::sort(nb.begin(), nb.end());
        nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
        if (int(nb.size()) != GRID_NP[par]) return false;
        for (int sl = 0; sl < GRID_NP[par]; ++sl) {
            H.t.pdy[par][sl] = int8_t(nb[size_t(sl)].first);
            H.t.pdx[par][sl] = int8_t(nb[size_t(sl)].second);
        }
        if (H.t.pdx[par][GRID_DIAG[par]] != 0 || H.t.pdy[par][GRID_DIAG[par]] != 0) return false;
        int cnt[9] = {0};
        for (int ri = 0; ri < r; ++ri)
            for (int k = 1; k < 3; ++k) {
                int sl = 0;
                while (sl < GRID_NP[par] &&
                       !(H.t.pdx[par][sl] == H.t.ndx[par][ri][k] && H.t.pdy[par][sl] == H.t.ndy[par][ri][k]))
                    ++sl;
                if (sl == GRID_NP[par]) return false;
                H.t.sr[par][ri][k - 1] = int16_t(sl | (cnt[sl]++ << 6));
            }
        // every off-diagonal pair gets exactly ranks 0 and 1 (an interior edge
        // lies in two triangles), so the kernel needs no cleared table
        for (int sl = 0; sl < GRID_NP[par]; ++sl)
            if (cnt[sl] != (sl == GRID_DIAG[par] ? 0 : 2)) return false;
    }
    return true;
}

// Whether regular node a is a node of the structured grid whose incidences,
// element node order, pattern columns, scatter indices and dofs are exactly
// what GridTpl gives; then its k_pair_struct record and word.
bool grid_node(const pg_problem_desc* s, const GridTplHost& H, int64_t a, int32_t rec[4], int32_t& inf) {
    const int64_t t0 = s->n2e_ptr[a], ni = s->n2e_ptr[a + 1] - t0;
    const int64_t p0 = s->npat_ptr[a], np = s->npat_ptr[a + 1] - p0;
    const int par = ni == GRID_NI[0] ? 0 : (ni == GRID_NI[1] ? 1 : -1);
    if (par < 0 || np != GRID_NP[par]) return false;
    // the first column is (x - 1, y - 1) (even) or (x, y - 1) (odd): the stride
    const int64_t S = a + H.t.pdx[par][0] - s->npat_col[p0];
    if (S < 3 || S >= (int64_t(1) << 24)) return false;
    const int64_t n = S - 1, rl = ni - 1;
    const int64_t eb = int64_t(s->n2e[t0 + rl]) - (2 * (H.t.ecx[par][rl] + H.t.ecy[par][rl] * n) + H.t.eh[par][rl]);
    if (eb < 0) return false;
    for (int r = 0; r < ni; ++r) {
        const int64_t e = eb + 2 * (H.t.ecx[par][r] + H.t.ecy[par][r] * n) + H.t.eh[par][r];
        if (s->n2e[t0 + r] != e || s->cut_of_elem[e] >= 0) return false;
        for (int m = 0; m < 3; ++m)
            if (s->conn[e * 3 + m] != a + H.ldx[par][r][m] + H.ldy[par][r][m] * S) return false;
    }
    const int64_t kb = s->npat_map[p0 * 4];
    const int32_t dof = s->node_dof[a * 2];
    if (s->node_dof[a * 2 + 1] != dof + 1) return false;
    for (int sl = 0; sl < np; ++sl) {
        if (s->npat_col[p0 + sl] != a + H.t.pdx[par][sl] + H.t.pdy[par][sl] * S) return false;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                if (s->npat_map[(p0 + sl) * 4 + i * 2 + j] != kb + i * 2 * np + 2 * sl + j) return false;
    }
    rec[0] = int32_t(a);
    rec[1] = int32_t(kb);
    rec[2] = dof;
    rec[3] = int32_t(eb);
    inf = int32_t(S | (int64_t(par) << 24));
    return true;
}

// GridTpl3 from the tet4 cube split of generate_mesh (src/Mesh.cpp:540-583:
// corners 0 .. 7 = (0,0,0), (1,0,0), (1,1,0), (0,1,0), then the same at z + 1;
// the six tets around the diagonal 0 - 6), for a node at the origin.
bool build_grid_tpl3(GridTpl3& T) {
    static const int C[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    static const int TT[6][4] = {{0, 2, 3, 6}, {0, 3, 7, 6}, {0, 7, 4, 6}, {0, 5, 6, 4}, {1, 5, 6, 0}, {1, 6, 2, 0}};
    memset(&T, 0, sizeof T);
    struct Inc {
        int64_t key;
        int c[4], la;
        int v[4][3];
    };
    std::vector<Inc> inc;
    const int64_t N = 1000;  // element ids of any mesh with n > 2 sort the same way
    for (int dz = -1; dz <= 0; ++dz)
        for (int dy = -1; dy <= 0; ++dy)
            for (int dx = -1; dx <= 0; ++dx)
                for (int t = 0; t < 6; ++t) {
                    Inc I{};
                    I.la = -1;
                    for (int k = 0; k < 4; ++k) {
                        I.v[k][0] = dx + C[TT[t][k]][0];
                        I.v[k][1] = dy + C[TT[t][k]][1];
                        I.v[k][2] = dz + C[TT[t][k]][2];
                        if (I.v[k][0] == 0 && I.v[k][1] == 0 && I.v[k][2] == 0) I.la = k;
                    }
                    if (I.la < 0) continue;
                    I.key = 6 * (dx + dy * N + dz * N * N) + t;
                    I.c[0] = dx; I.c[1] = dy; I.c[2] = dz; I.c[3] = t;
                    inc.push_back(I);
                }
    if (int(inc.size()) != G3_NI) return false;
    std::sort(inc.begin(), inc.end(), [](const Inc& A, const Inc& B) { return A.key < B.key; });
    // the columns in ascending node id: (dz, dy, dx) for any row stride > 2
    std::vector<std::array<int, 3>> nb;
    for (const Inc& I : inc)
        for (int k = 0; k < 4; ++k) nb.push_back({I.v[k][2], I.v[k][1], I.v[k][0]});
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    if (int(nb.size()) != G3_NP) return false;
    T.diag = -1;
    for (int sl = 0; sl < G3_NP; ++sl) {
        T.pd[sl][0] = int8_t(nb[size_t(sl)][2]);
        T.pd[sl][1] = int8_t(nb[size_t(sl)][1]);
        T.pd[sl][2] = int8_t(nb[size_t(sl)][0]);
        if (nb[size_t(sl)][0] == 0 && nb[size_t(sl)][1] == 0 && nb[size_t(sl)][2] == 0) T.diag = int8_t(sl);
    }
    if (T.diag < 0) return false;
    memset(T.pc, 0xff, sizeof T.pc);
    int cnt[G3_NP] = {0};
    for (int r = 0; r < G3_NI; ++r) {
        const Inc& I = inc[size_t(r)];
        for (int k = 0; k < 4; ++k) T.ec[r][k] = int8_t(I.c[k]);
        T.la[r] = int8_t(I.la);
        for (int k = 0; k < 4; ++k)
            for (int c = 0; c < 3; ++c) T.nd[r][k][c] = int8_t(I.v[owner_first(4, I.la, k)][c]);
        for (int m = 1; m < 4; ++m) {
            int sl = 0;
            while (sl < G3_NP && !(T.pd[sl][0] == T.nd[r][m][0] && T.pd[sl][1] == T.nd[r][m][1] &&
                                   T.pd[sl][2] == T.nd[r][m][2]))
                ++sl;
            if (sl == G3_NP || cnt[sl] >= G3_CAP) return false;
            T.pc[sl][cnt[sl]++] = uint8_t(r << 2 | m);
        }
    }
    return true;
}

// GridTab3 from the template, the grid spacing h and the materials' lambda =
// D[0][1] and mu = D[5][5] (pg_asm_create accepts isotropic LEI materials
// only): the gradients of each tet on the unit grid are integers, so the
// blocks are computed in FP64 from exact inputs and rounded to FP32 once.
void build_grid_tab3(const GridTpl3& T, const pg_problem_desc* s, double h, GridTab3& G) {
    memset(&G, 0, sizeof G);
    const double QW[5] = {-0.8, 0.45, 0.45, 0.45, 0.45};
    double wsum = 0.0;
    for (int k = 0; k < 5; ++k) wsum += QW[k];
    for (int r = 0; r < G3_NI; ++r) {
        int d[3][3];
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) d[k][c] = T.nd[r][k + 1][c] - T.nd[r][0][c];
        int cf[3][3];
        cf[0][0] = d[1][1] * d[2][2] - d[1][2] * d[2][1];
        cf[0][1] = d[1][2] * d[2][0] - d[1][0] * d[2][2];
        cf[0][2] = d[1][0] * d[2][1] - d[1][1] * d[2][0];
        cf[1][0] = d[0][2] * d[2][1] - d[0][1] * d[2][2];
        cf[1][1] = d[0][0] * d[2][2] - d[0][2] * d[2][0];
        cf[1][2] = d[0][1] * d[2][0] - d[0][0] * d[2][1];
        cf[2][0] = d[0][1] * d[1][2] - d[0][2] * d[1][1];
        cf[2][1] = d[0][2] * d[1][0] - d[0][0] * d[1][2];
        cf[2][2] = d[0][0] * d[1][1] - d[0][1] * d[1][0];
        const int Jt = d[0][0] * cf[0][0] + d[0][1] * cf[0][1] + d[0][2] * cf[0][2];  // +-1 on the unit grid
        double G4[4][3];
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) G4[k + 1][c] = double(cf[k][c] * Jt);
        for (int c = 0; c < 3; ++c) G4[0][c] = -(G4[1][c] + G4[2][c] + G4[3][c]);
        for (int m = 0; m < s->n_mat && m < G3_MT; ++m) {
            const double lam = s->mats[m].D[1], mu = s->mats[m].D[35], w = double(Jt) * wsum * h;
            for (int lb = 1; lb < 4; ++lb) {
                const double* ga = G4[0];
                const double* gb = G4[lb];
                const double dot = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        G.tk[m][r][lb - 1][i * 3 + j] =
                            float(w * (lam * ga[i] * gb[j] + mu * ga[j] * gb[i] + (i == j ? mu * dot : 0.0)));
            }
        }
    }
}

// Whether regular node a of a tet4 mesh with row stride S is a node of the
// structured grid whose incidences, element nodes, pattern columns, scatter
// indices and dofs are exactly what GridTpl3 gives; then its k_pair_struct3
// record and word.  tab_ok is cleared unless the node's 26 neighbours lie at
// the grid offsets times h to 1e-10 h and its elements' materials are in
// GridTab3.
bool grid_node3(const pg_problem_desc* s, const GridTpl3& T, int64_t S, int64_t a, int32_t rec[4], int32_t& inf,
                double h, bool& tab_ok) {
    const int64_t t0 = s->n2e_ptr[a], ni = s->n2e_ptr[a + 1] - t0;
    const int64_t p0 = s->npat_ptr[a], np = s->npat_ptr[a + 1] - p0;
    if (ni != G3_NI || np != G3_NP) return false;
    const int64_t n = S - 1;
    auto rel = [&](int r) { return 6 * (T.ec[r][0] + T.ec[r][1] * n + T.ec[r][2] * n * n) + T.ec[r][3]; };
    auto off = [&](const int8_t* o) { return o[0] + o[1] * S + o[2] * S * S; };
    const int64_t eb = int64_t(s->n2e[t0 + G3_NI - 1]) - rel(G3_NI - 1);
    if (eb < 0 || eb >= (int64_t(1) << 31)) return false;
    for (int r = 0; r < G3_NI; ++r) {
        const int64_t e = eb + rel(r);
        if (s->n2e[t0 + r] != e || s->cut_of_elem[e] >= 0) return false;
        for (int k = 0; k < 4; ++k)
            if (s->conn[e * 4 + owner_first(4, T.la[r], k)] != a + off(T.nd[r][k])) return false;
        if (s->elem_mat[e] < 0 || s->elem_mat[e] >= G3_MT) tab_ok = false;
    }
    if (tab_ok) {
        const double* xa = &s->coords[a * 3];
        for (int k = -1; k <= 1 && tab_ok; ++k)
            for (int j = -1; j <= 1 && tab_ok; ++j)
                for (int i = -1; i <= 1 && tab_ok; ++i) {
                    const double* xb = &s->coords[(a + i + j * S + k * S * S) * 3];
                    if (!(std::fabs(xb[0] - (xa[0] + i * h)) <= 1e-10 * h && std::fabs(xb[1] - (xa[1] + j * h)) <= 1e-10 * h &&
                          std::fabs(xb[2] - (xa[2] + k * h)) <= 1e-10 * h))
                        tab_ok = false;
                }
    }
    const int64_t kb = s->npat_map[p0 * 9];
    const int32_t dof = s->node_dof[a * 3];
    for (int i = 1; i < 3; ++i)
        if (s->node_dof[a * 3 + i] != dof + i) return false;
    for (int sl = 0; sl < G3_NP; ++sl) {
        if (s->npat_col[p0 + sl] != a + off(T.pd[sl])) return false;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                if (s->npat_map[(p0 + sl) * 9 + i * 3 + j] != kb + i * 3 * np + 3 * sl + j) return false;
    }
    rec[0] = int32_t(a);
    rec[1] = int32_t(kb);
    rec[2] = dof;
    rec[3] = int32_t(eb);
    inf = int32_t(S);
    return true;
}

int check_desc(const pg_problem_desc* s) {
    if (!s) { g_err = "null description"; return PG_E_ARG; }
    if (s->dim != 2 && s->dim != 3) { g_err = "dim must be 2 or 3"; return PG_E_ARG; }
    if (s->nen != s->dim + 1) { g_err = "only tri3 (2D) and tet4 (3D) parents are supported"; return PG_E_UNSUPPORTED; }
    if (s->nndof != 0 && s->nndof != s->dim) {
        g_err = "a description with one dof per node is the thermal assembler's (pg_therm_create)";
        return PG_E_UNSUPPORTED;
    }
    if (s->n_nodes <= 0 || s->n_elem <= 0 || s->n_free + s->n_pres != (s->n_nodes + s->n_enr) * s->dim) {
        g_err = "inconsistent node / dof counts";
        return PG_E_ARG;
    }
    if (!s->coords || !s->conn || !s->elem_mat || !s->cut_of_elem || !s->mats || !s->node_dof || !s->npat_row ||
        !s->npat_col || !s->npat_map || !s->n2e_ptr || !s->n2e) {
        g_err = "missing array in description";
        return PG_E_ARG;
    }
    if (s->n_cut > 0 && (!s->cut || !s->cut_elem)) { g_err = "missing cut records"; return PG_E_ARG; }
    if (s->n_mat <= 0 || s->n_mat > MAX_LDS_MATS) { g_err = "1 to 8 materials are supported"; return PG_E_UNSUPPORTED; }
    for (int m = 0; m < s->n_mat; ++m) {
        // LEI materials are isotropic (LinearElasticIsotropicMaterial): the
        // binned kernels use the closed form of B_a^T D B_b in lambda = D[0][1]
        // and mu = D[last][last], so D must have the isotropic pattern
        if (s->mats[m].type != PG_MAT_LEI) continue;
        if (s->mats[m].thermal_exp_set && !std::isfinite(s->mats[m].thermal_exp)) {
            g_err = "thermal expansion ratio must be finite";
            return PG_E_ARG;
        }
        const int ns = s->dim == 2 ? 3 : 6, nd = s->dim;
        const double* Dm = s->mats[m].D;
        const double lam = Dm[1], mu = Dm[ns * ns - 1];
        double mx = 0.0;
        for (int k = 0; k < ns * ns; ++k) mx = std::max(mx, std::fabs(Dm[k]));
        for (int k = 0; k < ns; ++k)
            for (int l = 0; l < ns; ++l) {
                const double iso = (k < nd && l < nd) ? lam + (k == l ? 2.0 * mu : 0.0) : (k == l ? mu : 0.0);
                if (std::fabs(Dm[k * ns + l] - iso) > 1e-12 * mx) {
                    g_err = "LEI material D matrix is not isotropic";
                    return PG_E_UNSUPPORTED;
                }
            }
    }
    for (int c = 0; c < s->n_cut; ++c) {
        const pg_cut_rec& r = s->cut[c];
        const int lim_n = s->dim == 2 ? MAXN2 : MAXN3, lim_q = s->dim == 2 ? MAXQ2 : MAXQ3;
        if (r.n_enr > PG_MAX_ENR || s->nen + r.n_enr > lim_n || r.n_child > PG_MAX_CHILD || r.n_coh > PG_MAX_COH ||
            r.n_vqp > lim_q) {
            g_err = "cut record exceeds the supported class sizes";
            return PG_E_UNSUPPORTED;
        }
        if (s->dim == 2 && r.n_child > 4) { g_err = "more than 4 children in a 2D cut record"; return PG_E_UNSUPPORTED; }
        int nq = 0;
        for (int k = 0; k < r.n_child; ++k) {
            int t = r.child_type[k];
            if (s->dim == 2 && t != PG_CHILD_TRI3 && t != PG_CHILD_QUAD4 && t != PG_CHILD_POLY5 && t != PG_CHILD_POLY6) {
                g_err = "bad 2D child";
                return PG_E_UNSUPPORTED;
            }
            if (s->dim == 3 && t != PG_CHILD_TET4) { g_err = "bad 3D child"; return PG_E_UNSUPPORTED; }
            nq += PG_CHILD_NQP(t);
        }
        if (nq != r.n_vqp) { g_err = "cut record qp count mismatch"; return PG_E_ARG; }
        if (r.n_coh && (s->dim == 2 ? r.coh_nn != 4 : (r.coh_nn != 6 && r.coh_nn != 8))) {
            g_err = "bad cohesive child";
            return PG_E_UNSUPPORTED;
        }
        if (r.isv_off < 0 || r.isv_off + r.n_coh * (r.coh_nn == 4 ? 2 : 4) > s->n_cqp) {
            g_err = "cut record ISV range out of bounds";
            return PG_E_ARG;
        }
        int nloc = s->nen + r.n_enr;
        for (int k = 0; k < r.n_child; ++k)
            for (int n = 0; n < PG_CHILD_NN(r.child_type[k]); ++n)
                if (r.child_node[PG_CHILD_STRIDE(s->dim) * k + n] >= nloc) {
                    g_err = "child node out of range";
                    return PG_E_ARG;
                }
        for (int k = 0; k < r.n_coh; ++k)
            for (int n = 0; n < r.coh_nn; ++n)
                if (r.coh_nodes[k][n] >= nloc) { g_err = "cohesive node out of range"; return PG_E_ARG; }
        for (int k = 0; k < r.n_enr; ++k)
            if (r.enr[k] < s->n_nodes || r.enr[k] >= s->n_nodes + s->n_enr) { g_err = "enrichment node out of range"; return PG_E_ARG; }
        for (int k = 0; k < r.n_child; ++k)
            if (r.child_mat[k] >= s->n_mat || s->mats[r.child_mat[k]].type != PG_MAT_LEI) { g_err = "bad child material"; return PG_E_ARG; }
        for (int k = 0; k < r.n_coh; ++k)
            if (r.coh_mat[k] >= s->n_mat || s->mats[r.coh_mat[k]].type == PG_MAT_LEI) { g_err = "bad cohesive material"; return PG_E_ARG; }
    }
    const int nall = s->n_nodes + s->n_enr;
    for (int e = 0; e < s->n_elem; ++e) {
        if (s->elem_mat[e] < 0 || s->elem_mat[e] >= s->n_mat || s->mats[s->elem_mat[e]].type != PG_MAT_LEI) {
            g_err = "bad element material";
            return PG_E_ARG;
        }
        int c = s->cut_of_elem[e];
        if (c < -1 || c >= s->n_cut || (c >= 0 && s->cut_elem[c] != e)) { g_err = "bad cut index"; return PG_E_ARG; }
    }
    for (int64_t k = 0; k < int64_t(s->n_elem) * s->nen; ++k)
        if (s->conn[k] < 0 || s->conn[k] >= s->n_nodes) { g_err = "connectivity out of range"; return PG_E_ARG; }
    for (int m = 0; m < s->n_mat; ++m)
        if (s->mats[m].type == PG_MAT_LEI && s->mats[m].nstr != (s->dim == 2 ? 3 : 6)) { g_err = "bad D size"; return PG_E_ARG; }
    const int64_t nnzt = s->nnz[0] + s->nnz[1] + s->nnz[2] + s->nnz[3];
    for (int64_t k = 0; k < s->n_pairs; ++k) {
        if (s->npat_row[k] < 0 || s->npat_row[k] >= nall || s->npat_col[k] < 0 || s->npat_col[k] >= nall) {
            g_err = "node pair out of range";
            return PG_E_ARG;
        }
    }
    for (int64_t k = 0; k < s->n_pairs * s->dim * s->dim; ++k)
        if (s->npat_map[k] < 0 || s->npat_map[k] >= nnzt) { g_err = "scatter index out of range"; return PG_E_ARG; }
    for (int64_t k = 0; k < int64_t(nall) * s->dim; ++k)
        if (s->node_dof[k] < 0 || s->node_dof[k] >= s->n_free + s->n_pres) { g_err = "dof out of range"; return PG_E_ARG; }
    if (s->n2e_ptr[0] != 0) { g_err = "bad incidence"; return PG_E_ARG; }
    for (int a = 0; a < nall; ++a)
        if (s->n2e_ptr[a + 1] < s->n2e_ptr[a]) { g_err = "bad incidence"; return PG_E_ARG; }
    for (int64_t k = 0; k < s->n2e_ptr[nall]; ++k)
        if (s->n2e[k] < 0 || s->n2e[k] >= s->n_elem) { g_err = "incidence out of range"; return PG_E_ARG; }
    return PG_OK;
}

template <typename T>
int alloc(pg_asm* h, T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    HIP_OK(hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T)));
    h->allocs.push_back(*p);
    h->dev_bytes += int64_t(n * sizeof(T));
    return PG_OK;
}

// Calls on one handle share its scratch buffers: a call on a stream other than
// the last call's waits for that call to complete, and every call records its
// own completion.
int order_after_last(pg_asm* h, hipStream_t st) {
    if (h->done_valid && st != h->done_st) HIP_OK(hipStreamWaitEvent(st, h->done, 0));
    return PG_OK;
}

int record_done(pg_asm* h, hipStream_t st) {
    HIP_OK(hipEventRecord(h->done, st));
    h->done_st = st;
    h->done_valid = true;
    return PG_OK;
}

// Node-group kernel launch: the block-table form for tri3 meshes unless the
// list form is asked for; tet4 meshes always use the list form.
template <int D, typename TC, typename TKC, int NG, int TV>
void launch_group(bool lists, int gbase, int64_t ibase, int ngroups, hipStream_t st, const DevProb& d,
                  const double* Un, const TKC* Kc, const double* Pc, double* K, double* P, unsigned flags) {
    constexpr int BS = GroupThreads<D, NG, TV>::value;
    if (ngroups <= 0) return;
    const unsigned grid = unsigned(ngroups);
    if constexpr (D == 2 && std::is_same<TC, TKC>::value) {
        if (!lists) {
            k_pair_group<D, TC, NG, TV><<<grid, BS, 0, st>>>(d, Un, Kc, Pc, K, P, flags, gbase, ibase);
            return;
        }
    }
    k_pair_lists<D, TC, TKC, NG, TV><<<grid, BS, 0, st>>>(d, Un, Kc, Pc, K, P, flags, gbase, ibase);
}

// Schedules: no PG_FP32 / PG_BATCHED -> plain FP64 (element kernel over all
// elements + staged node-pair reduction); PG_BATCHED -> class-binned FP64;
// PG_FP32 (with or without PG_BATCHED) -> class-binned mixed precision.
int launch(pg_asm* h, const double* dU, const double* dIc, double* dIt, unsigned flags, double* dK, double* dP,
           hipStream_t st) {
    const DevProb& d = h->d;
    const int64_t nall = int64_t(d.n_nodes) + d.n_enr;
    const int64_t nu = nall * d.dim;
    const bool fp32 = (flags & PG_FP32) != 0;
    const bool binned = fp32 || (flags & PG_BATCHED) != 0;
    if (int rc = order_after_last(h, st)) return rc;
    if (h->timing) HIP_OK(hipEventRecord(h->ev[0], st));
    // the error flag describes this assembly only (pg_asm_check_error)
    k_fill_int<<<1, 64, 0, st>>>(h->err, 0);
    k_permute_u<<<dim3(unsigned((nu + 255) / 256)), dim3(256), 0, st>>>(nu, d.node_dof, dU, h->Un);
    const unsigned gp = unsigned((d.n_pairs + h->pair_block - 1) / h->pair_block);
    if (!binned) {
        const unsigned gb = unsigned((d.n_elem + 255) / 256);
        if (d.dim == 2) {
            k_elem_plain<2><<<gb, 256, 0, st>>>(d, h->Un, h->stage);
            if (d.n_cut) k_cut_plain<2, false><<<unsigned(d.n_cut), 64, 0, st>>>(d, h->Un, dIc, dIt, h->stage_k, h->stage_p, h->err, nullptr);
            if (h->n_poly) k_cut_plain<2, true><<<unsigned(h->n_poly), 64, 0, st>>>(d, h->Un, dIc, dIt, h->stage_k, h->stage_p, h->err, h->poly_list);
        } else {
            k_elem_plain<3><<<gb, 256, 0, st>>>(d, h->Un, h->stage);
            if (d.n_cut) k_cut_plain<3, false><<<unsigned(d.n_cut), 64, 0, st>>>(d, h->Un, dIc, dIt, h->stage_k, h->stage_p, h->err, nullptr);
        }
        if (h->timing) HIP_OK(hipEventRecord(h->ev[1], st));
        if (h->timing) HIP_OK(hipEventRecord(h->ev[5], st));
        if (d.dim == 2) k_pair_staged<2><<<gp, h->pair_block, 0, st>>>(d, h->stage, h->stage_k, h->stage_p, dK, dP, flags);
        else k_pair_staged<3><<<gp, h->pair_block, 0, st>>>(d, h->stage, h->stage_k, h->stage_p, dK, dP, flags);
    } else {
        const int nall = d.n_nodes + d.n_enr;
        const bool cut64 = fp32 && (flags & PG_CUT64) != 0;  // cut elements staged in FP64
        int o = 0;
        while (o < 2 && h->gopt_ng[o] != h->group_nodes) ++o;
        // the groups that wait for the cut kernels: the same size, or the one
        // that group variant bits 5-6 name (1 + option index)
        const int rsel = (h->group_variant >> 5) & 3;
        const int orr = rsel ? rsel - 1 : o;
        const int nclean = h->nclean[o], nrest = h->ngroups[o] - h->nclean[o];
        const bool overlap = !(h->group_variant & 4) && nclean > 0 && nrest > 0;
        const bool tv0 = (h->group_variant & 1) != 0;
        // structured-grid nodes through k_pair_struct (off with group variant
        // bit 7, and with bit 3, which sends every node to the global path);
        // the generic groups of list 0 are then skipped
        const bool grid = h->n_struct > 0 && !(h->group_variant & (128 | 8));
        auto skip = [&](int oo) { return grid ? h->nstruct[oo] : 0; };
        auto grid_groups = [&](int oo, hipStream_t s, bool tv) {
            const int ng = h->gopt_ng[oo];
            const int ngr = int((h->n_struct + ng - 1) / ng);
            // group variant bits 8-9: each block takes 2^k consecutive groups
            const int per = 1 << ((h->group_variant >> 8) & 3);
            const unsigned grid_n = unsigned((ngr + per - 1) / per);
#define STRUCT1(TC, NGV, TVV) \
    k_pair_struct<TC, NGV, TVV><<<grid_n, GroupThreads<2, NGV, TVV>::value, 0, s>>>(d, h->stpl, h->srec, h->sinf, h->Un, dK, dP, flags, ngr, per)
#define STRUCT2(TC, NGV) \
    do { if (tv) STRUCT1(TC, NGV, 1); else STRUCT1(TC, NGV, 0); } while (0)
#define STRUCT3(TC) \
    do { if (ng == 8) STRUCT2(TC, 8); else if (ng == 16) STRUCT2(TC, 16); else STRUCT2(TC, 32); } while (0)
#define STRUCT1T(TC, NGV, TVV) \
    do { \
        if (std::is_same<TC, float>::value && h->stab3 && !(h->group_variant & 1024)) \
            k_pair_struct3<float, NGV, TVV, true><<<grid_n, GroupThreads<3, NGV, TVV>::value, 0, s>>>(d, h->stpl, h->stab3, h->srec, h->sinf, h->Un, dK, dP, flags, ngr, per); \
        else \
            k_pair_struct3<TC, NGV, TVV, false><<<grid_n, GroupThreads<3, NGV, TVV>::value, 0, s>>>(d, h->stpl, nullptr, h->srec, h->sinf, h->Un, dK, dP, flags, ngr, per); \
    } while (0)
#define STRUCT2T(TC, NGV) \
    do { if (tv) STRUCT1T(TC, NGV, 1); else STRUCT1T(TC, NGV, 0); } while (0)
#define STRUCT3T(TC) \
    do { if (ng == 2) STRUCT2T(TC, 2); else if (ng == 4) STRUCT2T(TC, 4); else STRUCT2T(TC, 8); } while (0)
            if (d.dim == 2) {
                if (fp32) STRUCT3(float);
                else STRUCT3(double);
            } else {
                if (fp32) STRUCT3T(float);
                else STRUCT3T(double);
            }
#undef STRUCT3T
#undef STRUCT2T
#undef STRUCT1T
#undef STRUCT3
#undef STRUCT2
#undef STRUCT1
        };
        // LEN node groups on stream S: node-table entries from BASE, incidence slots from ibase
#define GROUP1(DD, TC, TKC, KC, NGV, TVV, S, BASE, LEN) \
    launch_group<DD, TC, TKC, NGV, TVV>((h->group_variant & 2) != 0, BASE, ibase, LEN, S, d, h->Un, KC, h->Pc, dK, dP, \
                                        flags | ((h->group_variant & 8) ? PG_GLOBAL_PATH : 0u))
#define GROUP(DD, TC, TKC, KC, NGV, S, BASE, LEN)                     \
    do {                                                             \
        if (tv) GROUP1(DD, TC, TKC, KC, NGV, 1, S, BASE, LEN);       \
        else GROUP1(DD, TC, TKC, KC, NGV, 0, S, BASE, LEN);          \
    } while (0)
#define GROUPS2(TC, TKC, KC, S, BASE, LEN)                              \
    do {                                                                \
        if (ng == 8) GROUP(2, TC, TKC, KC, 8, S, BASE, LEN);            \
        else if (ng == 16) GROUP(2, TC, TKC, KC, 16, S, BASE, LEN);     \
        else GROUP(2, TC, TKC, KC, 32, S, BASE, LEN);                   \
    } while (0)
#define GROUPS3(TC, TKC, KC, S, BASE, LEN)                              \
    do {                                                                \
        if (ng == 2) GROUP(3, TC, TKC, KC, 2, S, BASE, LEN);            \
        else if (ng == 4) GROUP(3, TC, TKC, KC, 4, S, BASE, LEN);       \
        else GROUP(3, TC, TKC, KC, 8, S, BASE, LEN);                    \
    } while (0)
        // node groups first .. first + len - 1 of the list of this group size,
        // tv: two threads per node pair
        auto groups = [&](int oo, hipStream_t s, int first, int len, bool tv) {
            const int ng = h->gopt_ng[oo];
            const int base = h->goff[oo] + first * ng;
            const int64_t ibase = h->gioff[oo] + int64_t(first) * (d.dim == 2 ? 7 : 24) * ng;
            if (d.dim == 2) {
                if (!fp32) GROUPS2(double, double, h->Kc64, s, base, len);
                else if (cut64) GROUPS2(float, double, h->Kc64, s, base, len);
                else GROUPS2(float, float, h->Kc32, s, base, len);
            } else {
                if (!fp32) GROUPS3(double, double, h->Kc64, s, base, len);
                else if (cut64) GROUPS3(float, double, h->Kc64, s, base, len);
                else GROUPS3(float, float, h->Kc32, s, base, len);
            }
        };
#undef GROUPS2
#undef GROUPS3
#undef GROUP
#undef GROUP1
        if (overlap) {
            // nodes without cut incidences need no staged cut tangents: their
            // groups run on the side stream while the cut kernels run on st
            HIP_OK(hipEventRecord(h->sync[0], st));
            HIP_OK(hipStreamWaitEvent(h->side, h->sync[0], 0));
            if (h->timing) HIP_OK(hipEventRecord(h->ev[3], h->side));
            if (grid) grid_groups(o, h->side, tv0);
            groups(o, h->side, skip(o), nclean - skip(o), tv0);
            if (h->timing) HIP_OK(hipEventRecord(h->ev[4], h->side));
            HIP_OK(hipEventRecord(h->sync[1], h->side));
        }
        // cut elements, binned by class: small single-inclusion records share a
        // wavefront (EPW per wave); the general class keeps the full limits
        const int n0 = h->n_class[0], n1 = h->n_class[1];
        const int l0 = 0, l1 = n0;  // class positions of the two classes
        const int cv = h->cut_variant;
        // a mesh without cohesive qps (n_cqp == 0: check_desc then admits no
        // cohesive child) launches the instances with the cohesive code compiled out
        const bool coh = d.n_cqp > 0;
#define CUT_SMALL2(TK, KST, EPW, OCC)                                                                             \
    (coh ? k_cut<2, TK, EPW, 7, 8, 2, OCC><<<(n0 + EPW - 1) / EPW, 64, 0, st>>>(d, l0, n0, h->Un, dIc, dIt, KST, \
                                                                               h->Pc, h->err)                    \
         : k_cut<2, TK, EPW, 7, 8, 0, OCC><<<(n0 + EPW - 1) / EPW, 64, 0, st>>>(d, l0, n0, h->Un, dIc, dIt, KST, \
                                                                               h->Pc, h->err))
#define CUT_GEN(D_, TK, KST, EPW, MAXN, MAXQ)                                                                       \
    (coh ? k_cut<D_, TK, EPW, MAXN, MAXQ, PG_MAX_COH * 4, 1><<<(n1 + EPW - 1) / EPW, 64, 0, st>>>(                 \
               d, l1, n1, h->Un, dIc, dIt, KST, h->Pc, h->err)                                                     \
         : k_cut<D_, TK, EPW, MAXN, MAXQ, 0, 1><<<(n1 + EPW - 1) / EPW, 64, 0, st>>>(d, l1, n1, h->Un, dIc, dIt, KST, \
                                                                                   h->Pc, h->err))
        if (d.dim == 2) {
            if (n0) {
                if (fp32 && !cut64) {
                    if (cv == 1) CUT_SMALL2(float, h->Kc32, 4, 3);
                    else if (cv == 2) CUT_SMALL2(float, h->Kc32, 2, 1);
                    else if (cv == 3) CUT_SMALL2(float, h->Kc32, 8, 1);
                    else CUT_SMALL2(float, h->Kc32, 4, 1);
                } else {
                    if (cv == 1) CUT_SMALL2(double, h->Kc64, 4, 3);
                    else if (cv == 2) CUT_SMALL2(double, h->Kc64, 2, 1);
                    else if (cv == 3) CUT_SMALL2(double, h->Kc64, 8, 1);
                    else CUT_SMALL2(double, h->Kc64, 4, 1);
                }
            }
            if (n1) {
                if (fp32 && !cut64) CUT_GEN(2, float, h->Kc32, 2, MAXN2, MAXQ2);
                else CUT_GEN(2, double, h->Kc64, 2, MAXN2, MAXQ2);
            }
        } else {
            // small class: elements per wave and the register cap (waves per
            // EU) by the cut variant
#define CUT_SMALL3(TK, KST, EPW, OCC)                                                                              \
    (coh ? k_cut<3, TK, EPW, 8, 6, 4, OCC><<<(n0 + EPW - 1) / EPW, 64, 0, st>>>(d, l0, n0, h->Un, dIc, dIt, KST,   \
                                                                               h->Pc, h->err)                      \
         : k_cut<3, TK, EPW, 8, 6, 0, OCC><<<(n0 + EPW - 1) / EPW, 64, 0, st>>>(d, l0, n0, h->Un, dIc, dIt, KST,   \
                                                                               h->Pc, h->err))
            if (n0) {
                if (fp32 && !cut64) {
                    if (cv == 1) CUT_SMALL3(float, h->Kc32, 4, 3);
                    else if (cv == 2) CUT_SMALL3(float, h->Kc32, 2, 1);
                    else if (cv == 3) CUT_SMALL3(float, h->Kc32, 8, 1);
                    else CUT_SMALL3(float, h->Kc32, 4, 1);
                } else {
                    if (cv == 1) CUT_SMALL3(double, h->Kc64, 4, 3);
                    else if (cv == 2) CUT_SMALL3(double, h->Kc64, 2, 1);
                    else if (cv == 3) CUT_SMALL3(double, h->Kc64, 8, 1);
                    else CUT_SMALL3(double, h->Kc64, 4, 1);
                }
            }
#undef CUT_SMALL3
            if (n1) {
                if (fp32 && !cut64) CUT_GEN(3, float, h->Kc32, 1, MAXN3, PG_MAX_CHILD);
                else CUT_GEN(3, double, h->Kc64, 1, MAXN3, PG_MAX_CHILD);
            }
        }
#undef CUT_SMALL2
#undef CUT_GEN
        if (h->timing) HIP_OK(hipEventRecord(h->ev[1], st));
        if (h->timing) HIP_OK(hipEventRecord(h->ev[5], st));
        // the groups of enrichment nodes (the tail of the list) may take twice
        // the threads per pair (group variant bit 4): they have few incidences
        // and many pairs.  With overlap they start as soon as the cut kernels
        // end, while the side stream's groups may still run (the two sets of
        // groups write disjoint rows of K and entries of P); st joins the side
        // stream after them.
        if (!overlap && grid) grid_groups(orr, st, tv0);
        const int first = overlap ? h->nclean[orr] : skip(orr), nenr = h->nenrg[orr];
        const int nreg = h->ngroups[orr] - first - nenr;
        groups(orr, st, first, nreg, tv0);
        groups(orr, st, first + nreg, nenr, tv0 || (h->group_variant & 16) != 0);
        if (h->timing) HIP_OK(hipEventRecord(h->ev[6], st));
        if (overlap) HIP_OK(hipStreamWaitEvent(st, h->sync[1], 0));
    }
    HIP_OK(hipGetLastError());
    if (h->timing) HIP_OK(hipEventRecord(h->ev[2], st));
    return record_done(h, st);
}

}  // namespace

extern "C" {

const char* pg_asm_last_error(void) { return g_err.c_str(); }

int pg_asm_create(const pg_problem_desc* s, pg_asm** out) {
    if (!out) { g_err = "null output handle"; return PG_E_ARG; }
    *out = nullptr;
    int rc = check_desc(s);
    if (rc) return rc;
    pg_asm* h = new (std::nothrow) pg_asm();
    if (!h) { g_err = "out of memory"; return PG_E_NOMEM; }
    DevProb& d = h->d;
    d.dim = s->dim; d.nen = s->nen; d.n_nodes = s->n_nodes; d.n_enr = s->n_enr; d.n_elem = s->n_elem;
    d.n_cut = s->n_cut; d.n_free = s->n_free; d.n_pres = s->n_pres; d.cohesive = s->cohesive; d.n_cqp = s->n_cqp;
    d.n_pairs = s->n_pairs;
    h->n_dof = int64_t(s->n_free) + s->n_pres;
    for (int b = 0; b < 4; ++b) { h->nnz[b] = s->nnz[b]; h->nnz_total += s->nnz[b]; }
    const int64_t nall = int64_t(s->n_nodes) + s->n_enr;
    std::vector<int64_t> koff(size_t(s->n_cut) + 1), poff(size_t(s->n_cut) + 1);
    koff[0] = poff[0] = 0;
    for (int c = 0; c < s->n_cut; ++c) {
        int64_t nd = int64_t(s->nen + s->cut[c].n_enr) * s->dim;
        koff[c + 1] = koff[c] + nd * nd;
        poff[c + 1] = poff[c] + nd;
    }
    h->cut_k_total = koff[s->n_cut];
    h->cut_p_total = poff[s->n_cut];
    if (h->cut_p_total >= (int64_t(1) << 31)) {
        delete h;
        g_err = "cut-element dofs exceed 2^31";
        return PG_E_UNSUPPORTED;
    }
    d.n_mat = s->n_mat;
    for (int m = 0; m < s->n_mat; ++m) {
        h->thermal_set[m] = s->mats[m].type == PG_MAT_LEI && s->mats[m].thermal_exp_set != 0;
        h->thermal_exp[m] = h->thermal_set[m] ? s->mats[m].thermal_exp : 0.0;
        d.eth[m] = 0.0;
        h->mat_type[m] = s->mats[m].type;
    }
    if (s->cohesive)
        for (int c = 0; c < s->n_cut; ++c)
            for (int k = 0; k < s->cut[c].n_coh; ++k) h->coh_law_mask |= 1u << s->mats[s->cut[c].coh_mat[k]].type;
    h->n_cells = int64_t(s->n_elem) - s->n_cut;
    for (int c = 0; c < s->n_cut; ++c) h->n_cells += s->cut[c].n_child;
    // incidence records in n2e order (host only; the kernels read them regrouped
    // per node group below), nen + 1 words, the last a tag.  Uncut
    // element: its nodes, tag -1 - material.  Cut element: the staged row of
    // the owner node a -- block index (in d x d blocks) of block (la, 0) of the
    // element's tangent, the offset of a's residual entries, la | nloc << 8
    // (la = 255: a is not a node of the element) -- tag = cut index
    if (h->cut_k_total / (int64_t(s->dim) * s->dim) >= (int64_t(1) << 32)) {
        delete h;
        g_err = "cut-element tangent blocks exceed 2^32";
        return PG_E_UNSUPPORTED;
    }
    const int64_t n_inc = s->n2e_ptr[nall];
    std::vector<int32_t> incv(size_t(n_inc) * (s->nen + 1), 0);
    for (int64_t a = 0; a < nall; ++a)
        for (int64_t t = s->n2e_ptr[a]; t < s->n2e_ptr[a + 1]; ++t) {
            const int e = s->n2e[t];
            int32_t* r = &incv[size_t(t) * (s->nen + 1)];
            const int c = s->cut_of_elem[e];
            if (c < 0) {
                for (int n = 0; n < s->nen; ++n) r[n] = s->conn[int64_t(e) * s->nen + n];
                r[s->nen] = -1 - s->elem_mat[e];
                continue;
            }
            const int nloc = s->nen + s->cut[c].n_enr;
            int la = -1;
            for (int n = 0; n < s->nen && la < 0; ++n)
                if (s->conn[int64_t(e) * s->nen + n] == a) la = n;
            for (int k = 0; k < s->cut[c].n_enr && la < 0; ++k)
                if (s->cut[c].enr[k] == a) la = s->nen + k;
            const int lr = la < 0 ? 0 : la;
            r[0] = int32_t(uint32_t(uint64_t(koff[c] / (int64_t(s->dim) * s->dim) + int64_t(lr) * nloc)));
            r[1] = int32_t(poff[c] + int64_t(lr) * s->dim);
            r[2] = (la < 0 ? 255 : la) | (nloc << 8);
            r[s->nen] = c;
        }
    // Static contribution slots of the node-group kernel.  For incidence t of
    // node a (element e) and each other local node b of e, in local order: the
    // slot of pair (a, b) within a's row of the node-pair pattern, and the rank
    // of e among the incidences of a that contain b (= position in ascending
    // element order).  9 bits each, packed from bit 0; bit 63 marks an
    // incidence that does not fit (it then takes the global-memory path).
    std::vector<uint64_t> slotv(size_t(n_inc), 0);
    {
        std::vector<int> cntv;
        for (int64_t a = 0; a < nall; ++a) {
            const int p0 = s->npat_ptr[a], p1 = s->npat_ptr[a + 1];
            cntv.assign(size_t(p1 - p0), 0);
            for (int64_t t = s->n2e_ptr[a]; t < s->n2e_ptr[a + 1]; ++t) {
                const int e = s->n2e[t];
                const int c = s->cut_of_elem[e];
                int32_t ln[4 + PG_MAX_ENR];
                int nloc = 0;
                for (int n = 0; n < s->nen; ++n) ln[nloc++] = s->conn[int64_t(e) * s->nen + n];
                if (c >= 0)
                    for (int k = 0; k < s->cut[c].n_enr; ++k) ln[nloc++] = s->cut[c].enr[k];
                uint64_t w = 0;
                bool ok = nloc <= 8;
                int k = 0;
                for (int m = 0; m < nloc && ok; ++m) {
                    if (ln[m] == a) continue;
                    const int32_t* b0 = s->npat_col + p0;
                    const int32_t* it = std::lower_bound(b0, s->npat_col + p1, ln[m]);
                    const int slot = int(it - b0);
                    if (it == s->npat_col + p1 || *it != ln[m]) { ok = false; break; }
                    const int rank = cntv[size_t(slot)]++;
                    if (slot >= 64 || rank >= 8) { ok = false; break; }
                    w |= uint64_t(slot | (rank << 6)) << (9 * k);
                    ++k;
                }
                slotv[size_t(t)] = ok ? w : (uint64_t(1) << 63);
            }
        }
    }
    auto fail = [&](int code) { pg_asm_destroy(h); return code; };
#define UP(field, src, n, T)                                   \
    do {                                                       \
        T* tmp_;                                               \
        if ((rc = alloc(h, &tmp_, size_t(n)))) return fail(rc); \
        if (size_t(n) > 0 && hipMemcpy(tmp_, src, size_t(n) * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) { \
            g_err = "hipMemcpy of " #field " failed"; return fail(PG_E_HIP); } \
        d.field = tmp_;                                        \
    } while (0)
    UP(coords, s->coords, nall * s->dim, double);
    UP(conn, s->conn, int64_t(s->n_elem) * s->nen, int32_t);
    UP(elem_mat, s->elem_mat, s->n_elem, int32_t);
    UP(cut_of_elem, s->cut_of_elem, s->n_elem, int32_t);
    UP(cut_elem, s->cut_elem, s->n_cut, int32_t);
    UP(cut, s->cut, s->n_cut, pg_cut_rec);
    UP(mats, s->mats, s->n_mat, pg_material);
    UP(node_dof, s->node_dof, nall * s->dim, int32_t);
    UP(npat_ptr, s->npat_ptr, nall + 1, int32_t);
    UP(npat_row, s->npat_row, s->n_pairs, int32_t);
    UP(npat_col, s->npat_col, s->n_pairs, int32_t);
    UP(npat_map, s->npat_map, s->n_pairs * s->dim * s->dim, int32_t);
    UP(n2e_ptr, s->n2e_ptr, nall + 1, int32_t);
    UP(n2e, s->n2e, s->n2e_ptr[nall], int32_t);
    UP(cut_koff, koff.data(), s->n_cut + 1, int64_t);
    UP(cut_poff, poff.data(), s->n_cut + 1, int64_t);
    {
        std::vector<int32_t> pl;
        for (int32_t c = 0; c < s->n_cut; ++c)
            for (int k = 0; k < s->cut[c].n_child; ++k)
                if (s->cut[c].child_type[k] == PG_CHILD_POLY5 || s->cut[c].child_type[k] == PG_CHILD_POLY6) {
                    pl.push_back(c);
                    break;
                }
        h->n_poly = int32_t(pl.size());
        if (!pl.empty()) {
            if ((rc = alloc(h, &h->poly_list, pl.size()))) return fail(rc);
            if (hipMemcpy(h->poly_list, pl.data(), pl.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
                g_err = "hipMemcpy of the polygon list failed";
                return fail(PG_E_HIP);
            }
        }
    }
    {
        // Node lists of the node-group kernels, each in mesh order: (0) the
        // nodes of the structured grid whose incidences are all uncut (tri3;
        // k_pair_struct takes them from their records), (1) the other regular
        // nodes none of whose elements is cut, (2) the other regular nodes,
        // (3) the enrichment nodes.  Lists 0 and 1 need no staged cut tangent,
        // so their groups run while the cut kernels run.  Each list is cut into
        // blocks of B nodes (B = the smallest group size; the last block
        // padded) and a group of NG nodes takes NG / B consecutive blocks of one
        // list, so every group size covers the same node sets, and the groups
        // that wait for the cut kernels may use another size than the groups
        // that overlap them.  The generic tables cover list 0 too (the
        // structured path can be switched off).
        std::vector<char> dirty(size_t(nall), 0);
        for (int64_t a = 0; a < nall; ++a)
            for (int64_t t = s->n2e_ptr[a]; t < s->n2e_ptr[a + 1] && !dirty[a]; ++t)
                dirty[a] = s->cut_of_elem[s->n2e[t]] >= 0;
        std::vector<int32_t> tab(size_t(nall) * 4);
        for (int64_t a = 0; a < nall; ++a) {
            const int32_t ic = s->n2e_ptr[a + 1] - s->n2e_ptr[a], qc = s->npat_ptr[a + 1] - s->npat_ptr[a];
            if (ic > 0xffff || qc > 0x7fff) {
                g_err = "node with too many incidences or node pairs";
                return fail(PG_E_UNSUPPORTED);
            }
            tab[size_t(a) * 4 + 0] = int32_t(a);
            tab[size_t(a) * 4 + 1] = s->n2e_ptr[a];
            tab[size_t(a) * 4 + 2] = s->npat_ptr[a];
            tab[size_t(a) * 4 + 3] = ic | (qc << 16);
        }
        std::vector<int32_t> srec_v, sinf_v;
        std::vector<char> grid(size_t(nall), 0);
        GridTplHost T;
        GridTpl3 T3;
        if (s->dim == 2 && build_grid_tpl(T)) {
            for (int64_t a = 0; a < s->n_nodes; ++a) {
                int32_t rec[4], inf = 0;
                if (dirty[a] || !grid_node(s, T, a, rec, inf)) continue;
                grid[a] = 1;
                srec_v.insert(srec_v.end(), rec, rec + 4);
                sinf_v.push_back(inf);
            }
        }
        if (s->dim == 3 && build_grid_tpl3(T3)) {
            // the generator's cube has S^3 nodes, S = n + 1
            int64_t S = int64_t(std::llround(std::cbrt(double(s->n_nodes))));
            // the table of element tangents needs one spacing h (the first cube's)
            const double h3 = s->n_nodes > 1 ? s->coords[3] - s->coords[0] : 0.0;
            bool tab3_ok = h3 > 0.0;
            if (S >= 3 && S * S * S == s->n_nodes && S < (int64_t(1) << 24))
                for (int64_t a = 0; a < s->n_nodes; ++a) {
                    int32_t rec[4], inf = 0;
                    if (dirty[a] || !grid_node3(s, T3, S, a, rec, inf, h3, tab3_ok)) continue;
                    grid[a] = 1;
                    srec_v.insert(srec_v.end(), rec, rec + 4);
                    sinf_v.push_back(inf);
                }
            if (tab3_ok && !sinf_v.empty()) {
                GridTab3 G3;
                build_grid_tab3(T3, s, h3, G3);
                int4* g_;
                if ((rc = alloc(h, &g_, sizeof(GridTab3) / 16))) return fail(rc);
                if (hipMemcpy(g_, &G3, sizeof(GridTab3), hipMemcpyHostToDevice) != hipSuccess) {
                    g_err = "hipMemcpy of the grid tangent table failed";
                    return fail(PG_E_HIP);
                }
                h->stab3 = g_;
            }
        }
        h->n_struct = int64_t(sinf_v.size());
        // padding records (node -1) up to a whole number of the largest group
        while (sinf_v.size() % 32) {
            srec_v.insert(srec_v.end(), {-1, 0, 0, 0});
            sinf_v.push_back(0);
        }
        const int opts2[3] = {8, 16, 32}, opts3[3] = {2, 4, 8};
        const int B = s->dim == 2 ? opts2[0] : opts3[0];
        std::vector<int64_t> lst[4];
        for (int64_t a = 0; a < nall; ++a)
            lst[grid[a] ? 0 : (a >= s->n_nodes ? 3 : (dirty[a] ? 2 : 1))].push_back(a);
        std::vector<int32_t> gl;
        for (int o = 0; o < 3; ++o) {
            const int NGo = s->dim == 2 ? opts2[o] : opts3[o], per = NGo / B;
            int64_t ngl[4], ngt = 0;
            for (int l = 0; l < 4; ++l) {
                const int64_t nb = (int64_t(lst[l].size()) + B - 1) / B;
                ngl[l] = (nb + per - 1) / per;
                ngt += ngl[l];
            }
            if (int64_t(gl.size()) / 4 + ngt * NGo >= (int64_t(1) << 31)) {
                g_err = "node-group tables exceed 2^31 entries";
                return fail(PG_E_UNSUPPORTED);
            }
            h->gopt_ng[o] = NGo;
            h->goff[o] = int(gl.size() / 4);
            h->ngroups[o] = int(ngt);
            h->nstruct[o] = int(ngl[0]);
            h->nclean[o] = int(ngl[0] + ngl[1]);
            h->nenrg[o] = int(ngl[3]);
            for (int l = 0; l < 4; ++l)
                for (int64_t k = 0; k < ngl[l] * NGo; ++k) {
                    if (k < int64_t(lst[l].size())) {
                        const int64_t a = lst[l][size_t(k)];
                        gl.insert(gl.end(), &tab[size_t(a) * 4], &tab[size_t(a) * 4] + 4);
                    } else {
                        gl.insert(gl.end(), {0, 0, 0, 0});  // padding: no incidence, no pair
                    }
                }
        }
        {
            int32_t* t_;
            const size_t tb = s->dim == 2 ? sizeof(GridTpl) : sizeof(GridTpl3);
            const void* ts = s->dim == 2 ? static_cast<const void*>(&T.t) : static_cast<const void*>(&T3);
            if ((rc = alloc(h, &t_, tb / 4))) return fail(rc);
            if (hipMemcpy(t_, ts, tb, hipMemcpyHostToDevice) != hipSuccess) {
                g_err = "hipMemcpy of the grid template failed";
                return fail(PG_E_HIP);
            }
            h->stpl = t_;
            int4* r_;
            if ((rc = alloc(h, &r_, sinf_v.size()))) return fail(rc);
            if (!sinf_v.empty() && hipMemcpy(r_, srec_v.data(), srec_v.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
                g_err = "hipMemcpy of the grid records failed";
                return fail(PG_E_HIP);
            }
            h->srec = r_;
            int32_t* i_;
            if ((rc = alloc(h, &i_, sinf_v.size()))) return fail(rc);
            if (!sinf_v.empty() && hipMemcpy(i_, sinf_v.data(), sinf_v.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
                g_err = "hipMemcpy of the grid records failed";
                return fail(PG_E_HIP);
            }
            h->sinf = i_;
        }
        UP(gtab, gl.data(), gl.size(), int32_t);
        {
            const int W = s->nen + 1;  // words per record
            std::vector<int32_t> gi;
            std::vector<uint64_t> gs;
            for (int o = 0; o < 3; ++o) {
                const int NGo = h->gopt_ng[o], MAXIo = (s->dim == 2 ? 7 : 24) * NGo;
                h->gioff[o] = int64_t(gs.size());
                gi.resize(gi.size() + size_t(h->ngroups[o]) * MAXIo * W, 0);
                gs.resize(gs.size() + size_t(h->ngroups[o]) * MAXIo, 0);
                for (int g = 0; g < h->ngroups[o]; ++g) {
                    const size_t base = size_t(h->gioff[o]) + size_t(g) * MAXIo;
                    int i = 0;
                    for (int lo = 0; lo < NGo; ++lo) {
                        const int32_t* e4 = &gl[(size_t(h->goff[o]) + size_t(g) * NGo + lo) * 4];
                        const int ic = e4[3] & 0xffff;
                        if (ic == 0) continue;  // padding entry past the last node
                        const int64_t a = e4[0];
                        for (int r = 0; r < ic; ++r, ++i) {
                            if (i >= MAXIo) break;  // the group overflows: it takes the global path
                            const int64_t t = e4[1] + r;
                            const int32_t* src = &incv[size_t(t) * W];
                            int32_t* dst = &gi[(base + i) * W];
                            const int rk = std::min(r, 31);
                            if (src[s->nen] < 0) {
                                // uncut: the element's nodes owner first, in the
                                // even permutation owner_first(la, .) of its node
                                // order, and the slot fields in that order
                                int la = 0;
                                for (int n = 0; n < s->nen; ++n)
                                    if (src[n] == a) la = n;
                                const uint64_t w0 = slotv[size_t(t)];
                                uint64_t w1 = w0 & (uint64_t(1) << 63);
                                for (int n = 0; n < s->nen; ++n) {
                                    const int m = owner_first(s->nen, la, n);
                                    dst[n] = src[m];
                                    if (n > 0 && !(w0 >> 63)) {
                                        const int k0 = m < la ? m : m - 1;  // field of m in local order
                                        w1 |= ((w0 >> (9 * k0)) & 511u) << (9 * (n - 1));
                                    }
                                }
                                dst[s->nen] = int32_t(0x80000000u | uint32_t(-1 - src[s->nen]) | uint32_t(lo) << 3 |
                                                      uint32_t(rk) << 8);
                                gs[base + i] = w1;
                                continue;
                            } else {
                                dst[0] = src[0]; dst[1] = src[1];
                                dst[2] = (src[2] & 0xffff) | (lo << 16) | (rk << 21);
                                dst[s->nen] = src[s->nen];
                            }
                            gs[base + i] = slotv[size_t(t)];
                        }
                    }
                }
            }
            UP(incg, gi.data(), gi.size(), int32_t);
            UP(slotg, gs.data(), gs.size(), uint64_t);
        }
    }
#undef UP
    {
        std::vector<int32_t> small, general;
        for (int c = 0; c < s->n_cut; ++c) {
            const pg_cut_rec& r = s->cut[c];
            const int nloc = s->nen + r.n_enr;
            const int ncq = r.n_coh * (r.coh_nn == 4 ? 2 : 4);
            const bool is_small = s->dim == 2 ? (nloc <= 7 && r.n_vqp <= 8 && ncq <= 2)
                                              : (nloc <= 8 && r.n_vqp <= 30 && ncq <= 4);
            (is_small ? small : general).push_back(c);
        }
        h->n_class[0] = int(small.size());
        h->n_class[1] = int(general.size());
        small.insert(small.end(), general.begin(), general.end());
        int32_t* tmp_;
        if ((rc = alloc(h, &tmp_, small.size()))) return fail(rc);
        if (!small.empty() && hipMemcpy(tmp_, small.data(), small.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
            g_err = "hipMemcpy of cut classes failed";
            return fail(PG_E_HIP);
        }
        h->class_list = tmp_;
        // the records, node lists and offsets in class order
        std::vector<pg_cut_rec> cr(small.size());
        std::vector<int32_t> cn(small.size() * CUT_NODES, -1);
        std::vector<int64_t> ck(small.size()), cp(small.size());
        for (size_t k = 0; k < small.size(); ++k) {
            const int c = small[k];
            cr[k] = s->cut[c];
            const int e = s->cut_elem[c];
            for (int n = 0; n < s->nen; ++n) cn[k * CUT_NODES + n] = s->conn[int64_t(e) * s->nen + n];
            for (int n = 0; n < s->cut[c].n_enr && s->nen + n < CUT_NODES; ++n)
                cn[k * CUT_NODES + s->nen + n] = s->cut[c].enr[n];
            ck[k] = koff[c];
            cp[k] = poff[c];
        }
        auto upload = [&](auto** dst, const auto& v) -> int {
            using T = typename std::remove_reference<decltype(v)>::type::value_type;
            T* q;
            if (int r = alloc(h, &q, v.size())) return r;
            if (!v.empty() && hipMemcpy(q, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
                g_err = "hipMemcpy of the class-ordered cut data failed";
                return PG_E_HIP;
            }
            *dst = q;
            return PG_OK;
        };
        pg_cut_rec* q0 = nullptr;
        int32_t* q1 = nullptr;
        int64_t *q2 = nullptr, *q3 = nullptr;
        if ((rc = upload(&q0, cr)) || (rc = upload(&q1, cn)) || (rc = upload(&q2, ck)) || (rc = upload(&q3, cp)))
            return fail(rc);
        d.ccut = q0;
        d.cnodes = q1;
        d.ckoff = q2;
        d.cpoff = q3;
    }
    const int ndp = s->nen * s->dim;
    if ((rc = alloc(h, &h->Un, size_t(nall * s->dim)))) return fail(rc);
    (void)ndp;
    {
        const int64_t nblk = (int64_t(s->n_elem) + 63) / 64;
        const int64_t W = s->dim == 2 ? PlainStage<2>::W : PlainStage<3>::W;
        if ((rc = alloc(h, &h->stage, size_t(nblk * W)))) return fail(rc);
        if ((rc = alloc(h, &h->stage_k, size_t(h->cut_k_total)))) return fail(rc);
        if ((rc = alloc(h, &h->stage_p, size_t(h->cut_p_total)))) return fail(rc);
    }
    // the node-group kernels read a staged row's first 8 blocks at once: pad
    if ((rc = alloc(h, &h->Kc64, size_t(h->cut_k_total) + KC_PAD))) return fail(rc);
    if ((rc = alloc(h, &h->Kc32, size_t(h->cut_k_total) + KC_PAD))) return fail(rc);
    if ((rc = alloc(h, &h->Pc, size_t(h->cut_p_total)))) return fail(rc);
    if ((rc = alloc(h, &h->err, 1))) return fail(rc);
    if (hipMemset(h->err, 0, sizeof(int)) != hipSuccess) { g_err = "hipMemset failed"; return fail(PG_E_HIP); }
    if (hipStreamCreate(&h->stream) != hipSuccess) { g_err = "hipStreamCreate failed"; return fail(PG_E_HIP); }
    if (hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) != hipSuccess) {
        g_err = "hipStreamCreateWithFlags failed";
        return fail(PG_E_HIP);
    }
    for (auto& e : h->ev)
        if (hipEventCreate(&e) != hipSuccess) { g_err = "hipEventCreate failed"; return fail(PG_E_HIP); }
    for (auto& e : h->sync)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) { g_err = "hipEventCreate failed"; return fail(PG_E_HIP); }
    if (hipEventCreateWithFlags(&h->done, hipEventDisableTiming) != hipSuccess) {
        g_err = "hipEventCreate failed";
        return fail(PG_E_HIP);
    }
    // algorithmic bytes of one K+R assembly
    int64_t B = 4LL * s->nen * s->n_elem + 8LL * s->dim * nall + 8LL * h->n_dof + int64_t(sizeof(pg_cut_rec)) * s->n_cut +
                16LL * s->n_cqp + 12LL * h->nnz_total + 4LL * (int64_t(s->n_free) + s->n_pres + 2) + 8LL * h->n_dof;
    h->alg_bytes = B;
    *out = h;
    return PG_OK;
}

int pg_asm_destroy(pg_asm* h) {
    if (!h) return PG_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void* p : h->allocs) (void)hipFree(p);
    for (void* p : {(void*)h->dU, (void*)h->dK, (void*)h->dP, (void*)h->dIc, (void*)h->dIt})
        if (p) (void)hipFree(p);
    if (h->side) (void)hipStreamSynchronize(h->side);
    for (auto& e : h->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : h->sync)
        if (e) (void)hipEventDestroy(e);
    if (h->done) {
        (void)hipEventSynchronize(h->done);
        (void)hipEventDestroy(h->done);
    }
    if (h->side) (void)hipStreamDestroy(h->side);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PG_OK;
}

int pg_asm_sizes(const pg_asm* h, int64_t* n_dof, int64_t* nnz_total, int64_t* n_cqp) {
    if (!h) { g_err = "null handle"; return PG_E_ARG; }
    if (n_dof) *n_dof = h->n_dof;
    if (nnz_total) *nnz_total = h->nnz_total;
    if (n_cqp) *n_cqp = h->d.n_cqp;
    return PG_OK;
}

int64_t pg_asm_algorithmic_bytes(const pg_asm* h) { return h ? h->alg_bytes : -1; }

int64_t pg_asm_device_bytes(const pg_asm* h) { return h ? h->dev_bytes : -1; }

int64_t pg_asm_grid_nodes(const pg_asm* h) { return h ? h->n_struct : -1; }

int pg_asm_set_timing(pg_asm* h, int enable) {
    if (!h) { g_err = "null handle"; return PG_E_ARG; }
    h->timing = enable != 0;
    return PG_OK;
}

int pg_asm_cohesive_openings(pg_asm* h, const double* dU, const double* centers, double* out, int64_t n_out,
                             void* stream) {
    if (!h) { g_err = "null handle"; return PG_E_ARG; }
    if (h->d.dim != 2) { g_err = "cohesive openings are defined for 2D only"; return PG_E_UNSUPPORTED; }
    if (n_out != int64_t(h->d.n_cqp) / 2) { g_err = "n_out must be n_cqp / 2"; return PG_E_ARG; }
    if (n_out == 0) return PG_OK;
    if (!dU || !centers || !out) { g_err = "missing device buffer"; return PG_E_ARG; }
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : h->stream;
    const int64_t nu = int64_t(h->d.n_nodes + h->d.n_enr) * h->d.dim;
    if (int rc = order_after_last(h, st)) return rc;
    k_fill_int<<<1, 64, 0, st>>>(h->err, 0);
    k_permute_u<<<dim3(unsigned((nu + 255) / 256)), dim3(256), 0, st>>>(nu, h->d.node_dof, dU, h->Un);
    k_coh_openings<<<dim3(unsigned((h->d.n_cut + 255) / 256)), dim3(256), 0, st>>>(h->d, h->Un, centers, out, h->err);
    HIP_OK(hipGetLastError());
    if (int rc = record_done(h, st)) return rc;
    int herr = 0;
    HIP_OK(hipMemcpyAsync(&herr, h->err, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (herr & 4) { g_err = "cohesive openings need Ortiz-Pandolfi laws"; return PG_E_UNSUPPORTED; }
    return PG_OK;
}

int pg_asm_cohesive_failure(pg_asm* h, const double* dU, double* out, int64_t n_out, void* stream) {
    if (!h) { g_err = "null handle"; return PG_E_ARG; }
    if (h->d.dim != 2) { g_err = "cohesive failure lengths are defined for 2D only"; return PG_E_UNSUPPORTED; }
    if (n_out != int64_t(h->d.n_cqp) / 2) { g_err = "n_out must be n_cqp / 2"; return PG_E_ARG; }
    if (n_out == 0) return PG_OK;
    if (!dU || !out) { g_err = "missing device buffer"; return PG_E_ARG; }
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : h->stream;
    const int64_t nu = int64_t(h->d.n_nodes + h->d.n_enr) * h->d.dim;
    if (int rc = order_after_last(h, st)) return rc;
    k_permute_u<<<dim3(unsigned((nu + 255) / 256)), dim3(256), 0, st>>>(nu, h->d.node_dof, dU, h->Un);
    k_coh_failure<<<dim3(unsigned((h->d.n_cut + 255) / 256)), dim3(256), 0, st>>>(h->d, h->Un, out);
    HIP_OK(hipGetLastError());
    return record_done(h, st);
}

int pg_asm_set_tiles(pg_asm* h, int group_nodes, int cut_variant, int group_variant) {
    if (!h) { g_err = "null handle"; return PG_E_ARG; }
    const bool ok_g = group_nodes == 0 || (h->d.dim == 2 ? (group_nodes == 8 || group_nodes == 16 || group_nodes == 32)
                                                        : (group_nodes == 2 || group_nodes == 4 || group_nodes == 8));
    if (!ok_g) { g_err = "node-group size must be 8, 16 or 32 (2D), 2, 4 or 8 (3D)"; return PG_E_ARG; }
    if (cut_variant < 0 || cut_variant > 3) { g_err = "cut variant must be 0..3"; return PG_E_ARG; }
    if (group_variant < 0 || group_variant > 2047) { g_err = "group variant must be 0..2047"; return PG_E_ARG; }
    h->group_nodes = group_nodes;
    h->cut_variant = cut_variant;
    h->group_variant = group_variant;
    return PG_OK;
}

int pg_asm_last_timing(const pg_asm* h, double ms[5]) {
    if (!h || !ms) { g_err = "null argument"; return PG_E_ARG; }
    for (int i = 0; i < 5; ++i) ms[i] = h->last_ms[i];
    return PG_OK;
}

int pg_asm_assemble_device(pg_asm* h, const double* dU, const double* dIc, double* dIt, double delta_t,
                           unsigned flags, double* dK, double* dP, void* stream) {
    (void)delta_t;
    if (!h) { g_err = "null handle"; return PG_E_ARG; }
    if (flags & ~(PG_FUNC | PG_JAC | PG_FP32 | PG_BATCHED | PG_CUT64)) { g_err = "unknown flag bits"; return PG_E_ARG; }
    if (!(flags & (PG_FUNC | PG_JAC))) return PG_OK;
    if (!dU || ((flags & PG_JAC) && !dK) || ((flags & PG_FUNC) && !dP)) { g_err = "missing device buffer"; return PG_E_ARG; }
    if (h->d.n_cqp > 0 && (!dIc || !dIt)) { g_err = "cohesive problem needs ISV buffers"; return PG_E_ARG; }
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : h->stream;
    int rc = launch(h, dU, dIc, dIt, flags, dK, dP, st);
    if (rc) return rc;
    if (h->timing) {
        HIP_OK(hipEventSynchronize(h->ev[2]));
        int o = 0;
        while (o < 2 && h->gopt_ng[o] != h->group_nodes) ++o;
        const bool ov = (flags & (PG_FP32 | PG_BATCHED)) && !(h->group_variant & 4) && h->nclean[o] > 0 &&
                        h->nclean[o] < h->ngroups[o];
        float tot = 0, cut = 0, dirty = 0, clean = 0;
        HIP_OK(hipEventElapsedTime(&tot, h->ev[0], h->ev[2]));
        HIP_OK(hipEventElapsedTime(&cut, h->ev[0], h->ev[1]));
        HIP_OK(hipEventElapsedTime(&dirty, h->ev[5], (flags & (PG_FP32 | PG_BATCHED)) ? h->ev[6] : h->ev[2]));
        if (ov) HIP_OK(hipEventElapsedTime(&clean, h->ev[3], h->ev[4]));
        h->last_ms[0] = tot;
        h->last_ms[1] = cut;
        h->last_ms[2] = double(clean) + double(dirty);
        h->last_ms[3] = clean;
        h->last_ms[4] = dirty;
    }
    return PG_OK;
}

int pg_asm_assemble(pg_asm* h, const double* U, const double* Ic, double* It, double delta_t, unsigned flags,
                    double* Kff, double* Kfp, double* Kpf, double* Kpp, double* Pf, double* Pp) {
    if (!h) { g_err = "null handle"; return PG_E_ARG; }
    if (flags & ~(PG_FUNC | PG_JAC | PG_FP32 | PG_BATCHED | PG_CUT64)) { g_err = "unknown flag bits"; return PG_E_ARG; }
    if (!(flags & (PG_FUNC | PG_JAC))) return PG_OK;
    double* Kb[4] = {Kff, Kfp, Kpf, Kpp};
    bool missing = !U;
    if (flags & PG_JAC)
        for (int b = 0; b < 4; ++b) missing |= h->nnz[b] > 0 && !Kb[b];
    if (flags & PG_FUNC) missing |= (h->d.n_free > 0 && !Pf) || (h->d.n_pres > 0 && !Pp);
    if (missing) { g_err = "missing host buffer"; return PG_E_ARG; }
    const int64_t ncq = h->d.n_cqp;
    if (ncq > 0 && (!Ic || !It)) { g_err = "cohesive problem needs ISV buffers"; return PG_E_ARG; }
    if (!h->dU) {
        HIP_OK(hipMalloc(&h->dU, size_t(h->n_dof) * sizeof(double)));
        HIP_OK(hipMalloc(&h->dK, size_t(std::max<int64_t>(h->nnz_total, 1)) * sizeof(double)));
        HIP_OK(hipMalloc(&h->dP, size_t(h->n_dof) * sizeof(double)));
        HIP_OK(hipMalloc(&h->dIc, size_t(std::max<int64_t>(ncq, 1)) * sizeof(double)));
        HIP_OK(hipMalloc(&h->dIt, size_t(std::max<int64_t>(ncq, 1)) * sizeof(double)));
        h->dev_bytes += 8 * (2 * h->n_dof + std::max<int64_t>(h->nnz_total, 1) + 2 * std::max<int64_t>(ncq, 1));
    }
    hipStream_t st = h->stream;
    HIP_OK(hipMemcpyAsync(h->dU, U, size_t(h->n_dof) * sizeof(double), hipMemcpyHostToDevice, st));
    if (ncq > 0) HIP_OK(hipMemcpyAsync(h->dIc, Ic, size_t(ncq) * sizeof(double), hipMemcpyHostToDevice, st));
    int rc = pg_asm_assemble_device(h, h->dU, h->dIc, h->dIt, delta_t, flags, h->dK, h->dP, st);
    if (rc) return rc;
    if (flags & PG_JAC) {
        int64_t off = 0;
        for (int b = 0; b < 4; ++b) {
            if (h->nnz[b] > 0)
                HIP_OK(hipMemcpyAsync(Kb[b], h->dK + off, size_t(h->nnz[b]) * sizeof(double), hipMemcpyDeviceToHost, st));
            off += h->nnz[b];
        }
    }
    if (flags & PG_FUNC) {
        const int64_t nf = h->d.n_free, np = h->d.n_pres;
        if (nf > 0) HIP_OK(hipMemcpyAsync(Pf, h->dP, size_t(nf) * sizeof(double), hipMemcpyDeviceToHost, st));
        if (np > 0) HIP_OK(hipMemcpyAsync(Pp, h->dP + nf, size_t(np) * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    if (ncq > 0) HIP_OK(hipMemcpyAsync(It, h->dIt, size_t(ncq) * sizeof(double), hipMemcpyDeviceToHost, st));
    int herr = 0;
    HIP_OK(hipMemcpyAsync(&herr, h->err, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (herr) { g_err = "negative child Jacobian in a cut element"; return PG_E_MESH; }
    return PG_OK;
}

int pg_asm_set_temperature_change(pg_asm* h, double temp_change) {
    if (!h) { g_err = "null handle"; return PG_E_ARG; }
    if (!std::isfinite(temp_change)) { g_err = "temperature change must be finite"; return PG_E_ARG; }
    h->temp_change = temp_change;
    // LinearElasticIsotropicMaterial::Constitutive: strain_val = -1.0 *
    // input.temp_change * _thermal_exp (src/material_lei.cpp:276)
    for (int m = 0; m < h->d.n_mat; ++m)
        h->d.eth[m] = h->thermal_set[m] ? -1.0 * temp_change * h->thermal_exp[m] : 0.0;
    return PG_OK;
}

int pg_asm_sensitivity_dPdd(pg_asm* h, const pg_sens_param* prm, unsigned flags, const double* dU,
                            const double* disv, const double* disv_sens, double* ddPdd, void* stream) {
    if (!h || !prm) { g_err = "null argument"; return PG_E_ARG; }
    if (flags & ~PG_SENS_OWN_CHILD_B) { g_err = "unknown flag bits"; return PG_E_ARG; }
    const DevProb& d = h->d;
    if (prm->mat < 0 || prm->mat >= d.n_mat) { g_err = "material index out of range"; return PG_E_ARG; }
    const int mt = h->mat_type[prm->mat];
    if (prm->kind == PG_SENS_LEI) {
        if (mt != PG_MAT_LEI) { g_err = "an LEI parameter needs an LEI material"; return PG_E_UNSUPPORTED; }
        const int ns = d.dim == 2 ? 3 : 6;
        for (int k = 0; k < ns * ns; ++k)
            if (!std::isfinite(prm->dD[k])) { g_err = "dD/dd must be finite"; return PG_E_ARG; }
    } else if (prm->kind == PG_SENS_SIGMA_C || prm->kind == PG_SENS_DELTA_C) {
        if (mt == PG_MAT_BL || mt == PG_MAT_BL_NU) {
            g_err = "Sensitivity is not implemented for the bilinear cohesive model";
            return PG_E_UNSUPPORTED;
        }
        if (mt != PG_MAT_OP && mt != PG_MAT_OP_NU) {
            g_err = "a cohesive parameter needs an Ortiz-Pandolfi material";
            return PG_E_UNSUPPORTED;
        }
    } else {
        g_err = "unknown parameter kind";
        return PG_E_ARG;
    }
    if (h->coh_law_mask & (1u << PG_MAT_BL)) {
        // every element with a bilinear cohesive child holds internal variables and is
        // visited; the law's SensitivityConstitutive stops there
        g_err = "Sensitivity is not implemented for the bilinear cohesive model";
        return PG_E_UNSUPPORTED;
    }
    if (!dU || !ddPdd) { g_err = "missing device buffer"; return PG_E_ARG; }
    if ((h->coh_law_mask & (1u << PG_MAT_OP)) && (!disv || !disv_sens)) {
        g_err = "Ortiz-Pandolfi cohesive children need the internal variables and their sensitivities";
        return PG_E_ARG;
    }
    SensDev sd{};
    sd.mat = prm->mat;
    sd.kind = prm->kind;
    sd.own_b = (flags & PG_SENS_OWN_CHILD_B) ? 1 : 0;
    for (int k = 0; k < 36; ++k) sd.dD[k] = prm->kind == PG_SENS_LEI ? prm->dD[k] : 0.0;
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : h->stream;
    const int64_t nu = int64_t(d.n_nodes + d.n_enr) * d.dim;
    if (int rc = order_after_last(h, st)) return rc;
    k_permute_u<<<dim3(unsigned((nu + 255) / 256)), dim3(256), 0, st>>>(nu, d.node_dof, dU, h->Un);
    const unsigned gb = unsigned((d.n_elem + 255) / 256);
    if (d.dim == 2) {
        k_sens_elem<2><<<gb, 256, 0, st>>>(d, h->Un, sd, h->stage);
        if (d.n_cut) k_sens_cut<2, false><<<unsigned(d.n_cut), 64, 0, st>>>(d, h->Un, sd, disv, disv_sens, h->stage_p, nullptr);
        if (h->n_poly) k_sens_cut<2, true><<<unsigned(h->n_poly), 64, 0, st>>>(d, h->Un, sd, disv, disv_sens, h->stage_p, h->poly_list);
    } else {
        k_sens_elem<3><<<gb, 256, 0, st>>>(d, h->Un, sd, h->stage);
        if (d.n_cut) k_sens_cut<3, false><<<unsigned(d.n_cut), 64, 0, st>>>(d, h->Un, sd, disv, disv_sens, h->stage_p, nullptr);
    }
    const unsigned gp = unsigned((d.n_pairs + h->pair_block - 1) / h->pair_block);
    if (d.dim == 2) k_pair_staged<2><<<gp, h->pair_block, 0, st>>>(d, h->stage, h->stage_k, h->stage_p, nullptr, ddPdd, PG_FUNC);
    else k_pair_staged<3><<<gp, h->pair_block, 0, st>>>(d, h->stage, h->stage_k, h->stage_p, nullptr, ddPdd, PG_FUNC);
    HIP_OK(hipGetLastError());
    return record_done(h, st);
}

int pg_asm_plot_cells(const pg_asm* h, int64_t* n_cells) {
    if (!h || !n_cells) { g_err = "null argument"; return PG_E_ARG; }
    *n_cells = h->n_cells;
    return PG_OK;
}

int pg_asm_plot_fields(pg_asm* h, const double* dU, const int32_t* denr_edge, const double* denr_alpha,
                       double* dnodes, double* dcells, void* stream) {
    if (!h) { g_err = "null handle"; return PG_E_ARG; }
    const DevProb& d = h->d;
    if (!dU || !dnodes || !dcells || (d.n_enr > 0 && (!denr_edge || !denr_alpha))) {
        g_err = "missing device buffer";
        return PG_E_ARG;
    }
    if (!h->plot_cells) {
        // cell order of Writer_VTK::store_elements: elements in order, each cut
        // element's integration children in child order
        std::vector<int32_t> coe(size_t(d.n_elem));
        std::vector<pg_cut_rec> cut(size_t(d.n_cut));
        HIP_OK(hipMemcpy(coe.data(), d.cut_of_elem, coe.size() * 4, hipMemcpyDeviceToHost));
        if (d.n_cut) HIP_OK(hipMemcpy(cut.data(), d.cut, cut.size() * sizeof(pg_cut_rec), hipMemcpyDeviceToHost));
        std::vector<int2> plain, poly;
        std::vector<int64_t> order, qorder;  // each listed cell's position in store_elements order
        int64_t pos = 0;
        for (int32_t e = 0; e < d.n_elem; ++e) {
            if (coe[e] < 0) {
                plain.push_back(make_int2(e, -1));
                order.push_back(pos++);
                continue;
            }
            const pg_cut_rec& r = cut[coe[e]];
            for (int k = 0; k < r.n_child; ++k) {
                const bool pc = r.child_type[k] == PG_CHILD_POLY5 || r.child_type[k] == PG_CHILD_POLY6;
                (pc ? poly : plain).push_back(make_int2(e, k));
                (pc ? qorder : order).push_back(pos++);
            }
        }
        order.insert(order.end(), qorder.begin(), qorder.end());
        const size_t nc = order.size();
        if (int rc = alloc(h, &h->plot_cells, nc)) return rc;
        if (!plain.empty())
            HIP_OK(hipMemcpy(h->plot_cells, plain.data(), plain.size() * sizeof(int2), hipMemcpyHostToDevice));
        if (!poly.empty())
            HIP_OK(hipMemcpy(h->plot_cells + plain.size(), poly.data(), poly.size() * sizeof(int2), hipMemcpyHostToDevice));
        h->n_plot[0] = int32_t(plain.size());
        h->n_plot[1] = int32_t(poly.size());
        h->n_cells = int64_t(nc);
        if (int rc = alloc(h, &h->plot_pos, nc)) return rc;
        HIP_OK(hipMemcpy(h->plot_pos, order.data(), nc * sizeof(int64_t), hipMemcpyHostToDevice));
        constexpr int S2 = 1 + 2 * 3, S3 = 1 + 2 * 6;
        if (int rc = alloc(h, &h->plot_tmp, nc * size_t(d.dim == 2 ? S2 : S3))) return rc;
    }
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : h->stream;
    const int64_t nall = int64_t(d.n_nodes) + d.n_enr;
    const int64_t nu = nall * d.dim;
    if (int rc = order_after_last(h, st)) return rc;
    k_permute_u<<<dim3(unsigned((nu + 255) / 256)), dim3(256), 0, st>>>(nu, d.node_dof, dU, h->Un);
    k_plot_nodes<<<dim3(unsigned((nall + 255) / 256)), dim3(256), 0, st>>>(d, h->Un, denr_edge, denr_alpha, dnodes);
    const int S = d.dim == 2 ? 1 + 2 * 3 : 1 + 2 * 6;
    const int n0 = h->n_plot[0], n1 = h->n_plot[1];
    if (d.dim == 2) {
        if (n0) k_plot_cells<2, false><<<unsigned((n0 + 127) / 128), 128, 0, st>>>(d, h->Un, h->plot_cells, n0, h->plot_tmp);
        if (n1)
            k_plot_cells<2, true><<<unsigned((n1 + 127) / 128), 128, 0, st>>>(d, h->Un, h->plot_cells + n0, n1,
                                                                            h->plot_tmp + int64_t(n0) * S);
    } else if (n0) {
        k_plot_cells<3, false><<<unsigned((n0 + 127) / 128), 128, 0, st>>>(d, h->Un, h->plot_cells, n0, h->plot_tmp);
    }
    k_scatter_rows<<<dim3(unsigned((h->n_cells * S + 255) / 256)), dim3(256), 0, st>>>(h->n_cells, S, h->plot_pos,
                                                                                       h->plot_tmp, dcells);
    HIP_OK(hipGetLastError());
    return record_done(h, st);
}

int pg_asm_check_error(pg_asm* h) {
    if (!h) { g_err = "null handle"; return PG_E_ARG; }
    if (!h->done_valid) return PG_OK;
    HIP_OK(hipEventSynchronize(h->done));
    int herr = 0;
    HIP_OK(hipMemcpy(&herr, h->err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr & 1) { g_err = "negative child Jacobian in a cut element"; return PG_E_MESH; }
    return PG_OK;
}

}  // extern "C"

namespace {

// Workspace of the PCG solve in doubles: r, z, p, q, the inverse diagonal, the
// dot products' per-block partial sums and the scalars.
int64_t pcg_words(int64_t n) { return 5 * n + 2 * PCG_B + 8; }

// The PCG solve on a caller-given workspace of pcg_words(n) doubles.
int pcg_solve_ws(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* vals, const double* b,
                 double* x, double rtol, int max_iter, int* iters, double* rel_res, void* stream, double* ws) {
    const size_t nv = size_t(n);
    double* r = ws;
    double *z = r + nv, *pv = z + nv, *q = pv + nv, *dinv = q + nv, *part = dinv + nv, *sc = part + 2 * PCG_B;
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : nullptr;
    const dim3 g(unsigned((n + 255) / 256)), t(256);
    const int nb = int(std::min<int64_t>(PCG_B, (n + PCG_T - 1) / PCG_T));
    auto dot2 = [&](const double* a, const double* bb, const double* c, const double* d, int o0, int o1) {
        k_dot2<<<dim3(nb), dim3(PCG_T), 0, st>>>(n, a, bb, c, d, part);
        k_dot_fin<<<dim3(1), dim3(PCG_T), 0, st>>>(part, nb, sc, o0, o1);
    };
    // the scalars sc[0..3] come back together once per iteration
    double hs[4] = {0.0, 0.0, 0.0, 0.0};
    auto read = [&]() -> int {
        HIP_OK(hipMemcpyAsync(hs, sc, sizeof(hs), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        return PG_OK;
    };
    k_csr_diag_inv<<<g, t, 0, st>>>(n, row_ptr, col_idx, vals, dinv);
    dot2(b, b, b, b, 3, 4);
    if (int e = read()) return e;
    const double bb = hs[3];
    if (!std::isfinite(bb)) { g_err = "right-hand side is not finite"; return PG_E_ARG; }
    if (bb == 0.0) {
        k_fill_double<<<g, t, 0, st>>>(n, x, 0.0);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(st));
        return PG_OK;
    }
    k_csr_spmv<<<g, t, 0, st>>>(n, row_ptr, col_idx, vals, x, q);
    k_pcg_init<<<g, t, 0, st>>>(n, b, q, dinv, r, z, pv);
    dot2(r, z, r, r, 0, 3);
    if (int e = read()) return e;
    double rr = hs[3];
    if (!std::isfinite(rr) || !std::isfinite(hs[0])) {
        g_err = "matrix or initial guess is not finite";
        return PG_E_ARG;
    }
    int it = 0, cur = 0;
    while (it < max_iter && std::sqrt(rr) > rtol * std::sqrt(bb)) {
        // r.z > 0 and p.Ap > 0 hold for a symmetric positive definite matrix
        // and a positive Jacobi diagonal; anything else is a breakdown
        if (!(hs[cur] > 0.0)) { g_err = "PCG broke down: r.z <= 0 (matrix not positive definite)"; return PG_E_ARG; }
        k_csr_spmv<<<g, t, 0, st>>>(n, row_ptr, col_idx, vals, pv, q);
        dot2(pv, q, pv, q, 2, 4);
        k_pcg_xr<<<g, t, 0, st>>>(n, x, r, z, pv, q, dinv, sc, cur);
        dot2(r, z, r, r, 1 - cur, 3);
        k_pcg_p<<<g, t, 0, st>>>(n, pv, z, sc, 1 - cur, cur);
        HIP_OK(hipGetLastError());
        cur = 1 - cur;
        ++it;
        if (int e = read()) return e;
        if (!(hs[2] > 0.0)) { g_err = "PCG broke down: p.Ap <= 0 (matrix not positive definite)"; return PG_E_ARG; }
        rr = hs[3];
        if (!std::isfinite(rr)) { g_err = "PCG broke down: residual is not finite"; return PG_E_ARG; }
    }
    if (iters) *iters = it;
    if (rel_res) *rel_res = std::sqrt(rr / bb);
    return PG_OK;
}

// Workspace of pg_csr_pcg_noalloc: a file-scope device array, one per GPU the
// library runs on, held until the process ends.  Its address is taken once
// per device with hipGetSymbolAddress; host code uses only that address.
constexpr int64_t PCG_NOALLOC_MAX_N = int64_t(1) << 20;
static __device__ __align__(256) double scratch[5 * PCG_NOALLOC_MAX_N + 2 * PCG_B + 8];
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

int scratch_for_current_device(double** out) {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) { g_err = "device index beyond the supported range"; return PG_E_ARG; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        HIP_OK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) { g_err = "hipGetSymbolAddress returned no address"; return PG_E_HIP; }
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return PG_OK;
}

}  // namespace

extern "C" {

int pg_csr_pcg(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* vals, const double* b,
               double* x, double rtol, int max_iter, int* iters, double* rel_res, void* stream) {
    if (n < 0 || max_iter < 0 || !(rtol >= 0.0)) { g_err = "bad size, iteration limit or tolerance"; return PG_E_ARG; }
    if (iters) *iters = 0;
    if (rel_res) *rel_res = 0.0;
    if (n == 0) return PG_OK;
    if (!row_ptr || !col_idx || !vals || !b || !x) { g_err = "missing device buffer"; return PG_E_ARG; }
    struct Mem {
        void* p = nullptr;
        ~Mem() { if (p) (void)hipFree(p); }
    } mem;
    if (hipMalloc(&mem.p, size_t(pcg_words(n)) * sizeof(double)) != hipSuccess) {
        g_err = "device allocation failed";
        return PG_E_NOMEM;
    }
    return pcg_solve_ws(n, row_ptr, col_idx, vals, b, x, rtol, max_iter, iters, rel_res, stream,
                        static_cast<double*>(mem.p));
}

int pg_csr_pcg_noalloc(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                       const double* b, double* x, double rtol, int max_iter, int* iters, double* rel_res,
                       void* stream) {
    if (n < 0 || max_iter < 0 || !(rtol >= 0.0)) { g_err = "bad size, iteration limit or tolerance"; return PG_E_ARG; }
    if (iters) *iters = 0;
    if (rel_res) *rel_res = 0.0;
    if (n > PCG_NOALLOC_MAX_N) {
        g_err = "pg_csr_pcg_noalloc: n exceeds its fixed workspace (at most 1048576 rows)";
        return PG_E_ARG;
    }
    if (n == 0) return PG_OK;
    if (!row_ptr || !col_idx || !vals || !b || !x) { g_err = "missing device buffer"; return PG_E_ARG; }
    double* ws = nullptr;
    if (int rc = scratch_for_current_device(&ws)) return rc;
    return pcg_solve_ws(n, row_ptr, col_idx, vals, b, x, rtol, max_iter, iters, rel_res, stream, ws);
}

int pg_dev_alloc(void** ptr, int64_t bytes) {
    if (!ptr || bytes < 0) { g_err = "bad argument"; return PG_E_ARG; }
    *ptr = nullptr;
    HIP_OK(hipMalloc(ptr, size_t(bytes > 0 ? bytes : 1)));
    return PG_OK;
}

int pg_dev_free(void* ptr) {
    if (ptr) HIP_OK(hipFree(ptr));
    return PG_OK;
}

int pg_dev_upload(void* dst, const void* src, int64_t bytes) {
    if (bytes < 0 || (bytes > 0 && (!dst || !src))) { g_err = "bad argument"; return PG_E_ARG; }
    if (bytes) HIP_OK(hipMemcpy(dst, src, size_t(bytes), hipMemcpyHostToDevice));
    return PG_OK;
}

int pg_dev_download(void* dst, const void* src, int64_t bytes) {
    if (bytes < 0 || (bytes > 0 && (!dst || !src))) { g_err = "bad argument"; return PG_E_ARG; }
    if (bytes) HIP_OK(hipMemcpy(dst, src, size_t(bytes), hipMemcpyDeviceToHost));
    return PG_OK;
}

int pg_dev_copy(void* dst, const void* src, int64_t bytes) {
    if (bytes < 0 || (bytes > 0 && (!dst || !src))) { g_err = "bad argument"; return PG_E_ARG; }
    if (bytes) HIP_OK(hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyDeviceToDevice, nullptr));
    return PG_OK;
}

int pg_dev_synchronize(void) {
    HIP_OK(hipDeviceSynchronize());
    return PG_OK;
}

int pg_dev_axpy(int64_t n, double a, const double* x, double* y) {
    if (n < 0 || (n > 0 && (!x || !y))) { g_err = "bad argument"; return PG_E_ARG; }
    if (n == 0) return PG_OK;
    k_axpy<<<dim3(unsigned((n + 255) / 256)), dim3(256)>>>(n, a, x, y);
    HIP_OK(hipGetLastError());
    HIP_OK(hipDeviceSynchronize());
    return PG_OK;
}

int pg_dev_dot(int64_t n, const double* x, const double* y, double* out) {
    if (n < 0 || !out || (n > 0 && (!x || !y))) { g_err = "bad argument"; return PG_E_ARG; }
    *out = 0.0;
    if (n == 0) return PG_OK;
    struct Mem {
        void* p = nullptr;
        ~Mem() { if (p) (void)hipFree(p); }
    } mem;
    HIP_OK(hipMalloc(&mem.p, (2 * PCG_B + 8) * sizeof(double)));
    double* part = static_cast<double*>(mem.p);
    double* sc = part + 2 * PCG_B;
    const int nb = int(std::min<int64_t>(PCG_B, (n + PCG_T - 1) / PCG_T));
    k_dot2<<<dim3(nb), dim3(PCG_T)>>>(n, x, y, x, y, part);
    k_dot_fin<<<dim3(1), dim3(PCG_T)>>>(part, nb, sc, 0, 1);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(out, sc, sizeof(double), hipMemcpyDeviceToHost));
    return PG_OK;
}

}  // extern "C"

// --------------------------------------------------------------------------
// Heat-conduction handle (AssemblerThermal): the scalar field's description in
// HBM, the FP64 element and node-pair kernels above.
// --------------------------------------------------------------------------
struct pg_therm {
    DevProb d{};
    int64_t n_dof = 0, nnz[4] = {0, 0, 0, 0}, nnz_total = 0;
    std::vector<void*> allocs;
    double* Tn = nullptr;       // T per node
    double* stage = nullptr;    // uncut elements' K_el + P_el (ThermStage)
    double* stage_k = nullptr;  // cut elements' K_el at cut_koff
    double* stage_p = nullptr;  // cut elements' P_el at cut_poff
    int* err = nullptr;
    int32_t n_poly = 0;
    int32_t* poly_list = nullptr;
    double *dT = nullptr, *dK = nullptr, *dP = nullptr;  // host-API buffers
    hipStream_t stream = nullptr;
};

namespace {

template <typename T>
int talloc(pg_therm* h, T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    HIP_OK(hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T)));
    h->allocs.push_back(*p);
    return PG_OK;
}

int therm_launch(pg_therm* h, const double* dT, unsigned flags, double* dK, double* dP, hipStream_t st) {
    const DevProb& d = h->d;
    const int64_t nall = int64_t(d.n_nodes) + d.n_enr;
    k_fill_int<<<1, 64, 0, st>>>(h->err, 0);
    k_permute_u<<<dim3(unsigned((nall + 255) / 256)), dim3(256), 0, st>>>(nall, d.node_dof, dT, h->Tn);
    const unsigned gb = unsigned((d.n_elem + 255) / 256);
    if (d.dim == 2) {
        k_therm_elem<2><<<gb, 256, 0, st>>>(d, h->Tn, h->stage);
        if (d.n_cut) k_therm_cut<2, false><<<unsigned(d.n_cut), 64, 0, st>>>(d, h->Tn, h->stage_k, h->stage_p, h->err, nullptr);
        if (h->n_poly) k_therm_cut<2, true><<<unsigned(h->n_poly), 64, 0, st>>>(d, h->Tn, h->stage_k, h->stage_p, h->err, h->poly_list);
    } else {
        k_therm_elem<3><<<gb, 256, 0, st>>>(d, h->Tn, h->stage);
        if (d.n_cut) k_therm_cut<3, false><<<unsigned(d.n_cut), 64, 0, st>>>(d, h->Tn, h->stage_k, h->stage_p, h->err, nullptr);
    }
    const unsigned gp = unsigned((d.n_pairs + 255) / 256);
    if (d.dim == 2) k_therm_pair<2><<<gp, 256, 0, st>>>(d, h->stage, h->stage_k, h->stage_p, dK, dP, flags);
    else k_therm_pair<3><<<gp, 256, 0, st>>>(d, h->stage, h->stage_k, h->stage_p, dK, dP, flags);
    HIP_OK(hipGetLastError());
    return PG_OK;
}

}  // namespace

extern "C" {

int pg_therm_destroy(pg_therm* h) {
    if (!h) return PG_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void* p : h->allocs) (void)hipFree(p);
    for (void* p : {(void*)h->dT, (void*)h->dK, (void*)h->dP})
        if (p) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PG_OK;
}

int pg_therm_create(const pg_problem_desc* s, pg_therm** out) {
    if (!out) { g_err = "null output handle"; return PG_E_ARG; }
    *out = nullptr;
    if (!s) { g_err = "null description"; return PG_E_ARG; }
    if (s->nndof != 1) { g_err = "the thermal assembler takes one dof per node (nndof = 1)"; return PG_E_ARG; }
    if (s->dim != 2 && s->dim != 3) { g_err = "dim must be 2 or 3"; return PG_E_ARG; }
    if (s->nen != s->dim + 1) { g_err = "only tri3 (2D) and tet4 (3D) parents are supported"; return PG_E_UNSUPPORTED; }
    if (s->cohesive) {
        g_err = "cohesive meshes are not supported by the thermal assembler (AssemblerThermal::KernelCohesive)";
        return PG_E_UNSUPPORTED;
    }
    const int64_t nall = int64_t(s->n_nodes) + s->n_enr;
    if (s->n_nodes <= 0 || s->n_elem <= 0 || s->n_enr < 0 || s->n_cut < 0 || int64_t(s->n_free) + s->n_pres != nall) {
        g_err = "inconsistent node / dof counts";
        return PG_E_ARG;
    }
    if (!s->coords || !s->conn || !s->elem_mat || !s->cut_of_elem || !s->mats || !s->node_dof || !s->npat_row ||
        !s->npat_col || !s->npat_map || !s->n2e_ptr || !s->n2e || (s->n_cut > 0 && (!s->cut || !s->cut_elem))) {
        g_err = "missing array in description";
        return PG_E_ARG;
    }
    if (s->n_mat <= 0 || s->n_mat > MAX_LDS_MATS) { g_err = "1 to 8 materials are supported"; return PG_E_UNSUPPORTED; }
    for (int m = 0; m < s->n_mat; ++m)
        if (s->mats[m].type != PG_MAT_LT || s->mats[m].nstr != s->dim) {
            g_err = "the thermal assembler takes LinearThermalMaterial (PG_MAT_LT, nstr = dim) materials";
            return PG_E_ARG;
        }
    for (int32_t e = 0; e < s->n_elem; ++e) {
        if (s->elem_mat[e] < 0 || s->elem_mat[e] >= s->n_mat) { g_err = "element material out of range"; return PG_E_ARG; }
        if (s->cut_of_elem[e] >= s->n_cut) { g_err = "cut index out of range"; return PG_E_ARG; }
        for (int n = 0; n < s->nen; ++n) {
            const int32_t v = s->conn[int64_t(e) * s->nen + n];
            if (v < 0 || v >= s->n_nodes) { g_err = "connectivity out of range"; return PG_E_ARG; }
        }
    }
    std::vector<int64_t> koff(size_t(s->n_cut) + 1, 0), poff(size_t(s->n_cut) + 1, 0);
    for (int32_t c = 0; c < s->n_cut; ++c) {
        const pg_cut_rec& r = s->cut[c];
        if (s->cut_elem[c] < 0 || s->cut_elem[c] >= s->n_elem || r.n_enr > PG_MAX_ENR || r.n_coh != 0 ||
            r.n_child == 0 || r.n_child > (s->dim == 2 ? 4 : PG_MAX_CHILD)) {
            g_err = "bad cut record";
            return PG_E_ARG;
        }
        int nq = 0;
        for (int k = 0; k < r.n_child; ++k) {
            const int t = r.child_type[k];
            const bool ok = s->dim == 2 ? (t == PG_CHILD_TRI3 || t == PG_CHILD_QUAD4 || t == PG_CHILD_POLY5 ||
                                           t == PG_CHILD_POLY6)
                                        : t == PG_CHILD_TET4;
            if (!ok || r.child_mat[k] >= s->n_mat) { g_err = "bad child type or material"; return PG_E_ARG; }
            for (int n = 0; n < PG_CHILD_NN(t); ++n)
                if (r.child_node[PG_CHILD_STRIDE(s->dim) * k + n] >= s->nen + r.n_enr) {
                    g_err = "child node out of range";
                    return PG_E_ARG;
                }
            nq += PG_CHILD_NQP(t);
        }
        if (nq != r.n_vqp || nq > (s->dim == 2 ? MAXQ2 : MAXQ3)) { g_err = "cut record qp count"; return PG_E_ARG; }
        for (int k = 0; k < r.n_enr; ++k)
            if (r.enr[k] < s->n_nodes || r.enr[k] >= nall) { g_err = "enrichment node out of range"; return PG_E_ARG; }
        const int nl = s->nen + r.n_enr;
        koff[c + 1] = koff[c] + int64_t(nl) * nl;
        poff[c + 1] = poff[c] + nl;
    }
    for (int64_t k = 0; k < nall; ++k)
        if (s->node_dof[k] < 0 || s->node_dof[k] >= s->n_free + s->n_pres) { g_err = "dof out of range"; return PG_E_ARG; }
    int64_t nnz_tot = 0;
    for (int b = 0; b < 4; ++b) nnz_tot += s->nnz[b];
    for (int64_t q = 0; q < s->n_pairs; ++q) {
        if (s->npat_row[q] < 0 || s->npat_row[q] >= nall || s->npat_col[q] < 0 || s->npat_col[q] >= nall ||
            s->npat_map[q] < 0 || s->npat_map[q] >= nnz_tot) {
            g_err = "node-pair pattern out of range";
            return PG_E_ARG;
        }
    }
    if (s->n2e_ptr[0] != 0) { g_err = "incidence pointer"; return PG_E_ARG; }
    for (int64_t a = 0; a < nall; ++a)
        if (s->n2e_ptr[a + 1] < s->n2e_ptr[a]) { g_err = "incidence pointer"; return PG_E_ARG; }
    for (int64_t k = 0; k < s->n2e_ptr[nall]; ++k)
        if (s->n2e[k] < 0 || s->n2e[k] >= s->n_elem) { g_err = "incidence out of range"; return PG_E_ARG; }

    pg_therm* h = new (std::nothrow) pg_therm();
    if (!h) { g_err = "out of memory"; return PG_E_NOMEM; }
    auto fail = [&](int code) { pg_therm_destroy(h); return code; };
    DevProb& d = h->d;
    d.dim = s->dim; d.nen = s->nen; d.n_nodes = s->n_nodes; d.n_enr = s->n_enr; d.n_elem = s->n_elem;
    d.n_cut = s->n_cut; d.n_free = s->n_free; d.n_pres = s->n_pres; d.cohesive = 0; d.n_cqp = 0;
    d.n_pairs = s->n_pairs;
    d.n_mat = s->n_mat;
    h->n_dof = int64_t(s->n_free) + s->n_pres;
    for (int b = 0; b < 4; ++b) {
        h->nnz[b] = s->nnz[b];
        h->nnz_total += s->nnz[b];
    }
    int rc;
#define TUP(field, src, n, T)                                   \
    do {                                                        \
        T* tmp_;                                                \
        if ((rc = talloc(h, &tmp_, size_t(n)))) return fail(rc); \
        if (size_t(n) > 0 && hipMemcpy(tmp_, src, size_t(n) * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) { \
            g_err = "hipMemcpy of " #field " failed"; return fail(PG_E_HIP); } \
        d.field = tmp_;                                         \
    } while (0)
    TUP(coords, s->coords, nall * s->dim, double);
    TUP(conn, s->conn, int64_t(s->n_elem) * s->nen, int32_t);
    TUP(elem_mat, s->elem_mat, s->n_elem, int32_t);
    TUP(cut_of_elem, s->cut_of_elem, s->n_elem, int32_t);
    TUP(cut_elem, s->cut_elem, s->n_cut, int32_t);
    TUP(cut, s->cut, s->n_cut, pg_cut_rec);
    TUP(mats, s->mats, s->n_mat, pg_material);
    TUP(node_dof, s->node_dof, nall, int32_t);
    TUP(npat_row, s->npat_row, s->n_pairs, int32_t);
    TUP(npat_col, s->npat_col, s->n_pairs, int32_t);
    TUP(npat_map, s->npat_map, s->n_pairs, int32_t);
    TUP(n2e_ptr, s->n2e_ptr, nall + 1, int32_t);
    TUP(n2e, s->n2e, s->n2e_ptr[nall], int32_t);
    TUP(cut_koff, koff.data(), s->n_cut + 1, int64_t);
    TUP(cut_poff, poff.data(), s->n_cut + 1, int64_t);
#undef TUP
    {
        std::vector<int32_t> pl;
        for (int32_t c = 0; c < s->n_cut; ++c)
            for (int k = 0; k < s->cut[c].n_child; ++k)
                if (s->cut[c].child_type[k] == PG_CHILD_POLY5 || s->cut[c].child_type[k] == PG_CHILD_POLY6) {
                    pl.push_back(c);
                    break;
                }
        h->n_poly = int32_t(pl.size());
        if (!pl.empty()) {
            if ((rc = talloc(h, &h->poly_list, pl.size()))) return fail(rc);
            if (hipMemcpy(h->poly_list, pl.data(), pl.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
                g_err = "hipMemcpy of the polygon list failed";
                return fail(PG_E_HIP);
            }
        }
    }
    const int64_t nblk = (int64_t(s->n_elem) + 63) / 64;
    const int64_t W = s->dim == 2 ? ThermStage<3>::W : ThermStage<4>::W;
    if ((rc = talloc(h, &h->Tn, size_t(nall)))) return fail(rc);
    if ((rc = talloc(h, &h->stage, size_t(nblk * W)))) return fail(rc);
    if ((rc = talloc(h, &h->stage_k, size_t(koff.back())))) return fail(rc);
    if ((rc = talloc(h, &h->stage_p, size_t(poff.back())))) return fail(rc);
    if ((rc = talloc(h, &h->err, 1))) return fail(rc);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        g_err = "hipStreamCreate failed";
        return fail(PG_E_HIP);
    }
    *out = h;
    return PG_OK;
}

int pg_therm_assemble_device(pg_therm* h, const double* dT, unsigned flags, double* dK, double* dP, void* stream) {
    if (!h) { g_err = "null handle"; return PG_E_ARG; }
    if (flags & ~(PG_FUNC | PG_JAC)) { g_err = "the thermal assembler takes PG_FUNC and PG_JAC only"; return PG_E_ARG; }
    if (!(flags & (PG_FUNC | PG_JAC))) return PG_OK;
    if (!dT || ((flags & PG_JAC) && !dK) || ((flags & PG_FUNC) && !dP)) { g_err = "missing device buffer"; return PG_E_ARG; }
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : h->stream;
    return therm_launch(h, dT, flags, dK, dP, st);
}

int pg_therm_assemble(pg_therm* h, const double* T, unsigned flags, double* Kff, double* Kfp, double* Kpf,
                      double* Kpp, double* Pf, double* Pp) {
    if (!h) { g_err = "null handle"; return PG_E_ARG; }
    if (flags & ~(PG_FUNC | PG_JAC)) { g_err = "the thermal assembler takes PG_FUNC and PG_JAC only"; return PG_E_ARG; }
    if (!(flags & (PG_FUNC | PG_JAC))) return PG_OK;
    double* Kb[4] = {Kff, Kfp, Kpf, Kpp};
    bool missing = !T;
    if (flags & PG_JAC)
        for (int b = 0; b < 4; ++b) missing |= h->nnz[b] > 0 && !Kb[b];
    if (flags & PG_FUNC) missing |= (h->d.n_free > 0 && !Pf) || (h->d.n_pres > 0 && !Pp);
    if (missing) { g_err = "missing host buffer"; return PG_E_ARG; }
    if (!h->dT) {
        HIP_OK(hipMalloc(&h->dT, size_t(h->n_dof) * sizeof(double)));
        HIP_OK(hipMalloc(&h->dK, size_t(std::max<int64_t>(h->nnz_total, 1)) * sizeof(double)));
        HIP_OK(hipMalloc(&h->dP, size_t(std::max<int64_t>(h->n_dof, 1)) * sizeof(double)));
    }
    hipStream_t st = h->stream;
    HIP_OK(hipMemcpyAsync(h->dT, T, size_t(h->n_dof) * sizeof(double), hipMemcpyHostToDevice, st));
    if (int rc = therm_launch(h, h->dT, flags, h->dK, h->dP, st)) return rc;
    int herr = 0;
    HIP_OK(hipMemcpyAsync(&herr, h->err, sizeof(int), hipMemcpyDeviceToHost, st));
    if (flags & PG_JAC) {
        int64_t off = 0;
        for (int b = 0; b < 4; ++b) {
            if (h->nnz[b])
                HIP_OK(hipMemcpyAsync(Kb[b], h->dK + off, size_t(h->nnz[b]) * sizeof(double), hipMemcpyDeviceToHost, st));
            off += h->nnz[b];
        }
    }
    if (flags & PG_FUNC) {
        if (h->d.n_free) HIP_OK(hipMemcpyAsync(Pf, h->dP, size_t(h->d.n_free) * sizeof(double), hipMemcpyDeviceToHost, st));
        if (h->d.n_pres)
            HIP_OK(hipMemcpyAsync(Pp, h->dP + h->d.n_free, size_t(h->d.n_pres) * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    HIP_OK(hipStreamSynchronize(st));
    if (herr & 1) { g_err = "negative child Jacobian in a cut element"; return PG_E_MESH; }
    return PG_OK;
}

}  // extern "C"
