// MI355X (gfx950) element assembly of K and R for IGFEM / cohesive RVEs.
//
// Two schedules behind one C ABI (pgfem_asm.h):
//   plain  (FP64, drop-in): one thread per parent element computes the element
//          matrix exactly as the reference does (per-qp B^T.D.B accumulation,
//          src/Assembler.cpp:300-520) and stages it in HBM; cut elements are
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
#include <string>
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
    const int32_t* npat_row;
    const int32_t* npat_col;
    const int32_t* npat_map;
    const int32_t* n2e_ptr;
    const int32_t* n2e;
    const int64_t* cut_koff;   // per cut element: offset of its ndof^2 block
    const int64_t* cut_poff;   // per cut element: offset of its ndof vector
};

// --------------------------------------------------------------------------
// U in global-dof order -> node-major (node*dim + comp).  The reference gathers
// U per element through NodalData (formElementalVector, src/Utilities.cpp:599-617).
// --------------------------------------------------------------------------
__global__ void k_permute_u(int64_t n, const int32_t* __restrict__ node_dof, const double* __restrict__ U,
                            double* __restrict__ Un) {
    int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) Un[k] = U[node_dof[k]];
}

// --------------------------------------------------------------------------
// Plain FP64 element kernel: one thread per uncut parent element.  Follows
// Assembler::assemble_elem's non-intersected branch (src/Assembler.cpp:327-372)
// with KernelVolumetric's B-matrix variant (src/AssemblerSmallStrainStructural.cpp:23-49):
// per qp K_tmp = (B^T D) B, P_tmp = B^T sigma, K_el += (J w) K_tmp.  For a linear
// simplex B, D and sigma are the same at every qp, so each entry is accumulated
// over the qps in reference order from one K_tmp value.
// Staged layout per element: Ke (nd*nd, row-major) then Pe (nd).
// --------------------------------------------------------------------------
template <int D>
__global__ void k_plain_fp64(DevProb p, const double* __restrict__ Un, double* __restrict__ stage) {
    constexpr int NEN = D + 1, ND = NEN * D, NS = (D == 2 ? 3 : 6), NQ = (D == 2 ? 4 : 5);
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= p.n_elem) return;
    if (p.cut_of_elem[e] >= 0) return;
    double x[NEN][3], g[NEN][3], dNdr[NEN][D];
    int nodes[NEN];
    for (int n = 0; n < NEN; ++n) {
        nodes[n] = p.conn[int64_t(e) * NEN + n];
        for (int i = 0; i < D; ++i) x[n][i] = p.coords[int64_t(nodes[n]) * D + i];
        for (int a = 0; a < D; ++a) dNdr[n][a] = simplex_dNdr(n, a);
    }
    const double J = transform<D>(NEN, dNdr, x, g);
    const pg_material& m = p.mats[p.elem_mat[e]];
    double wq[NQ];
    for (int q = 0; q < NQ; ++q) wq[q] = J * (D == 2 ? TRI_QW[q] : TET_QW[q]);
    // strain = B U (ProblemUtilities::B_U), stress (computeStress)
    double eps[NS], sig[NS];
    for (int k = 0; k < NS; ++k) eps[k] = 0.0;
    for (int n = 0; n < NEN; ++n)
        for (int c = 0; c < D; ++c) {
            double b[NS];
            bcol<D, double>(g[n], c, b);
            double u = Un[int64_t(nodes[n]) * D + c];
            for (int k = 0; k < NS; ++k) eps[k] += b[k] * u;
        }
    lei_stress<D>(m.D, eps, sig);
    double* Ke = stage + int64_t(e) * (ND * ND + ND);
    double* Pe = Ke + ND * ND;
    for (int i = 0; i < ND; ++i) {
        double bi[NS], btd[NS];
        bcol<D, double>(g[i / D], i % D, bi);
        for (int l = 0; l < NS; ++l) {
            double s = 0.0;
            for (int k = 0; k < NS; ++k) s += bi[k] * m.D[k * NS + l];
            btd[l] = s;
        }
        for (int j = 0; j < ND; ++j) {
            double bj[NS];
            bcol<D, double>(g[j / D], j % D, bj);
            double kt = 0.0;
            for (int l = 0; l < NS; ++l) kt += btd[l] * bj[l];
            double acc = 0.0;
            for (int q = 0; q < NQ; ++q) acc += wq[q] * kt;
            Ke[i * ND + j] = acc;
        }
        double pt = 0.0;
        for (int k = 0; k < NS; ++k) pt += bi[k] * sig[k];
        double acc = 0.0;
        for (int q = 0; q < NQ; ++q) acc += wq[q] * pt;
        Pe[i] = acc;
    }
}

// --------------------------------------------------------------------------
// Cut-element kernel: one wavefront (64 lanes, one block) per intersected
// element.  Restates Mesh::store_shape_functions for intersected elements
// (src/Mesh.cpp:1521-1641) and Assembler::assemble_elem's intersected branch
// (src/Assembler.cpp:375-517):
//   * per child qp: child shape gradients and J from the child's own Jacobian
//     (FP64: children at fiber interfaces can be slivers), parent gradients
//     (constant for linear parents), enrichment gradients = child gradient of
//     the child node that is the enrichment node (scale factor 1.0,
//     src/Mesh.cpp:2713-2725), weight = child J * w;
//   * K_el += wJ B^T D B, P_el += wJ B^T sigma, in TK precision for K
//     (double for the plain schedule, float for the mixed one), FP64 for P;
//   * cohesive children: delta = R sum N (U+ - U-), constitutive update of the
//     trial delta_max (FP64), K_coh = N N^T (x) R^T D R with +/- signs,
//     P_coh = -/+ N R^T t (src/Utilities.cpp:528-594), folded into K_el / P_el
//     through the cohesive element's local node list (src/Assembler.cpp:487-515).
// --------------------------------------------------------------------------
constexpr int CUT_MAXN = 4 + PG_MAX_ENR;   // local nodes
constexpr int CUT_MAXQ = 32;               // volumetric qps
constexpr int CUT_MAXC = PG_MAX_COH * 4;   // cohesive qps

template <int D, typename TK>
__global__ __launch_bounds__(64) void k_cut(DevProb p, const double* __restrict__ Un,
                                            const double* __restrict__ isv_c, double* __restrict__ isv_t,
                                            TK* __restrict__ Kst, double* __restrict__ Pst, int* __restrict__ err) {
    constexpr int NS = (D == 2 ? 3 : 6), NEN = D + 1;
    __shared__ double xs[CUT_MAXN][3];
    __shared__ double us[CUT_MAXN * 3];
    __shared__ double qg[CUT_MAXQ][CUT_MAXN][D];
    __shared__ double qw[CUT_MAXQ];
    __shared__ int qm[CUT_MAXQ];
    __shared__ double qs[CUT_MAXQ][NS];
    __shared__ double cN[CUT_MAXC][3];
    __shared__ double cw[CUT_MAXC];
    __shared__ double cRDR[CUT_MAXC][D][D];
    __shared__ double ctR[CUT_MAXC][D];
    __shared__ pg_cut_rec rec;
    __shared__ int lnode[CUT_MAXN];

    const int ci = blockIdx.x;
    const int lane = threadIdx.x;
    if (ci >= p.n_cut) return;
    const int e = p.cut_elem[ci];
    if (lane < int(sizeof(pg_cut_rec) / 4))
        reinterpret_cast<int32_t*>(&rec)[lane] = reinterpret_cast<const int32_t*>(&p.cut[ci])[lane];
    __syncthreads();
    const int nloc = NEN + rec.n_enr;
    const int nd = nloc * D;
    if (lane < nloc) {
        int nodeid = lane < NEN ? p.conn[int64_t(e) * NEN + lane] : rec.enr[lane - NEN];
        lnode[lane] = nodeid;
        for (int i = 0; i < D; ++i) xs[lane][i] = p.coords[int64_t(nodeid) * D + i];
        for (int i = 0; i < D; ++i) us[lane * D + i] = Un[int64_t(nodeid) * D + i];
    }
    __syncthreads();

    // ---- volumetric qps: one lane per (child, qp) ----
    const int nvq = rec.n_vqp;
    if (lane < nvq) {
        int k = 0, q = lane;
        for (; k < rec.n_child; ++k) {
            int nq = rec.child_type[k] == PG_CHILD_TET4 ? 5 : 4;
            if (q < nq) break;
            q -= nq;
        }
        const int ct = rec.child_type[k];
        const int cnn = ct == PG_CHILD_TRI3 ? 3 : 4;
        double cx[4][3], cdN[4][D], cg[4][3];
        for (int n = 0; n < cnn; ++n)
            for (int i = 0; i < D; ++i) cx[n][i] = xs[rec.child_nodes[k][n]][i];
        double w;
        if (ct == PG_CHILD_QUAD4) {
            double r = (q % 2 == 0) ? -PG_GAUSS2 : PG_GAUSS2;
            double s = (q / 2 == 0) ? -PG_GAUSS2 : PG_GAUSS2;
            double t4[4][2];
            quad4_dNdr(r, s, t4);
            for (int n = 0; n < 4; ++n)
                for (int a = 0; a < D; ++a) cdN[n][a] = t4[n][a];
            w = 1.0;
        } else {
            for (int n = 0; n < cnn; ++n)
                for (int a = 0; a < D; ++a) cdN[n][a] = simplex_dNdr(n, a);
            w = (D == 2) ? TRI_QW[q] : TET_QW[q];
        }
        const double Jc = transform<D>(cnn, cdN, cx, cg);
        if (Jc < 0.0) atomicOr(err, 1);  // store_shape_functions aborts (src/Mesh.cpp:1561)
        // parent gradients (linear parent: the same at every inverse-mapped point)
        double px[NEN][3], pdN[NEN][D], pgr[NEN][3];
        for (int n = 0; n < NEN; ++n) {
            for (int i = 0; i < D; ++i) px[n][i] = xs[n][i];
            for (int a = 0; a < D; ++a) pdN[n][a] = simplex_dNdr(n, a);
        }
        transform<D>(NEN, pdN, px, pgr);
        for (int l = 0; l < nloc; ++l) {
            int m = -1;
            if (l >= NEN)
                for (int n = 0; n < cnn; ++n)
                    if (rec.child_nodes[k][n] == l) m = n;
            for (int a = 0; a < D; ++a) qg[lane][l][a] = l < NEN ? pgr[l][a] : (m >= 0 ? cg[m][a] : 0.0);
        }
        qw[lane] = Jc * w;
        qm[lane] = rec.child_mat[k];
        // strain (B_U over all local dofs), stress
        double eps[NS], sig[NS];
        for (int s = 0; s < NS; ++s) eps[s] = 0.0;
        for (int l = 0; l < nloc; ++l)
            for (int c = 0; c < D; ++c) {
                double b[NS];
                bcol<D, double>(qg[lane][l], c, b);
                for (int s = 0; s < NS; ++s) eps[s] += b[s] * us[l * D + c];
            }
        lei_stress<D>(p.mats[qm[lane]].D, eps, sig);
        for (int s = 0; s < NS; ++s) qs[lane][s] = sig[s];
    }

    // ---- cohesive qps: one lane per (cohesive child, qp) ----
    const int cnn = rec.coh_nn / 2;  // nodes per surface
    const int cnq = (rec.coh_nn == 4) ? 2 : 4;
    const int ncq = rec.n_coh * cnq;
    if (lane < ncq) {
        const int k = lane / cnq, q = lane % cnq;
        double N[3], dNr[3][2];
        double xi0 = 0, xi1 = 0, w = 1.0;
        if (rec.coh_nn == 4) {  // CohEdge2, src/CohEdge2.cpp:70-78,186-209
            xi0 = (q == 0) ? -PG_GAUSS2 : PG_GAUSS2;
            N[0] = 0.5 * (1.0 - xi0); N[1] = 0.5 * (1 + xi0);
            dNr[0][0] = -0.5; dNr[1][0] = 0.5;
        } else {  // CohTri3: tri3 order-3 rule, src/CohTri3.cpp:43-204
            xi0 = TRI_QX[q]; xi1 = TRI_QY[q]; w = TRI_QW[q];
            N[0] = 1.0 - xi0 - xi1; N[1] = xi0; N[2] = xi1;
            dNr[0][0] = -1.0; dNr[0][1] = -1.0; dNr[1][0] = 1.0; dNr[1][1] = 0.0; dNr[2][0] = 0.0; dNr[2][1] = 1.0;
        }
        // CohesiveElem::compute_rotation_general on the lower surface (src/CohesiveElem.cpp:157-238)
        double Jc[2][3] = {{0, 0, 0}, {0, 0, 0}};
        for (int n = 0; n < cnn; ++n)
            for (int a = 0; a < D - 1; ++a)
                for (int i = 0; i < D; ++i) Jc[a][i] += dNr[n][a] * xs[rec.coh_nodes[k][n]][i];
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
        double R[D][D];
        for (int i = 0; i < D; ++i) R[D - 1][i] = nrm[i] / dA;
        for (int i = 0; i < D; ++i)
            for (int a = 0; a < D - 1; ++a) R[a][i] = Jc[a][i] / mag[a];
        // opening, cohesiveDeltaFast (src/Utilities.cpp:528-540)
        double dl[D], del[D];
        for (int i = 0; i < D; ++i) {
            double s = 0;
            for (int n = 0; n < cnn; ++n)
                s += N[n] * (us[rec.coh_nodes[k][n + cnn] * D + i] - us[rec.coh_nodes[k][n] * D + i]);
            dl[i] = s;
        }
        for (int i = 0; i < D; ++i) {
            double s = 0;
            for (int j = 0; j < D; ++j) s += R[i][j] * dl[j];
            del[i] = s;
        }
        const int iq = rec.isv_off + lane;
        double hist = isv_c ? isv_c[iq] : 0.0;
        double t[D], Dm[D][D];
        cohesive_law<D>(p.mats[rec.coh_mat[k]], del, hist, t, Dm);
        if (isv_t) isv_t[iq] = hist;
        // R^T D R and R^T t (cohesiveKFast / cohesiveInternalForceFast)
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
                cRDR[lane][i][j] = s;
            }
        for (int i = 0; i < D; ++i) {
            double s = 0;
            for (int l = 0; l < D; ++l) s += R[l][i] * t[l];
            ctR[lane][i] = s;
        }
        for (int n = 0; n < cnn; ++n) cN[lane][n] = N[n];
        cw[lane] = dA * w;
    }
    __syncthreads();

    TK* Ke = Kst + p.cut_koff[ci];
    double* Pe = Pst + p.cut_poff[ci];
    // ---- P_el ----
    for (int i = lane; i < nd; i += 64) {
        const int l = i / D, c = i % D;
        double acc = 0.0;
        for (int q = 0; q < nvq; ++q) {
            double b[NS];
            bcol<D, double>(qg[q][l], c, b);
            double pt = 0.0;
            for (int s = 0; s < NS; ++s) pt += b[s] * qs[q][s];
            acc += qw[q] * pt;
        }
        double pc = 0.0;
        for (int k = 0; k < rec.n_coh; ++k)
            for (int m = 0; m < rec.coh_nn; ++m)
                if (rec.coh_nodes[k][m] == l) {
                    const int n = m % cnn;
                    const double sgn = m < cnn ? -1.0 : 1.0;
                    double s = 0.0;
                    for (int q = 0; q < cnq; ++q) {
                        const int cq = k * cnq + q;
                        s += cw[cq] * (sgn * (cN[cq][n] * ctR[cq][c]));
                    }
                    pc += s;
                }
        Pe[i] = acc + pc;
    }
    // ---- K_el ----
    for (int idx = lane; idx < nd * nd; idx += 64) {
        const int i = idx / nd, j = idx % nd;
        const int li = i / D, c_i = i % D, lj = j / D, c_j = j % D;
        TK acc = TK(0);
        for (int q = 0; q < nvq; ++q) {
            TK gi[D], gj[D], bi[NS], bj[NS];
            for (int a = 0; a < D; ++a) { gi[a] = TK(qg[q][li][a]); gj[a] = TK(qg[q][lj][a]); }
            bcol<D, TK>(gi, c_i, bi);
            bcol<D, TK>(gj, c_j, bj);
            const double* Dm = p.mats[qm[q]].D;
            TK kt = TK(0);
            for (int l = 0; l < NS; ++l) {
                TK s = TK(0);
                for (int k = 0; k < NS; ++k) s += bi[k] * TK(Dm[k * NS + l]);
                kt += s * bj[l];
            }
            acc += TK(qw[q]) * kt;
        }
        // cohesive fold
        for (int k = 0; k < rec.n_coh; ++k) {
            int m1 = -1, m2 = -1;
            for (int m = 0; m < rec.coh_nn; ++m) {
                if (rec.coh_nodes[k][m] == li) m1 = m;
                if (rec.coh_nodes[k][m] == lj) m2 = m;
            }
            if (m1 < 0 || m2 < 0) continue;
            const int n1 = m1 % cnn, n2 = m2 % cnn;
            const TK sgn = ((m1 < cnn) == (m2 < cnn)) ? TK(1) : TK(-1);
            TK s = TK(0);
            for (int q = 0; q < cnq; ++q) {
                const int cq = k * cnq + q;
                s += TK(cw[cq]) * (sgn * (TK(cN[cq][n1]) * TK(cN[cq][n2]) * TK(cRDR[cq][c_i][c_j])));
            }
            acc += s;
        }
        Ke[idx] = acc;
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
// Node-pair kernel, plain schedule: thread per node pair (a, b); the d x d block
// K[a-dofs, b-dofs] is the sum over the elements incident to a (ascending, the
// reference's element loop order) that also contain b, of the staged element
// matrices.  Written once through the scatter indices (set_matrix,
// src/Assembler.cpp:811-831, without atomics).  Diagonal pairs also sum P
// (set_vec, src/Assembler.cpp:833-853).
// --------------------------------------------------------------------------
template <int D>
__global__ void k_pair_staged(DevProb p, const double* __restrict__ stage, const double* __restrict__ Kc,
                              const double* __restrict__ Pc, double* __restrict__ K, double* __restrict__ P,
                              unsigned flags) {
    constexpr int NEN = D + 1, ND = NEN * D;
    int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= p.n_pairs) return;
    const int a = p.npat_row[q], b = p.npat_col[q];
    double acc[D][D];
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) acc[i][j] = 0.0;
    double pa[D];
    for (int i = 0; i < D; ++i) pa[i] = 0.0;
    const bool diag = (a == b) && (flags & PG_FUNC);
    const int e0 = p.n2e_ptr[a], e1 = p.n2e_ptr[a + 1];
    for (int t = e0; t < e1; ++t) {
        const int e = p.n2e[t];
        const int ci = p.cut_of_elem[e];
        const int la = find_local(p, e, ci, a, NEN);
        const int lb = (a == b) ? la : find_local(p, e, ci, b, NEN);
        if (la < 0 || lb < 0) continue;
        const double* Ke;
        const double* Pe;
        int nd;
        if (ci < 0) {
            Ke = stage + int64_t(e) * (ND * ND + ND);
            Pe = Ke + ND * ND;
            nd = ND;
        } else {
            Ke = Kc + p.cut_koff[ci];
            Pe = Pc + p.cut_poff[ci];
            nd = (NEN + p.cut[ci].n_enr) * D;
        }
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) acc[i][j] += Ke[int64_t(la * D + i) * nd + lb * D + j];
        if (diag)
            for (int i = 0; i < D; ++i) pa[i] += Pe[la * D + i];
    }
    if (flags & PG_JAC) {
        const int32_t* mp = p.npat_map + q * D * D;
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) K[mp[i * D + j]] = acc[i][j];
    }
    if (diag)
        for (int i = 0; i < D; ++i) P[p.node_dof[int64_t(a) * D + i]] = pa[i];
}

// --------------------------------------------------------------------------
// Node-pair kernel, mixed schedule: uncut elements are recomputed here (FP32
// B^T D B block of the pair, weight sum over the reference qps in FP64, sum in
// FP64); cut elements contribute their staged FP32 tangent.  The internal force
// of diagonal pairs is recomputed in FP64.
// --------------------------------------------------------------------------
template <int D>
__global__ void k_pair_fused(DevProb p, const double* __restrict__ Un, const float* __restrict__ Kc,
                             const double* __restrict__ Pc, double* __restrict__ K, double* __restrict__ P,
                             unsigned flags) {
    constexpr int NEN = D + 1, NS = (D == 2 ? 3 : 6), NQ = (D == 2 ? 4 : 5);
    int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= p.n_pairs) return;
    const int a = p.npat_row[q], b = p.npat_col[q];
    double acc[D][D];
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) acc[i][j] = 0.0;
    double pa[D];
    for (int i = 0; i < D; ++i) pa[i] = 0.0;
    const bool diag = (a == b) && (flags & PG_FUNC);
    const bool jac = (flags & PG_JAC) != 0;
    const int e0 = p.n2e_ptr[a], e1 = p.n2e_ptr[a + 1];
    for (int t = e0; t < e1; ++t) {
        const int e = p.n2e[t];
        const int ci = p.cut_of_elem[e];
        if (ci < 0) {
            int nodes[NEN];
            int la = -1, lb = -1;
            for (int n = 0; n < NEN; ++n) {
                nodes[n] = p.conn[int64_t(e) * NEN + n];
                if (nodes[n] == a) la = n;
                if (nodes[n] == b) lb = n;
            }
            if (la < 0 || lb < 0) continue;
            // FP64 coordinates -> Jacobian, FP64 gradients (cheap), FP32 block
            double x[NEN][3], dNdr[NEN][D], g[NEN][3];
            for (int n = 0; n < NEN; ++n) {
                for (int i = 0; i < D; ++i) x[n][i] = p.coords[int64_t(nodes[n]) * D + i];
                for (int c = 0; c < D; ++c) dNdr[n][c] = simplex_dNdr(n, c);
            }
            const double J = transform<D>(NEN, dNdr, x, g);
            double wsum = 0.0;
            for (int k = 0; k < NQ; ++k) wsum += J * (D == 2 ? TRI_QW[k] : TET_QW[k]);
            const double* Dm = p.mats[p.elem_mat[e]].D;
            if (jac) {
                float ga[D], gb[D];
                for (int c = 0; c < D; ++c) { ga[c] = float(g[la][c]); gb[c] = float(g[lb][c]); }
                const float wf = float(wsum);
                for (int i = 0; i < D; ++i) {
                    float bi[NS];
                    bcol<D, float>(ga, i, bi);
                    float btd[NS];
                    for (int l = 0; l < NS; ++l) {
                        float s = 0.f;
                        for (int k = 0; k < NS; ++k) s += bi[k] * float(Dm[k * NS + l]);
                        btd[l] = s;
                    }
                    for (int j = 0; j < D; ++j) {
                        float bj[NS];
                        bcol<D, float>(gb, j, bj);
                        float kt = 0.f;
                        for (int l = 0; l < NS; ++l) kt += btd[l] * bj[l];
                        acc[i][j] += double(wf * kt);
                    }
                }
            }
            if (diag) {
                double eps[NS], sig[NS];
                for (int k = 0; k < NS; ++k) eps[k] = 0.0;
                for (int n = 0; n < NEN; ++n)
                    for (int c = 0; c < D; ++c) {
                        double bb[NS];
                        bcol<D, double>(g[n], c, bb);
                        const double u = Un[int64_t(nodes[n]) * D + c];
                        for (int k = 0; k < NS; ++k) eps[k] += bb[k] * u;
                    }
                lei_stress<D>(Dm, eps, sig);
                for (int i = 0; i < D; ++i) {
                    double bi[NS];
                    bcol<D, double>(g[la], i, bi);
                    double pt = 0.0;
                    for (int k = 0; k < NS; ++k) pt += bi[k] * sig[k];
                    pa[i] += wsum * pt;
                }
            }
        } else {
            const int la = find_local(p, e, ci, a, NEN);
            const int lb = (a == b) ? la : find_local(p, e, ci, b, NEN);
            if (la < 0 || lb < 0) continue;
            const int nd = (NEN + p.cut[ci].n_enr) * D;
            if (jac) {
                const float* Ke = Kc + p.cut_koff[ci];
                for (int i = 0; i < D; ++i)
                    for (int j = 0; j < D; ++j) acc[i][j] += double(Ke[int64_t(la * D + i) * nd + lb * D + j]);
            }
            if (diag) {
                const double* Pe = Pc + p.cut_poff[ci];
                for (int i = 0; i < D; ++i) pa[i] += Pe[la * D + i];
            }
        }
    }
    if (jac) {
        const int32_t* mp = p.npat_map + q * D * D;
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) K[mp[i * D + j]] = acc[i][j];
    }
    if (diag)
        for (int i = 0; i < D; ++i) P[p.node_dof[int64_t(a) * D + i]] = pa[i];
}

template <typename T>
int dup(const T* src, size_t n, T** dst) {
    *dst = nullptr;
    if (n == 0) return PG_OK;
    HIP_OK(hipMalloc(reinterpret_cast<void**>(dst), n * sizeof(T)));
    HIP_OK(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return PG_OK;
}

}  // namespace

struct pg_asm {
    DevProb d{};
    int64_t n_dof = 0, nnz_total = 0;
    int64_t nnz[4] = {0, 0, 0, 0};
    int64_t cut_k_total = 0, cut_p_total = 0;
    std::vector<void*> allocs;
    double* Un = nullptr;        // node-major U
    double* stage = nullptr;     // plain schedule: per uncut element Ke + Pe (FP64)
    double* Kc64 = nullptr;      // cut tangents (FP64)
    float* Kc32 = nullptr;       // cut tangents (FP32)
    double* Pc = nullptr;        // cut residuals (FP64)
    int* err = nullptr;
    // host-API buffers
    double *dU = nullptr, *dK = nullptr, *dP = nullptr, *dIc = nullptr, *dIt = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool timing = false;
    double last_ms[3] = {0, 0, 0};
    int pair_block = 256;
    int elem_block = 256;
    int64_t alg_bytes = 0;
};

namespace {

int check_desc(const pg_problem_desc* s) {
    if (!s) { g_err = "null description"; return PG_E_ARG; }
    if (s->dim != 2 && s->dim != 3) { g_err = "dim must be 2 or 3"; return PG_E_ARG; }
    if (s->nen != s->dim + 1) { g_err = "only tri3 (2D) and tet4 (3D) parents are supported"; return PG_E_UNSUPPORTED; }
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
    for (int c = 0; c < s->n_cut; ++c) {
        const pg_cut_rec& r = s->cut[c];
        if (r.n_enr > PG_MAX_ENR || r.n_child > PG_MAX_CHILD || r.n_coh > PG_MAX_COH || r.n_vqp > CUT_MAXQ) {
            g_err = "cut record exceeds the supported class sizes";
            return PG_E_UNSUPPORTED;
        }
        int nq = 0;
        for (int k = 0; k < r.n_child; ++k) {
            int t = r.child_type[k];
            if (s->dim == 2 && t != PG_CHILD_TRI3 && t != PG_CHILD_QUAD4) { g_err = "bad 2D child"; return PG_E_UNSUPPORTED; }
            if (s->dim == 3 && t != PG_CHILD_TET4) { g_err = "bad 3D child"; return PG_E_UNSUPPORTED; }
            nq += t == PG_CHILD_TET4 ? 5 : 4;
        }
        if (nq != r.n_vqp) { g_err = "cut record qp count mismatch"; return PG_E_ARG; }
        if (r.n_coh && r.coh_nn != (s->dim == 2 ? 4 : 6)) { g_err = "bad cohesive child"; return PG_E_UNSUPPORTED; }
        if (r.isv_off < 0 || r.isv_off + r.n_coh * (r.coh_nn == 4 ? 2 : 4) > s->n_cqp) {
            g_err = "cut record ISV range out of bounds";
            return PG_E_ARG;
        }
        int nloc = s->nen + r.n_enr;
        for (int k = 0; k < r.n_child; ++k)
            for (int n = 0; n < (r.child_type[k] == PG_CHILD_TRI3 ? 3 : 4); ++n)
                if (r.child_nodes[k][n] >= nloc) { g_err = "child node out of range"; return PG_E_ARG; }
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
    return PG_OK;
}

int launch(pg_asm* h, const double* dU, const double* dIc, double* dIt, unsigned flags, double* dK, double* dP,
           hipStream_t st) {
    const DevProb& d = h->d;
    const int64_t nall = int64_t(d.n_nodes) + d.n_enr;
    const int64_t nu = nall * d.dim;
    const bool mixed = (flags & (PG_FP32 | PG_BATCHED)) != 0;
    if (h->timing) HIP_OK(hipEventRecord(h->ev[0], st));
    k_permute_u<<<dim3(unsigned((nu + 255) / 256)), dim3(256), 0, st>>>(nu, d.node_dof, dU, h->Un);
    if (d.n_cut > 0) {
        if (d.dim == 2) {
            if (mixed) k_cut<2, float><<<d.n_cut, 64, 0, st>>>(d, h->Un, dIc, dIt, h->Kc32, h->Pc, h->err);
            else k_cut<2, double><<<d.n_cut, 64, 0, st>>>(d, h->Un, dIc, dIt, h->Kc64, h->Pc, h->err);
        } else {
            if (mixed) k_cut<3, float><<<d.n_cut, 64, 0, st>>>(d, h->Un, dIc, dIt, h->Kc32, h->Pc, h->err);
            else k_cut<3, double><<<d.n_cut, 64, 0, st>>>(d, h->Un, dIc, dIt, h->Kc64, h->Pc, h->err);
        }
    }
    if (!mixed) {
        const unsigned gb = unsigned((d.n_elem + h->elem_block - 1) / h->elem_block);
        if (d.dim == 2) k_plain_fp64<2><<<gb, h->elem_block, 0, st>>>(d, h->Un, h->stage);
        else k_plain_fp64<3><<<gb, h->elem_block, 0, st>>>(d, h->Un, h->stage);
    }
    if (h->timing) HIP_OK(hipEventRecord(h->ev[1], st));
    const unsigned gp = unsigned((d.n_pairs + h->pair_block - 1) / h->pair_block);
    if (mixed) {
        if (d.dim == 2) k_pair_fused<2><<<gp, h->pair_block, 0, st>>>(d, h->Un, h->Kc32, h->Pc, dK, dP, flags);
        else k_pair_fused<3><<<gp, h->pair_block, 0, st>>>(d, h->Un, h->Kc32, h->Pc, dK, dP, flags);
    } else {
        if (d.dim == 2) k_pair_staged<2><<<gp, h->pair_block, 0, st>>>(d, h->stage, h->Kc64, h->Pc, dK, dP, flags);
        else k_pair_staged<3><<<gp, h->pair_block, 0, st>>>(d, h->stage, h->Kc64, h->Pc, dK, dP, flags);
    }
    HIP_OK(hipGetLastError());
    if (h->timing) HIP_OK(hipEventRecord(h->ev[2], st));
    return PG_OK;
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
    UP(npat_row, s->npat_row, s->n_pairs, int32_t);
    UP(npat_col, s->npat_col, s->n_pairs, int32_t);
    UP(npat_map, s->npat_map, s->n_pairs * s->dim * s->dim, int32_t);
    UP(n2e_ptr, s->n2e_ptr, nall + 1, int32_t);
    UP(n2e, s->n2e, s->n2e_ptr[nall], int32_t);
    UP(cut_koff, koff.data(), s->n_cut + 1, int64_t);
    UP(cut_poff, poff.data(), s->n_cut + 1, int64_t);
#undef UP
    const int ndp = s->nen * s->dim;
    if ((rc = alloc(h, &h->Un, size_t(nall * s->dim)))) return fail(rc);
    if ((rc = alloc(h, &h->stage, size_t(s->n_elem) * size_t(ndp * ndp + ndp)))) return fail(rc);
    if ((rc = alloc(h, &h->Kc64, size_t(h->cut_k_total)))) return fail(rc);
    if ((rc = alloc(h, &h->Kc32, size_t(h->cut_k_total)))) return fail(rc);
    if ((rc = alloc(h, &h->Pc, size_t(h->cut_p_total)))) return fail(rc);
    if ((rc = alloc(h, &h->err, 1))) return fail(rc);
    if (hipMemset(h->err, 0, sizeof(int)) != hipSuccess) { g_err = "hipMemset failed"; return fail(PG_E_HIP); }
    if (hipStreamCreate(&h->stream) != hipSuccess) { g_err = "hipStreamCreate failed"; return fail(PG_E_HIP); }
    for (auto& e : h->ev)
        if (hipEventCreate(&e) != hipSuccess) { g_err = "hipEventCreate failed"; return fail(PG_E_HIP); }
    // algorithmic bytes of one K+R assembly
    int64_t B = 4LL * s->nen * s->n_elem + 8LL * s->dim * nall + 8LL * h->n_dof + int64_t(sizeof(pg_cut_rec)) * s->n_cut +
                16LL * s->n_cqp + 12LL * h->nnz_total + 4LL * (int64_t(s->n_free) + s->n_pres + 2) + 8LL * h->n_dof;
    h->alg_bytes = B;
    *out = h;
    return PG_OK;
}

int pg_asm_destroy(pg_asm* h) {
    if (!h) return PG_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    for (void* p : h->allocs) hipFree(p);
    if (h->dU) hipFree(h->dU);
    if (h->dK) hipFree(h->dK);
    if (h->dP) hipFree(h->dP);
    if (h->dIc) hipFree(h->dIc);
    if (h->dIt) hipFree(h->dIt);
    for (auto& e : h->ev)
        if (e) hipEventDestroy(e);
    if (h->stream) hipStreamDestroy(h->stream);
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

int pg_asm_set_timing(pg_asm* h, int enable) {
    if (!h) { g_err = "null handle"; return PG_E_ARG; }
    h->timing = enable != 0;
    return PG_OK;
}

int pg_asm_set_tiles(pg_asm* h, int pair_block, int cut_block) {
    if (!h) { g_err = "null handle"; return PG_E_ARG; }
    auto ok = [](int b) { return b == 0 || (b >= 64 && b <= 1024 && b % 64 == 0); };
    if (!ok(pair_block) || !ok(cut_block)) { g_err = "block sizes must be multiples of 64 in [64, 1024]"; return PG_E_ARG; }
    if (pair_block) h->pair_block = pair_block;
    if (cut_block) h->elem_block = cut_block;
    return PG_OK;
}

int pg_asm_last_timing(const pg_asm* h, double ms[3]) {
    if (!h || !ms) { g_err = "null argument"; return PG_E_ARG; }
    for (int i = 0; i < 3; ++i) ms[i] = h->last_ms[i];
    return PG_OK;
}

int pg_asm_assemble_device(pg_asm* h, const double* dU, const double* dIc, double* dIt, double delta_t,
                           unsigned flags, double* dK, double* dP, void* stream) {
    (void)delta_t;
    if (!h) { g_err = "null handle"; return PG_E_ARG; }
    if (!(flags & (PG_FUNC | PG_JAC))) return PG_OK;
    if (!dU || ((flags & PG_JAC) && !dK) || ((flags & PG_FUNC) && !dP)) { g_err = "missing device buffer"; return PG_E_ARG; }
    if (h->d.n_cqp > 0 && (!dIc || !dIt)) { g_err = "cohesive problem needs ISV buffers"; return PG_E_ARG; }
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : h->stream;
    int rc = launch(h, dU, dIc, dIt, flags, dK, dP, st);
    if (rc) return rc;
    if (h->timing) {
        HIP_OK(hipEventSynchronize(h->ev[2]));
        float a = 0, b = 0;
        HIP_OK(hipEventElapsedTime(&a, h->ev[0], h->ev[1]));
        HIP_OK(hipEventElapsedTime(&b, h->ev[1], h->ev[2]));
        h->last_ms[0] = double(a) + double(b);
        h->last_ms[1] = a;
        h->last_ms[2] = b;
    }
    return PG_OK;
}

int pg_asm_assemble(pg_asm* h, const double* U, const double* Ic, double* It, double delta_t, unsigned flags,
                    double* K, double* P) {
    if (!h) { g_err = "null handle"; return PG_E_ARG; }
    if (!(flags & (PG_FUNC | PG_JAC))) return PG_OK;
    if (!U || ((flags & PG_JAC) && !K) || ((flags & PG_FUNC) && !P)) { g_err = "missing host buffer"; return PG_E_ARG; }
    const int64_t ncq = h->d.n_cqp;
    if (ncq > 0 && (!Ic || !It)) { g_err = "cohesive problem needs ISV buffers"; return PG_E_ARG; }
    if (!h->dU) {
        HIP_OK(hipMalloc(&h->dU, size_t(h->n_dof) * sizeof(double)));
        HIP_OK(hipMalloc(&h->dK, size_t(std::max<int64_t>(h->nnz_total, 1)) * sizeof(double)));
        HIP_OK(hipMalloc(&h->dP, size_t(h->n_dof) * sizeof(double)));
        HIP_OK(hipMalloc(&h->dIc, size_t(std::max<int64_t>(ncq, 1)) * sizeof(double)));
        HIP_OK(hipMalloc(&h->dIt, size_t(std::max<int64_t>(ncq, 1)) * sizeof(double)));
    }
    hipStream_t st = h->stream;
    HIP_OK(hipMemsetAsync(h->err, 0, sizeof(int), st));
    HIP_OK(hipMemcpyAsync(h->dU, U, size_t(h->n_dof) * sizeof(double), hipMemcpyHostToDevice, st));
    if (ncq > 0) HIP_OK(hipMemcpyAsync(h->dIc, Ic, size_t(ncq) * sizeof(double), hipMemcpyHostToDevice, st));
    int rc = pg_asm_assemble_device(h, h->dU, h->dIc, h->dIt, delta_t, flags, h->dK, h->dP, st);
    if (rc) return rc;
    if (flags & PG_JAC)
        HIP_OK(hipMemcpyAsync(K, h->dK, size_t(h->nnz_total) * sizeof(double), hipMemcpyDeviceToHost, st));
    if (flags & PG_FUNC) HIP_OK(hipMemcpyAsync(P, h->dP, size_t(h->n_dof) * sizeof(double), hipMemcpyDeviceToHost, st));
    if (ncq > 0) HIP_OK(hipMemcpyAsync(It, h->dIt, size_t(ncq) * sizeof(double), hipMemcpyDeviceToHost, st));
    int herr = 0;
    HIP_OK(hipMemcpyAsync(&herr, h->err, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (herr) { g_err = "negative child Jacobian in a cut element"; return PG_E_MESH; }
    return PG_OK;
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

}  // extern "C"
