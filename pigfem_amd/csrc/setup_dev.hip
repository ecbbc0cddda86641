// Device build of the set-up's detection and enrichment stages:
//   node detection (Mesh::detect_node, src/Mesh.cpp:2883-2930): one thread per
//     node over the inclusions of the 3^dim bins around it; the largest
//     inclusion index that contains the node wins (the reference checks the
//     inclusions from the last to the first and stops at the first hit);
//   element analysis (Mesh::analyze_elements, src/Mesh.cpp:2962-3071): one
//     thread per element runs the detection tables; the cut elements are
//     compacted in ascending element order and each builds its record;
//   enrichment nodes (Mesh::add_enrichments, src/Mesh.cpp:3104-3251): the
//     (sorted edge, inclusion) keys of every cut slot, radix-sorted and made
//     unique, give the enrichment nodes in the reference's std::set order; one
//     thread per key finds the edge intersection (Inclusion::find_intersection);
//     each cut element then finds its slots' keys by binary search
//     (add_enrich_node_to_elem, src/Mesh.cpp:3661-3720);
//   ISV offsets: an exclusive prefix of the cohesive qps per cut element.
// The arithmetic is cut_tables.h's, shared with the host harness, so the
// outputs equal the host build's array for array.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>

#include "cut_tables.h"
#include "dev_pool.h"
#include "setup_dev.h"

namespace {

using pgh::blocks;
using pgh::down;
using pgh::Pool;
using pgh::prefix;

constexpr int SLOTS = PG_MAX_ENR / 2;  // cut-edge slots per element (mirrors excluded)
constexpr uint64_t NO_KEY = ~uint64_t(0);

// Uniform bins of inclusion centres (the host harness's inclusion grid).
struct Bins {
    int dim, nc;
    double cell;
    const int32_t* ptr;    // nc^dim + 1
    const int32_t* items;  // inclusion indices by bin
};

__device__ inline int bin_idx(double v, const Bins& B) {
    const int i = int(floor(v / B.cell));
    return min(B.nc - 1, max(0, i));
}

__global__ void k_detect_nodes(int32_t n, const double* __restrict__ x, Bins B, const double* __restrict__ ic,
                               const double* __restrict__ ir, int32_t* __restrict__ det, int32_t* __restrict__ bad) {
    const int32_t a = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (a >= n) return;
    const int dim = B.dim;
    const double* p = x + int64_t(a) * dim;
    const int c0 = bin_idx(p[0], B), c1 = bin_idx(p[1], B), c2 = dim == 3 ? bin_idx(p[2], B) : 0;
    int best = -1, res = 0;
    const int z0 = dim == 3 ? max(0, c2 - 1) : 0, z1 = dim == 3 ? min(B.nc - 1, c2 + 1) : 0;
    for (int k = z0; k <= z1; ++k)
        for (int j = max(0, c1 - 1); j <= min(B.nc - 1, c1 + 1); ++j)
            for (int i = max(0, c0 - 1); i <= min(B.nc - 1, c0 + 1); ++i) {
                const int64_t b = int64_t(i) + int64_t(j) * B.nc + int64_t(k) * B.nc * B.nc;
                for (int32_t t = B.ptr[b]; t < B.ptr[b + 1]; ++t) {
                    const int32_t o = B.items[t];
                    if (o <= best) continue;
                    const int r = pgh::detect_one(dim, ic + int64_t(o) * dim, ir[o], p);
                    if (r) {
                        best = o;
                        res = r;
                    }
                }
            }
    det[a] = res == 2 ? best : -1;
    if (res == 1) atomicMin(bad, a);
}

// detection of element e from its nodes' detection values
__device__ inline int detect_elem(int nen, const int32_t* __restrict__ conn, const int32_t* __restrict__ det,
                                  int32_t e, int* nd, pgh::CutDetect& D, bool* cut) {
    const int32_t* c = conn + int64_t(e) * nen;
    for (int k = 0; k < nen; ++k) nd[k] = det[c[k]];
    int rc = pgh::CUT_OK;
    if (nen == 3) *cut = pgh::detect_tri3(nd, D);
    else rc = pgh::detect_tet4(nd, c, D, cut);
    if (!rc && *cut && D.nchild > (nen == 3 ? 4 : PG_MAX_CHILD)) rc = pgh::CUT_E_CHILDREN;
    return rc;
}

__global__ void k_elem_flags(int32_t n, int nen, const int32_t* __restrict__ conn, const int32_t* __restrict__ det,
                             int32_t* __restrict__ elem_mat, int32_t* __restrict__ flag,
                             unsigned long long* __restrict__ bad) {
    const int32_t e = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= n) return;
    int nd[4];
    pgh::CutDetect D;
    bool cut = false;
    const int rc = detect_elem(nen, conn, det, e, nd, D, &cut);
    if (rc) {
        atomicMin(bad, (static_cast<unsigned long long>(e) << 8) | unsigned(rc));
        cut = false;
    }
    flag[e] = cut ? 1 : 0;
    elem_mat[e] = (!cut && nd[0] != -1) ? 1 : 0;  // a cut element keeps the matrix
}

__global__ void k_compact(int32_t n, const int32_t* __restrict__ flag, const int32_t* __restrict__ pos,
                          int32_t* __restrict__ cut_elem, int32_t* __restrict__ cut_of_elem) {
    const int32_t e = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= n) return;
    cut_of_elem[e] = flag[e] ? pos[e] : -1;
    if (flag[e]) cut_elem[pos[e]] = e;
}

// the cut record (children, cohesive children), the element's cohesive qps and
// its slots' keys (edge nodes ascending | inclusion); unused slots hold NO_KEY
__global__ void k_cut_records(int32_t n_cut, int dim, int nen, bool coh, const int32_t* __restrict__ cut_elem,
                              const int32_t* __restrict__ conn, const int32_t* __restrict__ det,
                              pg_cut_rec* __restrict__ recs, int32_t* __restrict__ ncq, uint64_t* __restrict__ klohi,
                              uint32_t* __restrict__ kinc) {
    const int32_t i = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n_cut) return;
    const int32_t e = cut_elem[i];
    int nd[4];
    pgh::CutDetect D;
    bool cut = false;
    (void)detect_elem(nen, conn, det, e, nd, D, &cut);  // checked by k_elem_flags
    pg_cut_rec R;
    ncq[i] = pgh::cut_record(D, dim, nen, coh, R);
    recs[i] = R;
    const int32_t* c = conn + int64_t(e) * nen;
    for (int s = 0; s < SLOTS; ++s) {
        uint64_t key = NO_KEY;
        uint32_t inc = 0x7fffffffu;
        if (s < D.ncut) {
            const int le = D.cut_edge[s];
            const int a = nen == 3 ? pgh::tri_edge(le, 0) : pgh::tet_edge(le, 0);
            const int b = nen == 3 ? pgh::tri_edge(le, 1) : pgh::tet_edge(le, 1);
            const uint32_t lo = uint32_t(min(c[a], c[b])), hi = uint32_t(max(c[a], c[b]));
            key = (uint64_t(lo) << 32) | hi;
            inc = uint32_t(D.slot_incl[s]);
        }
        klohi[int64_t(i) * SLOTS + s] = key;
        kinc[int64_t(i) * SLOTS + s] = inc;
    }
}

__global__ void k_unique_flags(int64_t n, const uint64_t* __restrict__ lohi, const uint32_t* __restrict__ inc,
                               int32_t* __restrict__ uflag) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uflag[k] = lohi[k] != NO_KEY && (k == 0 || lohi[k] != lohi[k - 1] || inc[k] != inc[k - 1]);
}

__global__ void k_unique_out(int64_t n, const int32_t* __restrict__ uflag, const int32_t* __restrict__ upos,
                             const uint64_t* __restrict__ lohi, const uint32_t* __restrict__ inc,
                             uint64_t* __restrict__ ulohi, uint32_t* __restrict__ uinc) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n || !uflag[k]) return;
    ulohi[upos[k]] = lohi[k];
    uinc[upos[k]] = inc[k];
}

// enrichment node k (and its mirror n_init + k): intersection, position, edge
__global__ void k_intersect(int32_t n_init, int dim, bool coh, const double* __restrict__ x,
                            const uint64_t* __restrict__ ulohi, const uint32_t* __restrict__ uinc,
                            const double* __restrict__ ic, const double* __restrict__ ir, double* __restrict__ alpha,
                            double* __restrict__ xe, int32_t* __restrict__ enr_incl, int32_t* __restrict__ enr_edge,
                            int32_t* __restrict__ bad) {
    const int32_t k = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= n_init) return;
    const int32_t lo = int32_t(ulohi[k] >> 32), hi = int32_t(ulohi[k] & 0xffffffffu), in = int32_t(uinc[k]);
    const double* x0 = x + int64_t(lo) * dim;
    const double* x1 = x + int64_t(hi) * dim;
    int rc = 0;
    const double a = pgh::intersect_edge(dim, x0, x1, ic + int64_t(in) * dim, ir[in], &rc);
    if (rc) atomicOr(bad, 1);
    for (int m = 0; m < (coh ? 2 : 1); ++m) {
        const int32_t j = k + m * n_init;
        alpha[j] = a;
        for (int d = 0; d < dim; ++d) xe[int64_t(j) * dim + d] = x0[d] + a * (x1[d] - x0[d]);
        enr_incl[j] = in;
        enr_edge[int64_t(j) * 2] = lo;
        enr_edge[int64_t(j) * 2 + 1] = hi;
    }
}

// each slot's enrichment node (mirror at slot + ncut) and the ISV offset
__global__ void k_attach(int32_t n_cut, int32_t n_nodes, int32_t n_init, bool coh, pg_cut_rec* __restrict__ recs,
                         const int32_t* __restrict__ isv_off, const uint64_t* __restrict__ klohi,
                         const uint32_t* __restrict__ kinc, const uint64_t* __restrict__ ulohi,
                         const uint32_t* __restrict__ uinc) {
    const int32_t i = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n_cut) return;
    pg_cut_rec& R = recs[i];
    const int ncut = coh ? R.n_enr / 2 : R.n_enr;
    for (int s = 0; s < ncut; ++s) {
        const uint64_t key = klohi[int64_t(i) * SLOTS + s];
        const uint32_t in = kinc[int64_t(i) * SLOTS + s];
        int32_t lo = 0, hi = n_init;  // first unique key not less than (key, in)
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (ulohi[mid] < key || (ulohi[mid] == key && uinc[mid] < in)) lo = mid + 1;
            else hi = mid;
        }
        R.enr[s] = n_nodes + lo;
        if (coh) R.enr[s + ncut] = n_nodes + n_init + lo;
    }
    R.isv_off = isv_off[i];
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

namespace pgh {

int setup_device(const SetupIn& in, SetupOut& out, std::string& err) {
    const int dim = in.dim, nen = in.nen;
    const int32_t nn = in.n_nodes, ne = in.n_elem;
    if ((dim != 2 && dim != 3) || nen != dim + 1 || nn <= 0 || ne <= 0 || in.n_incl < 0 || !(in.cell > 0.0)) {
        err = "bad set-up input";
        return PG_E_ARG;
    }
    const double t0 = now_s();
    Pool pool;
    // inclusion bins (host: a counting pass over the inclusions)
    Bins B{};
    B.dim = dim;
    B.cell = in.cell;
    B.nc = int(ceil(1.0 / in.cell));
    if (B.nc < 1) B.nc = 1;
    const int64_t nb = int64_t(B.nc) * B.nc * (dim == 3 ? B.nc : 1);
    std::vector<int32_t> bptr(size_t(nb) + 1, 0), bitems(size_t(in.n_incl));
    auto hidx = [&](double v) { return std::min(B.nc - 1, std::max(0, int(std::floor(v / in.cell)))); };
    auto hbin = [&](int32_t j) {
        const double* c = in.incl_c + int64_t(j) * dim;
        return int64_t(hidx(c[0])) + int64_t(hidx(c[1])) * B.nc + (dim == 3 ? int64_t(hidx(c[2])) * B.nc * B.nc : 0);
    };
    for (int32_t j = 0; j < in.n_incl; ++j) bptr[size_t(hbin(j)) + 1]++;
    for (int64_t b = 0; b < nb; ++b) bptr[size_t(b) + 1] += bptr[size_t(b)];
    {
        std::vector<int32_t> fill(bptr.begin(), bptr.end() - 1);
        for (int32_t j = 0; j < in.n_incl; ++j) bitems[size_t(fill[size_t(hbin(j))]++)] = j;
    }
    double *dx, *dic, *dir;
    int32_t *dconn, *dbp, *dbi, *ddet, *dbad;
    DEV_OK(pool.get(&dx, size_t(nn) * dim));
    DEV_OK(pool.get(&dconn, size_t(ne) * nen));
    DEV_OK(pool.get(&dic, size_t(in.n_incl) * dim));
    DEV_OK(pool.get(&dir, size_t(in.n_incl)));
    DEV_OK(pool.get(&dbp, size_t(nb) + 1));
    DEV_OK(pool.get(&dbi, size_t(in.n_incl)));
    DEV_OK(pool.get(&ddet, size_t(nn)));
    DEV_OK(pool.get(&dbad, 1));
    DEV_OK(hipMemcpy(dx, in.coords, size_t(nn) * dim * 8, hipMemcpyHostToDevice));
    DEV_OK(hipMemcpy(dconn, in.conn, size_t(ne) * nen * 4, hipMemcpyHostToDevice));
    if (in.n_incl) {
        DEV_OK(hipMemcpy(dic, in.incl_c, size_t(in.n_incl) * dim * 8, hipMemcpyHostToDevice));
        DEV_OK(hipMemcpy(dir, in.incl_r, size_t(in.n_incl) * 8, hipMemcpyHostToDevice));
        DEV_OK(hipMemcpy(dbi, bitems.data(), size_t(in.n_incl) * 4, hipMemcpyHostToDevice));
    }
    DEV_OK(hipMemcpy(dbp, bptr.data(), (size_t(nb) + 1) * 4, hipMemcpyHostToDevice));
    B.ptr = dbp;
    B.items = dbi;
    const int32_t no_bad = 0x7fffffff;
    DEV_OK(hipMemcpy(dbad, &no_bad, 4, hipMemcpyHostToDevice));
    k_detect_nodes<<<blocks(nn), 256>>>(nn, dx, B, dic, dir, ddet, dbad);
    DEV_OK(hipGetLastError());
    int32_t hbad = 0;
    DEV_OK(hipMemcpy(&hbad, dbad, 4, hipMemcpyDeviceToHost));
    if (hbad != no_bad) {
        err = "node " + std::to_string(hbad) + " lies on an inclusion surface";
        return PG_E_MESH;
    }
    const double t1 = now_s();
    out.seconds[0] = t1 - t0;

    // element analysis
    int32_t *dmat, *dflag, *dpos, *dcoe;
    unsigned long long* debad;
    DEV_OK(pool.get(&dmat, size_t(ne)));
    DEV_OK(pool.get(&dflag, size_t(ne)));
    DEV_OK(pool.get(&dpos, size_t(ne) + 1));
    DEV_OK(pool.get(&dcoe, size_t(ne)));
    DEV_OK(pool.get(&debad, 1));
    const unsigned long long no_ebad = ~0ull;
    DEV_OK(hipMemcpy(debad, &no_ebad, 8, hipMemcpyHostToDevice));
    k_elem_flags<<<blocks(ne), 256>>>(ne, nen, dconn, ddet, dmat, dflag, debad);
    DEV_OK(hipGetLastError());
    unsigned long long hebad = 0;
    DEV_OK(hipMemcpy(&hebad, debad, 8, hipMemcpyDeviceToHost));
    if (hebad != no_ebad) {
        const int code = int(hebad & 255u);
        err = cut_error_text(code);
        return cut_error_code(code);
    }
    if (int rc = prefix(pool, dflag, dpos, ne, err)) return rc;
    int32_t n_cut = 0;
    DEV_OK(hipMemcpy(&n_cut, dpos + ne, 4, hipMemcpyDeviceToHost));
    int32_t* dce;
    DEV_OK(pool.get(&dce, size_t(n_cut)));
    k_compact<<<blocks(ne), 256>>>(ne, dflag, dpos, dce, dcoe);
    DEV_OK(hipGetLastError());

    // cut records and slot keys
    const int64_t nk = int64_t(n_cut) * SLOTS;
    pg_cut_rec* drec;
    int32_t *dncq, *disv;
    uint64_t *klohi, *slohi1, *slohi2;
    uint32_t *kinc, *sinc1, *sinc2;
    DEV_OK(pool.get(&drec, size_t(n_cut)));
    DEV_OK(pool.get(&dncq, size_t(n_cut)));
    DEV_OK(pool.get(&disv, size_t(n_cut) + 1));
    DEV_OK(pool.get(&klohi, size_t(nk)));
    DEV_OK(pool.get(&kinc, size_t(nk)));
    DEV_OK(pool.get(&slohi1, size_t(nk)));
    DEV_OK(pool.get(&sinc1, size_t(nk)));
    DEV_OK(pool.get(&slohi2, size_t(nk)));
    DEV_OK(pool.get(&sinc2, size_t(nk)));
    if (n_cut) {
        k_cut_records<<<blocks(n_cut, 128), 128>>>(n_cut, dim, nen, in.cohesive, dce, dconn, ddet, drec, dncq, klohi, kinc);
        DEV_OK(hipGetLastError());
    }
    if (int rc = prefix(pool, dncq, disv, n_cut, err)) return rc;
    // (edge, inclusion) order: a stable sort by inclusion, then by edge
    int32_t n_init = 0;
    uint64_t* ulohi = nullptr;
    uint32_t* uinc = nullptr;
    if (nk) {
        size_t tb1 = 0, tb2 = 0;
        DEV_OK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, kinc, sinc1, klohi, slohi1, int(nk)));
        DEV_OK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, slohi1, slohi2, sinc1, sinc2, int(nk)));
        char* tmp;
        DEV_OK(pool.get(&tmp, std::max(tb1, tb2)));
        DEV_OK(hipcub::DeviceRadixSort::SortPairs(tmp, tb1, kinc, sinc1, klohi, slohi1, int(nk)));
        DEV_OK(hipcub::DeviceRadixSort::SortPairs(tmp, tb2, slohi1, slohi2, sinc1, sinc2, int(nk)));
        int32_t *uflag, *upos;
        DEV_OK(pool.get(&uflag, size_t(nk)));
        DEV_OK(pool.get(&upos, size_t(nk) + 1));
        k_unique_flags<<<blocks(nk), 256>>>(nk, slohi2, sinc2, uflag);
        DEV_OK(hipGetLastError());
        if (int rc = prefix(pool, uflag, upos, nk, err)) return rc;
        DEV_OK(hipMemcpy(&n_init, upos + nk, 4, hipMemcpyDeviceToHost));
        DEV_OK(pool.get(&ulohi, size_t(n_init)));
        DEV_OK(pool.get(&uinc, size_t(n_init)));
        k_unique_out<<<blocks(nk), 256>>>(nk, uflag, upos, slohi2, sinc2, ulohi, uinc);
        DEV_OK(hipGetLastError());
    }
    const int32_t n_enr = in.cohesive ? 2 * n_init : n_init;
    double *dalpha, *dxe;
    int32_t *dei, *dee;
    DEV_OK(pool.get(&dalpha, size_t(n_enr)));
    DEV_OK(pool.get(&dxe, size_t(n_enr) * dim));
    DEV_OK(pool.get(&dei, size_t(n_enr)));
    DEV_OK(pool.get(&dee, size_t(n_enr) * 2));
    if (n_init) {
        const int32_t zero = 0;
        DEV_OK(hipMemcpy(dbad, &zero, 4, hipMemcpyHostToDevice));
        k_intersect<<<blocks(n_init, 128), 128>>>(n_init, dim, in.cohesive, dx, ulohi, uinc, dic, dir, dalpha, dxe,
                                                  dei, dee, dbad);
        DEV_OK(hipGetLastError());
        DEV_OK(hipMemcpy(&hbad, dbad, 4, hipMemcpyDeviceToHost));
        if (hbad != 0) {
            err = cut_error_text(CUT_E_NOROOT);
            return cut_error_code(CUT_E_NOROOT);
        }
    }
    if (n_cut) {
        k_attach<<<blocks(n_cut), 256>>>(n_cut, nn, n_init, in.cohesive, drec, disv, klohi, kinc, ulohi, uinc);
        DEV_OK(hipGetLastError());
    }
    DEV_OK(hipMemcpy(&out.n_cqp, disv + n_cut, 4, hipMemcpyDeviceToHost));
    out.n_enr = n_enr;
    if (int rc = down(out.node_detect, ddet, nn, err)) return rc;
    if (int rc = down(out.elem_mat, dmat, ne, err)) return rc;
    if (int rc = down(out.cut_of_elem, dcoe, ne, err)) return rc;
    if (int rc = down(out.cut_elem, dce, n_cut, err)) return rc;
    if (int rc = down(out.cut, drec, n_cut, err)) return rc;
    if (int rc = down(out.enr_incl, dei, n_enr, err)) return rc;
    if (int rc = down(out.enr_edge, dee, int64_t(n_enr) * 2, err)) return rc;
    if (int rc = down(out.enr_alpha, dalpha, n_enr, err)) return rc;
    if (int rc = down(out.enr_coords, dxe, int64_t(n_enr) * dim, err)) return rc;
    out.seconds[1] = now_s() - t1;
    return PG_OK;
}

}  // namespace pgh
