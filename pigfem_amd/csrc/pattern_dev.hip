// Device build of the set-up's pattern stage: the node -> element incidence,
// the node-pair pattern, the 4-block AIJ pattern and the scatter indices
// (Solver::preallocate_matrix, src/Solver.cpp:414-490, with the insertion
// pattern of set_matrix, src/Assembler.cpp:659-679: the union of the full
// element dof blocks), each as a radix sort of packed keys:
//   incidences (node, element), sorted: n2e in ascending element order;
//   node pairs (a, b) of every element, sorted and made unique: the node-pair
//     pattern, each row's columns ascending;
//   every K entry keyed (block, row, column), sorted: the four CSR patterns
//     back to back, so an entry's rank in the sort is its index in the
//     concatenated [ff | fp | pf | pp] values -- its scatter index.
// The outputs equal the host build's (the sorted, unique keys fix them).
#include <algorithm>

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "dev_pool.h"
#include "pattern_dev.h"

namespace {

using pgh::Pool;
using pgh::blocks;
using pgh::down;
using pgh::prefix;

struct Mesh {
    int nen;
    int32_t n_elem;
    const int32_t* conn;
    const int32_t* coe;
    const pg_cut_rec* cut;
};

__device__ inline int local_count(const Mesh& m, int32_t e) {
    const int32_t c = m.coe[e];
    return m.nen + (c >= 0 ? int(m.cut[c].n_enr) : 0);
}

// local node k of element e: parent nodes, then enrichment nodes
__device__ inline int32_t local_node(const Mesh& m, int32_t e, int k) {
    return k < m.nen ? m.conn[int64_t(e) * m.nen + k] : m.cut[m.coe[e]].enr[k - m.nen];
}

__global__ void k_counts(Mesh m, int64_t* cnt, int64_t* cnt2) {
    const int32_t e = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= m.n_elem) return;
    const int64_t k = local_count(m, e);
    cnt[e] = k;
    cnt2[e] = k * k;
}

__global__ void k_emit_incidences(Mesh m, const int64_t* off, uint64_t* keys) {
    const int32_t e = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= m.n_elem) return;
    const int k = local_count(m, e);
    for (int i = 0; i < k; ++i) keys[off[e] + i] = (uint64_t(uint32_t(local_node(m, e, i))) << 32) | uint32_t(e);
}

__global__ void k_emit_pairs(Mesh m, const int64_t* off, uint64_t* keys) {
    const int32_t e = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= m.n_elem) return;
    const int k = local_count(m, e);
    int32_t ln[4 + PG_MAX_ENR];
    for (int i = 0; i < k; ++i) ln[i] = local_node(m, e, i);
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j)
            keys[off[e] + int64_t(i) * k + j] = (uint64_t(uint32_t(ln[i])) << 32) | uint32_t(ln[j]);
}

// count per high word (row node) and split the keys into rows and columns
__global__ void k_split(int64_t n, const uint64_t* keys, int32_t* hi, int32_t* lo, int32_t* cnt) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    const int32_t h = int32_t(k >> 32);
    if (hi) hi[i] = h;
    lo[i] = int32_t(uint32_t(k));
    atomicAdd(&cnt[h], 1);
}

// every K entry of every node pair: key (block << 62 | row << 31 | col), value
// = its position q d^2 + i d + j in the scatter-index array
__global__ void k_emit_entries(int64_t npairs, int d, const int32_t* prow, const int32_t* pcol, const int32_t* node_dof,
                               int32_t nf, uint64_t* keys, int32_t* vals) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= npairs) return;
    const int32_t a = prow[q], b = pcol[q];
    for (int i = 0; i < d; ++i) {
        const int32_t r = node_dof[int64_t(a) * d + i];
        for (int j = 0; j < d; ++j) {
            const int32_t c = node_dof[int64_t(b) * d + j];
            const uint64_t blk = uint64_t((r >= nf ? 2 : 0) + (c >= nf ? 1 : 0));
            const uint64_t row = uint64_t(r >= nf ? r - nf : r), col = uint64_t(c >= nf ? c - nf : c);
            const int64_t v = q * d * d + i * d + j;
            keys[v] = (blk << 62) | (row << 31) | col;
            vals[v] = int32_t(v);
        }
    }
}

// after the sort: scatter index of each entry, its column, and the row counts
// of its block
__global__ void k_entries_out(int64_t n, const uint64_t* keys, const int32_t* vals, int32_t* map, int32_t* col,
                              int32_t* cnt0, int32_t* cnt1, int32_t* cnt2, int32_t* cnt3) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint64_t k = keys[p];
    map[vals[p]] = int32_t(p);
    col[p] = int32_t(k & 0x7fffffffu);
    const int blk = int(k >> 62);
    const int32_t row = int32_t((k >> 31) & 0x7fffffffu);
    int32_t* c = blk == 0 ? cnt0 : blk == 1 ? cnt1 : blk == 2 ? cnt2 : cnt3;
    atomicAdd(&c[row], 1);
}

int bits_for(int64_t v) {
    int b = 1;
    while ((int64_t(1) << b) <= v) ++b;
    return b;
}

}  // namespace

namespace pgh {

int build_pattern_device(const PatternIn& in, PatternOut& out, std::string& err) {
    const int d = in.nndof ? in.nndof : in.dim;  // dofs per node
    const int64_t nall = int64_t(in.n_nodes) + in.n_enr, ndof = int64_t(in.n_free) + in.n_pres;
    if (in.n_elem <= 0 || nall <= 0 || (in.dim != 2 && in.dim != 3) || (d != 1 && d != in.dim)) {
        err = "empty or malformed mesh";
        return PG_E_ARG;
    }
    if (ndof >= (int64_t(1) << 31)) { err = "dofs exceed 2^31"; return PG_E_UNSUPPORTED; }
    Pool pool;
    // the mesh on the device
    int32_t *conn, *coe, *ndof_d;
    pg_cut_rec* cut;
    DEV_OK(pool.get(&conn, size_t(in.n_elem) * in.nen));
    DEV_OK(pool.get(&coe, size_t(in.n_elem)));
    DEV_OK(pool.get(&cut, size_t(in.n_cut)));
    DEV_OK(pool.get(&ndof_d, size_t(nall) * d));
    DEV_OK(hipMemcpy(conn, in.conn, size_t(in.n_elem) * in.nen * 4, hipMemcpyHostToDevice));
    DEV_OK(hipMemcpy(coe, in.cut_of_elem, size_t(in.n_elem) * 4, hipMemcpyHostToDevice));
    if (in.n_cut) DEV_OK(hipMemcpy(cut, in.cut, size_t(in.n_cut) * sizeof(pg_cut_rec), hipMemcpyHostToDevice));
    DEV_OK(hipMemcpy(ndof_d, in.node_dof, size_t(nall) * d * 4, hipMemcpyHostToDevice));
    const Mesh m{in.nen, in.n_elem, conn, coe, cut};
    const int64_t E = in.n_elem;
    int64_t *cnt, *cnt2, *off, *off2;
    DEV_OK(pool.get(&cnt, size_t(E)));
    DEV_OK(pool.get(&cnt2, size_t(E)));
    DEV_OK(pool.get(&off, size_t(E) + 1));
    DEV_OK(pool.get(&off2, size_t(E) + 1));
    k_counts<<<blocks(E), 256>>>(m, cnt, cnt2);
    DEV_OK(hipGetLastError());
    if (int rc = prefix(pool, cnt, off, E, err)) return rc;
    if (int rc = prefix(pool, cnt2, off2, E, err)) return rc;
    int64_t n_inc = 0, n_pk = 0;
    DEV_OK(hipMemcpy(&n_inc, off + E, 8, hipMemcpyDeviceToHost));
    DEV_OK(hipMemcpy(&n_pk, off2 + E, 8, hipMemcpyDeviceToHost));
    if (n_pk >= (int64_t(1) << 31)) { err = "element dof blocks exceed 2^31 node pairs"; return PG_E_UNSUPPORTED; }
    const int nb = bits_for(nall);

    // ---- node -> element incidence ----
    {
        uint64_t *k0, *k1;
        int32_t *lo, *cn, *ptr;
        DEV_OK(pool.get(&k0, size_t(n_inc)));
        DEV_OK(pool.get(&k1, size_t(n_inc)));
        DEV_OK(pool.get(&lo, size_t(n_inc)));
        DEV_OK(pool.get(&cn, size_t(nall)));
        DEV_OK(pool.get(&ptr, size_t(nall) + 1));
        k_emit_incidences<<<blocks(E), 256>>>(m, off, k0);
        DEV_OK(hipGetLastError());
        size_t tb = 0;
        DEV_OK(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, k0, k1, int(n_inc), 0, 32 + nb));
        char* tmp;
        DEV_OK(pool.get(&tmp, tb));
        DEV_OK(hipcub::DeviceRadixSort::SortKeys(tmp, tb, k0, k1, int(n_inc), 0, 32 + nb));
        DEV_OK(hipMemset(cn, 0, size_t(nall) * 4));
        k_split<<<blocks(n_inc), 256>>>(n_inc, k1, nullptr, lo, cn);
        DEV_OK(hipGetLastError());
        if (int rc = prefix(pool, cn, ptr, nall, err)) return rc;
        if (int rc = down(out.n2e, lo, n_inc, err)) return rc;
        if (int rc = down(out.n2e_ptr, ptr, nall + 1, err)) return rc;
    }

    // ---- node-pair pattern ----
    int32_t *prow, *pcol;
    int64_t npairs = 0;
    {
        uint64_t *k0, *k1, *k2;
        int32_t *nsel, *cn, *ptr;
        DEV_OK(pool.get(&k0, size_t(n_pk)));
        DEV_OK(pool.get(&k1, size_t(n_pk)));
        DEV_OK(pool.get(&k2, size_t(n_pk)));
        DEV_OK(pool.get(&nsel, 1));
        k_emit_pairs<<<blocks(E), 256>>>(m, off2, k0);
        DEV_OK(hipGetLastError());
        size_t tb = 0, tu = 0;
        DEV_OK(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, k0, k1, int(n_pk), 0, 32 + nb));
        DEV_OK(hipcub::DeviceSelect::Unique(nullptr, tu, k1, k2, nsel, int(n_pk)));
        char* tmp;
        DEV_OK(pool.get(&tmp, std::max(tb, tu)));
        DEV_OK(hipcub::DeviceRadixSort::SortKeys(tmp, tb, k0, k1, int(n_pk), 0, 32 + nb));
        DEV_OK(hipcub::DeviceSelect::Unique(tmp, tu, k1, k2, nsel, int(n_pk)));
        int32_t ns = 0;
        DEV_OK(hipMemcpy(&ns, nsel, 4, hipMemcpyDeviceToHost));
        npairs = ns;
        DEV_OK(pool.get(&prow, size_t(npairs)));
        DEV_OK(pool.get(&pcol, size_t(npairs)));
        DEV_OK(pool.get(&cn, size_t(nall)));
        DEV_OK(pool.get(&ptr, size_t(nall) + 1));
        DEV_OK(hipMemset(cn, 0, size_t(nall) * 4));
        k_split<<<blocks(npairs), 256>>>(npairs, k2, prow, pcol, cn);
        DEV_OK(hipGetLastError());
        if (int rc = prefix(pool, cn, ptr, nall, err)) return rc;
        if (int rc = down(out.npat_ptr, ptr, nall + 1, err)) return rc;
        if (int rc = down(out.npat_row, prow, npairs, err)) return rc;
        if (int rc = down(out.npat_col, pcol, npairs, err)) return rc;
    }

    // ---- 4-block AIJ pattern and scatter indices ----
    {
        const int64_t nnz = npairs * d * d;
        if (nnz >= (int64_t(1) << 31)) { err = "total nnz exceeds 2^31"; return PG_E_NOMEM; }
        uint64_t *k0, *k1;
        int32_t *v0, *v1, *map, *col;
        DEV_OK(pool.get(&k0, size_t(nnz)));
        DEV_OK(pool.get(&k1, size_t(nnz)));
        DEV_OK(pool.get(&v0, size_t(nnz)));
        DEV_OK(pool.get(&v1, size_t(nnz)));
        DEV_OK(pool.get(&map, size_t(nnz)));
        DEV_OK(pool.get(&col, size_t(nnz)));
        k_emit_entries<<<blocks(npairs), 256>>>(npairs, d, prow, pcol, ndof_d, in.n_free, k0, v0);
        DEV_OK(hipGetLastError());
        size_t tb = 0;
        DEV_OK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, v0, v1, int(nnz), 0, 64));
        char* tmp;
        DEV_OK(pool.get(&tmp, tb));
        DEV_OK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, v1, int(nnz), 0, 64));
        const int64_t rows[4] = {in.n_free, in.n_free, in.n_pres, in.n_pres};
        int32_t* rc_[4];
        for (int b = 0; b < 4; ++b) {
            DEV_OK(pool.get(&rc_[b], size_t(rows[b]) + 1));
            DEV_OK(hipMemset(rc_[b], 0, (size_t(rows[b]) + 1) * 4));
        }
        k_entries_out<<<blocks(nnz), 256>>>(nnz, k1, v1, map, col, rc_[0], rc_[1], rc_[2], rc_[3]);
        DEV_OK(hipGetLastError());
        int64_t total = 0;
        for (int b = 0; b < 4; ++b) {
            int32_t* ptr;
            DEV_OK(pool.get(&ptr, size_t(rows[b]) + 1));
            if (int rc = prefix(pool, rc_[b], ptr, rows[b], err)) return rc;
            if (int rc = down(out.rp[b], ptr, rows[b] + 1, err)) return rc;
            total += out.rp[b].back();
        }
        if (total != nnz) { err = "block entry counts do not add up"; return PG_E_MESH; }
        for (int b = 0, base = 0; b < 4; base += out.rp[b].back(), ++b)
            if (int rc = down(out.ci[b], col + base, out.rp[b].back(), err)) return rc;
        if (int rc = down(out.npat_map, map, nnz, err)) return rc;
    }
    DEV_OK(hipDeviceSynchronize());
    return PG_OK;
}

}  // namespace pgh
