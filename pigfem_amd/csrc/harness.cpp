// Host set-up of synthetic IGFEM RVEs: the part of PIGFEM's Problem::init that
// produces the static data the element assembly reads, restated on flat arrays.
// Citations are to brandyb2/PIGFEM.
#include "pgfem_harness.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "cut_tables.h"
#include "pattern_dev.h"
#include "setup_dev.h"

namespace {

thread_local std::string g_err;

struct HarnessError : std::runtime_error {
    int code;
    HarnessError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const std::string& msg) { throw HarnessError(code, msg); }

// Uniform double in [a, b) from the top 53 bits of one mt19937_64 draw.
struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t s) : g(s) {}
    double uni(double a, double b) {
        double u = double(g() >> 11) * 0x1.0p-53;
        return a + (b - a) * u;
    }
};

struct Incl {
    double c[3];
    double r;
};

// Per cut element, the detection output before enrichment nodes are attached.
struct CutInfo {
    int32_t elem;
    int ncut;                       // cut edges (enrichment slots before mirrors)
    uint8_t cut_edge[PG_MAX_ENR];   // local edge index per slot
    int32_t slot_incl[PG_MAX_ENR];  // inclusion per slot
    int ncq;                        // cohesive qps
    pg_cut_rec rec;                 // children and cohesive children (pgh::cut_record)
};

}  // namespace

struct pgh_problem {
    pgh_rve_params prm{};
    int dim = 0, nen = 0;
    int nndof = 0;                       // dofs per node: dim, or 1 for a scalar field
    int cohesive = 0;
    int32_t n_nodes = 0, n_enr = 0, n_elem = 0;
    std::vector<double> coords;          // (n_nodes + n_enr) * dim
    std::vector<int32_t> conn, elem_mat, cut_of_elem, cut_elem;
    std::vector<pg_cut_rec> cut;
    std::vector<pg_material> mats;
    std::vector<int32_t> node_detect;
    std::vector<Incl> incl;
    std::vector<double> incl_c, incl_r;
    int32_t n_cqp = 0;
    std::vector<int32_t> node_dof;
    int32_t n_free = 0, n_pres = 0;
    std::vector<double> bc_val;
    std::vector<int32_t> rp[4], ci[4];
    std::vector<int32_t> npat_ptr, npat_col, npat_row, npat_map;
    std::vector<int32_t> n2e_ptr, n2e;
    std::vector<int32_t> enr_incl;        // inclusion of each enrichment node (mirrors included)
    std::vector<int32_t> enr_edge;        // per enrichment node: the cut edge's two nodes (ascending)
    std::vector<double> enr_alpha;        // per enrichment node: its fraction along that edge
    // batch of independent RVEs: per instance the first element, regular node,
    // enrichment node, free dof, prescribed dof, cohesive qp and cut element,
    // plus one closing row (PGH_BATCH_COLS columns)
    int32_t n_inst = 0;
    std::vector<int64_t> inst_off;
    double stage_s[PGH_N_STAGES] = {};
};

namespace {

// --------------------------------------------------------------------------
// Structured mesh, src/Mesh.cpp:236-630.  Node id = x + y(Nx+1) + z(Nx+1)(Ny+1);
// coordinate = min + iter * ((max - min) / N).
// --------------------------------------------------------------------------
void generate_mesh(pgh_problem& P) {
    const int n = P.prm.n;
    const double h = (1.0 - 0.0) / n;
    if (P.dim == 2) {
        P.nen = 3;
        P.n_nodes = (n + 1) * (n + 1);
        P.coords.resize(size_t(P.n_nodes) * 2);
        for (int y = 0; y <= n; ++y)
            for (int x = 0; x <= n; ++x) {
                size_t id = size_t(x) + size_t(y) * (n + 1);
                P.coords[id * 2 + 0] = 0.0 + unsigned(x) * h;
                P.coords[id * 2 + 1] = 0.0 + unsigned(y) * h;
            }
        P.n_elem = 2 * n * n;
        P.conn.resize(size_t(P.n_elem) * 3);
        size_t e = 0;
        auto id = [n](int x, int y) { return int32_t(x + y * (n + 1)); };
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) {
                int32_t* c = &P.conn[e * 3];
                if ((x + y) % 2 == 0) {  // src/Mesh.cpp:418-429
                    c[0] = id(x, y); c[1] = id(x + 1, y); c[2] = id(x + 1, y + 1);
                    c[3] = id(x, y); c[4] = id(x + 1, y + 1); c[5] = id(x, y + 1);
                } else {                 // src/Mesh.cpp:432-441
                    c[0] = id(x, y); c[1] = id(x + 1, y); c[2] = id(x, y + 1);
                    c[3] = id(x + 1, y); c[4] = id(x + 1, y + 1); c[5] = id(x, y + 1);
                }
                e += 2;
            }
    } else {
        P.nen = 4;
        const size_t nx1 = n + 1;
        P.n_nodes = int32_t(nx1 * nx1 * nx1);
        P.coords.resize(size_t(P.n_nodes) * 3);
        for (int z = 0; z <= n; ++z)
            for (int y = 0; y <= n; ++y)
                for (int x = 0; x <= n; ++x) {
                    size_t id = size_t(x) + size_t(y) * nx1 + size_t(z) * nx1 * nx1;
                    P.coords[id * 3 + 0] = 0.0 + unsigned(x) * h;
                    P.coords[id * 3 + 1] = 0.0 + unsigned(y) * h;
                    P.coords[id * 3 + 2] = 0.0 + unsigned(z) * h;
                }
        P.n_elem = 6 * n * n * n;
        P.conn.resize(size_t(P.n_elem) * 4);
        auto id = [nx1](int x, int y, int z) {
            return int32_t(size_t(x) + size_t(y) * nx1 + size_t(z) * nx1 * nx1);
        };
        size_t e = 0;
        for (int z = 0; z < n; ++z)
            for (int y = 0; y < n; ++y)
                for (int x = 0; x < n; ++x) {
                    // cube corners 0..7 as in src/Mesh.cpp:540-583
                    int32_t v[8] = {id(x, y, z),         id(x + 1, y, z),
                                    id(x + 1, y + 1, z), id(x, y + 1, z),
                                    id(x, y, z + 1),     id(x + 1, y, z + 1),
                                    id(x + 1, y + 1, z + 1), id(x, y + 1, z + 1)};
                    static const int T[6][4] = {{0, 2, 3, 6}, {0, 3, 7, 6}, {0, 7, 4, 6},
                                                {0, 5, 6, 4}, {1, 5, 6, 0}, {1, 6, 2, 0}};
                    for (int t = 0; t < 6; ++t, ++e)
                        for (int k = 0; k < 4; ++k) P.conn[e * 4 + k] = v[T[t][k]];
                }
    }
}

// --------------------------------------------------------------------------
// Inclusion placement (synthetic input generator): radius U[r_min, r_max],
// centre U[r, 1-r]^dim, rejected if closer than `gap` to an earlier inclusion,
// if a mesh node falls within 1e-9 r of its surface (the reference aborts on
// nodes within 1e-12 r, src/Mesh.cpp:2906-2912), and, for spheres, if the
// centre's z exceeds its y: Sphere_Inclusion::get_domain_lims bounds z by
// center[1] + r (src/InclusionSphere.cpp:163-167), so such spheres would be
// detected only partially and the reference's edge intersection would abort.
// --------------------------------------------------------------------------
struct Grid {
    double cell;
    int nc;
    std::vector<std::vector<int32_t>> bins;
    int dim;
    Grid(double cell_, int dim_) : cell(cell_), dim(dim_) {
        nc = std::max(1, int(std::ceil(1.0 / cell)));
        size_t nb = size_t(nc) * nc * (dim == 3 ? nc : 1);
        bins.resize(nb);
    }
    int idx(double v) const { return std::min(nc - 1, std::max(0, int(std::floor(v / cell)))); }
    size_t bin(int i, int j, int k) const { return size_t(i) + size_t(j) * nc + size_t(k) * nc * nc; }
};

bool node_near_surface(const pgh_problem& P, const Incl& I) {
    const int n = P.prm.n;
    const double h = 1.0 / n;
    int lo[3], hi[3];
    for (int d = 0; d < P.dim; ++d) {
        lo[d] = std::max(0, int(std::floor((I.c[d] - I.r) / h)) - 1);
        hi[d] = std::min(n, int(std::ceil((I.c[d] + I.r) / h)) + 1);
    }
    if (P.dim == 2) { lo[2] = 0; hi[2] = 0; }
    for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int x = lo[0]; x <= hi[0]; ++x) {
                double X = 0.0 + unsigned(x) * h, Y = 0.0 + unsigned(y) * h, Z = 0.0 + unsigned(z) * h;
                double val;
                if (P.dim == 2) {
                    double dx = X - I.c[0], dy = Y - I.c[1];
                    val = (std::sqrt(dx * dx + dy * dy) - I.r) / I.r;
                } else {
                    double gx = (X - I.c[0]) / I.r, gy = (Y - I.c[1]) / I.r, gz = (Z - I.c[2]) / I.r;
                    val = gx * gx + gy * gy + gz * gz - 1.0;
                }
                if (std::fabs(val) <= 1e-9) return true;
            }
    return false;
}

void place_inclusions(pgh_problem& P) {
    const pgh_rve_params& p = P.prm;
    Rng rng(p.seed);
    const double reach = 2.0 * p.r_max + p.gap;
    Grid G(std::max(reach, 1e-3), P.dim);
    long budget = p.max_tries > 0 ? p.max_tries : 50000000L;
    long tries = 0;
    while (int(P.incl.size()) < p.n_incl) {
        if (++tries > budget) {
            if ((p.flags & PGH_PLACE_BEST_EFFORT) && !P.incl.empty()) break;  // keep what fits
            fail(PG_E_ARG, "inclusion placement did not converge (" +
                               std::to_string(P.incl.size()) + " of " + std::to_string(p.n_incl) + " placed)");
        }
        Incl I{};
        I.r = rng.uni(p.r_min, p.r_max);
        if (p.flags & PGH_PLACE_AT_NODES) {
            // an interior mesh node; r < h / 2 keeps every other node outside, so
            // the only elements cut are those of the node's star, one node inside
            const double h = 1.0 / p.n;
            if (p.r_max >= 0.5 * h || p.n < 2) fail(PG_E_ARG, "node-centred inclusions need n >= 2 and r_max < h / 2");
            for (int d = 0; d < P.dim; ++d) {
                int k = 1 + int(rng.uni(0.0, 1.0) * (p.n - 1));
                if (k > p.n - 1) k = p.n - 1;
                I.c[d] = k * h;
            }
        } else {
            for (int d = 0; d < P.dim; ++d) I.c[d] = rng.uni(I.r, 1.0 - I.r);
        }
        // the reference's sphere box tests z against the y centre: keep z <= y
        if (P.dim == 3 && I.c[2] > I.c[1]) continue;
        int ci[3] = {G.idx(I.c[0]), G.idx(I.c[1]), P.dim == 3 ? G.idx(I.c[2]) : 0};
        bool ok = true;
        int kz0 = P.dim == 3 ? std::max(0, ci[2] - 1) : 0, kz1 = P.dim == 3 ? std::min(G.nc - 1, ci[2] + 1) : 0;
        for (int k = kz0; k <= kz1 && ok; ++k)
            for (int j = std::max(0, ci[1] - 1); j <= std::min(G.nc - 1, ci[1] + 1) && ok; ++j)
                for (int i = std::max(0, ci[0] - 1); i <= std::min(G.nc - 1, ci[0] + 1) && ok; ++i)
                    for (int32_t o : G.bins[G.bin(i, j, k)]) {
                        const Incl& J = P.incl[o];
                        double d2 = 0;
                        for (int d = 0; d < P.dim; ++d) d2 += (I.c[d] - J.c[d]) * (I.c[d] - J.c[d]);
                        double need = I.r + J.r + p.gap;
                        if (d2 < need * need) { ok = false; break; }
                    }
        if (!ok) continue;
        if (node_near_surface(P, I)) continue;
        G.bins[G.bin(ci[0], ci[1], ci[2])].push_back(int32_t(P.incl.size()));
        P.incl.push_back(I);
    }
    for (auto& I : P.incl) {
        for (int d = 0; d < P.dim; ++d) P.incl_c.push_back(I.c[d]);
        P.incl_r.push_back(I.r);
    }
}

// --------------------------------------------------------------------------
// Node detection, Mesh::detect_node (src/Mesh.cpp:2883-2930): inclusions are
// checked from the last to the first; the first one containing the node wins
// (pgh::detect_one per inclusion).
// --------------------------------------------------------------------------
void detect_nodes(pgh_problem& P) {
    const double reach = 2.0 * P.prm.r_max + P.prm.gap;
    Grid G(std::max(reach, 1e-3), P.dim);
    for (size_t j = 0; j < P.incl.size(); ++j) {
        const Incl& I = P.incl[j];
        G.bins[G.bin(G.idx(I.c[0]), G.idx(I.c[1]), P.dim == 3 ? G.idx(I.c[2]) : 0)].push_back(int32_t(j));
    }
    P.node_detect.assign(P.n_nodes, -1);
    std::vector<int32_t> cand;
    for (int32_t nd = 0; nd < P.n_nodes; ++nd) {
        const double* x = &P.coords[size_t(nd) * P.dim];
        int ci[3] = {G.idx(x[0]), G.idx(x[1]), P.dim == 3 ? G.idx(x[2]) : 0};
        cand.clear();
        int kz0 = P.dim == 3 ? std::max(0, ci[2] - 1) : 0, kz1 = P.dim == 3 ? std::min(G.nc - 1, ci[2] + 1) : 0;
        for (int k = kz0; k <= kz1; ++k)
            for (int j = std::max(0, ci[1] - 1); j <= std::min(G.nc - 1, ci[1] + 1); ++j)
                for (int i = std::max(0, ci[0] - 1); i <= std::min(G.nc - 1, ci[0] + 1); ++i)
                    for (int32_t o : G.bins[G.bin(i, j, k)]) cand.push_back(o);
        std::sort(cand.begin(), cand.end(), std::greater<int32_t>());
        for (int32_t j : cand) {
            int res = pgh::detect_one(P.dim, P.incl[j].c, P.incl[j].r, x);
            if (res == 0) continue;
            if (res == 1) fail(PG_E_MESH, "node " + std::to_string(nd) + " lies on an inclusion surface");
            P.node_detect[nd] = j;
            break;
        }
    }
}

// Element detection tables (pgh::detect_tri3, pgh::detect_tet4) and the edge
// intersection (pgh::intersect_edge): cut_tables.h, shared with the device
// set-up.

void build_materials(pgh_problem& P) {
    const pgh_rve_params& p = P.prm;
    auto lei = [&](double E, double nu) {
        pg_material m{};
        m.type = PG_MAT_LEI;
        // LinearElasticIsotropicMaterial::computeLinearDmat, src/material_lei.cpp:52-97
        if (P.dim == 2) {
            m.nstr = 3;
            double coef = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
            m.D[0] = coef * (1.0 - nu); m.D[1] = coef * nu;
            m.D[3] = coef * nu;         m.D[4] = coef * (1.0 - nu);
            m.D[8] = coef * (1.0 - 2.0 * nu) / 2.0;
        } else {
            m.nstr = 6;
            double coef = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) m.D[i * 6 + j] = coef * (i == j ? (1.0 - nu) : nu);
            for (int i = 3; i < 6; ++i) m.D[i * 6 + i] = coef * (1.0 - 2.0 * nu) / 2;
        }
        return m;
    };
    if (P.nndof == 1) {
        // LinearThermalMaterial::computeLinearDmat, src/material_lt.cpp:36-47
        auto lt = [&](double kappa) {
            pg_material m{};
            m.type = PG_MAT_LT;
            m.nstr = P.dim;
            for (int i = 0; i < P.dim; ++i) m.D[i * P.dim + i] = kappa;
            return m;
        };
        P.mats.push_back(lt(p.kappa_matrix));  // 0: matrix
        P.mats.push_back(lt(p.kappa_fiber));   // 1: fiber
        return;
    }
    P.mats.push_back(lei(p.E_matrix, p.nu_matrix));  // 0: matrix
    P.mats.push_back(lei(p.E_fiber, p.nu_fiber));    // 1: fiber
    if (p.cohesive != PGH_COH_NONE) {
        pg_material m{};
        m.sigma_c = p.sigma_c;
        m.beta2 = p.beta * p.beta;
        switch (p.cohesive) {
            case PGH_COH_BL: m.type = PG_MAT_BL; break;
            case PGH_COH_BL_NU: m.type = PG_MAT_BL_NU; break;
            case PGH_COH_OP: m.type = PG_MAT_OP; break;
            case PGH_COH_OP_NU: m.type = PG_MAT_OP_NU; break;
            default: fail(PG_E_ARG, "unknown cohesive law");
        }
        if (m.type == PG_MAT_BL || m.type == PG_MAT_BL_NU) {
            m.delta_c1 = p.delta_c1;
            m.delta_c2 = p.delta_c2;
            m.alpha = p.alpha;
            if (!(m.delta_c2 >= m.delta_c1)) fail(PG_E_ARG, "bilinear law needs delta_c2 >= delta_c1");
        } else {
            m.delta_c = p.delta_c;
        }
        P.mats.push_back(m);  // 2: cohesive
    }
}

void analyze_and_enrich(pgh_problem& P) {
    const bool coh = P.prm.cohesive != PGH_COH_NONE;
    const int nen = P.nen;
    P.elem_mat.assign(P.n_elem, 0);
    P.cut_of_elem.assign(P.n_elem, -1);
    std::vector<CutInfo> infos;
    int nd[4];
    // Mesh::analyze_elements, src/Mesh.cpp:2962-3071
    for (int32_t e = 0; e < P.n_elem; ++e) {
        const int32_t* c = &P.conn[size_t(e) * nen];
        for (int k = 0; k < nen; ++k) nd[k] = P.node_detect[c[k]];
        pgh::CutDetect D;
        bool cut = false;
        if (nen == 3) {
            cut = pgh::detect_tri3(nd, D);
        } else if (int rc = pgh::detect_tet4(nd, c, D, &cut)) {
            fail(pgh::cut_error_code(rc), pgh::cut_error_text(rc));
        }
        if (!cut) {
            if (nd[0] != -1) P.elem_mat[e] = 1;
            continue;
        }
        CutInfo ci{};
        ci.elem = e;
        ci.ncut = D.ncut;
        for (int s = 0; s < D.ncut; ++s) { ci.cut_edge[s] = D.cut_edge[s]; ci.slot_incl[s] = D.slot_incl[s]; }
        ci.ncq = pgh::cut_record(D, P.dim, nen, coh, ci.rec);
        if (ci.ncq < 0) fail(pgh::cut_error_code(-ci.ncq), pgh::cut_error_text(-ci.ncq));
        infos.push_back(ci);
    }

    // Mesh::add_enrichments serial branch, src/Mesh.cpp:3104-3251: unique
    // (sorted edge, inclusion) triples in std::set order.
    struct Key { int32_t lo, hi, inc; };
    std::vector<Key> keys;
    auto edge_of = [&](const CutInfo& ci, int s, int32_t& lo, int32_t& hi) {
        const int32_t* c = &P.conn[size_t(ci.elem) * nen];
        int le = ci.cut_edge[s];
        int a = nen == 3 ? pgh::tri_edge(le, 0) : pgh::tet_edge(le, 0);
        int b = nen == 3 ? pgh::tri_edge(le, 1) : pgh::tet_edge(le, 1);
        lo = std::min(c[a], c[b]);
        hi = std::max(c[a], c[b]);
    };
    for (auto& ci : infos)
        for (int s = 0; s < ci.ncut; ++s) {
            Key k; edge_of(ci, s, k.lo, k.hi); k.inc = ci.slot_incl[s];
            keys.push_back(k);
        }
    auto less = [](const Key& a, const Key& b) {
        if (a.lo != b.lo) return a.lo < b.lo;
        if (a.hi != b.hi) return a.hi < b.hi;
        return a.inc < b.inc;
    };
    std::sort(keys.begin(), keys.end(), less);
    keys.erase(std::unique(keys.begin(), keys.end(),
                           [](const Key& a, const Key& b) { return a.lo == b.lo && a.hi == b.hi && a.inc == b.inc; }),
               keys.end());
    const int32_t n_init = int32_t(keys.size());
    P.n_enr = coh ? 2 * n_init : n_init;
    P.enr_incl.resize(size_t(P.n_enr));
    for (int32_t k = 0; k < P.n_enr; ++k) P.enr_incl[k] = keys[k % n_init].inc;
    P.enr_edge.resize(size_t(P.n_enr) * 2);
    P.enr_alpha.resize(size_t(P.n_enr));
    for (int32_t k = 0; k < P.n_enr; ++k) {
        P.enr_edge[size_t(k) * 2] = keys[k % n_init].lo;
        P.enr_edge[size_t(k) * 2 + 1] = keys[k % n_init].hi;
    }
    P.coords.resize(size_t(P.n_nodes + P.n_enr) * P.dim);
    for (int32_t k = 0; k < n_init; ++k) {
        const double* x0 = &P.coords[size_t(keys[k].lo) * P.dim];
        const double* x1 = &P.coords[size_t(keys[k].hi) * P.dim];
        const Incl& I = P.incl[keys[k].inc];
        int rc = 0;
        double a = pgh::intersect_edge(P.dim, x0, x1, I.c, I.r, &rc);
        if (rc) fail(pgh::cut_error_code(rc), pgh::cut_error_text(rc));
        P.enr_alpha[k] = a;
        if (coh) P.enr_alpha[n_init + k] = a;
        double* xe = &P.coords[size_t(P.n_nodes + k) * P.dim];
        for (int d = 0; d < P.dim; ++d) xe[d] = x0[d] + a * (x1[d] - x0[d]);
        if (coh) {
            double* xm = &P.coords[size_t(P.n_nodes + n_init + k) * P.dim];
            for (int d = 0; d < P.dim; ++d) xm[d] = xe[d];
        }
    }

    // Records: attach enrichment nodes per cut-edge slot (add_enrich_node_to_elem,
    // src/Mesh.cpp:3661-3720), mirrors at slot + ncut; ISV offsets.
    P.cut.resize(infos.size());
    P.cut_elem.resize(infos.size());
    int32_t isv = 0;
    for (size_t i = 0; i < infos.size(); ++i) {
        const CutInfo& ci = infos[i];
        pg_cut_rec& R = P.cut[i];
        R = ci.rec;
        for (int s = 0; s < ci.ncut; ++s) {
            Key k; edge_of(ci, s, k.lo, k.hi); k.inc = ci.slot_incl[s];
            auto it = std::lower_bound(keys.begin(), keys.end(), k, less);
            int32_t idx = int32_t(it - keys.begin());
            R.enr[s] = P.n_nodes + idx;
            if (coh) R.enr[s + ci.ncut] = P.n_nodes + n_init + idx;
        }
        R.isv_off = isv;
        isv += ci.ncq;
        P.cut_elem[i] = ci.elem;
        P.cut_of_elem[ci.elem] = int32_t(i);
        // base material of a cut element: the matrix (analyze_elements leaves it)
        P.elem_mat[ci.elem] = 0;
    }
    P.n_cqp = isv;
}

// Detection, element analysis and enrichment on the GPU (setup_dev.hip): the
// same outputs as detect_nodes + analyze_and_enrich.
void setup_on_device(pgh_problem& P) {
    pgh::SetupIn in;
    in.dim = P.dim;
    in.nen = P.nen;
    in.cohesive = P.prm.cohesive != PGH_COH_NONE;
    in.n_nodes = P.n_nodes;
    in.n_elem = P.n_elem;
    in.n_incl = int32_t(P.incl.size());
    in.coords = P.coords.data();
    in.conn = P.conn.data();
    in.incl_c = P.incl_c.data();
    in.incl_r = P.incl_r.data();
    in.cell = std::max(2.0 * P.prm.r_max + P.prm.gap, 1e-3);  // detect_nodes' inclusion grid
    pgh::SetupOut o;
    std::string err;
    if (int rc = pgh::setup_device(in, o, err)) fail(rc, err);
    P.node_detect = std::move(o.node_detect);
    P.elem_mat = std::move(o.elem_mat);
    P.cut_of_elem = std::move(o.cut_of_elem);
    P.cut_elem = std::move(o.cut_elem);
    P.cut = std::move(o.cut);
    P.n_enr = o.n_enr;
    P.enr_incl = std::move(o.enr_incl);
    P.enr_edge = std::move(o.enr_edge);
    P.enr_alpha = std::move(o.enr_alpha);
    P.coords.resize(size_t(P.n_nodes + P.n_enr) * P.dim);
    std::copy(o.enr_coords.begin(), o.enr_coords.end(), P.coords.begin() + size_t(P.n_nodes) * P.dim);
    P.n_cqp = o.n_cqp;
    P.stage_s[2] += o.seconds[0];
    P.stage_s[3] += o.seconds[1];
}

// --------------------------------------------------------------------------
// Boundary conditions (pattern of Tests/Main.cpp:562-570) and free-first dof
// numbering (DofObject::set_local_owned_normal_dofs / _enrich_dofs,
// src/DofObject.cpp:336-388, 498-538).
// --------------------------------------------------------------------------
void number_dofs(pgh_problem& P) {
    const int d = P.nndof;
    const int n = P.prm.n;
    const int32_t nall = P.n_nodes + P.n_enr;
    std::vector<int8_t> has_bc(size_t(nall) * d, 0);
    std::vector<double> val(size_t(nall) * d, 0.0);
    const double app = (1.0 - 0.0) * P.prm.strain;
    for (int32_t nd = 0; nd < P.n_nodes; ++nd) {
        int x = int(nd % (n + 1));
        if (x == 0) { has_bc[size_t(nd) * d] = 1; val[size_t(nd) * d] = 0.0; }     // LEFT, dof 0
        if (x == n) { has_bc[size_t(nd) * d] = 1; val[size_t(nd) * d] = app; }     // RIGHT, dof 0
    }
    for (int j = 0; j < d; ++j) { has_bc[j] = 1; val[j] = 0.0; }                  // node 0, all dofs
    int32_t nfree = 0, npres = 0;
    for (size_t k = 0; k < has_bc.size(); ++k) (has_bc[k] ? npres : nfree)++;
    P.n_free = nfree;
    P.n_pres = npres;
    P.node_dof.resize(size_t(nall) * d);
    P.bc_val.assign(npres, 0.0);
    int32_t f = 0, p = 0;
    for (int32_t nd = 0; nd < nall; ++nd)
        for (int j = 0; j < d; ++j) {
            size_t k = size_t(nd) * d + j;
            if (has_bc[k]) { P.node_dof[k] = nfree + p; P.bc_val[p] = val[k]; ++p; }
            else P.node_dof[k] = f++;
        }
}

// Element-local node list: parents, then enrichment nodes of a cut record.
inline int local_nodes(const pgh_problem& P, int32_t e, int32_t* out) {
    const int32_t* c = &P.conn[size_t(e) * P.nen];
    int n = 0;
    for (int k = 0; k < P.nen; ++k) out[n++] = c[k];
    int32_t ci = P.cut_of_elem[e];
    if (ci >= 0) {
        const pg_cut_rec& R = P.cut[ci];
        for (int k = 0; k < R.n_enr; ++k) out[n++] = R.enr[k];
    }
    return n;
}

// --------------------------------------------------------------------------
// Node incidence, node-pair pattern, 4-block CSR (union of full element dof
// blocks: set_matrix inserts every (i,j) of every element, explicit zeros
// included, src/Assembler.cpp:659-679) and the scatter indices.
// --------------------------------------------------------------------------
void build_pattern(pgh_problem& P) {
    const int d = P.nndof;
    const int32_t nall = P.n_nodes + P.n_enr;
    int32_t ln[PG_MAX_ENR + 4];
    P.n2e_ptr.assign(size_t(nall) + 1, 0);
    for (int32_t e = 0; e < P.n_elem; ++e) {
        int m = local_nodes(P, e, ln);
        for (int k = 0; k < m; ++k) P.n2e_ptr[ln[k] + 1]++;
    }
    for (int32_t a = 0; a < nall; ++a) P.n2e_ptr[a + 1] += P.n2e_ptr[a];
    P.n2e.resize(P.n2e_ptr[nall]);
    {
        std::vector<int32_t> fill(P.n2e_ptr.begin(), P.n2e_ptr.end() - 1);
        for (int32_t e = 0; e < P.n_elem; ++e) {  // ascending element order
            int m = local_nodes(P, e, ln);
            for (int k = 0; k < m; ++k) P.n2e[fill[ln[k]]++] = e;
        }
    }
    // node-pair pattern
    P.npat_ptr.assign(size_t(nall) + 1, 0);
    std::vector<int32_t> nb;
    std::vector<std::vector<int32_t>> tmp;  // not stored; two passes instead
    for (int pass = 0; pass < 2; ++pass) {
        for (int32_t a = 0; a < nall; ++a) {
            nb.clear();
            for (int32_t q = P.n2e_ptr[a]; q < P.n2e_ptr[a + 1]; ++q) {
                int m = local_nodes(P, P.n2e[q], ln);
                nb.insert(nb.end(), ln, ln + m);
            }
            std::sort(nb.begin(), nb.end());
            nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
            if (pass == 0) P.npat_ptr[a + 1] = int32_t(nb.size());
            else {
                std::copy(nb.begin(), nb.end(), P.npat_col.begin() + P.npat_ptr[a]);
                std::fill(P.npat_row.begin() + P.npat_ptr[a], P.npat_row.begin() + P.npat_ptr[a + 1], a);
            }
        }
        if (pass == 0) {
            for (int32_t a = 0; a < nall; ++a) P.npat_ptr[a + 1] += P.npat_ptr[a];
            P.npat_col.resize(P.npat_ptr[nall]);
            P.npat_row.resize(P.npat_ptr[nall]);
        }
    }
    // dof -> (node, comp)
    const int32_t ndof = P.n_free + P.n_pres;
    std::vector<int32_t> dof_node(ndof), dof_comp(ndof);
    for (int32_t a = 0; a < nall; ++a)
        for (int j = 0; j < d; ++j) {
            int32_t g = P.node_dof[size_t(a) * d + j];
            dof_node[g] = a;
            dof_comp[g] = j;
        }
    // row lengths per block
    const int32_t nf = P.n_free;
    for (int b = 0; b < 4; ++b) P.rp[b].clear();
    P.rp[0].assign(size_t(nf) + 1, 0); P.rp[1].assign(size_t(nf) + 1, 0);
    P.rp[2].assign(size_t(P.n_pres) + 1, 0); P.rp[3].assign(size_t(P.n_pres) + 1, 0);
    for (int32_t r = 0; r < ndof; ++r) {
        int32_t a = dof_node[r];
        int64_t cf = 0, cp = 0;
        for (int32_t q = P.npat_ptr[a]; q < P.npat_ptr[a + 1]; ++q) {
            int32_t b = P.npat_col[q];
            for (int j = 0; j < d; ++j) (P.node_dof[size_t(b) * d + j] < nf ? cf : cp)++;
        }
        if (r < nf) { P.rp[0][r + 1] = int32_t(cf); P.rp[1][r + 1] = int32_t(cp); }
        else { P.rp[2][r - nf + 1] = int32_t(cf); P.rp[3][r - nf + 1] = int32_t(cp); }
    }
    int64_t base[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < 4; ++b) {
        int64_t acc = 0;
        for (size_t r = 1; r < P.rp[b].size(); ++r) { acc += P.rp[b][r]; }
        if (acc >= (int64_t(1) << 31)) fail(PG_E_NOMEM, "block nnz exceeds 2^31");
        for (size_t r = 1; r < P.rp[b].size(); ++r) P.rp[b][r] += P.rp[b][r - 1];
        P.ci[b].resize(P.rp[b].back());
        base[b + 1] = base[b] + P.rp[b].back();
    }
    if (base[4] >= (int64_t(1) << 31)) fail(PG_E_NOMEM, "total nnz exceeds 2^31");
    std::vector<int32_t> cols;
    for (int32_t r = 0; r < ndof; ++r) {
        int32_t a = dof_node[r];
        cols.clear();
        for (int32_t q = P.npat_ptr[a]; q < P.npat_ptr[a + 1]; ++q) {
            int32_t b = P.npat_col[q];
            for (int j = 0; j < d; ++j) cols.push_back(P.node_dof[size_t(b) * d + j]);
        }
        std::sort(cols.begin(), cols.end());
        int bl = r < nf ? 0 : 2, br = r < nf ? 1 : 3;
        int32_t row = r < nf ? r : r - nf;
        int32_t pl = P.rp[bl][row], pr = P.rp[br][row];
        for (int32_t c : cols) {
            if (c < nf) P.ci[bl][pl++] = c;
            else P.ci[br][pr++] = c - nf;
        }
    }
    // scatter indices per node pair and (i, j)
    const int64_t npairs = int64_t(P.npat_col.size());
    P.npat_map.resize(size_t(npairs) * d * d);
    for (int64_t q = 0; q < npairs; ++q) {
        int32_t a = P.npat_row[q], b = P.npat_col[q];
        for (int i = 0; i < d; ++i) {
            int32_t r = P.node_dof[size_t(a) * d + i];
            for (int j = 0; j < d; ++j) {
                int32_t c = P.node_dof[size_t(b) * d + j];
                int blk = (r < nf ? 0 : 2) + (c < nf ? 0 : 1);
                int32_t row = r < nf ? r : r - nf;
                int32_t col = c < nf ? c : c - nf;
                const int32_t* beg = P.ci[blk].data() + P.rp[blk][row];
                const int32_t* end = P.ci[blk].data() + P.rp[blk][row + 1];
                const int32_t* it = std::lower_bound(beg, end, col);
                if (it == end || *it != col) fail(PG_E_MESH, "scatter index not found in pattern");
                P.npat_map[size_t(q) * d * d + i * d + j] = int32_t(base[blk] + (it - P.ci[blk].data()));
            }
        }
    }
}

// The pattern stage on the GPU (pattern_dev.hip): same outputs as build_pattern.
void build_pattern_device(pgh_problem& P) {
    pgh::PatternIn in;
    in.dim = P.dim;
    in.nndof = P.nndof;
    in.nen = P.nen;
    in.n_nodes = P.n_nodes;
    in.n_enr = P.n_enr;
    in.n_elem = P.n_elem;
    in.n_cut = int32_t(P.cut.size());
    in.n_free = P.n_free;
    in.n_pres = P.n_pres;
    in.conn = P.conn.data();
    in.cut_of_elem = P.cut_of_elem.data();
    in.cut = P.cut.data();
    in.node_dof = P.node_dof.data();
    pgh::PatternOut o;
    std::string err;
    if (int rc = pgh::build_pattern_device(in, o, err)) fail(rc, "device pattern build: " + err);
    P.n2e_ptr = std::move(o.n2e_ptr);
    P.n2e = std::move(o.n2e);
    P.npat_ptr = std::move(o.npat_ptr);
    P.npat_col = std::move(o.npat_col);
    P.npat_row = std::move(o.npat_row);
    P.npat_map = std::move(o.npat_map);
    for (int b = 0; b < 4; ++b) {
        P.rp[b] = std::move(o.rp[b]);
        P.ci[b] = std::move(o.ci[b]);
    }
}

void build_pattern_any(pgh_problem& P, bool device) {
    if (device) build_pattern_device(P);
    else build_pattern(P);
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// runs one set-up stage and adds its wall time to P.stage_s[k]
template <typename F>
void timed(pgh_problem& P, int k, F&& f) {
    const double t0 = now_s();
    f();
    P.stage_s[k] += now_s() - t0;
}

constexpr int32_t kKnownFlags =
    PGH_PLACE_BEST_EFFORT | PGH_PLACE_AT_NODES | PGH_PATTERN_DEVICE | PGH_SETUP_DEVICE | PGH_SCALAR_DOFS;

}  // namespace

extern "C" {

const char* pgh_last_error(void) { return g_err.c_str(); }

int pgh_build_rve(const pgh_rve_params* p, pgh_problem** out) {
    if (!p || !out) { g_err = "null argument"; return PG_E_ARG; }
    *out = nullptr;
    pgh_problem* P = new (std::nothrow) pgh_problem();
    if (!P) { g_err = "out of memory"; return PG_E_NOMEM; }
    try {
        P->prm = *p;
        P->dim = p->dim;
        if (p->dim != 2 && p->dim != 3) fail(PG_E_ARG, "dim must be 2 or 3");
        if (p->n < 1) fail(PG_E_ARG, "n must be >= 1");
        if (p->n_incl < 0 || p->r_min <= 0 || p->r_max < p->r_min || 2 * p->r_max >= 1.0)
            fail(PG_E_ARG, "bad inclusion parameters");
        if (p->flags & ~kKnownFlags) fail(PG_E_ARG, "unknown flag bits");
        P->cohesive = p->cohesive != PGH_COH_NONE;
        P->nndof = (p->flags & PGH_SCALAR_DOFS) ? 1 : p->dim;
        if (P->nndof == 1 && P->cohesive)
            fail(PG_E_UNSUPPORTED, "a scalar field takes no cohesive law (AssemblerThermal::KernelCohesive)");
        pgh_problem& Q = *P;
        timed(Q, 0, [&] { generate_mesh(Q); });
        timed(Q, 1, [&] { place_inclusions(Q); });
        if (p->flags & PGH_SETUP_DEVICE) {
            build_materials(Q);
            setup_on_device(Q);  // adds its own stage times
        } else {
            timed(Q, 2, [&] { detect_nodes(Q); });
            timed(Q, 3, [&] {
                build_materials(Q);
                analyze_and_enrich(Q);
            });
        }
        timed(Q, 4, [&] { number_dofs(Q); });
        timed(Q, 5, [&] { build_pattern_any(Q, (p->flags & PGH_PATTERN_DEVICE) != 0); });
    } catch (const HarnessError& e) {
        g_err = e.what();
        delete P;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "out of memory";
        delete P;
        return PG_E_NOMEM;
    }
    *out = P;
    return PG_OK;
}

int pgh_build_pattern(const pg_problem_desc* m, pgh_problem** out) { return pgh_build_pattern_flags(m, 0, out); }

int pgh_build_pattern_flags(const pg_problem_desc* m, int32_t flags, pgh_problem** out) {
    if (!m || !out) { g_err = "null argument"; return PG_E_ARG; }
    *out = nullptr;
    if (flags & ~PGH_PATTERN_DEVICE) { g_err = "unknown flag bits"; return PG_E_ARG; }
    const int nn = m->nndof ? m->nndof : m->dim;  // dofs per node
    if ((m->dim != 2 && m->dim != 3) || m->nen != m->dim + 1 || m->n_nodes <= 0 || m->n_elem <= 0 || m->n_enr < 0 ||
        m->n_cut < 0 || m->n_mat <= 0 || m->n_free < 0 || m->n_pres < 0 || (nn != 1 && nn != m->dim) ||
        m->n_free + m->n_pres != (m->n_nodes + m->n_enr) * nn) {
        g_err = "inconsistent mesh description";
        return PG_E_ARG;
    }
    if (!m->coords || !m->conn || !m->elem_mat || !m->cut_of_elem || !m->mats || !m->node_dof ||
        (m->n_cut > 0 && (!m->cut || !m->cut_elem))) {
        g_err = "missing mesh array";
        return PG_E_ARG;
    }
    pgh_problem* P = new (std::nothrow) pgh_problem();
    if (!P) { g_err = "out of memory"; return PG_E_NOMEM; }
    try {
        const int64_t nall = int64_t(m->n_nodes) + m->n_enr;
        P->dim = m->dim;
        P->nndof = nn;
        P->nen = m->nen;
        P->cohesive = m->cohesive != 0;
        P->n_nodes = m->n_nodes;
        P->n_enr = m->n_enr;
        P->n_elem = m->n_elem;
        P->n_cqp = m->n_cqp;
        P->n_free = m->n_free;
        P->n_pres = m->n_pres;
        P->coords.assign(m->coords, m->coords + nall * m->dim);
        P->conn.assign(m->conn, m->conn + int64_t(m->n_elem) * m->nen);
        P->elem_mat.assign(m->elem_mat, m->elem_mat + m->n_elem);
        P->cut_of_elem.assign(m->cut_of_elem, m->cut_of_elem + m->n_elem);
        P->cut_elem.assign(m->cut_elem, m->cut_elem + m->n_cut);
        P->cut.assign(m->cut, m->cut + m->n_cut);
        P->mats.assign(m->mats, m->mats + m->n_mat);
        P->node_dof.assign(m->node_dof, m->node_dof + nall * nn);
        for (int64_t k = 0; k < int64_t(m->n_elem) * m->nen; ++k)
            if (P->conn[k] < 0 || P->conn[k] >= m->n_nodes) fail(PG_E_ARG, "connectivity out of range");
        for (int32_t c = 0; c < m->n_cut; ++c) {
            const pg_cut_rec& R = P->cut[c];
            if (P->cut_elem[c] < 0 || P->cut_elem[c] >= m->n_elem || R.n_enr > PG_MAX_ENR)
                fail(PG_E_ARG, "bad cut record");
            for (int k = 0; k < R.n_enr; ++k)
                if (R.enr[k] < m->n_nodes || R.enr[k] >= nall) fail(PG_E_ARG, "enrichment node out of range");
        }
        {
            // every dof belongs to exactly one (node, component)
            std::vector<uint8_t> seen(size_t(m->n_free) + m->n_pres, 0);
            for (int64_t k = 0; k < nall * nn; ++k) {
                const int32_t g = P->node_dof[k];
                if (g < 0 || g >= m->n_free + m->n_pres) fail(PG_E_ARG, "dof out of range");
                if (seen[g]++) fail(PG_E_ARG, "dof numbered twice");
            }
        }
        pgh_problem& Q = *P;
        timed(Q, 5, [&] { build_pattern_any(Q, (flags & PGH_PATTERN_DEVICE) != 0); });
    } catch (const HarnessError& e) {
        g_err = e.what();
        delete P;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "out of memory";
        delete P;
        return PG_E_NOMEM;
    }
    *out = P;
    return PG_OK;
}

int pgh_build_batch(const pgh_rve_params* params, int32_t n, pgh_problem** out) {
    if (!params || !out || n <= 0) { g_err = "bad argument"; return PG_E_ARG; }
    *out = nullptr;
    for (int32_t i = 0; i < n; ++i)
        if (params[i].flags & ~kKnownFlags) { g_err = "unknown flag bits"; return PG_E_ARG; }
    const double t_start = now_s();
    std::vector<pgh_problem*> inst(size_t(n), nullptr);
    std::vector<int> code(size_t(n), 0);
    std::vector<std::string> msg((size_t)n);
    std::atomic<int> next(0);
    auto worker = [&]() {
        for (int i = next++; i < n; i = next++) {
            pgh_problem* P = new (std::nothrow) pgh_problem();
            if (!P) { code[i] = PG_E_NOMEM; msg[i] = "out of memory"; continue; }
            try {
                const pgh_rve_params& p = params[i];
                P->prm = p;
                P->dim = p.dim;
                P->nndof = p.dim;
                if (p.dim != 2) fail(PG_E_UNSUPPORTED, "batches are 2D");
                if (p.flags & PGH_SCALAR_DOFS) fail(PG_E_UNSUPPORTED, "batches are structural");
                if (p.n < 1 || p.n_incl < 0 || p.r_min <= 0 || p.r_max < p.r_min || 2 * p.r_max >= 1.0)
                    fail(PG_E_ARG, "bad instance parameters");
                P->cohesive = p.cohesive != PGH_COH_NONE;
                generate_mesh(*P);
                place_inclusions(*P);
                build_materials(*P);
                if (p.flags & PGH_SETUP_DEVICE) {
                    setup_on_device(*P);
                } else {
                    detect_nodes(*P);
                    analyze_and_enrich(*P);
                }
                number_dofs(*P);
            } catch (const HarnessError& e) {
                code[i] = e.code;
                msg[i] = e.what();
            } catch (const std::bad_alloc&) {
                code[i] = PG_E_NOMEM;
                msg[i] = "out of memory";
            }
            inst[i] = P;
        }
    };
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) th.emplace_back(worker);
    for (auto& t : th) t.join();
    auto cleanup = [&]() { for (auto* P : inst) delete P; };
    for (int i = 0; i < n; ++i)
        if (code[i]) {
            g_err = "instance " + std::to_string(i) + ": " + msg[i];
            const int c = code[i];
            cleanup();
            return c;
        }
    for (int i = 1; i < n; ++i)
        if (inst[i]->mats.size() != inst[0]->mats.size() || inst[i]->cohesive != inst[0]->cohesive ||
            std::memcmp(inst[i]->mats.data(), inst[0]->mats.data(), inst[0]->mats.size() * sizeof(pg_material)) != 0) {
            g_err = "instances of a batch must share their materials";
            cleanup();
            return PG_E_UNSUPPORTED;
        }
    pgh_problem* B = new (std::nothrow) pgh_problem();
    if (!B) { cleanup(); g_err = "out of memory"; return PG_E_NOMEM; }
    try {
        // merged numbering: regular nodes instance by instance, then every
        // instance's enrichment nodes; free dofs instance by instance, then the
        // prescribed dofs -- so each instance owns contiguous rows and columns of
        // all four blocks and the merged pattern is block diagonal
        constexpr int C = PGH_BATCH_COLS;
        B->n_inst = n;
        B->inst_off.assign(size_t(n + 1) * C, 0);
        for (int i = 0; i < n; ++i) {
            const pgh_problem& P = *inst[i];
            int64_t* o = &B->inst_off[size_t(i) * C];
            int64_t* q = o + C;
            q[0] = o[0] + P.n_elem;
            q[1] = o[1] + P.n_nodes;
            q[2] = o[2] + P.n_enr;
            q[3] = o[3] + P.n_free;
            q[4] = o[4] + P.n_pres;
            q[5] = o[5] + P.n_cqp;
            q[6] = o[6] + int64_t(P.cut.size());
            q[7] = o[7] + int64_t(P.incl_r.size());
        }
        const int64_t* T = &B->inst_off[size_t(n) * C];
        if (T[0] >= (int64_t(1) << 31) || T[1] + T[2] >= (int64_t(1) << 31) || (T[3] + T[4]) >= (int64_t(1) << 31))
            fail(PG_E_UNSUPPORTED, "batch exceeds 32-bit indices");
        B->dim = 2;
        B->nndof = 2;
        B->nen = 3;
        B->cohesive = inst[0]->cohesive;
        B->prm = inst[0]->prm;
        B->mats = inst[0]->mats;
        B->n_elem = int32_t(T[0]);
        B->n_nodes = int32_t(T[1]);
        B->n_enr = int32_t(T[2]);
        B->n_free = int32_t(T[3]);
        B->n_pres = int32_t(T[4]);
        B->n_cqp = int32_t(T[5]);
        const int64_t nall = T[1] + T[2];
        B->coords.resize(size_t(nall) * 2);
        B->conn.resize(size_t(T[0]) * 3);
        B->elem_mat.resize(size_t(T[0]));
        B->cut_of_elem.resize(size_t(T[0]));
        B->cut_elem.resize(size_t(T[6]));
        B->cut.resize(size_t(T[6]));
        B->node_dof.resize(size_t(nall) * 2);
        B->bc_val.resize(size_t(T[4]));
        B->node_detect.resize(size_t(T[1]));
        B->enr_incl.resize(size_t(T[2]));
        B->enr_edge.resize(size_t(T[2]) * 2);
        B->enr_alpha.resize(size_t(T[2]));
        for (int i = 0; i < n; ++i) {
            const pgh_problem& P = *inst[i];
            const int64_t* o = &B->inst_off[size_t(i) * C];
            const int32_t e0 = int32_t(o[0]), r0 = int32_t(o[1]), c0 = int32_t(o[6]);
            const int32_t x0 = int32_t(T[1] + o[2]);  // first merged id of this instance's enrichment nodes
            auto node = [&](int32_t a) { return a < P.n_nodes ? r0 + a : x0 + (a - P.n_nodes); };
            for (int32_t a = 0; a < P.n_nodes + P.n_enr; ++a) {
                const int32_t m = node(a);
                for (int c = 0; c < 2; ++c) {
                    B->coords[size_t(m) * 2 + c] = P.coords[size_t(a) * 2 + c];
                    const int32_t dl = P.node_dof[size_t(a) * 2 + c];
                    B->node_dof[size_t(m) * 2 + c] =
                        dl < P.n_free ? int32_t(o[3] + dl) : int32_t(T[3] + o[4] + (dl - P.n_free));
                }
            }
            for (int32_t a = 0; a < P.n_nodes; ++a)
                B->node_detect[size_t(r0) + a] = P.node_detect[a] < 0 ? -1 : P.node_detect[a];
            for (int32_t k = 0; k < P.n_enr; ++k) {
                B->enr_incl[size_t(o[2]) + k] = int32_t(o[7] + P.enr_incl[k]);
                B->enr_edge[(size_t(o[2]) + k) * 2] = r0 + P.enr_edge[size_t(k) * 2];
                B->enr_edge[(size_t(o[2]) + k) * 2 + 1] = r0 + P.enr_edge[size_t(k) * 2 + 1];
                B->enr_alpha[size_t(o[2]) + k] = P.enr_alpha[k];
            }
            for (int32_t k = 0; k < P.n_pres; ++k) B->bc_val[size_t(o[4]) + k] = P.bc_val[k];
            for (int32_t e = 0; e < P.n_elem; ++e) {
                for (int k = 0; k < 3; ++k) B->conn[size_t(e0 + e) * 3 + k] = r0 + P.conn[size_t(e) * 3 + k];
                B->elem_mat[size_t(e0) + e] = P.elem_mat[e];
                B->cut_of_elem[size_t(e0) + e] = P.cut_of_elem[e] < 0 ? -1 : c0 + P.cut_of_elem[e];
            }
            for (size_t c = 0; c < P.cut.size(); ++c) {
                pg_cut_rec R = P.cut[c];
                for (int k = 0; k < R.n_enr; ++k) R.enr[k] = node(R.enr[k]);
                R.isv_off += int32_t(o[5]);
                B->cut[size_t(c0) + c] = R;
                B->cut_elem[size_t(c0) + c] = e0 + P.cut_elem[c];
            }
            for (size_t k = 0; k < P.incl_r.size(); ++k) {
                B->incl_c.push_back(P.incl_c[k * 2]);
                B->incl_c.push_back(P.incl_c[k * 2 + 1]);
                B->incl_r.push_back(P.incl_r[k]);
            }
        }
        cleanup();
        inst.clear();
        B->stage_s[6] = now_s() - t_start;
        pgh_problem& Q = *B;
        timed(Q, 5, [&] { build_pattern_any(Q, (params[0].flags & PGH_PATTERN_DEVICE) != 0); });
    } catch (const HarnessError& e) {
        g_err = e.what();
        cleanup();
        delete B;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "out of memory";
        cleanup();
        delete B;
        return PG_E_NOMEM;
    }
    *out = B;
    return PG_OK;
}

int pgh_batch_info(const pgh_problem* h, int32_t* n_inst, const int64_t** offsets) {
    if (!h || !n_inst || !offsets) { g_err = "null argument"; return PG_E_ARG; }
    *n_inst = h->n_inst;
    *offsets = h->n_inst ? h->inst_off.data() : nullptr;
    return PG_OK;
}

int pgh_get_desc(const pgh_problem* h, pg_problem_desc* o) {
    if (!h || !o) { g_err = "null argument"; return PG_E_ARG; }
    std::memset(o, 0, sizeof(*o));
    o->dim = h->dim;
    o->nen = h->nen;
    o->n_nodes = h->n_nodes;
    o->n_enr = h->n_enr;
    o->n_elem = h->n_elem;
    o->n_cut = int32_t(h->cut.size());
    o->n_mat = int32_t(h->mats.size());
    o->n_cqp = h->n_cqp;
    o->n_free = h->n_free;
    o->n_pres = h->n_pres;
    o->cohesive = h->cohesive;
    o->nndof = h->nndof == h->dim ? 0 : h->nndof;
    for (int b = 0; b < 4; ++b) {
        o->nnz[b] = int64_t(h->ci[b].size());
        o->rp[b] = h->rp[b].data();
        o->ci[b] = h->ci[b].data();
    }
    o->n_pairs = int64_t(h->npat_col.size());
    o->coords = h->coords.data();
    o->conn = h->conn.data();
    o->elem_mat = h->elem_mat.data();
    o->cut_of_elem = h->cut_of_elem.data();
    o->cut_elem = h->cut_elem.data();
    o->cut = h->cut.data();
    o->mats = h->mats.data();
    o->node_dof = h->node_dof.data();
    o->npat_ptr = h->npat_ptr.data();
    o->npat_col = h->npat_col.data();
    o->npat_row = h->npat_row.data();
    o->npat_map = h->npat_map.data();
    o->n2e_ptr = h->n2e_ptr.data();
    o->n2e = h->n2e.data();
    return PG_OK;
}

void pgh_free(pgh_problem* h) { delete h; }

int pgh_inclusions(const pgh_problem* h, const double** centers, const double** radii, int32_t* n) {
    if (!h) return PG_E_ARG;
    if (centers) *centers = h->incl_c.data();
    if (radii) *radii = h->incl_r.data();
    if (n) *n = int32_t(h->incl_r.size());
    return PG_OK;
}

const int32_t* pgh_node_detect(const pgh_problem* h) { return h ? h->node_detect.data() : nullptr; }

const double* pgh_bc_values(const pgh_problem* h) { return h ? h->bc_val.data() : nullptr; }

int pgh_enrich_edges(const pgh_problem* h, const int32_t** edges, const double** alpha) {
    if (!h || !edges || !alpha) { g_err = "null argument"; return PG_E_ARG; }
    if (h->enr_alpha.size() != size_t(h->n_enr)) { g_err = "enrichment edges unknown for this problem"; return PG_E_ARG; }
    *edges = h->enr_edge.data();
    *alpha = h->enr_alpha.data();
    return PG_OK;
}

const int32_t* pgh_enrich_inclusion(const pgh_problem* h) {
    return h && !h->enr_incl.empty() ? h->enr_incl.data() : nullptr;
}

int pgh_stage_seconds(const pgh_problem* h, double* out, int32_t n) {
    if (!h || !out || n < PGH_N_STAGES) { g_err = "bad argument"; return PG_E_ARG; }
    std::copy(h->stage_s, h->stage_s + PGH_N_STAGES, out);
    return PG_OK;
}

int pgh_elem_qps(const pgh_problem* h, int32_t e, int32_t* n_vqp, int32_t* n_cqp) {
    if (!h || e < 0 || e >= h->n_elem) return PG_E_ARG;
    int32_t ci = h->cut_of_elem[e];
    if (ci < 0) {
        if (n_vqp) *n_vqp = h->nen == 3 ? 4 : 5;
        if (n_cqp) *n_cqp = 0;
    } else {
        const pg_cut_rec& R = h->cut[ci];
        if (n_vqp) *n_vqp = R.n_vqp;
        if (n_cqp) *n_cqp = R.n_coh * (R.coh_nn == 4 ? 2 : 4);
    }
    return PG_OK;
}

}  // extern "C"
