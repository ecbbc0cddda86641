// C++ host-mirror test: AssemblerGPU (pgfem_assembler.hpp) against the oracle.
//
//   test_host_mirror cpu      host-side checks only (no GPU): ISV layout and
//                             flatten / unflatten, NodalData gather, pattern
//   test_host_mirror gpu      assemble through the C++ API with every schedule
//                             and compare K (four blocks), P_int and the trial
//                             ISVs with the oracle
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <random>
#include <string>
#include <vector>

#include "../../include/pgfem_assembler.hpp"
#include "../../include/pgfem_harness.h"

extern "C" {
struct oracle;
oracle* oracle_new(const pg_problem_desc* d, int e0, int e1);
void oracle_free(oracle* o);
const char* oracle_error(const oracle* o);
void oracle_nnz(const oracle* o, int64_t* nnz);
int oracle_assemble(oracle* o, const double* U, const double* isv_c, double* isv_t, int func, int jac, double* K,
                    double* P);
}

using namespace pgfem;

static int fails = 0;
#define CHECK(c, ...)                                   \
    do {                                                \
        if (!(c)) {                                     \
            printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            printf(__VA_ARGS__);                        \
            printf("\n");                               \
            ++fails;                                    \
        }                                               \
    } while (0)

static pgh_problem* build(int cohesive, int dim, int n, int n_incl, uint64_t seed) {
    pgh_rve_params p;
    memset(&p, 0, sizeof(p));
    p.dim = dim;
    p.n = n;
    p.n_incl = n_incl;
    p.cohesive = cohesive;
    p.r_min = dim == 2 ? 0.08 : 0.12;
    p.r_max = dim == 2 ? 0.14 : 0.2;
    p.gap = 2.0 / n * (dim == 3 ? sqrt(3.0) : 1.0);
    p.seed = seed;
    p.E_matrix = 3.5e3;
    p.nu_matrix = 0.35;
    p.E_fiber = 7.0e4;
    p.nu_fiber = 0.2;
    p.sigma_c = 50.0;
    p.delta_c = 1e-3;
    p.delta_c1 = 50.0 / 1e6;
    p.delta_c2 = 1e-3;
    p.beta = 1.0;
    p.alpha = 50.0;
    p.strain = 0.02;
    pgh_problem* h = nullptr;
    if (pgh_build_rve(&p, &h) != PG_OK) {
        printf("pgh_build_rve: %s\n", pgh_last_error());
        return nullptr;
    }
    return h;
}

static double rel(const std::vector<double>& a, const std::vector<double>& b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return den > 0 ? sqrt(num / den) : sqrt(num);
}

// host-side parts: no device call is made
static void cpu_checks(const pgh_problem* h, const pg_problem_desc& d) {
    InternalVars iv(d);
    ISVs s;
    iv.preallocate_internal_vars_object(s);
    CHECK(s.size() == size_t(d.n_elem), "ISV storage has one entry per element");
    size_t ncq = 0;
    for (int32_t e = 0; e < d.n_elem; ++e) {
        int32_t nv = 0, nc = 0;
        pgh_elem_qps(h, e, &nv, &nc);
        CHECK(s[e].size() == size_t(nv + nc), "element %d: %zu qps, harness %d + %d", e, s[e].size(), nv, nc);
        for (auto& q : s[e]) ncq += q.size();
    }
    CHECK(ncq == size_t(d.n_cqp), "one delta_max per cohesive qp (%zu vs %d)", ncq, d.n_cqp);
    std::vector<double> flat(size_t(d.n_cqp)), back;
    for (size_t k = 0; k < flat.size(); ++k) flat[k] = 1.0 + double(k);
    iv.unflatten(flat, s);
    iv.flatten(s, back);
    CHECK(back == flat, "flatten(unflatten(x)) == x");
    iv.copy_data_in(s);
    ISVs t;
    iv.copy_data(t);
    CHECK(t == s, "copy_data returns what copy_data_in stored");
}

static int gpu_checks(const pg_problem_desc& d, const char* name) {
    oracle* o = oracle_new(&d, 0, -1);
    if (!o || oracle_error(o)[0]) {  // the oracle reports "" when set-up succeeded
        printf("oracle_new: %s\n", o ? oracle_error(o) : "null");
        return 1;
    }
    int64_t onnz[4];
    oracle_nnz(o, onnz);
    for (int b = 0; b < 4; ++b) CHECK(onnz[b] == d.nnz[b], "block %d nnz: oracle %lld, harness %lld", b, (long long)onnz[b], (long long)d.nnz[b]);

    const int32_t nall = d.n_nodes + d.n_enr;
    NodalData sol(nall, d.dim);
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> ud(-1e-5, 1e-5);
    std::vector<double> U(size_t(d.n_free + d.n_pres));
    for (int32_t a = 0; a < nall; ++a)
        for (int c = 0; c < d.dim; ++c) {
            const double v = ud(g);
            sol.get_value_local(a, c) = v;
            U[d.node_dof[int64_t(a) * d.dim + c]] = v;
        }
    InternalVars iv(d);
    std::vector<double> ic(size_t(d.n_cqp));
    std::uniform_real_distribution<double> uh(0.0, 5e-4);
    for (auto& x : ic) x = uh(g);
    iv.unflatten(ic, iv.data());

    int64_t nnz_total = d.nnz[0] + d.nnz[1] + d.nnz[2] + d.nnz[3];
    std::vector<double> Kr((size_t)nnz_total), Pr(U.size()), Ir((size_t)d.n_cqp + 1);
    if (oracle_assemble(o, U.data(), d.n_cqp ? ic.data() : nullptr, Ir.data(), 1, 1, Kr.data(), Pr.data()) != 0) {
        printf("oracle_assemble: %s\n", oracle_error(o));
        return 1;
    }
    Ir.resize(size_t(d.n_cqp));

    const AssemblerGPU::Schedule sch[3] = {AssemblerGPU::Schedule::FP64, AssemblerGPU::Schedule::Binned64,
                                           AssemblerGPU::Schedule::Mixed};
    const char* sname[3] = {"fp64", "binned64", "mixed"};
    for (int s = 0; s < 3; ++s) {
        AssemblerGPU asmb(d, sch[s]);
        asmb.attach_internal_vars(&iv);
        pmatrix K;
        pvector P;
        asmb.preallocate_matrix(K);
        asmb.preallocate_vector(P);
        for (int b = 0; b < 4; ++b) std::fill(K.block(b).values.begin(), K.block(b).values.end(), 9.0);
        ISVs upd;
        const int rc = asmb.assemble(K, P, 0.0, upd, &sol, true, true);
        CHECK(rc == PG_OK, "%s assemble rc %d: %s", sname[s], rc, asmb.last_error());
        if (rc != PG_OK) continue;
        std::vector<double> Kc, Pc, It;
        for (int b = 0; b < 4; ++b) Kc.insert(Kc.end(), K.block(b).values.begin(), K.block(b).values.end());
        Pc = P.f;
        Pc.insert(Pc.end(), P.p.begin(), P.p.end());
        iv.flatten(upd, It);
        const double rk = rel(Kc, Kr), rp = rel(Pc, Pr);
        double di = 0, mi = 1e-30;
        for (size_t k = 0; k < It.size(); ++k) {
            di = fmax(di, fabs(It[k] - Ir[k]));
            mi = fmax(mi, fabs(Ir[k]));
        }
        const double tk = s == 2 ? 1e-6 : 1e-12;
        printf("%s %-8s relK %.3e relP %.3e max|isv diff| %.3e\n", name, sname[s], rk, rp, di);
        CHECK(rk < tk, "%s %s relK %.3e", name, sname[s], rk);
        CHECK(rp < 1e-12, "%s %s relP %.3e", name, sname[s], rp);
        CHECK(di <= 1e-12 * mi, "%s %s isv", name, sname[s]);
        // the committed variables are untouched by an assembly
        std::vector<double> ic2;
        iv.flatten(iv.data(), ic2);
        CHECK(ic2 == ic, "%s committed ISVs unchanged", sname[s]);
    }
    oracle_free(o);
    return 0;
}

int main(int argc, char** argv) {
    const bool gpu = argc > 1 && std::string(argv[1]) == "gpu";
    struct Case { const char* name; int coh, dim, n, ni; } cases[] = {
        {"c0", PGH_COH_NONE, 2, 24, 3}, {"bl", PGH_COH_BL, 2, 24, 3}, {"op", PGH_COH_OP, 2, 24, 3}, {"tet", PGH_COH_NONE, 3, 8, 1}};
    for (const Case& c : cases) {
        pgh_problem* h = build(c.coh, c.dim, c.n, c.ni, 11);
        if (!h) return 2;
        pg_problem_desc d;
        pgh_get_desc(h, &d);
        printf("case %s: %d elements, %d cut, %d cohesive qps\n", c.name, d.n_elem, d.n_cut, d.n_cqp);
        fflush(stdout);
        cpu_checks(h, d);
        if (gpu && gpu_checks(d, c.name)) fails++;
        pgh_free(h);
    }
    printf("%s: %d failure(s)\n", fails ? "FAILED" : "OK", fails);
    return fails ? 1 : 0;
}
