// Newton's linear solve on the device: K_ff x = b with a Krylov method and a
// multicolour symmetric Gauss-Seidel preconditioner (rocALUTION, AMD's
// iterative solver library), in place of the KSP solve of SolverNonlinearKSP
// (src/SolverNonlinearKSP.cpp:94; MUMPS through PETSc, src/Solver.cpp:108-160).
// (Measured on tangents of these RVEs, rocALUTION's aggregation and
// Ruge-Stueben multigrid preconditioners did not converge within 3000
// iterations where multicolour SGS did; Jacobi needs 2-3x the iterations.)
//
// The structure of K_ff comes from the problem description once; the values
// are taken at every solve from the device buffer that pg_asm_assemble_device
// writes (the ff block leads [ff | fp | pf | pp]) by a device-to-device copy,
// so K never crosses PCIe.  The preconditioner is built at the first solve
// and rebuilt numerically (same pattern) at the later ones.  The workspace
// (matrix, vectors, preconditioner) belongs to the solver object.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>
#include <string>

#include <hip/hip_runtime_api.h>
#include <rocalution/rocalution.hpp>

#include "pgfem_asm.h"

namespace {

using Mat = rocalution::LocalMatrix<double>;
using Vec = rocalution::LocalVector<double>;
using Solver = rocalution::IterativeLinearSolver<Mat, Vec, double>;
using Sgs = rocalution::MultiColoredSGS<Mat, Vec, double>;

thread_local std::string k_err;
std::once_flag g_init;
int g_init_rc = 0;

}  // namespace

struct pg_ksp {
    int64_t n = 0, nnz = 0;
    int method = 0;
    Mat A;
    Vec x, b;
    Solver* solver = nullptr;
    Sgs* pre = nullptr;
    bool built = false;
};

extern "C" {

const char* pg_ksp_last_error(void) { return k_err.c_str(); }

int pg_ksp_create(const pg_problem_desc* d, int method, pg_ksp** out) {
    if (!out) { k_err = "null output"; return PG_E_ARG; }
    *out = nullptr;
    if (!d || !d->rp[0] || (d->nnz[0] > 0 && !d->ci[0])) { k_err = "missing K_ff structure"; return PG_E_ARG; }
    if (method != PG_KSP_CG && method != PG_KSP_FGMRES && method != PG_KSP_BICGSTAB) {
        k_err = "unknown solver method";
        return PG_E_ARG;
    }
    const int64_t n = d->n_free, nnz = d->nnz[0];
    if (n <= 0) { k_err = "no free dofs"; return PG_E_ARG; }
    if (d->rp[0][0] != 0 || d->rp[0][n] != nnz) { k_err = "K_ff row pointers do not match its nnz"; return PG_E_ARG; }
    std::call_once(g_init, [] {
        g_init_rc = rocalution::init_rocalution();
        if (g_init_rc == 0 && !rocalution::_rocalution_available_accelerator()) g_init_rc = -1;
    });
    if (g_init_rc != 0) { k_err = "rocALUTION found no accelerator"; return PG_E_HIP; }
    pg_ksp* k = new (std::nothrow) pg_ksp();
    if (!k) { k_err = "out of memory"; return PG_E_NOMEM; }
    k->n = n;
    k->nnz = nnz;
    k->method = method;
    // structure on the host, then the whole matrix on the device
    k->A.AllocateCSR("K_ff", nnz, n, n);
    {
        double* zeros = new (std::nothrow) double[size_t(nnz > 0 ? nnz : 1)]();
        if (!zeros) { delete k; k_err = "out of memory"; return PG_E_NOMEM; }
        k->A.CopyFromCSR(d->rp[0], d->ci[0], zeros);
        delete[] zeros;
    }
    k->A.MoveToAccelerator();
    k->x.Allocate("x", n);
    k->b.Allocate("b", n);
    k->x.MoveToAccelerator();
    k->b.MoveToAccelerator();
    if (method == PG_KSP_CG) k->solver = new rocalution::CG<Mat, Vec, double>();
    else if (method == PG_KSP_FGMRES) {
        auto* g = new rocalution::FGMRES<Mat, Vec, double>();
        g->SetBasisSize(50);
        k->solver = g;
    } else k->solver = new rocalution::BiCGStab<Mat, Vec, double>();
    k->pre = new Sgs();
    k->solver->Verbose(0);
    *out = k;
    return PG_OK;
}

int pg_ksp_solve(pg_ksp* k, const double* dK, const double* db, double* dx, double rtol, int max_iter, int* iters,
                 double* rel_res) {
    if (!k) { k_err = "null solver"; return PG_E_ARG; }
    if (!dK || !db || !dx) { k_err = "missing device buffer"; return PG_E_ARG; }
    if (!(rtol > 0.0) || max_iter <= 0) { k_err = "bad tolerance or iteration limit"; return PG_E_ARG; }
    if (iters) *iters = 0;
    if (rel_res) *rel_res = 0.0;
    // the values of the ff block, device to device into the matrix's own
    // value array (the matrix lives on the accelerator: its arrays are device
    // memory), then the right-hand side and the initial guess
    {
        int32_t* rp = nullptr;
        int* ci = nullptr;
        double* val = nullptr;
        k->A.LeaveDataPtrCSR(&rp, &ci, &val);
        const hipError_t e = hipMemcpy(val, dK, size_t(k->nnz) * sizeof(double), hipMemcpyDeviceToDevice);
        k->A.SetDataPtrCSR(&rp, &ci, &val, "K_ff", k->nnz, k->n, k->n);
        if (e != hipSuccess) { k_err = std::string("hipMemcpy of K_ff: ") + hipGetErrorString(e); return PG_E_HIP; }
    }
    k->b.CopyFromData(db);
    k->x.CopyFromData(dx);
    const double bn = k->b.Norm();
    if (!std::isfinite(bn)) { k_err = "right-hand side is not finite"; return PG_E_ARG; }
    if (bn == 0.0) {
        k->x.Zeros();
        k->x.CopyToData(dx);
        return PG_OK;
    }
    if (!k->built) {
        k->solver->SetOperator(k->A);
        k->solver->SetPreconditioner(*k->pre);
        k->solver->Build();
        k->built = true;
    } else {
        k->solver->ReBuildNumeric();
    }
    k->solver->Init(0.0, rtol, 1e8, max_iter);
    k->solver->Solve(k->b, &k->x);
    const int st = k->solver->GetSolverStatus();
    const int it = k->solver->GetIterationCount();
    // the true residual, |b - A x| / |b|
    Vec r;
    r.CloneBackend(k->b);
    r.Allocate("r", k->n);
    k->A.Apply(k->x, &r);
    r.ScaleAdd(-1.0, k->b);
    const double rel = r.Norm() / bn;
    if (iters) *iters = it;
    if (rel_res) *rel_res = rel;
    if (!std::isfinite(rel)) { k_err = "the solve diverged (non-finite residual)"; return PG_E_ARG; }
    k->x.CopyToData(dx);
    if (st == 3) { k_err = "the solve diverged"; return PG_E_ARG; }
    return PG_OK;
}

int pg_ksp_destroy(pg_ksp* k) {
    if (!k) return PG_OK;
    if (k->solver) k->solver->Clear();
    delete k->solver;
    delete k->pre;
    delete k;
    return PG_OK;
}

}  // extern "C"
