/*
 * pgfem_assembler.hpp -- C++ host mirror of PIGFEM's Assembler plug-in API over
 * the C ABI of pgfem_asm.h (header only; link with -lpgfem_asm).
 *
 * Keeps the reference's names, argument meaning and error behaviour:
 *   pmatrix / pvector                 include/common.h:32-101 (four SeqAIJ
 *                                     blocks ff, fp, pf, pp; vectors f, p)
 *   NodalData                         include/NodalData.h (per node, nndof values)
 *   InternalVars (ISV storage, copy_data / copy_data_in /
 *     preallocate_internal_vars_object) src/InternalVars.cpp:43-170
 *   Assembler::assemble(K, P_int, delta_t, update_ISVs, solution,
 *                       assembleFunc, assembleJac)
 *                                     include/Assembler.h:45-47,
 *                                     src/Assembler.cpp:65-146
 *   Solver::preallocate_matrix        src/Solver.cpp:414-490
 * `assemble` returns 0 or a PG_E_* status, as the reference returns a
 * PetscErrorCode; the message is in last_error().  Construction failures throw
 * std::runtime_error (the reference calls err_message and exits).
 *
 * The CSR blocks hold PETSc's SeqAIJ arrays (row pointers, sorted column
 * indices, values); an adapter inside the reference would hand
 * MatSeqAIJGetArray's value arrays to pg_asm_assemble instead.
 */
#ifndef PGFEM_ASSEMBLER_HPP
#define PGFEM_ASSEMBLER_HPP

#include <stdint.h>

#include <cctype>
#include <stdexcept>
#include <string>
#include <vector>

#include "pgfem_asm.h"

namespace pgfem {

using ISVs = std::vector<std::vector<std::vector<double>>>;  // [element][qp][internal variable]

struct CSRBlock {
    int32_t n_rows = 0, n_cols = 0;
    std::vector<int32_t> row_ptr, col_idx;
    std::vector<double> values;
};

struct pmatrix {
    CSRBlock ff, fp, pf, pp;
    CSRBlock& block(int b) { return b == 0 ? ff : b == 1 ? fp : b == 2 ? pf : pp; }
    const CSRBlock& block(int b) const { return b == 0 ? ff : b == 1 ? fp : b == 2 ? pf : pp; }
};

struct pvector {
    std::vector<double> f, p;
};

/* Nodal solution: regular nodes, then enrichment nodes, nndof values each. */
class NodalData {
public:
    NodalData(int32_t n_nodes, int32_t nndof) : _nndof(nndof), _data(size_t(n_nodes) * nndof, 0.0) {}
    double& get_value_local(int32_t node, int32_t dof) { return _data[size_t(node) * _nndof + dof]; }
    double get_value_local(int32_t node, int32_t dof) const { return _data[size_t(node) * _nndof + dof]; }
    int32_t nndof() const { return _nndof; }
    int32_t n_nodes() const { return int32_t(_data.size() / size_t(_nndof)); }

private:
    int32_t _nndof;
    std::vector<double> _data;
};

/* Committed internal variables of every quadrature point: an element holds its
 * volumetric qps (LEI: no variables) and, if intersected and cohesive, one
 * delta_max per cohesive qp after them (src/InternalVars.cpp:122-170). */
class InternalVars {
public:
    explicit InternalVars(const pg_problem_desc& d) : _d(d) { preallocate_internal_vars_object(_iv); }

    void preallocate_internal_vars_object(ISVs& storage) const {
        const int nq_plain = _d.dim == 2 ? 4 : 5;  // tri3 / tet4 order-3 rules
        storage.assign(size_t(_d.n_elem), {});
        for (int32_t e = 0; e < _d.n_elem; ++e) {
            const int32_t ci = _d.cut_of_elem[e];
            if (ci < 0) {
                storage[e].assign(nq_plain, {});
                continue;
            }
            const pg_cut_rec& r = _d.cut[ci];
            const int ncq = r.n_coh * (r.coh_nn == 4 ? 2 : 4);  // CohEdge2: 2 Gauss points, CohTri3: 4
            storage[e].assign(size_t(r.n_vqp) + ncq, {});
            for (int k = 0; k < ncq; ++k) storage[e][r.n_vqp + k].assign(1, 0.0);
        }
    }
    /* committed -> storage (the reset at the start of every assembly) */
    void copy_data(ISVs& storage) const { storage = _iv; }
    /* storage -> committed (after a converged load step) */
    void copy_data_in(const ISVs& storage) { _iv = storage; }
    ISVs& data() { return _iv; }

    /* flat cohesive-qp order of the C ABI <-> nested storage */
    void flatten(const ISVs& s, std::vector<double>& flat) const {
        flat.assign(size_t(_d.n_cqp), 0.0);
        for (int32_t c = 0; c < _d.n_cut; ++c) {
            const pg_cut_rec& r = _d.cut[c];
            const auto& q = s[_d.cut_elem[c]];
            for (size_t k = r.n_vqp; k < q.size(); ++k) flat[r.isv_off + (k - r.n_vqp)] = q[k].at(0);
        }
    }
    void unflatten(const std::vector<double>& flat, ISVs& s) const {
        for (int32_t c = 0; c < _d.n_cut; ++c) {
            const pg_cut_rec& r = _d.cut[c];
            auto& q = s[_d.cut_elem[c]];
            for (size_t k = r.n_vqp; k < q.size(); ++k) q[k].at(0) = flat[r.isv_off + (k - r.n_vqp)];
        }
    }

private:
    const pg_problem_desc& _d;
    ISVs _iv;
};

class AssemblerGPU {
public:
    enum class Schedule { FP64, Binned64, Mixed };

    AssemblerGPU(const pg_problem_desc& d, Schedule s = Schedule::FP64) : _d(d), _sched(s) {
        if (pg_asm_create(&d, &_h) != PG_OK) throw std::runtime_error(std::string("pg_asm_create: ") + pg_asm_last_error());
        int64_t nd = 0, nnz = 0, ncq = 0;
        pg_asm_sizes(_h, &nd, &nnz, &ncq);
        _n_dof = nd;
    }
    ~AssemblerGPU() { pg_asm_destroy(_h); }
    AssemblerGPU(const AssemblerGPU&) = delete;
    AssemblerGPU& operator=(const AssemblerGPU&) = delete;

    void attach_internal_vars(InternalVars* iv) { _iv = iv; }
    void set_schedule(Schedule s) { _sched = s; }
    const char* last_error() const { return pg_asm_last_error(); }

    /* Solver::preallocate_matrix: the four blocks with the assembly pattern */
    void preallocate_matrix(pmatrix& K) const {
        const int32_t rows[4] = {_d.n_free, _d.n_free, _d.n_pres, _d.n_pres};
        const int32_t cols[4] = {_d.n_free, _d.n_pres, _d.n_free, _d.n_pres};
        for (int b = 0; b < 4; ++b) {
            CSRBlock& B = K.block(b);
            B.n_rows = rows[b];
            B.n_cols = cols[b];
            B.row_ptr.assign(_d.rp[b], _d.rp[b] + rows[b] + 1);
            B.col_idx.assign(_d.ci[b], _d.ci[b] + _d.nnz[b]);
            B.values.assign(size_t(_d.nnz[b]), 0.0);
        }
    }
    void preallocate_vector(pvector& P) const {
        P.f.assign(size_t(_d.n_free), 0.0);
        P.p.assign(size_t(_d.n_pres), 0.0);
    }

    /* Assembler::assemble: K and P_int overwritten, update_ISVs reset from the
     * committed variables and then updated by the cohesive laws. */
    int assemble(pmatrix& K, pvector& P_int, double delta_t, ISVs& update_ISVs, NodalData* solution,
                 bool assembleFunc = true, bool assembleJac = true) {
        if (!assembleFunc && !assembleJac) return PG_OK;
        if (!solution || solution->nndof() != _d.dim || solution->n_nodes() != _d.n_nodes + _d.n_enr) return PG_E_ARG;
        // formElementalVector's gather source in global-dof order
        _U.assign(size_t(_n_dof), 0.0);
        const int32_t nall = _d.n_nodes + _d.n_enr;
        for (int32_t a = 0; a < nall; ++a)
            for (int32_t c = 0; c < _d.dim; ++c) _U[_d.node_dof[int64_t(a) * _d.dim + c]] = solution->get_value_local(a, c);
        if (_iv) {
            _iv->copy_data(update_ISVs);
            _iv->flatten(update_ISVs, _ic);
        } else {
            _ic.assign(size_t(_d.n_cqp), 0.0);
        }
        _it.assign(size_t(_d.n_cqp), 0.0);
        if (assembleJac)
            for (int b = 0; b < 4; ++b)
                if (K.block(b).values.size() != size_t(_d.nnz[b])) return PG_E_ARG;
        if (assembleFunc && (P_int.f.size() != size_t(_d.n_free) || P_int.p.size() != size_t(_d.n_pres))) return PG_E_ARG;
        unsigned flags = (assembleFunc ? PG_FUNC : 0u) | (assembleJac ? PG_JAC : 0u);
        if (_sched == Schedule::Binned64) flags |= PG_BATCHED;
        if (_sched == Schedule::Mixed) flags |= PG_FP32 | PG_BATCHED;
        const bool ncq = _d.n_cqp > 0;
        const int rc = pg_asm_assemble(_h, _U.data(), ncq ? _ic.data() : nullptr, ncq ? _it.data() : nullptr, delta_t, flags,
                                       assembleJac ? K.ff.values.data() : nullptr, assembleJac ? K.fp.values.data() : nullptr,
                                       assembleJac ? K.pf.values.data() : nullptr, assembleJac ? K.pp.values.data() : nullptr,
                                       assembleFunc ? P_int.f.data() : nullptr, assembleFunc ? P_int.p.data() : nullptr);
        if (rc != PG_OK) return rc;
        if (_iv && ncq) _iv->unflatten(_it, update_ISVs);
        return PG_OK;
    }

protected:
    const pg_problem_desc& _d;
    Schedule _sched;
    pg_asm* _h = nullptr;
    InternalVars* _iv = nullptr;
    int64_t _n_dof = 0;
    std::vector<double> _U, _ic, _it;
};

/* AssemblerSmallStrainStructural_ThermalPreLoad
 * (src/AssemblerSmallStrainStructural_ThermalPreLoad.cpp): the structural
 * assembler with a temperature change that strains the LEI materials whose
 * thermal_exp_set is nonzero.  set_parameter / get_parameter / clear and the
 * load-step rule of assemble_new_load_step follow :25-33,197-267; the current
 * change reaches the kernels through pg_asm_set_temperature_change. */
class AssemblerGPU_ThermalPreLoad : public AssemblerGPU {
public:
    using AssemblerGPU::AssemblerGPU;

    void clear() {
        _delta_T = _curr_delta_T = 0.0;
        _temp_set = _relaxation_complete = false;
        pg_asm_set_temperature_change(_h, 0.0);
    }
    /* returns PG_E_ARG for an unknown name (the reference calls err_message) */
    int set_parameter(std::string name, double val) {
        for (auto& c : name) c = char(std::toupper(static_cast<unsigned char>(c)));
        if (name == "TEMP" || name == "DELTA_T" || name == "DELTA_TEMP") {
            _delta_T = val;
            _curr_delta_T = 0.0;
            _temp_set = true;
        } else if (name == "COMPLETE" || name == "CONVERGED") {
            _relaxation_complete = val != 0.0;
        } else {
            return PG_E_ARG;
        }
        return pg_asm_set_temperature_change(_h, _curr_delta_T);
    }
    double get_parameter(std::string name) const {
        for (auto& c : name) c = char(std::toupper(static_cast<unsigned char>(c)));
        if (name == "TEMP" || name == "DELTA_T" || name == "DELTA_TEMP") {
            if (!_temp_set) throw std::runtime_error("Please set a temperature change value before attempting to get it.");
            return _delta_T;
        }
        if (name == "COMPLETE" || name == "CONVERGED") return _relaxation_complete ? 1.0 : 0.0;
        if (name == "CURRENT TEMP" || name == "CURRENT_TEMP") return _curr_delta_T;
        throw std::runtime_error(name + " is not a valid parameter name");
    }
    /* the temperature update of assemble_new_load_step; the prescribed
     * displacements and loads of the base class stay with the caller */
    int assemble_new_load_step(double current_time_frac) {
        if (!_temp_set) return PG_E_ARG;
        _curr_delta_T = _relaxation_complete ? _delta_T : _delta_T * current_time_frac;
        return pg_asm_set_temperature_change(_h, _curr_delta_T);
    }

private:
    double _delta_T = 0.0, _curr_delta_T = 0.0;
    bool _temp_set = false, _relaxation_complete = false;
};

}  // namespace pgfem

#endif /* PGFEM_ASSEMBLER_HPP */
