// Device build of the set-up's pattern stage (internal interface between the
// host harness and pattern_dev.hip).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "pgfem_asm.h"

namespace pgh {

struct PatternIn {
    int dim = 0, nen = 0;
    int nndof = 0;                         // dofs per node (dim, or 1 for a scalar field)
    int32_t n_nodes = 0, n_enr = 0, n_elem = 0, n_cut = 0, n_free = 0, n_pres = 0;
    const int32_t* conn = nullptr;         // n_elem * nen
    const int32_t* cut_of_elem = nullptr;  // n_elem
    const pg_cut_rec* cut = nullptr;       // n_cut
    const int32_t* node_dof = nullptr;     // (n_nodes + n_enr) * nndof
};

struct PatternOut {
    std::vector<int32_t> n2e_ptr, n2e, npat_ptr, npat_col, npat_row, npat_map;
    std::vector<int32_t> rp[4], ci[4];
};

// The node incidence (ascending elements), the node-pair pattern (sorted
// neighbours), the 4-block AIJ pattern and the scatter indices, equal to the
// host build.  Returns PG_OK or a PG_E_* code with a message in err.
int build_pattern_device(const PatternIn& in, PatternOut& out, std::string& err);

}  // namespace pgh
