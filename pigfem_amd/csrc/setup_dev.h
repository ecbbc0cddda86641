// Device build of the set-up's detection and enrichment stages (internal
// interface between the host harness and setup_dev.hip).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "pgfem_asm.h"

namespace pgh {

struct SetupIn {
    int dim = 0, nen = 0;
    bool cohesive = false;
    int32_t n_nodes = 0, n_elem = 0, n_incl = 0;
    const double* coords = nullptr;  // n_nodes * dim (regular nodes)
    const int32_t* conn = nullptr;   // n_elem * nen
    const double* incl_c = nullptr;  // n_incl * dim
    const double* incl_r = nullptr;  // n_incl
    double cell = 0.0;               // bin size of the inclusion grid (>= the largest radius)
};

struct SetupOut {
    std::vector<int32_t> node_detect, elem_mat, cut_of_elem, cut_elem, enr_incl, enr_edge;
    std::vector<double> enr_alpha, enr_coords;  // enr_coords: n_enr * dim
    std::vector<pg_cut_rec> cut;
    int32_t n_enr = 0, n_cqp = 0;
    double seconds[2] = {0.0, 0.0};  // node detection; element analysis and enrichment
};

// Node detection (Mesh::detect_node), element detection and child topology
// (Mesh::analyze_elements), enrichment nodes (Mesh::add_enrichments) and the
// cut records, equal to the host build.  Returns PG_OK or a PG_E_* code with a
// message in err.
int setup_device(const SetupIn& in, SetupOut& out, std::string& err);

}  // namespace pgh
