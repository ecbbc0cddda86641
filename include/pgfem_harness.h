/*
 * pgfem_harness.h -- host set-up of synthetic IGFEM RVEs (C ABI).
 *
 * The harness restates, on flat arrays, the part of PIGFEM's Problem::init that
 * produces what the assembly reads (src/Problem.cpp:100-318):
 *   structured tri3 / tet4 mesh generation       src/Mesh.cpp:236-630
 *   node detection against circles / spheres      src/Mesh.cpp:2883-2930,
 *                                                 src/InclusionCircle.cpp:150-163,
 *                                                 src/InclusionSphere.cpp:163-177
 *   element cut classification + child topology   src/tri3.cpp:287-596,
 *                                                 src/tet4.cpp:228-520,
 *                                                 src/elem.cpp:428-559
 *   enrichment nodes (set order, mirrors after)   src/Mesh.cpp:3104-3251,3661-3720
 *   edge intersection (Steffensen, |da|<1e-7)     src/Inclusion.cpp:32-120
 *   Dirichlet sets, free-first dof numbering      src/BoundaryObject.cpp:32-99,
 *                                                 src/DofObject.cpp:241-278,498-538
 *   4-block AIJ pattern (union of element blocks) src/Solver.cpp:414-490,
 *                                                 src/Assembler.cpp:659-679
 * It fills a pg_problem_desc (pgfem_asm.h) that the assembler consumes.
 */
#ifndef PGFEM_HARNESS_H
#define PGFEM_HARNESS_H

#include <stdint.h>
#include "pgfem_asm.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PGH_COH_NONE 0
#define PGH_COH_BL   1   /* Material_BilinearCohesive  */
#define PGH_COH_OP   2   /* Material_OPCohesive        */
#define PGH_COH_BL_NU 3
#define PGH_COH_OP_NU 4

typedef struct pgh_rve_params {
    int32_t dim;            /* 2: unit square of tri3, 3: unit cube of tet4   */
    int32_t n;              /* elements per side (n x n, or n x n x n cubes)   */
    int32_t n_incl;         /* circles / spheres                               */
    int32_t cohesive;       /* PGH_COH_*                                       */
    double  r_min, r_max;   /* radius ~ U[r_min, r_max]                        */
    double  gap;            /* minimum surface-to-surface gap between inclusions */
    uint64_t seed;          /* std::mt19937_64 seed                             */
    double  E_matrix, nu_matrix, E_fiber, nu_fiber;
    double  sigma_c, delta_c, delta_c1, delta_c2, beta, alpha;
    double  strain;         /* applied u_x on RIGHT = strain * width            */
    int32_t max_tries;      /* rejection-sampling budget (0 = default)          */
    int32_t flags;          /* PGH_PLACE_* | PGH_PATTERN_DEVICE | PGH_SETUP_DEVICE | PGH_SCALAR_DOFS */
    double  kappa_matrix, kappa_fiber;  /* PGH_SCALAR_DOFS: conductivities      */
} pgh_rve_params;

/* Stop placing inclusions when the rejection budget is spent and keep those
 * already placed (at least one), instead of failing. */
#define PGH_PLACE_BEST_EFFORT 1
/* Centre each inclusion at a random interior mesh node, radius < h / 2: only
 * the elements around that node are cut, each with one node inside (in 3D the
 * interface is then made of CohTri3 triangles only). */
#define PGH_PLACE_AT_NODES 2
/* Build the pattern stage (node incidence, node-pair pattern, 4-block AIJ
 * pattern, scatter indices) on the GPU with radix sorts of packed keys instead
 * of on the host; the result is identical.  For a batch, the flag is read from
 * params[0]. */
#define PGH_PATTERN_DEVICE 4
/* Run node detection (Mesh::detect_node, src/Mesh.cpp:2883-2930), element
 * analysis and child topology (Mesh::analyze_elements, src/Mesh.cpp:2962-3071)
 * and enrichment (Mesh::add_enrichments, src/Mesh.cpp:3104-3251: node
 * positions by the edge intersection, cut records, ISV offsets) on the GPU;
 * the result is identical to the host build, array for array, and so are the
 * error codes and messages.  Mesh generation and inclusion placement (the
 * synthetic input) stay on the host.  For a batch, each instance reads its own
 * flag. */
#define PGH_SETUP_DEVICE 8
/* A scalar field (heat conduction, ProblemLinearThermal): one dof per node,
 * LinearThermalMaterial materials (kappa_matrix, kappa_fiber), T prescribed on
 * the LEFT (0) and RIGHT (strain) sides, perfect bond only (no cohesive law).
 * The description then has nndof = 1 and is for pg_therm_*. */
#define PGH_SCALAR_DOFS 16

typedef struct pgh_problem pgh_problem;

/* Build one RVE.  Returns PG_OK or a negative code (pg_asm_last_error-style
 * message in pgh_last_error). */
int  pgh_build_rve(const pgh_rve_params* p, pgh_problem** out);
int  pgh_get_desc(const pgh_problem* h, pg_problem_desc* out);
/* Pattern and scatter indices for a mesh that was set up elsewhere (an
 * integration inside PIGFEM): reads the mesh part of `mesh` (counts, coords,
 * conn, elem_mat, cut_of_elem, cut_elem, cut, mats, node_dof, n_free, n_pres,
 * n_cqp, cohesive) and builds the 4-block AIJ pattern (union of full element
 * dof blocks, as set_matrix inserts them, src/Assembler.cpp:659-679), the
 * node-pair pattern, the scatter indices and the node incidence.  The full
 * description is then read with pgh_get_desc; the input arrays are copied. */
int  pgh_build_pattern(const pg_problem_desc* mesh, pgh_problem** out);
/* The same with flags (PGH_PATTERN_DEVICE; any other bit is rejected). */
int  pgh_build_pattern_flags(const pg_problem_desc* mesh, int32_t flags, pgh_problem** out);
void pgh_free(pgh_problem* h);
const char* pgh_last_error(void);

/* A batch of independent 2D RVEs (one parameter set each, same materials)
 * merged into one block-diagonal problem: instance i owns a contiguous range
 * of elements, regular nodes, enrichment nodes, free and prescribed dofs (so
 * contiguous rows and columns of all four K blocks), cohesive qps and cut
 * elements.  Instances are generated in parallel host threads. */
int  pgh_build_batch(const pgh_rve_params* params, int32_t n, pgh_problem** out);
/* Per-instance first indices, PGH_BATCH_COLS columns per row, n_inst + 1 rows
 * (the last row holds the totals): element, regular node, enrichment node,
 * free dof, prescribed dof, cohesive qp, cut element, inclusion.  n_inst is 0
 * for a problem that is not a batch. */
#define PGH_BATCH_COLS 8
int  pgh_batch_info(const pgh_problem* h, int32_t* n_inst, const int64_t** offsets);

/* Inclusions actually placed: centers (dim each) and radii. */
int  pgh_inclusions(const pgh_problem* h, const double** centers,
                    const double** radii, int32_t* n);
/* Inclusion of each enrichment node (n_enr values, mirrors included: the
 * inclusion of the cut-edge slot it was made for, the reference's
 * Elem::get_inclusion_from_enrich); NULL for a pattern built from an external
 * mesh, whose inclusions the harness does not know. */
const int32_t* pgh_enrich_inclusion(const pgh_problem* h);
/* The cut edge of each enrichment node (n_enr pairs of regular node ids,
 * ascending, mirrors included) and its fraction alpha along that edge: the
 * node sits at x_0 + alpha (x_1 - x_0) (Mesh::_enrich_node_to_edge and
 * getEnrichmentEdgeDistance, src/Mesh.cpp:3178-3217,3723-3735).  PG_E_ARG for
 * a pattern built from an external mesh. */
int pgh_enrich_edges(const pgh_problem* h, const int32_t** edges, const double** alpha);
/* Nodal detection value per regular node (-1 matrix, else inclusion index). */
const int32_t* pgh_node_detect(const pgh_problem* h);
/* Prescribed value at full load of each prescribed dof (length n_pres); the
 * load-step value is time_fraction * value (src/Assembler.cpp:394-423). */
const double* pgh_bc_values(const pgh_problem* h);
/* Wall seconds of the set-up stages, PGH_N_STAGES values: mesh, placement,
 * detection, materials + cut analysis + enrichment, dof numbering, pattern,
 * and (batches only, where stages 0-4 stay 0) the parallel per-instance build
 * plus the merge.  n must be at least PGH_N_STAGES. */
#define PGH_N_STAGES 7
int  pgh_stage_seconds(const pgh_problem* h, double* out, int32_t n);
/* Per element: number of volumetric + cohesive quadrature points. */
int  pgh_elem_qps(const pgh_problem* h, int32_t e, int32_t* n_vqp, int32_t* n_cqp);

#ifdef __cplusplus
}
#endif

#endif /* PGFEM_HARNESS_H */
