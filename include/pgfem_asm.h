/*
 * pgfem_asm.h -- C ABI of the MI355X element-assembly library.
 *
 * This header is the drop-in boundary for the tangent (K) and internal-force
 * (R = P_int) assembly of PIGFEM's AssemblerSmallStrainStructural.  Every entry
 * point takes plain pointers and sizes; nothing in a signature is a torch or
 * HIP type.  All functions return 0 on success and a negative PG_E* code on
 * failure; pg_asm_last_error() then holds a message (thread-local).
 *
 * Reference interface replaced (file:line in brandyb2/PIGFEM):
 *   Assembler::assemble(pmatrix& K, pvector& P_int, double delta_t,
 *       vector<vector<vector<double>>>& update_ISVs, NodalData* solution,
 *       bool assembleFunc, bool assembleJac)      include/Assembler.h:45-47,
 *                                                 src/Assembler.cpp:65-146
 *   Assembler::storeBmats / Mesh::store_shape_functions (set-up that the
 *       GPU path replaces by on-the-fly recomputation)
 *                                                 src/Assembler.cpp:602-621,
 *                                                 src/Mesh.cpp:1475-1643
 *   Assembler::set_matrix / set_vec (the scatter) src/Assembler.cpp:659-701
 */
#ifndef PGFEM_ASM_H
#define PGFEM_ASM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------------------------
 * Status codes
 * ------------------------------------------------------------------------- */
#define PG_OK            0
#define PG_E_ARG        -1   /* bad argument / inconsistent description      */
#define PG_E_HIP        -2   /* HIP runtime error                            */
#define PG_E_UNSUPPORTED -3  /* element class or flag combination unsupported */
#define PG_E_MESH       -4   /* invalid mesh (negative Jacobian, ...)         */
#define PG_E_NOMEM      -5

/* ----------------------------------------------------------------------------
 * Flags of pg_asm_assemble.  PG_FUNC / PG_JAC mirror assembleFunc / assembleJac
 * (src/Assembler.cpp:69-70); both clear is a no-op, as in the reference.
 * ------------------------------------------------------------------------- */
#define PG_FUNC     1u   /* assemble the internal-force vector P_int        */
#define PG_JAC      2u   /* assemble the tangent K (4 blocks)               */
#define PG_FP32     4u   /* mixed precision: FP32 B^T.D.B / cohesive tangent,
                            FP64 sliver Jacobians, damage update, CSR sums, P */
#define PG_BATCHED  8u   /* class-binned schedule (cut-element waves +
                            node-pair recompute); implied by PG_FP32         */
#define PG_CUT64   16u   /* with PG_FP32: cut elements (children, enrichment,
                            cohesive tangents) stay FP64 end to end; FP32 is
                            used only for uncut elements                     */
/* Any other bit in `flags` is rejected with PG_E_ARG. */

/* ----------------------------------------------------------------------------
 * Materials.  Reference: material_lei.cpp:52-97 (LEI D matrix),
 * Material_BilinearCohesive.cpp:38-125, Material_BilinearCohesiveNU.cpp:48-138,
 * Material_OPCohesive.cpp:38-98, Material_OPCohesiveNU.cpp:43-111.
 * ------------------------------------------------------------------------- */
#define PG_MAT_LEI        0  /* linear elastic isotropic (D precomputed)     */
#define PG_MAT_BL         1  /* BilinearCohesive (delta_max read, never set) */
#define PG_MAT_BL_NU      2  /* BilinearCohesiveNU                           */
#define PG_MAT_OP         3  /* OPCohesive (delta_max history)               */
#define PG_MAT_OP_NU      4  /* OPCohesiveNU                                 */
#define PG_MAT_LT         5  /* LinearThermalMaterial (material_lt.cpp:36-60):
                                D = the dim x dim conductivity (kappa I),
                                row-major with stride dim; nstr = dim        */

typedef struct pg_material {
    int32_t type;       /* PG_MAT_*                                           */
    int32_t nstr;       /* LEI: rows of D (3 in 2D plane strain, 6 in 3D)     */
    double  D[36];      /* LEI: row-major nstr x nstr constitutive matrix     */
    double  sigma_c;    /* cohesive: critical traction                        */
    double  delta_c;    /* OP: critical opening                               */
    double  delta_c1;   /* BL: opening at peak traction                       */
    double  delta_c2;   /* BL: opening at full separation                     */
    double  beta2;      /* cohesive: beta^2 (tangential weight)               */
    double  alpha;      /* BL: contact penalty sharpness                      */
    double  thermal_exp;     /* LEI: thermal expansion ratio (set_parameter
                                "THERMAL_EXPANSION", material_lei.cpp:384,690)  */
    int32_t thermal_exp_set; /* LEI: nonzero once thermal_exp is set; only then
                                does a temperature change strain the material
                                (material_lei.cpp:272-291)                      */
    int32_t pad1;
} pg_material;

/* ----------------------------------------------------------------------------
 * Cut (IGFEM-intersected) element record.  Element-local node numbering follows
 * the reference: 0..nen-1 parent nodes, then enrichment nodes in cut-edge order,
 * then (cohesive meshes) their mirror nodes (src/tri3.cpp:287-596,
 * src/tet4.cpp:228-520, src/elem.cpp:428-559).  A child is an integration
 * sub-element listed by local node index: tri3, quad4 or a 5- / 6-node polygon
 * (a tri3 cut by two or three inclusions, src/tri3.cpp:376-452,
 * src/Polygon.cpp:51-228) in 2D, tet4 in 3D.  Child k's nodes start at
 * child_node[PG_CHILD_STRIDE(dim) * k]: at most 4 children of at most 6 nodes
 * in 2D, at most 8 tet4 children in 3D.  A cohesive child is a CohEdge2
 * (4 local nodes: 2 lower then 2 upper), CohTri3 (6) or CohQuad4 (8: the
 * interface of a tet4 cut two nodes against two, src/CohQuad4.cpp).
 * ------------------------------------------------------------------------- */
#define PG_CHILD_TRI3   0
#define PG_CHILD_QUAD4  1
#define PG_CHILD_TET4   2
#define PG_CHILD_POLY5  3   /* pentagon: Wachspress shapes, 5 x 4 qps */
#define PG_CHILD_POLY6  4   /* hexagon: 6 x 4 qps */
#define PG_MAX_ENR     12
#define PG_MAX_CHILD    8
#define PG_MAX_COH      3
#define PG_CHILD_STRIDE(dim) ((dim) == 2 ? 8 : 4)
/* nodes and volumetric qps of a child type.  A polygon is integrated with the
 * tri3 order-3 rule on each of its sub-triangles: the reference calls
 * n_q_points() / q_point() through Elem*, whose default order (3) applies,
 * not Polygon's own default of 5 (include/elem.h:66-67, include/Polygon.h:42-43). */
#define PG_CHILD_NN(t)  ((t) == PG_CHILD_TRI3 ? 3 : (t) == PG_CHILD_POLY5 ? 5 : (t) == PG_CHILD_POLY6 ? 6 : 4)
#define PG_CHILD_NQP(t) ((t) == PG_CHILD_TET4 ? 5 : (t) == PG_CHILD_POLY5 ? 20 : (t) == PG_CHILD_POLY6 ? 24 : 4)

typedef struct pg_cut_rec {
    int32_t enr[PG_MAX_ENR];        /* global node index of each local enrichment node */
    int32_t isv_off;                /* first cohesive qp of this element in the ISV array */
    uint8_t n_enr;                  /* local enrichment nodes (incl. mirrors)           */
    uint8_t n_child;
    uint8_t n_coh;
    uint8_t n_vqp;                  /* volumetric qps (sum over children)               */
    uint8_t child_type[PG_MAX_CHILD];
    uint8_t child_mat[PG_MAX_CHILD];
    uint8_t child_node[32];         /* child k's nodes from PG_CHILD_STRIDE(dim) * k    */
    uint8_t coh_nodes[PG_MAX_COH][8];
    uint8_t coh_mat[PG_MAX_COH];
    uint8_t coh_nn;                 /* nodes per cohesive child: 4 (CohEdge2), 6 (CohTri3)
                                       or 8 (CohQuad4) */
    uint8_t pad[4];
} pg_cut_rec;                       /* 136 bytes */

/* ----------------------------------------------------------------------------
 * Static problem description (set-up output of the host harness; the
 * reference's Mesh + DofObject + Solver::preallocate_matrix state).
 *
 *   nodes: n_nodes regular nodes, then n_enr enrichment nodes; global node index
 *          n_nodes + k is enrichment node k (reference id ENRICH_START + k).
 *   node_dof[(node)*dim + j]: global dof; free dofs are < n_free
 *          (src/DofObject.cpp:241-278, 498-538).
 *   K blocks: CSR of ff (n_free x n_free), fp (n_free x n_pres),
 *          pf (n_pres x n_free), pp (n_pres x n_pres) with block-local column
 *          indices (src/Assembler.cpp:659-679).  The library stores the values
 *          of the four blocks back to back: [ff | fp | pf | pp].
 *   P: [Pf (n_free) | Pp (n_pres)].
 *   node-pair pattern: npat_ptr / npat_col list, per node a, the nodes b that
 *          share an element with a (sorted); npat_map holds, per pair and per
 *          (i,j) in dim x dim, the index into the concatenated K values of entry
 *          (node_dof[a*dim+i], node_dof[b*dim+j])  -- the scatter indices.
 *   n2e_ptr / n2e: elements incident to each node, ascending (summation order
 *          of the reference element loop, src/Assembler.cpp:100).
 * ------------------------------------------------------------------------- */
typedef struct pg_problem_desc {
    int32_t dim;            /* 2 or 3 */
    int32_t nen;            /* parent nodes per element: 3 (tri3) or 4 (tet4) */
    int32_t n_nodes;
    int32_t n_enr;
    int32_t n_elem;
    int32_t n_cut;
    int32_t n_mat;
    int32_t n_cqp;          /* cohesive quadrature points (ISV entries) */
    int32_t n_free;
    int32_t n_pres;
    int32_t cohesive;       /* mesh has cohesive interfaces */
    int32_t nndof;          /* dofs per node: 0 (= dim, the structural assemblers
                               and pg_asm_*) or 1 (a scalar field, the thermal
                               assembler pg_therm_*); node_dof then holds nndof
                               entries per node and npat_map nndof^2 per pair */
    int64_t nnz[4];         /* ff, fp, pf, pp */
    int64_t n_pairs;        /* length of npat_col */
    const double*   coords;     /* (n_nodes+n_enr)*dim */
    const int32_t*  conn;       /* n_elem*nen */
    const int32_t*  elem_mat;   /* n_elem */
    const int32_t*  cut_of_elem;/* n_elem: -1 or index into cut[] */
    const int32_t*  cut_elem;   /* n_cut: element id of each cut record */
    const pg_cut_rec* cut;      /* n_cut */
    const pg_material* mats;    /* n_mat */
    const int32_t*  node_dof;   /* (n_nodes+n_enr)*dim */
    const int32_t*  rp[4];      /* block row pointers */
    const int32_t*  ci[4];      /* block column indices */
    const int32_t*  npat_ptr;   /* n_nodes+n_enr+1 */
    const int32_t*  npat_col;   /* n_pairs */
    const int32_t*  npat_row;   /* n_pairs: row node of each pair */
    const int32_t*  npat_map;   /* n_pairs*dim*dim (total nnz < 2^31) */
    const int32_t*  n2e_ptr;    /* n_nodes+n_enr+1 */
    const int32_t*  n2e;        /* incidence list */
} pg_problem_desc;

/* ----------------------------------------------------------------------------
 * Assembler handle.  Owns all device buffers of one problem, one HIP stream.
 * One caller thread per handle.
 * ------------------------------------------------------------------------- */
typedef struct pg_asm pg_asm;

/* Upload the static description to HBM and build the per-class schedules.
 * Replaces Problem::init's storeBmats/store_shape_functions + preallocation. */
int pg_asm_create(const pg_problem_desc* desc, pg_asm** out);
int pg_asm_destroy(pg_asm* h);
const char* pg_asm_last_error(void);

/* Host-buffer drop-in for Assembler::assemble (src/Assembler.cpp:65-146).
 *   U_all         global-dof-ordered solution, length n_free+n_pres
 *   isv_committed committed cohesive history (delta_max), length n_cqp (may be
 *                 NULL when n_cqp == 0)
 *   isv_trial     out: trial history after this assembly (reset from committed
 *                 first, as InternalVars::copy_data does, src/Assembler.cpp:92)
 *   Kff..Kpp      out: CSR value arrays of the four blocks of pmatrix
 *                 (include/common.h:32-73), in the order of rp/ci of the
 *                 description -- what MatSeqAIJGetArray returns for K.ff,
 *                 K.fp, K.pf, K.pp (NULL unless PG_JAC; a block with nnz 0
 *                 may be NULL)
 *   Pf, Pp        out: P_int.f and P_int.p of pvector (include/common.h:74-101),
 *                 lengths n_free and n_pres (NULL unless PG_FUNC)
 * Blocks are overwritten, not added to: Assembler::assemble zeroes K and P_int
 * itself before its element loop (src/Assembler.cpp:80-89).
 * Copies host->device, assembles, copies device->host, synchronises. */
int pg_asm_assemble(pg_asm* h, const double* U_all, const double* isv_committed,
                    double* isv_trial, double delta_t, unsigned flags,
                    double* Kff, double* Kfp, double* Kpf, double* Kpp,
                    double* Pf, double* Pp);

/* Device-resident variant: every pointer is device memory (HBM).  dK holds the
 * four blocks back to back, [ff | fp | pf | pp] (nnz_total values, block b at
 * the sum of the nnz of the blocks before it), dP holds [Pf | Pp] (n_dof
 * values).  The launch is asynchronous on `stream` (a hipStream_t, or NULL for
 * the handle's own stream).  No host synchronisation, no allocation.
 * The negative-child-Jacobian check of the reference (store_shape_functions
 * aborts, src/Mesh.cpp:1561) lands in a device flag that each assembly clears
 * first; pg_asm_check_error reports it.
 * Calls on one handle share its scratch buffers: a call on a stream other
 * than the previous call's waits for the previous call to complete. */
int pg_asm_assemble_device(pg_asm* h, const double* dU_all,
                           const double* disv_committed, double* disv_trial,
                           double delta_t, unsigned flags,
                           double* dK, double* dP, void* stream);

/* Thermal pre-load (AssemblerSmallStrainStructural_ThermalPreLoad,
 * src/AssemblerSmallStrainStructural_ThermalPreLoad.cpp:40-98,213-227): the
 * temperature change that the following assemblies of this handle apply, i.e.
 * the assembler's current delta T that its assemble_new_load_step sets (the
 * full delta T once relaxation is complete, else delta T times the load-step
 * time fraction) and KernelVolumetric hands to the material as
 * input.temp_change.  Every LEI material whose thermal_exp_set is nonzero then
 * adds the thermal strain -temp_change * thermal_exp to its normal strains
 * before its stress (src/material_lei.cpp:272-291): P changes, K does not.
 * 0 (the default) is the plain structural assembler.  Applies to every
 * schedule and to calls made after it returns; PG_E_ARG if not finite. */
int pg_asm_set_temperature_change(pg_asm* h, double temp_change);

/* ----------------------------------------------------------------------------
 * Output fields of the VTK writers from device data: what
 * Writer_VTK::gather_nodal_data and gather_element_data compute
 * (src/Writer_VTK.cpp:642-1000) before Writer_ASCII_VTK / Writer_Binary_VTK
 * print them (the host writer in pigfem_amd/vtk.py keeps their formats).
 *   pg_asm_plot_cells   number of VTK cells: every uncut element, and every
 *                       integration child of a cut element (store_elements,
 *                       src/Writer_VTK.cpp:221-312; cohesive cells, the
 *                       writer's "COHESIVE" option, are not produced)
 *   pg_asm_plot_fields  dnodes: per node (regular, then enrichment) its
 *                       displacement, an enrichment node's being
 *                       U(n1)(1 - alpha) + U(n2) alpha + U_enr
 *                       (Problem::interpolate_nodal_enrichment,
 *                       src/Problem.cpp:538-570) with denr_edge (n_enr pairs
 *                       n1, n2) and denr_alpha (n_enr) from the set-up
 *                       (pgh_enrich_edges); dcells: per cell, in store_elements
 *                       order, [volume | integral of the strain (nstr) |
 *                       integral of the stress (nstr)], each the sum over the
 *                       cell's qps of (J w) times the volume functor, the
 *                       strain functor (SmallStrainFast) or the stress
 *                       functor (the material's stress at that strain,
 *                       include/AssemblerSmallStrainStructural.h:23-79,
 *                       include/Integrator.h:151-228); the writer divides by
 *                       the volume.  U: global-dof order.  Device pointers;
 *                       asynchronous on stream (NULL: the handle's), ordered
 *                       after the handle's previous call.  The first call
 *                       allocates the cell list (n_cells int2 + int64 + rows). */
int pg_asm_plot_cells(const pg_asm* h, int64_t* n_cells);
int pg_asm_plot_fields(pg_asm* h, const double* dU, const int32_t* denr_edge, const double* denr_alpha,
                       double* dnodes, double* dcells, void* stream);

/* ----------------------------------------------------------------------------
 * Heat conduction: AssemblerThermal (include/AssemblerThermal.h:16-45,
 * src/AssemblerThermal.cpp:16-107), the sibling plug-in of the structural
 * assembler behind the same Assembler::assemble.  One dof per node
 * (desc->nndof == 1, ProblemLinearThermal::nndof, src/ProblemLinearThermal.cpp:24-28),
 * the same mesh, cut records and IGFEM enrichment; PG_MAT_LT materials.  Per qp
 * B = the shape gradients (dim x nodes, Assemble_Thermal_B_Mat,
 * src/Utilities.cpp:147-160), K_el += (J w) (B^T D) B and P_el += (J w) B^T q,
 * summed in the reference's order (Assembler::assemble_elem, FP64).
 * LinearThermalMaterial::Constitutive returns D but never writes the flux
 * (src/material_lt.cpp:51-60), so the reference's Bt_F reads an empty vector;
 * here q = D . (B T), the flux of the linear law, which makes P = K T.
 * Cohesive meshes are refused (PG_E_UNSUPPORTED): AssemblerThermal::KernelCohesive
 * returns a K_coh smaller than the fold that reads it (:90-99).
 *   pg_therm_create      upload the description (nndof == 1) to HBM
 *   pg_therm_assemble_device  T (global-dof order), K [ff | fp | pf | pp], P
 *                        [f | p], device pointers; PG_FUNC / PG_JAC as in
 *                        pg_asm_assemble; asynchronous on stream (NULL: the
 *                        handle's own), no allocation
 *   pg_therm_assemble    the same on host buffers (K blocks, Pf, Pp as in
 *                        pg_asm_assemble); synchronous
 * Errors: pg_asm_last_error().  Results are bitwise repeatable. */
typedef struct pg_therm pg_therm;
int pg_therm_create(const pg_problem_desc* desc, pg_therm** out);
int pg_therm_assemble_device(pg_therm* h, const double* dT_all, unsigned flags, double* dK, double* dP,
                             void* stream);
int pg_therm_assemble(pg_therm* h, const double* T_all, unsigned flags, double* Kff, double* Kfp, double* Kpf,
                      double* Kpp, double* Pf, double* Pp);
int pg_therm_destroy(pg_therm* h);

/* ----------------------------------------------------------------------------
 * Material-parameter sensitivity: the partial derivative of the internal force
 * with respect to one material parameter d, dP/dd, replacing
 * SensitivityParameter::assemble_dPdd (src/SensitivityParameter.cpp:108-151)
 * with SensitivityMaterialParameter::assemble_elem_dPdd
 * (src/SensitivityMaterialParameter.cpp:54-264) and the structural kernels of
 * SensitivityMaterialParameterStructural (src/SensitivityMaterialParameterStructural.cpp:17-121).
 *
 *   kind PG_SENS_LEI: a parameter of LEI material `mat` (E, lambda, mu, K or
 *       nu); dD holds dD/dd (row-major nstr x nstr, the matrix of
 *       LinearElasticIsotropicMaterial::computeDmatSensitivities,
 *       src/material_lei.cpp:106-210, which the caller computes from E and nu).
 *       Each qp of that material adds (J w) B^T (dD . B U) (SensitivityConstitutive,
 *       src/material_lei.cpp:300-350); no thermal strain enters.
 *   kind PG_SENS_SIGMA_C / PG_SENS_DELTA_C: sigma_c or delta_c of the
 *       Ortiz-Pandolfi law `mat` (PG_MAT_OP or PG_MAT_OP_NU): each cohesive qp adds
 *       the fold of (J w) -+N (R^T dt), dt the traction derivative of
 *       OPCohesiveMaterial / OPCohesiveNUMaterial::SensitivityConstitutive
 *       (src/Material_OPCohesive.cpp:104-180, src/Material_OPCohesiveNU.cpp:117-170).
 * Cohesive qps of a PG_MAT_OP law that is not `mat` still add the
 * history term of OPCohesiveMaterial (unloading: d delta_max / dd times the
 * traction's derivative in delta_max), as the reference visits every element
 * whose materials hold internal variables.
 *
 * As in the reference (src/SensitivityMaterialParameter.cpp:138-150), the qps of
 * a cut element's child k take the B matrix of the element's qp numbered by the
 * child-local qp index (the first children's B) while J and w are their own;
 * flag PG_SENS_OWN_CHILD_B gives every child qp its own B instead (the
 * integral the reference's formula states).
 *
 *   dU_all      global-dof-ordered solution (device)
 *   disv        delta_max per cohesive qp (the committed internal variables,
 *               device; may be NULL without PG_MAT_OP cohesive children)
 *   disv_sens   d delta_max / dd per cohesive qp for this parameter (the
 *               sensitivity internal variables, device; same rule)
 *   ddPdd       out: [dP_f | dP_p], n_free + n_pres values, overwritten
 * Asynchronous on stream (NULL: the handle's stream), ordered after the
 * handle's previous call; uses the FP64 drop-in's staging buffers.
 * PG_E_UNSUPPORTED where the reference stops with an error: a bilinear cohesive
 * law on the mesh (its sensitivity is not implemented,
 * src/Material_BilinearCohesive.cpp:131-134), a cohesive kind on a non-OP
 * material, PG_SENS_LEI on a non-LEI one. */
#define PG_SENS_LEI          0
#define PG_SENS_SIGMA_C      1
#define PG_SENS_DELTA_C      2
#define PG_SENS_OWN_CHILD_B  1u
typedef struct pg_sens_param {
    int32_t mat;      /* index of the material differentiated against */
    int32_t kind;     /* PG_SENS_* */
    double  dD[36];   /* PG_SENS_LEI: dD/dd, row-major nstr x nstr */
} pg_sens_param;
int pg_asm_sensitivity_dPdd(pg_asm* h, const pg_sens_param* prm, unsigned flags, const double* dU_all,
                            const double* disv, const double* disv_sens, double* ddPdd, void* stream);

/* Status of the handle's last assembly (either API): waits for it to complete
 * and returns PG_E_MESH if a cut element had a child of negative Jacobian (the
 * reference aborts there, src/Mesh.cpp:1561), PG_OK otherwise.  The host API
 * pg_asm_assemble makes this check itself. */
int pg_asm_check_error(pg_asm* h);

/* Sizes a caller needs for buffers of the calls above. */
int pg_asm_sizes(const pg_asm* h, int64_t* n_dof, int64_t* nnz_total,
                 int64_t* n_cqp);

/* Timing of the last assembly while pg_asm_set_timing is on, in ms, measured
 * with HIP events on the streams the kernels run on:
 *   kernel_ms[0] the whole assembly,
 *   kernel_ms[1] U permutation + element (plain) or cut-element (binned) kernels,
 *   kernel_ms[2] the node-pair kernels ([3] + [4]),
 *   kernel_ms[3] binned: the node groups without cut incidences, which run on
 *                the handle's second stream alongside the cut kernels and the
 *                remaining groups (0 if none),
 *   kernel_ms[4] the remaining node groups, which start when the cut kernels
 *                end (plain: the staged node-pair kernel).
 * With the second stream, [3] and [4] overlap in time and [2] is their sum. */
int pg_asm_last_timing(const pg_asm* h, double kernel_ms[5]);
int pg_asm_set_timing(pg_asm* h, int enable);

/* Schedule parameters of the class-binned schedules, chosen by the autotuner:
 *   group_nodes  nodes per block of the node-group kernel: 8, 16 or 32 in 2D,
 *                2, 4 or 8 in 3D (0: default 32 / 8);
 *   cut_variant  small-class cut kernel: 0 = 4 elements per wavefront,
 *                1 = 4 per wavefront with a 3-waves/SIMD register cap,
 *                2 = 2 per wavefront, 3 = 8 per wavefront;
 *   group_variant node-group kernel: bit 0 clear = about one thread per node
 *                pair (8 per node in 2D, 16 in 3D), set = twice as many; bit 1
 *                set = the list form of the kernel in 2D (always used in 3D);
 *                bit 2 set = no overlap (all node groups after the cut kernels
 *                on one stream; by default the groups of nodes without cut
 *                incidences run on a second stream during the cut kernels);
 *                bit 3 set = every node group takes the global-memory path
 *                (the fallback of groups that overflow the LDS tables; slow,
 *                for testing); bit 4 set = the groups of enrichment nodes
 *                take twice the threads per node pair whatever bit 0 says;
 *                bits 5-6 = 0: the groups that run after the cut kernels have
 *                group_nodes nodes too, k = 1..3: they have the k-th size of
 *                the list above (8, 16, 32 in 2D; 2, 4, 8 in 3D); bit 7 set =
 *                the nodes of a structured tri3 grid (generate_mesh's pattern,
 *                checked at create) take the generic node-group kernel instead
 *                of the one that computes their records from the grid; bits
 *                8-9 = k: each block of that kernel takes 2^k consecutive
 *                groups, the next group's records loading during the current
 *                group's work; bit 10 set = with FP32 tangents, the
 *                structured tet4 nodes evaluate each incidence's blocks
 *                instead of reading them from the grid's table of element
 *                tangents (by material and incidence; built at create when the
 *                cube grid has one spacing and its elements at most two
 *                materials).
 * Every choice gives bit-identical results, except that the FP32 tangent
 * blocks that structured tet4 nodes read from the table (bits 3, 7 and 10
 * clear) are rounded once from exact values, where the other kernels round
 * products of FP32 gradients: those K values agree to FP32 rounding, and P and
 * the ISVs stay bit-identical. */
int pg_asm_set_tiles(pg_asm* h, int group_nodes, int cut_variant, int group_variant);

/* Algorithmic bytes of one K+R assembly of this problem (the roofline
 * numerator): connectivity, coordinates, U, cut records, ISVs, CSR values +
 * column indices + row pointers, R. */
int64_t pg_asm_algorithmic_bytes(const pg_asm* h);

/* Device memory (bytes) held by the handle: the static description, the class
 * lists, the staging buffers, and the host-API buffers once pg_asm_assemble
 * has been called. */
int64_t pg_asm_device_bytes(const pg_asm* h);

/* Nodes that the class-binned schedules assemble from the structured-grid
 * formulas (tri3 meshes in the pattern of the reference's generate_mesh,
 * src/Mesh.cpp:418-441): regular nodes none of whose elements is cut and whose
 * incidences, element node order, pattern columns, scatter indices and dofs
 * pg_asm_create found equal to those formulas.  No counterpart in the
 * reference (a diagnostic of the schedule). */
int64_t pg_asm_grid_nodes(const pg_asm* h);

/* Cohesive openings of a 2D cohesive mesh from device data: the data of
 * WriterCohesiveOpenings::writeFromStream (src/WriterCohesiveOpenings.cpp:73-147)
 * before its sort.  For cohesive edge k of cut record c, at row isv_off / 2 + k
 * of out (3 doubles): the angle of the edge midpoint (its lower nodes) about
 * centers[isv_off / 2 + k] (2 doubles: the centre of the edge's inclusion, the
 * one of its first lower node's cut-edge slot, src/WriterCohesiveOpenings.cpp:128),
 * then the opening at the midpoint in the edge frame, tangential and normal,
 * each divided by the law's delta_c.  U: global-dof order.  All pointers are
 * device pointers; asynchronous on stream (0: the handle's stream).
 * Ortiz-Pandolfi laws only, as in the reference (PG_E_UNSUPPORTED otherwise);
 * n_out = rows of out = n_cqp / 2.  It uses the handle's scratch, so it is
 * ordered after the handle's previous call (as pg_asm_assemble_device is) and
 * returns once the rows are written. */
int pg_asm_cohesive_openings(pg_asm* h, const double* U, const double* centers, double* out, int64_t n_out,
                             void* stream);

/* Cohesive failure of a 2D cohesive mesh from device data: the per-edge data of
 * WriterCohesiveFailureLengths (store_mesh and writeFromStream,
 * src/WriterCohesiveFailureLengths.cpp:64-86,183-224).  For cohesive edge k of
 * cut record c, at row isv_off / 2 + k of out (2 doubles): the distance between
 * its two lower nodes, then the mean over its Gauss points of the law's
 * normalised failure (delta_eff / delta_c for PG_MAT_OP_NU, delta_eff / delta_c1
 * for PG_MAT_BL_NU, 0 for PG_MAT_OP and PG_MAT_BL, as the reference's
 * getNormalizedCohesiveFailure).  U: global-dof order.  Device pointers;
 * asynchronous on stream (0: the handle's stream), ordered after the handle's
 * previous call as pg_asm_assemble_device is.  n_out = n_cqp / 2. */
int pg_asm_cohesive_failure(pg_asm* h, const double* U, double* out, int64_t n_out, void* stream);

/* Linear solve of a Newton iteration on the device: Jacobi-preconditioned
 * conjugate gradients on an n x n CSR matrix (row_ptr n + 1, col_idx and vals
 * nnz entries, 0-based), replacing the KSP solve of SolverNonlinearKSP
 * (src/SolverNonlinearKSP.cpp:94, src/Solver.cpp:108-160) for symmetric
 * positive definite tangents.  x holds the initial guess on entry and the
 * solution on return.  Stops when |b - A x| <= rtol |b| (the recursive
 * residual) or after max_iter iterations; *iters and *rel_res report where it
 * stopped (a caller checks rel_res against rtol).  Device pointers; ordered on
 * stream (0: the null stream) and synchronous: the host reads |r| once per
 * iteration.  Allocates 5 n + 1032 doubles of workspace for the call.  Results
 * are bitwise repeatable.  PG_E_ARG if b, the matrix or x is not finite, or as
 * soon as the iteration breaks down: r.z <= 0 or p.Ap <= 0, which an
 * indefinite matrix (or a non-positive diagonal) produces. */
int pg_csr_pcg(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* vals, const double* b,
               double* x, double rtol, int max_iter, int* iters, double* rel_res, void* stream);

/* pg_csr_pcg without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * arithmetic and results (bit for bit) as pg_csr_pcg, but the workspace is a
 * fixed device array inside the library, one per GPU, held until the process
 * ends: 5 * 1048576 + 1032 doubles (about 42 MB), so n may be at most 1048576
 * rows (PG_E_ARG beyond that, before any device work).  Every call on one GPU
 * uses that GPU's array: calls that overlap on one GPU (from several threads
 * or streams) race and must be serialised by the caller.  Nothing in the
 * array needs clearing between calls (each call writes every word it reads). */
int pg_csr_pcg_noalloc(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                       const double* b, double* x, double rtol, int max_iter, int* iters, double* rel_res,
                       void* stream);

/* ----------------------------------------------------------------------------
 * Newton's linear solve on the device: K_ff x = b, replacing the KSP solve
 * (MUMPS through PETSc) of SolverNonlinearKSP::nonlinearStep
 * (src/SolverNonlinearKSP.cpp:94, src/Solver.cpp:108-160).  The solver object
 * takes the structure of K_ff (rp[0], ci[0], n_free, nnz[0] of the
 * description) once; each solve takes the values from a device buffer laid out
 * as pg_asm_assemble_device writes K ([ff | fp | pf | pp]: the first nnz[0]
 * values are K_ff), copied device to device, so K never crosses PCIe.  A
 * multicolour symmetric Gauss-Seidel preconditioner (rocALUTION) serves CG
 * (symmetric positive definite tangents), FGMRES(50) (indefinite or
 * non-symmetric ones: softening cohesive laws, the bilinear contact branch) or
 * BiCGStab; it is built at the first solve and rebuilt numerically at the
 * others.  An iterative solve: its iteration count grows with the mesh
 * (hundreds at 100 x 100 cells, thousands at 200 x 200), so it suits meshes
 * far below the largest workloads, whose direct solve stays on the host.  dx
 * holds the initial guess on entry and the solution on return; *rel_res is the
 * true |b - K_ff x| / |b|, which the caller compares with its tolerance.
 * Synchronous.  PG_E_ARG on non-finite data or divergence.  The messages of
 * these calls come from pg_ksp_last_error(). */
#define PG_KSP_CG        0
#define PG_KSP_FGMRES    1
#define PG_KSP_BICGSTAB  2
typedef struct pg_ksp pg_ksp;
int pg_ksp_create(const pg_problem_desc* desc, int method, pg_ksp** out);
int pg_ksp_solve(pg_ksp* k, const double* dK, const double* db, double* dx, double rtol, int max_iter,
                 int* iters, double* rel_res);
int pg_ksp_destroy(pg_ksp* k);
const char* pg_ksp_last_error(void);

/* ----------------------------------------------------------------------------
 * Device buffers for callers that have no HIP allocator of their own (the
 * Python host layer uses these; a PETSc/hipSPARSE caller would pass its own
 * device pointers).  No counterpart in the reference (CPU-only).  All act on
 * HIP device 0 and are synchronous except pg_dev_copy, which is ordered on the
 * null stream.
 * ------------------------------------------------------------------------- */
int pg_dev_alloc(void** ptr, int64_t bytes);
int pg_dev_free(void* ptr);
int pg_dev_upload(void* dst_device, const void* src_host, int64_t bytes);
int pg_dev_download(void* dst_host, const void* src_device, int64_t bytes);
int pg_dev_copy(void* dst_device, const void* src_device, int64_t bytes);
int pg_dev_synchronize(void);
/* y += a x on n doubles (device pointers; synchronous). */
int pg_dev_axpy(int64_t n, double a, const double* x, double* y);
/* *out = x . y over n doubles (device pointers; synchronous, fixed summation
 * order, so bitwise repeatable). */
int pg_dev_dot(int64_t n, const double* x, const double* y, double* out);

#ifdef __cplusplus
}
#endif

#endif /* PGFEM_ASM_H */
