/*
 * curvami_gpu.h -- C ABI of the MI355X (gfx950) implementation of Curvami's
 * per-Gauss-point hot path: fixed-radius neighbour search, RKPM shape
 * functions up to second derivatives with crease-cut visibility, and the
 * Kirchhoff-Love residual / block-upper CSC tangent with the elastic hinge.
 *
 * Conventions
 *  - Plain C types only; every entry point returns a cvm_status (never throws).
 *  - Host pointers unless a flag says otherwise; outputs are caller-allocated.
 *  - One HIP stream per context; a context is not thread-safe (as the
 *    reference's assembler, plate/Assembly.h:18-45, is not re-entrant).
 *    Many contexts may coexist.
 *  - Reference citations are relative to the Curvami source tree.
 */
#ifndef CURVAMI_GPU_H
#define CURVAMI_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CVM_API __attribute__((visibility("default")))

typedef struct cvm_context cvm_context;

/* Status codes. SINGULAR_MOMENT and PIVOT_COLLAPSE are the two runtime_errors
 * of GlobalRKPM::solve (core/include/curvami/GlobalRKPM.h:241-256, :347-359);
 * PATTERN_MISS is the CSR-map throw of initSparsityPattern (plate/Assembly.h:148-152). */
typedef enum {
    CVM_OK = 0,
    CVM_E_INVALID = 1,
    CVM_E_HIP = 2,
    CVM_E_SINGULAR_MOMENT = 3,
    CVM_E_PIVOT_COLLAPSE = 4,
    CVM_E_PATTERN_MISS = 5,
    CVM_E_NO_MEMORY = 6,
    CVM_E_STATE = 7,
    CVM_E_BUDGET = 8
} cvm_status;

/* Integration-domain kinds (plate/Meshless.h:14-17: faces, edges, points, fold left/right). */
typedef enum {
    CVM_DOMAIN_FACE = 0,
    CVM_DOMAIN_EDGE = 1,
    CVM_DOMAIN_POINT = 2,
    CVM_DOMAIN_FOLD_LEFT = 3,
    CVM_DOMAIN_FOLD_RIGHT = 4
} cvm_domain_kind;

/* Window functions: the reference's quartic spline (GlobalRKPM.h:15-17) and an
 * added cubic B-spline. */
typedef enum { CVM_KERNEL_QUARTIC = 0, CVM_KERNEL_CUBIC_BSPLINE = 1 } cvm_kernel;

/* cvm_assemble flags. */
enum {
    CVM_RESIDUAL = 1,        /* compute F_int */
    CVM_TANGENT = 2,         /* compute K values */
    CVM_ELASTIC_FOLDS = 4,   /* add elastic hinges (Meshless.h:118-126, i.e. flagRig == false) */
    CVM_DEVICE_PTRS = 8,     /* x, F_int and K_values are device pointers */
    CVM_EXACT_HINGE = 16,    /* add the FD dihedral-angle Hessian term to the hinge tangent
                                (Meshless::fold_exact_tangent, Meshless.h:48-50; Constraints.h:394-416) */
    CVM_FP32_TANGENT = 32    /* mixed precision: face tangent blocks computed and accumulated in FP32
                                (inexact-Newton Jacobian); shape functions, kinematics, residual and
                                hinge stay FP64; K values are returned as doubles. No reference
                                counterpart (the reference is FP64 throughout). */
};

/* ---- context ------------------------------------------------------------ */

/* Create a context on HIP device `device`. No reference counterpart: the
 * reference keeps this state in Meshless (plate/Meshless.h:11-27). */
CVM_API int cvm_create(int device, cvm_context** out);
CVM_API void cvm_destroy(cvm_context* ctx);
/* Text of the last error (the reference's exception message where one exists). */
CVM_API const char* cvm_last_error(const cvm_context* ctx);

/* Node reference coordinates (Node::X_ref, core/include/curvami/MeshfreeData.h:7-14). */
CVM_API int cvm_set_nodes(cvm_context* ctx, int nn, const double* X, const double* Y);

/* Laminate from plies, 6 doubles per ply: E1, E2, nu12, G12, angle_deg, thickness
 * (Ply, plate/PlateData.h:11-15); computes A, B, D as Constitutive::computeABD
 * (plate/Constitutive.h:10-50). */
CVM_API int cvm_set_laminate(cvm_context* ctx, int nply, const double* plies);
/* Read back A, B, D (row-major 3x3 each). */
CVM_API int cvm_get_abd(const cvm_context* ctx, double* A, double* B, double* D);

/* Natural curvature (Kinematics::naturalCurvature(), plate/Kinematics.h:20-45;
 * set from the configuration's "natural_curvature" band, main.cpp:80-97): with
 * active != 0, every face GP's elastic bending strain is E1_el = E1 +
 * kappa_bar with kappa_bar(0, 0) = sign * theta0 * exp(-((X - x0) / width)^2) /
 * (sqrt(pi) width) at its reference X (theta0 in radians; sign -1 mountain, +1
 * valley); the residual, the tangent (through the resultants) and the bending
 * energy use E1_el. The field is formed per face GP here or at
 * cvm_init_assembly. CVM_E_INVALID for width <= 0 or sign other than +-1. */
CVM_API int cvm_set_natural_curvature(cvm_context* ctx, int active, double x0, double width, double theta0,
                                      double sign);

/* Batched contexts: many independent sheets in one context, so that each stage
 * (neighbour search, shape functions, assembly) runs as one launch over all of
 * them. No reference counterpart (the reference holds one sheet per Meshless).
 * Appends a sheet's nodes (ids node0 .. node0 + nn - 1) with its support radius;
 * creases and domains added afterwards belong to this sheet until the next
 * cvm_add_sheet. Face domains of all sheets share one face domain (the first
 * one's id is returned again). Neighbours and shape functions only ever couple
 * nodes and GPs of the same sheet, so K is block diagonal over sheets and each
 * sheet's results equal those of a context holding that sheet alone. Pass rho = 0
 * to cvm_neighbours in a batched context. Not combinable with cvm_set_nodes. */
CVM_API int cvm_add_sheet(cvm_context* ctx, int nn, const double* X, const double* Y, double rho, int* sheet_id,
                          int* node0);

/* Crease polyline (Enrichment::FoldLine, core/include/curvami/Enrichment.h:10-15). */
CVM_API int cvm_add_crease(cvm_context* ctx, int nv, const double* vx, const double* vy,
                           int extend_start, int extend_end, int is_closed);

/* Removes every crease (single-sheet contexts). Shape functions must be
 * recomputed (cvm_shape_functions / _all) and the constraint layout rebuilt
 * before the next use; neighbour lists and the assembly plan stay valid.
 * The crease re-cut of progressive creasing (SetupHelper::recomputeShapeFunctions,
 * plate/SetupHelper.h:112-137, called from Progressive::run,
 * plate/ProgressiveCreasing.h:592-605). */
CVM_API int cvm_clear_creases(cvm_context* ctx);

/* Integration domain: Gauss points (GaussPointData::X_ref, weight, normal,
 * tangent; MeshfreeData.h:17-35). Normal/tangent may be NULL (zero). */
CVM_API int cvm_add_domain(cvm_context* ctx, int kind, int ng, const double* gx, const double* gy,
                           const double* w, const double* nx, const double* ny,
                           const double* tx, const double* ty, int* domain_id);

/* Elastic fold between two domains of equal size (Fold, plate/PlateData.h:72-94;
 * Meshless::addFold, plate/Meshless.h:30-46). */
CVM_API int cvm_add_fold(cvm_context* ctx, int left_domain, int right_domain, double Kt, double Ang0,
                         int* fold_id);
/* Per-fold rigid hinge (no reference counterpart: there flagRig is per call and
 * covers every fold). A rigid fold adds no hinge terms to F_int and K, even with
 * CVM_ELASTIC_FOLDS, and its angle rows theta - theta0 are constraint rows
 * (cvm_constraint_layout) whatever the per-call flag. Invalidates the plan. */
CVM_API int cvm_set_fold_rigid(cvm_context* ctx, int fold_id, int rigid);
/* Per crease-GP rest angle and activity (progressive creasing; Fold::gp_theta0,
 * gp_active, PlateData.h:90-93). NULL clears. */
CVM_API int cvm_set_fold_gp_state(cvm_context* ctx, int fold_id, int n, const double* theta0,
                                  const char* active);

/* ---- hot path ----------------------------------------------------------- */

/* Fixed-radius neighbour search of one domain against the nodes.
 * Replaces GlobalFRNN::execute (core/include/curvami/FRNN.h:23-78): same
 * neighbour sets and order, s = sqrt(dist2) / rho. */
CVM_API int cvm_neighbours(cvm_context* ctx, int domain, double rho, int64_t* npairs);
/* The neighbour search of every searched domain again, on the device, with
 * the radius and node grid of its last cvm_neighbours call: the FRNN of
 * SetupHelper::recomputeShapeFunctions (plate/SetupHelper.h:112-138), which
 * searches again before every RKPM solve. Every list is computed again and
 * compared with the resident one; a domain whose lists all match keeps them.
 * *changed = 1 if any list differs from the previous search (nodes or GPs
 * moved): such a domain takes the full search (lists, offsets, host mirrors),
 * its shape functions must be recomputed and a changed face or fold domain
 * invalidates the assembly plan.
 * The crease indicators and node sides are evaluated again, as
 * GlobalRKPM::solve evaluates them on every call (GlobalRKPM.h:86-112): here,
 * beside the verification (on the context's side stream), or by the next
 * shape-function call for a domain that took the full search. */
CVM_API int cvm_refresh_neighbours(cvm_context* ctx, int* changed);
/* Copy the lists out: off[ng+1], ids[npairs], s[npairs] (any may be NULL). */
CVM_API int cvm_get_neighbours(cvm_context* ctx, int domain, int* off, int* ids, double* s);

/* RKPM shape functions of one domain, cut by every crease added.
 * Replaces GlobalRKPM::solve (core/include/curvami/GlobalRKPM.h:19-397).
 * Returns CVM_E_SINGULAR_MOMENT / CVM_E_PIVOT_COLLAPSE with *nbad and
 * *first_bad set as the reference's messages report them. */
CVM_API int cvm_shape_functions(cvm_context* ctx, int domain, int order, int kernel,
                                int* nbad, int* first_bad);
/* Shape functions of every domain with one status read-back (the per-domain
 * launches are queued back to back): SetupHelper::recomputeShapeFunctions
 * (plate/SetupHelper.h:112-138). On failure *failed_domain names the first
 * failing domain and the status is as above. */
CVM_API int cvm_shape_functions_all(cvm_context* ctx, int order, int kernel, int* failed_domain, int* nbad,
                                    int* first_bad);
/* SetupHelper::recomputeShapeFunctions (plate/SetupHelper.h:112-138) in one
 * call: the neighbour search of every domain again (cvm_refresh_neighbours:
 * *changed as there) and the shape functions of every domain
 * (cvm_shape_functions_all: *failed_domain, *nbad, *first_bad and the status
 * as there), with the same results as the two calls. While the lists are
 * verified, the shape functions are formed from the resident lists beside
 * the verification (on the context's side stream); if any list changed they
 * are formed again from the new lists. */
CVM_API int cvm_recompute_shape_functions(cvm_context* ctx, int order, int kernel, int* changed, int* failed_domain,
                                          int* nbad, int* first_bad);
/* Copy Psi, dPsi/dX, dPsi/dY, d2Psi/dX2, d2Psi/dY2, d2Psi/dXdY (npairs each,
 * FRNN order; GaussPointData::Psi/dPsi/ddPsi). Any pointer may be NULL. */
CVM_API int cvm_get_shape_functions(cvm_context* ctx, int domain, double* psi, double* dpx, double* dpy,
                                    double* dxx, double* dyy, double* dxy);

/* Sparsity pattern and assembly plan over the face domain and the folds.
 * Replaces NonLinearAssembler::initSparsityPattern (plate/Assembly.h:59-153).
 * nnz_face: values of the face pattern; nnz_union: with the fold blocks added
 * (the pattern of Meshless::assembleSystem with elastic folds, Meshless.h:118-126). */
CVM_API int cvm_init_assembly(cvm_context* ctx, int64_t* nnz_face, int64_t* nnz_union);
/* Eigen-style CSC of the block-upper K: outer[3nn+1], inner[nnz]. */
CVM_API int cvm_get_pattern(cvm_context* ctx, int with_folds, int* outer, int* inner);
/* NonLinearAssembler::csrMapping (Assembly.h:54): one position per local pair
 * (i <= j) of face GP `gp`; returns the count in *n. */
CVM_API int cvm_csr_mapping(cvm_context* ctx, int gp, int* out, int cap, int* n);

/* Residual and tangent at displacement x (3 DOFs per node).
 * Replaces NonLinearAssembler::updateStiffness (plate/Assembly.h:158-345)
 * together with Meshless::assembleSystem / assembleSystemInternalOnly
 * (plate/Meshless.h:102-127) and Constraints::evaluateElasticFold
 * (plate/Constraints.h:341-436). F_int[3nn] is overwritten; K_values[nnz]
 * follows cvm_get_pattern(with_folds = flags & CVM_ELASTIC_FOLDS). */
CVM_API int cvm_assemble(cvm_context* ctx, const double* x, int flags, double* F_int, double* K_values);

/* Bytes of the static device scratch of cvm_assemble_noalloc, per GPU. */
#define CVM_NOALLOC_SCRATCH_BYTES (256LL << 20)

/* cvm_assemble with device pointers (x_dev, F_dev, K_dev; CVM_DEVICE_PTRS is
 * implied) that allocates nothing: the workspace (tile scratch rows, residual
 * rows, hinge records) lies in a static device array of CVM_NOALLOC_SCRATCH_BYTES
 * that the library holds on each GPU until the process ends, and the call makes
 * no hipMalloc, hipFree or other allocating call. No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). CVM_E_BUDGET, with the
 * size needed in cvm_last_error, when the sheet's workspace exceeds the array
 * (up to about 48 x 48 nodes with the FP64 tangent, 66 x 66 with the FP32
 * tangent, at a 3.75 h support).
 * Every call on one GPU shares the array, which keeps the last call's contents:
 * calls that overlap on one GPU (other contexts or streams included) race.
 * Results equal cvm_assemble's bit for bit. */
CVM_API int cvm_assemble_noalloc(cvm_context* ctx, const double* x_dev, int flags, double* F_dev, double* K_dev);

/* ---- constraints and loads (the rest of PlateProblem, plate/PlateProblem.h:52-91) ---- */

/* BoundaryCondition of a domain (plate/PlateData.h:36-50): active bit 0 u, 1 v,
 * 2 w, 3 normal rotation (values val[4], may be NULL = 0); dead force per unit
 * (force[3], may be NULL); bending moment per unit length on edge domains. */
CVM_API int cvm_set_domain_bc(cvm_context* ctx, int domain, int active, const double* val, const double* force,
                              double moment);
/* Row and triplet layout of Meshless::evaluateConstraints (plate/Meshless.h:129-195):
 * Dirichlet rows of the point domains, then of the edge domains
 * (Constraints::evaluateDirichlet, plate/Constraints.h:24-52), then per fold and
 * active crease GP three tie rows u_L - u_R and, when rigid != 0 (flagRig), the
 * angle row theta - theta0 (Constraints::evaluateFold, plate/Constraints.h:247-308).
 * Needs the shape functions of those domains; rebuild after they change. */
CVM_API int cvm_constraint_layout(cvm_context* ctx, int rigid, int64_t* nrows, int64_t* ntriplets);
/* Triplet rows and columns in the reference's push order (duplicates are summed
 * when J is formed, as Eigen's setFromTriplets does). */
CVM_API int cvm_get_constraint_pattern(cvm_context* ctx, int* rows, int* cols);
/* c(x)[nrows] and the triplet values[ntriplets] at x. flags: CVM_DEVICE_PTRS. */
CVM_API int cvm_constraints(cvm_context* ctx, const double* x, int flags, double* c, double* j_values);
/* Reference load F_ext[3nn] of Meshless::computeExternalForce (plate/Meshless.h:60-100):
 * dead forces of point, edge and fold domains and edge moments. flags: CVM_DEVICE_PTRS. */
CVM_API int cvm_external_force(cvm_context* ctx, int flags, double* F_ext);

/* Tile shape of the assembly plan (no reference counterpart; chosen by the
 * autotuner): at most tile_gps (1..16) consecutive face GPs per tile, at most
 * union_max (8..255; 0 = automatic, 64 for up to 47 neighbours per GP and 4/3 of
 * the largest neighbour count above) distinct nodes per tile, and the tile kernel grid (0: one
 * workgroup per tile; > 0: a persistent grid of that many workgroups striding
 * over the tiles). Changing tile_gps or union_max invalidates the plan
 * (call cvm_init_assembly again); the grid applies to the next launch. */
CVM_API int cvm_set_tile_shape(cvm_context* ctx, int tile_gps, int union_max, int grid);
/* out4: tile_gps in effect, union_max, grid, number of tiles (0 without a plan). */
CVM_API int cvm_get_tile_shape(cvm_context* ctx, int* out4);

/* Capture one evaluation (cvm_assemble with device pointers x, F_int, K_values)
 * into a HIP graph held by the context; cvm_graph_launch replays it on the
 * context's stream (asynchronous; cvm_synchronize waits). The graph reads and
 * writes the captured addresses, so a Newton loop updates x in place and
 * replays. No reference counterpart (one launch per Newton iteration). */
CVM_API int cvm_graph_capture(cvm_context* ctx, int flags, const double* x_dev, double* F_dev, double* K_dev);
CVM_API int cvm_graph_launch(cvm_context* ctx);

/* Wait for the context's stream. */
CVM_API int cvm_synchronize(cvm_context* ctx);

/* ---- serving ------------------------------------------------------------- */
/* No counterpart in the reference, which evaluates one system at a time from
 * its Newton loops. Callers: request loops that evaluate many residual/tangent
 * requests (landscape sweeps, arc-length probes) on many sheets at once.
 *
 * A server owns max_inflight slots (HIP stream + workspace). Sheets are prepared
 * contexts (shape functions and assembly plan resident in HBM) registered with
 * the server while no request is pending or in flight (CVM_E_STATE otherwise);
 * they must not be modified while registered. The slots' workspaces are
 * allocated at the first request after a registration, within hbm_budget: the
 * first slots hold every registered sheet, later ones the smaller sheets that
 * the budget left still covers; evaluations never allocate device memory.
 * Requests are admitted in arrival order as soon as a free slot holds them;
 * admitted requests run concurrently. A request that no slot holds is refused
 * at submit with CVM_E_BUDGET. A request's result does not depend on what else
 * is in flight. */
typedef struct cvm_server cvm_server;
CVM_API int cvm_server_create(int device, int64_t hbm_budget, int max_inflight, cvm_server** out);
CVM_API void cvm_server_destroy(cvm_server* s);
CVM_API const char* cvm_server_last_error(cvm_server* s);
CVM_API int cvm_server_add_sheet(cvm_server* s, cvm_context* sheet, int* sheet_id);
/* Workspace bytes a request of this sheet and flags is charged (with_K: the
 * server holds K rather than a caller device buffer). */
CVM_API int cvm_server_request_bytes(cvm_server* s, int sheet, int flags, int with_K, int64_t* bytes);
/* Asynchronous evaluation at host x[3nn] (copied at submit). F_host[3nn] is
 * written when the request completes (keep it alive until cvm_server_wait);
 * K_dev: caller device buffer for the K values, or NULL to leave K in the slot.
 * flags as cvm_assemble (CVM_DEVICE_PTRS not allowed). CVM_E_BUDGET if one
 * request alone exceeds the budget. */
CVM_API int cvm_server_submit(cvm_server* s, int sheet, const double* x, int flags, double* F_host, double* K_dev,
                              int64_t* request_id);
CVM_API int cvm_server_poll(cvm_server* s, int64_t request_id, int* done);
/* Wait for a request; times[3] (may be NULL): queue ms, device ms, total ms.
 * The wait that returns a request forgets it: another wait on the same id
 * (later, or in another thread at the same time) returns CVM_E_INVALID. */
CVM_API int cvm_server_wait(cvm_server* s, int64_t request_id, double* times);
/* completed, in flight, pending, workspace bytes allocated */
CVM_API int cvm_server_stats(cvm_server* s, int64_t* out4);
/* Requests admitted in the same dispatcher pass share one tile launch (the
 * Gauss-point tiles of all their sheets in one kernel over a job table) when
 * they agree in tangent type, laminate class and neighbour width; on by
 * default. Results are the same bits either way. */
CVM_API int cvm_server_set_batching(cvm_server* s, int on);
/* out2 = shared tile launches so far, requests evaluated in them. */
CVM_API int cvm_server_batching_stats(cvm_server* s, int64_t* out2);

/* ---- energies and fold angles --------------------------------------------- */

/* out3 = E_mem, E_ben over the face GPs and E_hin over every fold's active
 * crease GPs at state x (Meshless::computeStrainEnergies,
 * computeTotalHingeEnergy, plate/Meshless.h:198-219; Constraints::computeHingeEnergy,
 * plate/Constraints.h:310-339). flags: CVM_DEVICE_PTRS (x and out3 on the device). */
CVM_API int cvm_energies(cvm_context* ctx, const double* x, int flags, double* out3);
/* Dihedral angle theta at every crease GP of a fold, active or not
 * (Constraints::foldAngleAndGrad, plate/Constraints.h:71-88, as used by
 * Progressive::activeFoldAngles, plate/ProgressiveCreasing.h:157-179). */
CVM_API int cvm_fold_angles(cvm_context* ctx, int fold_id, const double* x, int flags, double* theta);

/* Exporter columns per face GP (Exporter::writeBinaryStep, plate/Exporter.h:34-91):
 * out[g * (7 + nfields) + c] for c = X, Y, u, v, w, rotX (deg), rotY (deg), then
 * one column per field code (0 E_XX, 1 E_YY, 2 E_XY, 3 K_XX, 4 K_YY, 5 K_XY;
 * any other code writes zeros, as the reference does for an unknown name).
 * flags: CVM_DEVICE_PTRS (x and out on the device). */
CVM_API int cvm_export_fields(cvm_context* ctx, const double* x, int flags, int nfields, const int* field_codes,
                              double* out);

/* ---- restart record (host only) ------------------------------------------ */

/* RestartManager::save / load (core/include/curvami/RestartManager.h:11-117),
 * the same byte layout. load: expected_ns <= 0 skips the DOF check; with
 * x = dU_old = NULL only *ns is returned (after the header checks);
 * scalars4 = lambda, ds, dl_old, W_ext. Messages: cvm_restart_last_error. */
CVM_API const char* cvm_restart_last_error(void);
CVM_API int cvm_restart_save(const char* path, int k, double lambda, double ds, double dl_old, double W_ext, int ns,
                             const double* x, const double* dU_old);
CVM_API int cvm_restart_load(const char* path, int expected_ns, int* k, double* scalars4, int* ns, double* x,
                             double* dU_old, int cap);

/* ---- measurement -------------------------------------------------------- */

/* Per-kernel HIP-event timing on the context's stream (on/off). */
CVM_API int cvm_set_profiling(cvm_context* ctx, int on);
/* Number of timed kernels; name, total ms and launch count of kernel i. */
CVM_API int cvm_kernel_count(cvm_context* ctx);
CVM_API int cvm_kernel_stats(cvm_context* ctx, int i, char* name, int name_cap, double* total_ms,
                             int64_t* launches);
CVM_API int cvm_reset_profiling(cvm_context* ctx);

/* ---- mesh input (host only; no device work) ------------------------------ */

/* A Gmsh 4.1 ASCII mesh: nodes in file order (z dropped) and the point (15),
 * line (1), triangle (2) and quadrangle (3) elements with their entity tags. */
typedef struct cvm_mesh cvm_mesh;
/* Message of the last failed mesh-API call on this thread. */
CVM_API const char* cvm_mesh_last_error(void);
/* Replaces GmshReader::readMesh (core/include/curvami/GmshReader.h:32-146):
 * the same node indexing (file order) and element node indices; unsupported
 * element types, undefined node tags and empty meshes are errors as there.
 * Binary files and non-4.x formats are refused; parametric node blocks are read. */
CVM_API int cvm_mesh_read(const char* path, cvm_mesh** out);
CVM_API void cvm_mesh_destroy(cvm_mesh* m);
/* out3: nodes, elements, element node entries */
CVM_API int cvm_mesh_info(const cvm_mesh* m, int64_t* out3);
CVM_API int cvm_mesh_nodes(const cvm_mesh* m, double* X, double* Y);
/* type[nel], entity_tag[nel], node_off[nel + 1], node_idx[node_off[nel]]; NULL skips */
CVM_API int cvm_mesh_elements(const cvm_mesh* m, int* type, int* entity_tag, int64_t* node_off, int* node_idx);
/* Replaces Quadrature::generate (core/include/curvami/Quadrature.h:59-233) in
 * element order: active nodes (raw indices of face-element nodes, ascending:
 * the renumbering to 0..n_active-1), face GPs (x, y, w), edge GPs (x, y, w and
 * tx, ty, nx, ny), point GPs (x, y, w = 1), each with its entity tag.
 * out4: n_active, n_face, n_edge, n_point. Any output pointer may be NULL. */
CVM_API int cvm_mesh_quadrature_sizes(const cvm_mesh* m, int line_degree, int face_degree, int64_t* out4);
CVM_API int cvm_mesh_quadrature(const cvm_mesh* m, int line_degree, int face_degree, int* active,
                                double* face_xyw, int* face_tag, double* edge_xyw, double* edge_tn,
                                int* edge_tag, double* point_xyw, int* point_tag);
/* Replaces MeshUtils::extractCurve (core/include/curvami/MeshUtils.h:57-134):
 * the chain of the line elements with the given entity tags, from the lowest
 * end node (or the lowest node of a loop), a closed chain repeating its first
 * vertex; flags bit 0 closed, bit 1 start on the convex hull of all nodes
 * (extend start), bit 2 end on it; dropped = nodes not on the chain (branching
 * or disconnected tag sets). With vx = vy = NULL only *nvert is returned. */
CVM_API int cvm_mesh_extract_curve(const cvm_mesh* m, const int* tags, int ntags, int64_t cap, double* vx,
                                   double* vy, int64_t* nvert, int* flags, int64_t* dropped);
/* Replaces MeshUtils::computeSupportRadius (core/include/curvami/MeshUtils.h:144-209):
 * (order + 3) x the median (element n/2 in ascending order) over the nodes of
 * the distance to the 3rd-nearest other node; 0.1 for fewer than 4 nodes. */
CVM_API int cvm_support_radius(const double* X, const double* Y, int64_t n, int order, double* rho);

#ifdef __cplusplus
}
#endif
#endif /* CURVAMI_GPU_H */
