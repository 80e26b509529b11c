/*
 * prx.h -- C ABI of libprx, the MI355X (gfx950) screened-Poisson solve path.
 *
 * The entry points replace the public API of the reference crate
 * (Avanatiker/PoissonReconstruction); each declaration names the reference
 * item it stands in for.  All pointers are plain host pointers unless the
 * name ends in _device; no torch or HIP types appear in the signatures.
 * Every function returns a status code (PRX_OK = 0) and never unwinds; the
 * text of the last error on the calling thread is prx_last_error().
 *
 * Vector layout: every fine-level vector (solution, RHS, weights) is in the
 * reference's FEM node index order, i.e. the order of FEMTree::solution
 * (src/fem_tree.rs:59), which is the Morton code of the node offset with x in
 * bit 0.  Coarse-level vectors exported by prx_coarse_system use the
 * reference's row-major coarse order (src/fem_tree.rs:361).
 */
#ifndef PRX_H
#define PRX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
#define PRX_OK 0
#define PRX_ERR_INVALID 1       /* bad argument (depth out of range, null pointer, size) */
#define PRX_ERR_NONFINITE 2     /* NaN/Inf in the input points */
#define PRX_ERR_NOT_CONVERGED 3 /* iteration cap reached before the tolerance */
#define PRX_ERR_BREAKDOWN 4     /* CG breakdown: non-finite beta (src/solvers.rs:63) */
#define PRX_ERR_DEVICE 5        /* HIP runtime / kernel failure, or no GPU */
#define PRX_ERR_OOM 6           /* device allocation failed */
#define PRX_ERR_STATE 7         /* call order (e.g. solve before assemble) */

#define PRX_MAX_DEPTH 10 /* 8^10 nodes = 2^30: FEM indices stay 32-bit (octree.rs NodeIndex is i32) */

/* ------------------------------------------------------------ types */

/* OrientedPoint (src/fem_tree.rs:10) with Point3/Vec3 (src/geometry.rs:9-13,
 * 146-150) flattened; the Rust side marshals into this repr(C) layout. */
typedef struct prx_point {
    double px, py, pz;
    double nx, ny, nz;
} prx_point;

enum { PRX_PREC_FP64 = 0, PRX_PREC_MIXED = 1 };
/* Stopping test of the fine solve.  REFERENCE: ||r|| <= eps*||r0|| with r0
 * the residual after smoothing (src/solvers.rs:36-58).  TRUE: the FP64 true
 * residual ||b-Ax|| <= eps*||b||. */
enum { PRX_RESID_REFERENCE = 0, PRX_RESID_TRUE = 1 };
/* Fine solver.  NONE: the reference's cascadic solve (coarse CG, GS, CG).
 * MG: CG preconditioned with a multigrid V-cycle (an extension). */
enum { PRX_PRECOND_NONE = 0, PRX_PRECOND_MG = 1 };

/* ReconstructParams (src/reconstructor.rs:8-15): the reference fields in the
 * reference order, then extension fields.  prx_params_default() gives the
 * reference defaults (8, 1.5, 1.1, 1e-3, 500, 0.0, false). */
typedef struct prx_params {
    uint32_t depth;
    double samples_per_node; /* accepted and ignored, as in src/fem_tree.rs:83 */
    double scale;
    double cg_accuracy;
    uint64_t cg_iters;
    double point_weight;
    int32_t verbose; /* ignored inside reconstruct, as in the reference */
    /* extensions */
    int32_t precision;      /* PRX_PREC_FP64 (default) or PRX_PREC_MIXED */
    int32_t residual_mode;  /* PRX_RESID_REFERENCE (default) or PRX_RESID_TRUE */
    int32_t device;         /* HIP device ordinal (default 0) */
    int32_t gs_iters;       /* smoothing sweeps per level (reference: 20) */
    int32_t base_depth;     /* -1: D-2 rule of src/reconstructor.rs:51 */
    int32_t refresh_period; /* CG true-residual refresh period (reference: 10; 0 = never) */
    double inner_tol;       /* mixed: FP32 residual reduction between FP64 residual replacements (default 0.05) */
    uint64_t max_outer;     /* mixed: cap on FP64 residual replacements (default 64) */
    int32_t precond;        /* PRX_PRECOND_NONE (default: the reference's cascadic solve) or PRX_PRECOND_MG */
} prx_params;

/* ReconstructStats (src/reconstructor.rs:17-23) plus extension fields.
 * Unlike the reference (which leaves them 0), time_system and matrix_nnz are
 * filled: time_system is the fine+coarse assembly time, matrix_nnz the number
 * of stored entries the reference's CSR would hold. */
typedef struct prx_stats {
    double time_transform, time_tree, time_system, time_solve, time_extract, time_total;
    uint64_t fem_nodes, matrix_nnz, mesh_vertices, mesh_triangles;
    double iso_value;
    /* extensions */
    uint64_t coarse_iters, fine_iters, outer_iters;
    double final_rel_residual; /* FP64 ||b-Ax||/||b|| of the returned solution */
    double solve_ms;           /* GPU time of the solve (hipEvent) */
    int32_t solve_status;      /* PRX_OK / PRX_ERR_NOT_CONVERGED / PRX_ERR_BREAKDOWN */
} prx_stats;

/* Mesh (src/marching_cubes.rs:8-11), in model space; arrays owned by the
 * library until prx_mesh_free. */
typedef struct prx_mesh {
    double* vertices;    /* [3*n_vertices] */
    uint64_t n_vertices;
    uint64_t* triangles; /* [3*n_triangles] */
    uint64_t n_triangles;
} prx_mesh;

/* Options and report of one cascadic solve (FEMTree::solve_cascadic,
 * src/fem_tree.rs:458-487). */
typedef struct prx_solve_opts {
    int32_t precision;
    int32_t residual_mode;
    int32_t refresh_period;
    int32_t gs_iters;
    double inner_tol;
    uint64_t max_outer;
    double coarse_eps;      /* PRX_RESID_TRUE only: coarsest-level ||b-Ax|| <= coarse_eps ||b||; 0 = eps */
    /* PRX_PRECOND_MG replaces the cascadic GS + CG by CG preconditioned with
     * a geometric-multigrid V-cycle, from x = 0, stopping at
     * ||b-Ax|| <= eps ||b|| (no counterpart in the reference; same system,
     * same stopping test as PRX_RESID_TRUE) */
    int32_t precond;        /* PRX_PRECOND_NONE (default) or PRX_PRECOND_MG */
    int32_t mg_smooth;      /* Jacobi sweeps before and after each coarse correction (default 2) */
    double mg_omega;        /* smoothing: > 0 damped Jacobi with this weight; 0 (default) Chebyshev weights on [lmax/6, lmax]; < 0 Chebyshev on [lmax/-mg_omega, lmax] */
    double mg_inner_tol;    /* mixed MG: FP32 residual reduction per FP64 refinement step (default 1e-4) */
    /* CG iteration form (no counterpart in the reference; the same
     * iteration): PRX_CG_FUSED (default) one fused pass per iteration on
     * grids of depth >= 3 (the residual update r -= alpha q joins the next
     * operator pass, whose |r|^2 for the stopping test and beta is expanded
     * from the pass's d.q, r.q, q.q and r.r), in FP32 on grids of depth >= 8
     * without storing q = A d (the pass recomputes it inside each 16^3 brick
     * and reads it on the brick's boundary from the neighbours' faces; batched
     * solves keep q stored);
     * PRX_CG_TWO_PASS the two-pass form (operator pass, then the vector
     * update with its own |r|^2); PRX_CG_FUSED_STORED_Q the fused pass with
     * q stored as a full vector in every precision; PRX_CG_FUSED_FACES the
     * face form in every precision on grids of depth >= 4, batches included */
    int32_t cg_form;
    /* Batched cascadic solves only (no counterpart in the reference): how the
     * CG iterations of all instances are launched.  PRX_BATCH_LAUNCH_PERSISTENT
     * runs each level's CG of every instance in one persistent launch (a
     * cooperative kernel whose workgroups iterate the instances in lockstep
     * rounds and decide replacements, confirmations and stops on the device;
     * PRX_RESID_TRUE with the fused CG step); PRX_BATCH_LAUNCH_GRAPHS replays
     * captured chunks of iterations with host checks between them;
     * PRX_BATCH_LAUNCH_AUTO (0, the default) picks the graph replays, the
     * faster of the two on many-cloud batches.  An explicit persistent request
     * that does not apply returns PRX_ERR_INVALID. */
    int32_t batch_launch;
} prx_solve_opts;

enum { PRX_CG_FUSED = 0, PRX_CG_TWO_PASS = 1, PRX_CG_FUSED_STORED_Q = 2, PRX_CG_FUSED_FACES = 3 };
enum { PRX_BATCH_LAUNCH_AUTO = 0, PRX_BATCH_LAUNCH_GRAPHS = 1, PRX_BATCH_LAUNCH_PERSISTENT = 2 };

typedef struct prx_solve_report {
    uint64_t coarse_iters, fine_iters, outer_iters, applies;
    double coarse_rel;      /* coarse CG sqrt(delta/delta0) at exit */
    double fine_rel;        /* fine stopping-test value at exit */
    double true_rel;        /* FP64 ||b-Ax||/||b|| of the final solution */
    double ms_assemble, ms_coarse, ms_fine_gs, ms_fine_cg, ms_total;
    int32_t status;
    /* ||r0||/||b||, r0 the fine CG's starting residual after smoothing (single-tree
     * cascadic solves with PRX_RESID_TRUE; 0 otherwise).  The reference's stopping
     * quantity ||r||/||r0|| (src/solvers.rs:36-58) of the returned solution is
     * true_rel / fine_r0_rel. */
    double fine_r0_rel;
} prx_solve_report;

typedef struct prx_tree prx_tree;

/* ------------------------------------------------------------ errors */
const char* prx_last_error(void);
int prx_device_count(int* out);

/* ------------------------------------------------------------ pipeline */
void prx_params_default(prx_params* p);
void prx_solve_opts_default(prx_solve_opts* o);

/* reconstruct (src/reconstructor.rs:25-70).  Empty input gives PRX_OK and an
 * empty mesh with zeroed stats (src/reconstructor.rs:28). */
int prx_reconstruct(const prx_point* points, size_t n, const prx_params* params, prx_mesh* out_mesh,
                    prx_stats* out_stats);
void prx_mesh_free(prx_mesh* mesh);

/* ------------------------------------------------------------ FEMTree */

/* FEMTree::new + initialize + finalize (src/fem_tree.rs:67-126): the complete
 * tree to `depth`, nodes numbered in Morton order. */
int prx_tree_create(uint32_t depth, int device, prx_tree** out);
void prx_tree_destroy(prx_tree* t);
/* Frees the solver workspace a tree keeps between solves (CG vectors,
 * multigrid levels); the solution, points and lists stay and the next solve
 * allocates again.  No counterpart in the reference (for callers holding
 * many trees). */
int prx_tree_release_workspace(prx_tree* t);
/* fem_node_count (src/fem_tree.rs:54) */
int64_t prx_tree_node_count(const prx_tree* t);
/* idx_to_offset [3N] and grid_to_idx [res^3], row-major z*res*res+y*res+x
 * (src/fem_tree.rs:56-58, :117-118); either pointer may be NULL. */
int prx_tree_export_nodes(prx_tree* t, uint32_t* idx_to_offset, int64_t* grid_to_idx);

/* Upload points already in the unit cube (the output of the transform at
 * src/reconstructor.rs:36-40) and splat them: FEMTree::splat_normal_field
 * (src/fem_tree.rs:130-178).  Builds the per-node point lists used by
 * assembly and the operator. */
int prx_splat_normal_field(prx_tree* t, const prx_point* unit_points, size_t n);
/* The same with the points already in device memory of the tree's device:
 * pos_device and nrm_device are [n][3] (x, y, z per point), as a torch
 * tensor of shape (n, 3) holds them.  Used by the torch path. */
int prx_splat_normal_field_device(prx_tree* t, const double* pos_device, const double* nrm_device, size_t n);
/* normal_field [3N, AoS x,y,z] and node_weight [N] (src/fem_tree.rs:60-61). */
int prx_export_normal_field(prx_tree* t, double* normal_field, double* node_weight);

/* FEMTree::assemble_system + assemble_rhs (src/fem_tree.rs:180-299): sets up
 * the matrix-free fine operator and computes the RHS on the device. */
int prx_assemble(prx_tree* t, double point_weight);
int prx_export_rhs(prx_tree* t, double* rhs);
/* Verification mode: the CSR the reference would hold (src/sparse.rs:24-33),
 * row i sorted by column, values summed in the reference's push order.
 * prx_csr_nnz first, then prx_export_csr with off[N+1], col[nnz], val[nnz]. */
int prx_csr_nnz(prx_tree* t, int64_t* nnz);
int prx_export_csr(prx_tree* t, int64_t* off, int64_t* col, double* val);

/* y = A x with the matrix-free operator (SparseMatrix::multiply_vector,
 * src/sparse.rs:105-117, on the assembled system).  precision
 * PRX_PREC_MIXED applies the FP32 operator. */
int prx_apply(prx_tree* t, int precision, const double* x, double* y);

/* z = M in, one multigrid V-cycle of the PRX_PRECOND_MG solver (M ~ A^-1
 * without the per-point screening; smooth <= 0 selects 2 sweeps, omega as
 * prx_solve_opts.mg_omega).  MIXED applies the FP32 V-cycle of the scaled
 * system.  No counterpart in the reference: a diagnostic of the
 * preconditioner. */
int prx_mg_vcycle(prx_tree* t, int precision, int smooth, double omega, const double* in, double* out);

/* build_system_at_depth (src/fem_tree.rs:346-404) at depth cdepth: RHS in
 * row-major coarse order [8^cdepth] and the operator's constants
 * coef[4] = {K centre, K edge, K corner, diagonal screening wh3*1.005}. */
int prx_coarse_system(prx_tree* t, uint32_t cdepth, double* rhs_rowmajor, double* coef);
/* prolongate (src/fem_tree.rs:310-343): coarse row-major -> fine [N]. */
int prx_prolongate(prx_tree* t, uint32_t cdepth, const double* coarse_rowmajor, double* fine);

/* `sweeps` multicolour Gauss-Seidel sweeps on the fine system, in place on x
 * (the reference's gauss_seidel_sweep, src/solvers.rs:164-175, recoloured:
 * 8 colours by coordinate parity). */
int prx_gs(prx_tree* t, int precision, uint32_t sweeps, const double* b, double* x);
/* solve_cg (src/solvers.rs:20-69) on the fine system from x.
 * `iters` and `rel` may be NULL. */
int prx_cg(prx_tree* t, int precision, const double* b, double* x, uint64_t max_iters, double eps,
           int residual_mode, int refresh_period, uint64_t* iters, double* rel);

/* FEMTree::solve_cascadic (src/fem_tree.rs:458-487). The solution stays on
 * the device; prx_get_solution copies it out. */
int prx_solve_cascadic(prx_tree* t, uint32_t base_depth, uint32_t gs, uint64_t cg_iters, double eps,
                       const prx_solve_opts* opts, prx_solve_report* report);
int prx_get_solution(prx_tree* t, double* x);
int prx_set_solution(prx_tree* t, const double* x);

/* FEMTree::evaluate (src/fem_tree.rs:490-513) at m query points q[3m]. */
int prx_evaluate(prx_tree* t, const double* q, size_t m, double* out);
/* The same with q [m][3] and out [m] in device memory (the torch path). */
int prx_evaluate_device(prx_tree* t, const double* q_device, size_t m, double* out_device);
/* evaluate at every marching-cubes corner k/res, k in [0,res]^3
 * (src/fem_tree.rs:596-605), x fastest: out[(kz*(res+1)+ky)*(res+1)+kx]. */
int prx_evaluate_corners(prx_tree* t, double* out);
/* iso value: mean of evaluate over the uploaded points (src/reconstructor.rs:56). */
int prx_iso_value(prx_tree* t, double* iso);
/* FEMTree::extract_surface (src/fem_tree.rs:576-636) of the current solution
 * at `iso`, in unit-cube coordinates: the same vertices (to the reference's
 * 1e-9 position quantisation), numbered in the reference's insertion order,
 * and the same triangles in the same order.  The arrays are library-owned
 * until prx_mesh_free. */
int prx_extract_surface(prx_tree* t, double iso, prx_mesh* out);

/* BBox + unit_cube_transform + transform_point/vector + normalize
 * (src/geometry.rs:349-402, :256-269, :49-56) as applied at
 * src/reconstructor.rs:31-40.  m, minv: column-major 4x4 (m[col*4+row]) of
 * the model->unit map and its inverse (Mat4::inverse, src/geometry.rs:274-316);
 * out receives the n transformed points.  Any output may be NULL. */
int prx_unit_transform(const prx_point* points, size_t n, double scale, double* m, double* minv, prx_point* out);
/* The same on the device (bounding-box reduction and point map in kernels,
 * unfused arithmetic as the reference); bit-identical to prx_unit_transform.
 * prx_reconstruct uses this path. */
int prx_unit_transform_device(const prx_point* points, size_t n, double scale, int device, double* m, double* minv,
                              prx_point* out);

/* ------------------------------------------------------------ measurement
 * No counterpart in the reference (it has no benchmarks): device-side timing
 * used by bench.py.  prx_time_operator launches the operator `reps` times on
 * the tree's stream and returns the average hipEvent duration of the
 * screening gather kernel and of the stencil kernel.  prx_copy_bandwidth
 * times a streaming device copy of `bytes` (read + write) and returns the
 * average ms per copy. */
int prx_time_operator(prx_tree* t, int precision, int reps, double* ms_gather, double* ms_stencil);
/* As prx_time_operator for one pass of the CG iteration: op 0 = y = A x,
 * op 1 = the fused direction update and apply (d_new = r + beta d_old,
 * q = A d_new, beta = 0.5), op 2 = the fused CG step (r' = r - alpha q,
 * d_new = r' + beta d_old, q' = A d_new, alpha = 0.25, beta = 0.5; depth >= 3),
 * each with its screening gather; op 3 = the same step in face form (q
 * recomputed in each brick, faces of q' written; no gather; depth >= 4). */
int prx_time_stencil_op(prx_tree* t, int precision, int op, int reps, double* ms_gather, double* ms_stencil);
int prx_copy_bandwidth(size_t bytes, int reps, int device, double* ms);
/* Measurement: the tile kernels' memory access pattern without arithmetic.
 * Each 256-thread workgroup copies one 4096-node brick (FP32 or FP64 by
 * `precision`) from each of `streams` (1..3) input vectors to as many
 * output vectors, workgroups mapped to the XCDs as the tile kernels map them;
 * *ms = average ms per launch over nbricks bricks (a multiple of 8).  The
 * fused CG step reads and writes three vectors this way, so the 3-stream
 * rate is the ceiling its access pattern allows. */
int prx_brick_copy_bandwidth(int64_t nbricks, int streams, int precision, int reps, int device, double* ms);
/* Measurement: the phases of the last prx_splat_normal_field, ms[5] =
 * {host packing and upload of the points, spatial sort of the points, the
 * splat's per-node lists, the screening (truncated-cell) per-node lists, the
 * splat kernel}.  The screening lists stand in for the per-point blocks the
 * reference pushes in assemble_system (src/fem_tree.rs:199-227); they are
 * built once per point set and reused by every assembly and solve. */
int prx_splat_times(prx_tree* t, double* ms);
/* Measurement (no counterpart in the reference): the duration of every
 * launch of the finest level's fused CG pass (the face-form pass in the FP32
 * inner solves at depth >= 8, else the stored-q pass) in the solves that run
 * between prx_pass_timing(t, 1, ...) and prx_pass_timing(t, 0, ms, cap, &n),
 * timed with a hipEvent pair on the tree's stream around each launch.  The
 * second call returns n durations (ms, launch order; the first min(n, cap)
 * are written).  Passes enqueued after the solve converged within a chunk of
 * iterations return at once and appear as durations of a few microseconds. */
int prx_pass_timing(prx_tree* t, int on, float* ms, size_t cap, size_t* n);
/* Sizes of the screening data of the last splat (measurement; no
 * counterpart in the reference): out[3] = {touched nodes, (point, corner)
 * entries, occupied trunc cells}. */
int prx_screening_sizes(prx_tree* t, int64_t* out);
/* Measurement: evaluate the current solution at every marching-cubes corner
 * (as prx_evaluate_corners, src/fem_tree.rs:596-605) into a device buffer
 * `reps` times; *ms = average ms per evaluation, *sumsq = sum of the squared
 * corner values (a checksum).  general != 0 times the general per-point
 * evaluate() instead of the corner form. */
int prx_time_evaluate_corners(prx_tree* t, int reps, int general, double* ms, double* sumsq);

/* ------------------------------------------------------------ adjoint
 * No counterpart in the reference (it is not differentiable): the vjp of the
 * field at the splatted points, F_p = evaluate(x, q_p) with x the current
 * converged solution, through the splat, RHS, screening and solve.  Given
 * dL/dF [P] (input point order) it writes dL/dq [3P] and dL/dn [3P] (AoS,
 * unit-cube coordinates, the point transform held fixed).  The adjoint system
 * A lam = dL/dx is solved to ||r|| <= eps ||dL/dx|| in `precision`, with
 * the solver of the tree's last prx_solve_cascadic (multigrid-preconditioned
 * CG when that used PRX_PRECOND_MG, CG otherwise). */
typedef struct prx_vjp_report {
    uint64_t iters;
    double rel;
    double ms;
    int32_t status;
} prx_vjp_report;
int prx_vjp_points(prx_tree* t, const double* dL_dF, double* dL_dpos, double* dL_dnrm, int precision, double eps,
                   prx_vjp_report* report);
/* The same with dL_dF [P], dL_dpos [3P] and dL_dnrm [3P] in device memory
 * (the torch path). */
int prx_vjp_points_device(prx_tree* t, const double* dL_dF_device, double* dL_dpos_device, double* dL_dnrm_device,
                          int precision, double eps, prx_vjp_report* report);

/* ------------------------------------------------------------ batches
 * No counterpart in the reference (it reconstructs one cloud per call,
 * src/reconstructor.rs:25): many clouds, each a full FEMTree of its own depth
 * (6..10), solved together.  prx_batch_add transforms nothing (points are in
 * the unit cube, as for prx_splat_normal_field), splats and assembles one
 * instance; prx_solve_batched runs FEMTree::solve_cascadic
 * (src/fem_tree.rs:458-487, base depth D-2) for every instance over one
 * concatenated node array, with the CG iterations and GS sweeps replayed as
 * captured HIP graphs, in FP64 or (opts->precision = PRX_PREC_MIXED, fused
 * CG step only) in mixed precision: FP32 smoothing and CG on each instance's
 * correction with FP64 scalars, and batched FP64 passes that fold the
 * correction into x and replace each instance's residual by its FP64 true
 * residual, as the single-tree mixed solve does.  With autotune != 0 the first solve times
 * the operator brick size, the GS block size and the CG chunk length on the
 * batch and keeps the fastest (reported).  With opts->precond =
 * PRX_PRECOND_MG every instance is instead solved by multigrid-preconditioned
 * CG (FP64 or mixed) to ||b-Ax|| <= eps ||b||, the instances spread over up to
 * four streams; gs and autotune are then unused.  Instance trees (prx_batch_tree)
 * are borrowed: evaluate / get_solution / vjp work on them; never destroy
 * them.  per_instance (optional) receives one report per instance. */
typedef struct prx_batch prx_batch;
typedef struct prx_batch_report {
    uint32_t instances, converged;
    uint64_t total_nodes, total_points;
    uint64_t max_fine_iters, sum_fine_iters, max_coarse_iters, graph_launches;
    double ms_total, ms_coarse, ms_fine_gs, ms_fine_cg, ms_tune;
    double max_true_rel;   /* max over instances of ||b-Ax||/||b|| */
    int32_t tile, chunk, gs_block, status; /* operator brick edge, CG iterations per graph, GS brick edge */
    uint64_t persistent_launches;          /* CG levels run as one persistent launch each */
    uint64_t persistent_rounds;            /* lockstep rounds over all persistent launches */
    int32_t persistent_blocks;             /* workgroups of a persistent launch */
    int32_t persistent_xcd_match;          /* of those, how many run on XCD (launch index mod 8) */
    double persistent_phase_ms[4];         /* this solve's persistent launches: replacement, gather, pass and fold
                                              phases (each to the end of its grid barrier) */
} prx_batch_report;
int prx_batch_create(int device, prx_batch** out);
void prx_batch_destroy(prx_batch* b);
int prx_batch_add(prx_batch* b, uint32_t depth, const prx_point* unit_pts, size_t n, double point_weight, uint32_t* index);
int prx_batch_size(prx_batch* b, uint32_t* count);
int prx_batch_tree(prx_batch* b, uint32_t index, prx_tree** tree);
int prx_solve_batched(prx_batch* b, uint32_t gs, uint64_t cg_iters, double eps, const prx_solve_opts* opts, int autotune,
                      prx_batch_report* report, prx_solve_report* per_instance);

/* ------------------------------------------------------------ generic CSR
 * SparseMatrix::multiply_vector (src/sparse.rs:105-117) and solve_cg
 * (src/solvers.rs:20-69) on a caller CSR, run on the device. */
int prx_csr_spmv(int64_t n, const int64_t* off, const int64_t* col, const double* val, const double* x, double* y);
int prx_csr_cg(int64_t n, const int64_t* off, const int64_t* col, const double* val, const double* b, double* x,
               uint64_t max_iters, double eps, uint64_t* iters);

/* ------------------------------------------------------------ no allocation
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): the same
 * solve_cg as prx_csr_cg, with off[n+1], col, val, b and x (in/out) already in
 * device memory on the current device, launched on `stream` (a hipStream_t;
 * NULL = the null stream).  The solver's scratch (r, d, q, partial sums, CG
 * scalars) lives in one fixed array of the library per device (about 6.3 MB,
 * held until the process ends); the call makes no device allocation.  n above
 * PRX_CSR_CG_NOALLOC_MAX_ROWS returns PRX_ERR_INVALID.  Every call on a device
 * shares that array: calls that overlap on one device race.  The host reads
 * the stopping flag every 16 iterations, so the call cannot be graph-captured. */
#define PRX_CSR_CG_NOALLOC_MAX_ROWS 262144
int prx_csr_cg_noalloc_device(int64_t n, const int64_t* off_device, const int64_t* col_device, const double* val_device,
                              const double* b_device, double* x_device, uint64_t max_iters, double eps, uint64_t* iters,
                              void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PRX_H */
