/*
 * lss.h — C ABI of the MI355X hybrid solver for locally stiff SPD systems.
 *
 * Drop-in boundary for the reference's GPU hybrid solver
 * polysolve::linear::GPUHybrid (src/polysolve/linear/GPUHybrid.hpp:27-192,
 * src/polysolve/linear/GPUHybrid.cu:79-1543) and the polysolve::linear::Solver
 * surface it implements (src/polysolve/linear/Solver.hpp:32-149).  Plain
 * pointers and sizes only; every entry point returns an lss_status and never
 * exits the process (the reference calls exit(EXIT_FAILURE) on device errors,
 * GPUHybrid.cu:34-63).
 *
 * Matrix input: the compressed-sparse-column arrays of polysolve's
 * StiffnessMatrix (Eigen::SparseMatrix<double, ColMajor, int>,
 * src/polysolve/Types.hpp:162-166): 0-based int32 indices, full symmetric
 * storage, ascending row indices per column.  Like the reference (which hands
 * the CSC arrays to hypre as CSR rows, GPUHybrid.cu:401-433) the solver reads
 * column j as row j of the operator, so the matrix must be symmetric.
 *
 * Threading: one handle per host thread; calls on one handle must not overlap.
 * Device: the HIP device current at lss_create.
 */
#ifndef LSS_H
#define LSS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum lss_status {
    LSS_OK = 0,
    LSS_ERR_CONFIG = 1,          /* invalid JSON / key / type (Solver.cpp:157-165 throws)          */
    LSS_ERR_SIZE = 2,            /* size mismatch (MASSolver.cu:486-489 throws)                     */
    LSS_ERR_NOT_FACTORIZED = 3,  /* solve before factorize                                          */
    LSS_ERR_HIP = 4,             /* HIP runtime error (GPUHybrid.cu:34-63 exits instead)            */
    LSS_ERR_NOT_SPD = 5,         /* non-SPD stiff block or MAS bank (MASPreconditioner.cu:934)      */
    LSS_ERR_OOM = 6,             /* device allocation failed (CudaUtils.cuh:41-124 diagnoses)       */
    LSS_ERR_UNSUPPORTED = 7,     /* input outside what this build implements                        */
    LSS_ERR_ARG = 8              /* null handle / pointer, negative size                            */
} lss_status;

typedef struct lss_solver lss_solver;

/*
 * Create a solver from a JSON document shaped like spec/GPUHybrid.json:
 *   {"solver": "GPUHybrid" | ["GPUHybrid", ...], "GPUHybrid": {key: value, ...}}
 * Keys and defaults follow linear-solver-spec.json:646-747 (the defaults
 * Solver::create injects, Solver.cpp:143-173); keys parsed as in
 * GPUHybrid::set_parameters (GPUHybrid.cu:106-179).  Additional keys:
 * "precision" (fp64 | fp32 | bf16 | fp16 bank storage), "inner_spmv" (auto | fp32 |
 * fp64), "base_preconditioner" (MAS | Jacobi), "partitioner" (device | host),
 * "partition_local_max" (0-8192), "max_coarse_levels" (0-4), "damping",
 * "lanczos_steps", "check_every",
 * "max_dense_component", "true_residual_check", "profile_kernels"; unknown keys
 * are rejected.  json may be NULL or "" for all defaults.
 * Replaces: Solver::create(json, logger) + GPUHybrid::GPUHybrid() (GPUHybrid.cu:79-104).
 */
int lss_create(const char *json, lss_solver **out);

/* Re-apply parameters.  Replaces GPUHybrid::set_parameters (GPUHybrid.cu:106-179). */
int lss_set_parameters(lss_solver *s, const char *json);

/* User-supplied stiff DOFs, used when select_bad_dofs_from_l1_norm is false.
 * Replaces GPUHybrid::set_problematic_dofs (GPUHybrid.hpp:69). */
int lss_set_problematic_dofs(lss_solver *s, const int32_t *dofs, int64_t n);

/* No-op kept for the Solver surface.  Replaces Solver::analyze_pattern (Solver.hpp:97). */
int lss_analyze_pattern(lss_solver *s, int32_t n, int64_t nnz, const int32_t *colptr, const int32_t *rowidx);

/* Factorize from host CSC arrays (copied; the caller keeps ownership).
 * Replaces GPUHybrid::factorize (GPUHybrid.cu:195-282). */
int lss_factorize_csc(lss_solver *s, int32_t n, int64_t nnz, const int32_t *colptr, const int32_t *rowidx,
                      const double *vals);

/* Factorize from device-resident CSC arrays (not retained after return).
 * stream: hipStream_t, or NULL for the solver's own stream (a blocking stream,
 * ordered with the legacy default stream).  Same semantics as lss_factorize_csc
 * without the host-to-device copy. */
int lss_factorize_csc_device(lss_solver *s, int32_t n, int64_t nnz, const int32_t *d_colptr,
                             const int32_t *d_rowidx, const double *d_vals, void *stream);

/* Solve A x = b with host vectors; x is in/out (initial guess).
 * Replaces GPUHybrid::solve (GPUHybrid.cu:342-399) and pcg_solve (1413-1512). */
int lss_solve(lss_solver *s, const double *b, double *x);

/* Same with device vectors (n doubles each), enqueued on stream (or the solver's
 * own stream when NULL); returns after the solve has finished. */
int lss_solve_device(lss_solver *s, const double *d_b, double *d_x, void *stream);

/* Solve k right-hand sides (column-major n x k device arrays) as one fused
 * block PCG that shares every matrix and preconditioner read; each column keeps
 * its own scalars and stopping test.  No reference counterpart (the reference
 * solves one right-hand side per call, GPUHybrid.cu:342). */
int lss_solve_multi_device(lss_solver *s, int32_t k, const double *d_B, double *d_X, void *stream);

/* JSON info: {"num_iterations", "final_res_norm"} as GPUHybrid::get_info
 * (GPUHybrid.cu:181-185) plus phase timings and stiff-set statistics.
 * Writes at most cap bytes (NUL-terminated); *needed (optional) gets the full length. */
int lss_get_info(const lss_solver *s, char *buf, size_t cap, size_t *needed);

/* Solver name with the reference's rules (GPUHybrid.hpp:57-67):
 * "GPUHybrid", "GPUAMGF" or "GPUAMG". */
const char *lss_name(const lss_solver *s);

/* Stiff DOF components selected at factorize (ascending DOFs per component,
 * components ordered by smallest DOF).  dofs[cap], comp_off[comp_cap] (ncomp+1
 * entries).  The reference keeps them in bad_indices_arrays (GPUHybrid.cu:219-262). */
int lss_get_stiff_dofs(const lss_solver *s, int32_t *dofs, int64_t cap, int64_t *n_dofs, int64_t *comp_off,
                       int64_t comp_cap, int64_t *n_comp);

/* Last error message of this handle (empty string when none).  With s == NULL:
 * the message of the last failed lss_create or handle-free call (lss_adjoint_*,
 * lss_residual_norm*) on this thread. */
const char *lss_last_error(const lss_solver *s);

/* Release all device memory.  Replaces GPUHybrid::~GPUHybrid (GPUHybrid.cu:1514-1530). */
void lss_destroy(lss_solver *s);

/* ---- adjoint of the solve ------------------------------------------------
 * The reference has no adjoint.  For a loss L(x) with A x = b solved by a
 * handle, the caller solves A^T lam = dL/dx with the same handle (the operator
 * is symmetric, so A^T = A and the factorization is reused: lss_solve_device
 * with the right-hand side dL/dx); then dL/db = lam and the two calls below
 * give the matrix and per-element gradients.  Both are handle-free, run
 * asynchronously on `stream` (NULL = the null stream) and take device pointers.
 *
 * dL/dA on the stored pattern: d_grad[e] = -lam[j] * x[d_rowidx[e]] for every
 * entry e of CSC column j, i.e. the gradient with respect to operator entry
 * (j, rowidx[e]) (CSC column j is operator row j, GPUHybrid.cu:401-433).
 * d_grad has nnz entries in the order of d_rowidx. */
int lss_adjoint_pattern_grad(int32_t n, int64_t nnz, const int32_t *d_colptr, const int32_t *d_rowidx,
                             const double *d_lam, const double *d_x, double *d_grad, void *stream);
/* Chain rule to parameters p_s through the sparse Jacobian d vals / d p:
 * d_out[s] = sum over k in [d_seg_ptr[s], d_seg_ptr[s+1]) of d_coef[k] * d_grad_vals[d_slot[k]],
 * e.g. per-element coefficients with (slot, unit-coefficient value) pairs per element. */
int lss_adjoint_chain(int64_t nseg, const int64_t *d_seg_ptr, const int32_t *d_slot, const double *d_coef,
                      const double *d_grad_vals, double *d_out, void *stream);

/* ---- true residual --------------------------------------------------------
 * d_out[0] = ||b - A x||_2 and d_out[1] = ||b||_2 (two doubles in device memory)
 * for the CSC arrays read as operator rows, computed in double-double arithmetic
 * (exact products by fused multiply-add, compensated sums), so the norm is
 * accurate even where the FP64 evaluation of b - A x has an error floor of
 * eps * || |A| |x| || above the residual itself.  The reference reports the FP64
 * residual norm (final_res_norm, GPUHybrid.cu:383-386); this is its
 * extended-precision counterpart for checking a solution.  Handle-free and
 * asynchronous on `stream` (NULL = the null stream); allocates and frees its
 * partial sums (32 KiB at most) in stream order. */
int lss_residual_norm(int32_t n, int64_t nnz, const int32_t *d_colptr, const int32_t *d_rowidx, const double *d_vals,
                      const double *d_b, const double *d_x, double *d_out, void *stream);

/* lss_residual_norm without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The partial sums live in
 * a static device array of the library (64 KiB; each GPU the library runs on holds
 * its own copy until the process ends); every n fits it (the grid is capped at 2048
 * workgroups), and each call overwrites what it reads, so nothing needs clearing.
 * All calls on one GPU share that array: calls that overlap on one GPU (other
 * streams, other host threads) race and must be serialized by the caller.  Same
 * arguments, results and status codes as lss_residual_norm. */
int lss_residual_norm_noalloc(int32_t n, int64_t nnz, const int32_t *d_colptr, const int32_t *d_rowidx,
                              const double *d_vals, const double *d_b, const double *d_x, double *d_out, void *stream);

/* ---- inspection entry points (parity tests and tools) --------------------
 * Each runs the named device stage of the factorized solver on host data.  No
 * reference counterpart as separate calls: they expose the stages of
 * GPUHybrid.cu named in each comment. */

/* Row l1 norms computed on the device at factorize (GPUHybrid.cu:753-765). */
int lss_debug_row_norms(const lss_solver *s, double *out, int64_t n);
/* GMM result: params = {global_mean, global_var, mean0, mean1, var0, var1, w0, w1},
 * iters, jumped, selected (GPUHybrid.cu:767-883). */
int lss_debug_gmm(const lss_solver *s, double *params, int32_t *iters, int32_t *jumped, int64_t *selected);
/* MAS partition: perm[old_node] = new_node (nb entries), part_off (nparts+1). */
int lss_debug_partition(const lss_solver *s, int32_t *perm, int64_t nb, int32_t *part_off, int64_t cap,
                        int32_t *nparts);
/* y = A x with the device SpMV (FP64). */
int lss_debug_spmv(lss_solver *s, const double *x, double *y);
/* AMG hierarchy of base_preconditioner "AMG" (BoomerAMG setup, GPUHybrid.cu:357-376):
 * *nlevels levels; level `level` has *n rows, operator A (levels >= 1; level 0 is the
 * input) and interpolation P (n x rows of the next level, none on the coarsest) with
 * *nnz_a / *nnz_p entries, and the C/F splitting cf (1 C, -1 F, -3 no strong
 * connection; n entries, none on the coarsest).  Arrays (CSR, int64 row pointers)
 * are copied when non-NULL. */
int lss_debug_amg(const lss_solver *s, int32_t level, int32_t *nlevels, int32_t *n, int64_t *nnz_a, int64_t *nnz_p,
                  int8_t *cf, int64_t *a_ptr, int32_t *a_col, double *a_val, int64_t *p_ptr, int32_t *p_col,
                  double *p_val);
/* z = B r (base preconditioner, undamped) and z = M^{-1} r (full hybrid preconditioner). */
int lss_debug_base_apply(lss_solver *s, const double *r, double *z);
int lss_debug_precond_apply(lss_solver *s, const double *r, double *z);
/* Nested-dissection ordering and multifrontal structure that the sparse stiff
 * factorization (units above max_dense_component; the cuDSS analysis of
 * GPUHybrid.cu:1231-1249) builds for a symmetric graph: CSR adj_ptr[m + 1] /
 * adj_idx over m vertices, no self loops; subgraphs of at most `leaf` vertices become
 * one front.  Host only.  Always sets *nfronts and *nstruct; when order is non-NULL
 * also fills order[m] (position -> vertex), per front (postorder) its first pivot
 * position, pivot count and parent (-1 for a root), and struct_ptr[nfronts + 1] /
 * struct_idx[*nstruct]: the update positions of each front, ascending. */
int lss_debug_nested_dissection(int32_t m, const int32_t *adj_ptr, const int32_t *adj_idx, int32_t leaf,
                                int32_t *order, int32_t *nfronts, int64_t *nstruct, int32_t *front_p0,
                                int32_t *front_np, int32_t *front_parent, int64_t *struct_ptr, int32_t *struct_idx);

#ifdef __cplusplus
}
#endif
#endif /* LSS_H */
