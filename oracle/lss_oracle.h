/*
 * ORACLE — test infrastructure only.
 *
 * A CPU (C++17/OpenMP, FP64) restatement of the reference GPU hybrid solve
 * path of geometryprocessing/local-stiff-solver.  It is the checker for the
 * HIP implementation and the timed CPU baseline; it is never linked into, or
 * called by, the product library.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.
 *
 * What each function restates (reference file:line) and what pins it:
 *   orc_row_l1          src/polysolve/linear/GPUHybrid.cu:753-765   pinned: sequential
 *                       storage-order sum, checked bit-exact against a Python loop
 *   orc_canonical_sum   fixed reduction tree shared with the device kernels
 *                       (no reference counterpart; pinned: math.fsum bound + exact
 *                       emulation in Python on small n)
 *   orc_gmm_select      GPUHybrid.cu:767-883 (global stats, EM loop, jump test,
 *                       gamma0 < gamma1 selection)  pinned: independent numpy EM,
 *                       EM invariants, separated point masses
 *   orc_filter          GPUHybrid.cu:885-970 (union-find components of the
 *                       induced subgraph, size limits, Gershgorin bound)
 *                       pinned: scipy connected_components + brute force
 *   orc_expand          GPUHybrid.cu:972-991   pinned: one-ring closure in numpy
 *   orc_decompose       GPUHybrid.cu:691-735   pinned: scipy connected_components
 *   disjoint set        src/polysolve/linear/Hybrid.hpp:223-260
 *   orc_partition       the build's recursive breadth-first bisection partitioner
 *                       (the reference uses KaMinPar, MAS/GraphPartition.cpp:176-289,
 *                       whose output is not reproducible here: parity unpinned vs
 *                       the reference; pinned: part-size, count and coverage
 *                       invariants, and bit-equality with the product's partition)
 *   orc_mas_*           src/polysolve/linear/mas_utils/MASPreconditioner.cu:33-1164
 *                       (bank-local CCO coarse spaces, per-bank matrices, zero
 *                       diagonal padding, inverse, restrict/apply/prolong)
 *                       pinned: dense numpy R_l A R_l^T construction on small grids
 *   orc_hybrid_*        GPUHybrid.cu:195-282 (factorize phases), 521-675 (B-P-B
 *                       composition), 1413-1512 (PCG + stopping rule), 383-386
 *                       (true residual)  pinned: scipy spsolve solution, closed-form
 *                       preconditioner on tiny dense systems
 *
 * Operator convention: like the reference (GPUHybrid.cu:401-433, which hands the
 * CSC arrays to hypre as CSR rows), the arrays (ptr, idx, vals) are read as CSR
 * rows of the operator; for the symmetric matrices the solver accepts this is
 * the same matrix.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reductions -------------------------------------------------------- */
/* Device reduction tree: 1024 chunks x 256 lanes, sequential per lane, then
 * pairwise trees (stride 128..1 over lanes, 512..1 over chunks). */
double orc_canonical_sum(int64_t n, const double *v);

/* ---- stiff DOF selection ---------------------------------------------- */
void orc_row_l1(int32_t n, const int32_t *ptr, const double *vals, double *out);

typedef struct orc_gmm_info {
    double global_mean, global_var;
    double mean0, mean1, var0, var1, w0, w1;
    int iters;          /* value of the loop counter at exit (reference log field) */
    int jumped;         /* |mean1| / |mean0| > jump threshold */
    int64_t selected;   /* number of rows with gamma0 < gamma1 (0 unless jumped) */
    double min_margin;  /* min |log g1 - log g0| over rows (decision margin) */
} orc_gmm_info;

/* flags[i] = 1 for selected rows; gamma (2n, optional) receives gamma0, gamma1. */
int orc_gmm_select(int32_t n, const double *norms, double tol, int max_iter, double jump,
                   uint8_t *flags, double *gamma, orc_gmm_info *info);

/* set: ascending DOF list. out: ascending list; stats = {too_small, too_large, not_poorly_conditioned}. */
int orc_filter(int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals,
               const int32_t *set, int64_t nset, int64_t min_size, int64_t max_size,
               double cond_thr, int32_t *out, int64_t *nout, int64_t *stats);
int orc_expand(int32_t n, const int32_t *ptr, const int32_t *idx, const int32_t *set, int64_t nset,
               int32_t *out, int64_t *nout);
/* Components sorted by smallest DOF; comp_off has ncomp+1 entries. */
int orc_decompose(int32_t n, const int32_t *ptr, const int32_t *idx, const int32_t *set, int64_t nset,
                  int64_t max_size, int32_t *comp_dofs, int64_t *comp_off, int64_t *ncomp);

/* ---- graph partition (parts of at most max_part nodes) ------------------ */
int orc_partition(int32_t nb, const int64_t *bptr, const int32_t *bidx, int max_part,
                  int32_t *part_of, int32_t *nparts);

/* ---- matrix helpers ------------------------------------------------------ */
void orc_spmv(int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals, const double *x,
              double *y);

/* ---- MAS preconditioner -------------------------------------------------- */
typedef struct orc_mas orc_mas;
/* Operator rows (CSR view) in the ORIGINAL numbering; perm[old_node] = new_node;
 * part_off: parts are contiguous new-node ranges (<= 32 nodes each). */
orc_mas *orc_mas_create(int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals, int bs,
                        const int32_t *perm, const int32_t *part_off, int32_t nparts);
void orc_mas_destroy(orc_mas *m);
/* levels: number of matrix levels (1 + coarse levels); counts[l] = matrices at level l;
 * cco[l] = coarse nodes at coarse level l (l < levels-1). */
void orc_mas_info(const orc_mas *m, int *levels, int32_t *counts, int32_t *cco);
/* coarse map of permuted node i at coarse level l (l < levels-1) */
int32_t orc_mas_map(const orc_mas *m, int32_t node, int level);
/* Copy the dense inverse (dim x dim, row-major) of matrix `bank` at `level`. */
int orc_mas_inverse(const orc_mas *m, int level, int32_t bank, double *out);
/* z = M_MAS^{-1} r in the original numbering. */
void orc_mas_apply(const orc_mas *m, const double *r, double *z);

typedef struct orc_hybrid orc_hybrid;
/* ---- MAS solver stiff-DOF selection (MAS/BadDOF*Preconditioner.cu) ------- */
/* Indicator per permuted scalar row (perm[old block] = new block), in the
 * reference's loop order over the permuted block row: mode 0 sum |a_ij|
 * (BadDofGMMPreconditioner.cu:273-306), 1 |a_ii| (BadDOFPreconditioner.cu:117-143),
 * 2 max |a_ij|, j != i (BadDofKMeanPreconditioner.cu:187-222).  out: nb*bs values,
 * padding rows 0. */
void orc_mas_indicator(int mode, int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals, int bs,
                       const int32_t *perm, double *out);
/* Selected positions (ascending) of each strategy over n indicator values. */
int64_t orc_mas_gmm_positions(int64_t n, const double *x, double tol, int max_iter, double jump, int expand,
                              int64_t *out, double *stats /* mean0 mean1 var0 var1 gmean gvar ratio iters */);
int64_t orc_mas_elbow_positions(int64_t n, const double *x, int64_t *out, int64_t *cutoff);
int64_t orc_mas_kmeans_positions(int64_t n, const double *x, double fraction, double jump, int max_iter, int expand,
                                 int64_t *out, double *stats /* c0 c1 ratio iters */);
/* Solver status of the last MAS-mode solve: 1 relative, 2 absolute tolerance, 3 max iterations. */
int orc_hybrid_mas_status(const orc_hybrid *h, double *residual);

/* ---- hybrid solver (factorize + PCG) ------------------------------------- */
typedef struct orc_config {
    int max_iter;
    double rel_tol, abs_tol;
    int block_size;
    int do_hybrid_precond;
    int select_bad_dofs_from_l1_norm;
    int decompose_subdomains;
    int expand_subdomains;
    int64_t min_subdomain_size, max_subdomain_size;
    double gmm_jump_threshold, gmm_tol;
    int max_gmm_iterations;
    double conditioning_threshold;
    int base_precond;   /* 0 = MAS, 1 = Jacobi, 2 = identity */
    int threads;        /* OpenMP threads (0 = runtime default) */
    double damping;     /* omega = damping / lambda_max(BA) in the B-P-B sweep (0 -> 1) */
    int composition;    /* 0 = B-P-B (GPUHybrid.cu:521-542), 1 = P + (I-PA) B (I-AP), 2 = B + P */
    /* the standalone MAS solver (MASSolver.cu): solver_mode 1 */
    int solver_mode;            /* 0 = GPUHybrid, 1 = MAS */
    int use_bad_dof;            /* MAS: additive stiff correction */
    int bad_dof_strategy;       /* 0 = elbow, 1 = k-means, 2 = GMM */
    double mgmm_jump, mgmm_tol; /* GMM strategy (defaults 10, 1e-3) */
    int mgmm_max_iter, mgmm_expand;  /* (5, 32) */
    double km_fraction, km_jump;     /* k-means strategy (0.02, 10) */
    int km_max_iter, km_expand;      /* (32, 0) */
    int true_residual_period;        /* 4; 0 = never */
    int precond_norm;                /* use_preconditioned_residual_norm */
} orc_config;

typedef struct orc_hybrid orc_hybrid;
orc_hybrid *orc_hybrid_create(const orc_config *cfg);
void orc_hybrid_destroy(orc_hybrid *h);
/* Optional user stiff set (set_problematic_dofs); used when select_bad_dofs_from_l1_norm = 0. */
void orc_hybrid_set_problematic(orc_hybrid *h, const int32_t *dofs, int64_t n);
int orc_hybrid_factorize(orc_hybrid *h, int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals);
int orc_hybrid_solve(orc_hybrid *h, const double *b, double *x);
/* iters, final_res_norm, setup seconds, solve seconds, n stiff dofs, n components */
void orc_hybrid_info(const orc_hybrid *h, int *iters, double *final_res, double *t_setup, double *t_solve,
                     int64_t *n_stiff, int64_t *n_comp);
/* Use this partition (perm[old] = new, parts = ranges of part_off) instead of orc_partition. */
void orc_hybrid_set_partition(orc_hybrid *h, const int32_t *perm, int32_t nb, const int32_t *part_off,
                              int32_t nparts);
/* Damping omega = 1 / lambda_max(BA) of the base preconditioner in the B-P-B sweep. */
double orc_hybrid_omega(const orc_hybrid *h);
void orc_hybrid_set_omega(orc_hybrid *h, double omega);
/* 1 when the last solve broke down with the configured damping and was redone with damping 1.5 */
int orc_hybrid_damping_retry(const orc_hybrid *h);
/* Thread count (0: OpenMP default) and iteration cap of the following solves (timing of the
   CPU baseline at another thread count without a new factorization). */
void orc_hybrid_set_run(orc_hybrid *h, int threads, int max_iter);
/* Export the stiff components (sorted by smallest DOF). Returns count written. */
int64_t orc_hybrid_stiff(const orc_hybrid *h, int32_t *dofs, int64_t cap, int64_t *comp_off, int64_t ccap);
/* Apply the full hybrid preconditioner once: z = M^{-1} r (after factorize). */
void orc_hybrid_precond(orc_hybrid *h, const double *r, double *z);
/* Partition used by the MAS base preconditioner (perm[old] = new, part_off). */
int32_t orc_hybrid_partition(const orc_hybrid *h, int32_t *perm, int32_t *part_off);

#ifdef __cplusplus
}
#endif
