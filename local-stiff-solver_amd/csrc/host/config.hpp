// Solver configuration: the keys of spec/GPUHybrid.json with the defaults that
// Solver::create injects from linear-solver-spec.json:646-747, parsed as in
// GPUHybrid::set_parameters (GPUHybrid.cu:106-179), plus this build's own keys.
#pragma once
#include <cstdint>
#include <string>
#include <tuple>

namespace lss {

struct Config {
    // "GPUHybrid" (SRC/GPUHybrid.hpp) or "MAS" (SRC/MASSolver.cu:140-788)
    std::string solver = "GPUHybrid";
    // reference keys (linear-solver-spec.json defaults)
    int max_iter = 10000;
    double relative_tolerance = 1e-10;
    double absolute_tolerance = 0.0;
    int block_size = 1;
    double theta = 0.5;  // AMG strength threshold for block_size > 1 (0.25 for scalar problems, GPUHybrid.cu:285-339)
    bool do_hybrid_precond = true;
    bool select_bad_dofs_from_l1_norm = true;
    double conditioning_threshold = 100.0;
    double gmm_jump_threshold = 100.0;
    double gmm_tol = 1e-3;
    int max_gmm_iterations = 20;
    bool decompose_subdomains = true;
    int64_t min_subdomain_size = 1;
    int64_t max_subdomain_size = 10000000;
    bool expand_subdomains = true;
    int64_t max_dense_size = 0;           // dense branch: components of at most this many DOFs go to bins
    bool use_float_on_subdomains = false;  // accepted; the stiff factor stays FP64
    // keys of the "MAS" solver section (MASSolver.cu:148-172, 226-279); with solver
    // "MAS", max_iter / relative_tolerance / absolute_tolerance / block_size (its
    // "block_dim") take that solver's defaults 100000 / 1e-6 / 1e-20 / 3
    // reuse the first MAS partition while the block count is unchanged (MASSolver.cu:389-405;
    // default true for "MAS" as linear-solver-spec.json:776-781, false for GPUHybrid)
    bool lazy_partitioning = false;
    bool use_preconditioned_residual_norm = false;
    bool use_bad_dof_precond = false;
    std::string bad_dof_strategy = "elbow";  // "elbow" | "kmean" | "gmm" (other strings: elbow)
    double bad_dof_kmeans_search_fraction = 0.02;
    double bad_dof_kmeans_jump_threshold = 10.0;
    int bad_dof_kmeans_max_iterations = 32;
    int bad_dof_kmeans_expand_neighbors = 0;
    double bad_dof_gmm_jump_threshold = 10.0;
    double bad_dof_gmm_tol = 1e-3;
    int bad_dof_gmm_max_iterations = 5;
    int bad_dof_gmm_expand_neighbors = 32;
    // MASSolver.cu:150 (a constant there); "MAS"/"true_residual_period" sets it, 0 = never
    int true_residual_period = 4;
    bool mas_mode() const { return solver == "MAS"; }
    // keys of this build
    std::string precision = "fp64";     // MAS bank inverses: "fp64" | "fp32" | "bf16"
    std::string inner_spmv = "auto";    // preconditioner-internal SpMV values: "auto" | "fp64" | "fp32" | "bf16"
    std::string base_preconditioner = "MAS";  // "MAS" | "AMG" | "Jacobi"
    int check_every = 8;                // PCG iterations between host polls of the device stop flag
    double damping = 1.8;               // omega = damping / lambda_max(B A) inside the B-P-B sweep
    int lanczos_steps = 20;          // CG-Lanczos steps of the lambda_max(B A) estimate
    int64_t max_dense_component = 16384;  // largest stiff component for the dense Cholesky
    // units above max_dense_component: "auto" = multifrontal sparse Cholesky, "dense" = rejected;
    // "sparse" = every unit outside the dense branch takes the sparse path
    std::string stiff_factorization = "auto";
    int sparse_leaf_size = 128;  // nested dissection: subgraphs this small are one front
    int64_t sparse_batch_rows = 4096;                  // fronts up to this size are factored in level batches
    int64_t sparse_batch_bytes = int64_t(4) << 30;     // workspace of one batch (its fronts' m^2 doubles)
    bool profile_kernels = false;       // time SpMV / MAS launches with HIP events
    bool true_residual_check = false;   // on convergence verify ||b - A x|| and restart PCG from it if needed
    std::string partitioner = "device";  // MAS partition: "device" (HIP kernels) | "host" (OpenMP)
    int max_coarse_levels = 4;           // MAS coarse levels (MAS_MAX_COARSE_LEVEL, MASPreconditioner.hpp:10-12)
    int partition_local_max = 8192;      // device partition: subsets this small finish in one workgroup (0: never)
    // level-0 bank solve of FP32 factors: "vector" (FP64 accumulation on the vector units) or
    // "mfma" (matrix cores, FP32 accumulators added into FP64 every mfma_flush products)
    std::string bank_apply = "vector";
    int mfma_flush = 4;
    int mfma_flush_steps() const { return bank_apply == "mfma" && precision == "fp32" ? mfma_flush : 0; }

    // FP32 matrix values in the preconditioner-internal SpMV (by default with reduced-
    // precision banks).  The stiff rows are then recomputed in FP64: an FP32 rounding
    // of a 1e8 contact entry exceeds the soft entries and would make the sweep
    // inconsistent with the FP64 stiff solve.
    // reduced-precision inner SpMV (FP32 x copy, stiff rows recomputed in FP64); bf16_inner:
    // its values in bfloat16 instead of FP32 (opt-in: fewer bytes, workload-dependent iterations)
    bool fp32_inner() const
    {
        return inner_spmv == "fp32" || inner_spmv == "bf16" || (inner_spmv == "auto" && precision != "fp64");
    }
    bool bf16_inner() const { return inner_spmv == "bf16"; }

    // Every key that shapes what factorize builds (buffers, layouts, selections, omega).
    // The others (tolerances, iteration caps, polling, timing, residual checks) only
    // steer a solve and may change between factorize and solve.
    auto factor_keys() const
    {
        return std::tie(solver, block_size, theta, do_hybrid_precond, select_bad_dofs_from_l1_norm,
                        conditioning_threshold, gmm_jump_threshold, gmm_tol, max_gmm_iterations,
                        decompose_subdomains, min_subdomain_size, max_subdomain_size, expand_subdomains,
                        max_dense_size, use_float_on_subdomains, lazy_partitioning, use_bad_dof_precond,
                        bad_dof_strategy, bad_dof_kmeans_search_fraction, bad_dof_kmeans_jump_threshold,
                        bad_dof_kmeans_max_iterations, bad_dof_kmeans_expand_neighbors, bad_dof_gmm_jump_threshold,
                        bad_dof_gmm_tol, bad_dof_gmm_max_iterations, bad_dof_gmm_expand_neighbors, precision,
                        inner_spmv, base_preconditioner, damping, lanczos_steps, max_dense_component, partitioner,
                        max_coarse_levels, partition_local_max, stiff_factorization, sparse_leaf_size,
                        sparse_batch_rows, sparse_batch_bytes);
    }
    bool same_factorization(const Config &o) const { return factor_keys() == o.factor_keys(); }
};

// Parse a configuration document; throws lss::Error(LSS_ERR_CONFIG) on bad input.
// `base` supplies the values of keys that are absent.
Config parse_config(const char *json, const Config &base);

}  // namespace lss
