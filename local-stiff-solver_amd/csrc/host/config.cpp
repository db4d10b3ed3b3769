#include "config.hpp"

#include <cmath>
#include <string>

#include "../common/json.hpp"
#include "../common/runtime.hpp"

namespace lss {
namespace {

[[noreturn]] void bad(const std::string &m) { fail(LSS_ERR_CONFIG, "invalid configuration: " + m); }

int64_t as_int(const json::Value &v, const std::string &k)
{
    if (v.kind != json::Value::Number || !v.is_int || std::floor(v.num) != v.num) bad(k + " must be an integer");
    return static_cast<int64_t>(v.num);
}
double as_float(const json::Value &v, const std::string &k)
{
    if (v.kind != json::Value::Number) bad(k + " must be a number");
    return v.num;
}
bool as_bool(const json::Value &v, const std::string &k)
{
    if (v.kind != json::Value::Bool) bad(k + " must be a boolean");
    return v.b;
}
std::string as_str(const json::Value &v, const std::string &k)
{
    if (v.kind != json::Value::String) bad(k + " must be a string");
    return v.str;
}

}  // namespace

Config validate(Config c);

Config parse_config(const char *text, const Config &base)
{
    Config c = base;
    if (!text) return c;
    std::string s(text);
    bool blank = true;
    for (char ch : s)
        if (!std::isspace(static_cast<unsigned char>(ch))) blank = false;
    if (blank) return c;
    json::Value doc;
    try {
        doc = json::parse(s);
    } catch (const std::exception &e) {
        bad(e.what());
    }
    if (doc.kind != json::Value::Object) bad("document must be an object");
    // solver selection (Solver.cpp:99-141): a name, or a priority list whose first
    // entry this library implements (GPUHybrid or MAS)
    if (const json::Value *sv = doc.find("solver")) {
        std::string sel;
        auto known = [](const std::string &n) { return n == "GPUHybrid" || n == "MAS"; };
        if (sv->kind == json::Value::String && known(sv->str)) sel = sv->str;
        if (sv->kind == json::Value::Array)
            for (auto &e : sv->arr)
                if (sel.empty() && e.kind == json::Value::String && known(e.str)) sel = e.str;
        if (sel.empty()) bad("\"solver\" must select GPUHybrid or MAS");
        if (sel != c.solver) {
            // a different solver starts from its own defaults
            c = Config{};
            c.solver = sel;
            if (sel == "MAS") {
                c.max_iter = 100000;
                c.relative_tolerance = 1e-6;
                c.absolute_tolerance = 1e-20;
                c.block_size = 3;
                c.lazy_partitioning = true;  // linear-solver-spec.json:776-781
            }
        }
    }
    if (c.mas_mode()) {
        // MASSolver::set_parameters reads the "MAS" object (MASSolver.cu:226-279, 760-768)
        if (const json::Value *m = doc.find("MAS")) {
            if (m->kind != json::Value::Object) bad("\"MAS\" must be an object");
            for (auto &kv : m->obj) {
                const std::string &k = kv.first;
                const json::Value &v = kv.second;
                if (k == "block_dim") c.block_size = int(as_int(v, k));
                else if (k == "max_iter") c.max_iter = int(as_int(v, k));
                else if (k == "relative_tolerance") c.relative_tolerance = as_float(v, k);
                else if (k == "absolute_tolerance") c.absolute_tolerance = as_float(v, k);
                else if (k == "lazy_partitioning") c.lazy_partitioning = as_bool(v, k);
                else if (k == "use_preconditioned_residual_norm") c.use_preconditioned_residual_norm = as_bool(v, k);
                else if (k == "use_bad_dof_precond") c.use_bad_dof_precond = as_bool(v, k);
                else if (k == "bad_dof_strategy") {
                    const std::string st = as_str(v, k);
                    c.bad_dof_strategy = (st == "kmean" || st == "gmm") ? st : "elbow";
                } else if (k == "bad_dof_kmeans_search_fraction") c.bad_dof_kmeans_search_fraction = as_float(v, k);
                else if (k == "bad_dof_kmeans_jump_threshold") c.bad_dof_kmeans_jump_threshold = as_float(v, k);
                else if (k == "bad_dof_kmeans_max_iterations") c.bad_dof_kmeans_max_iterations = int(as_int(v, k));
                else if (k == "bad_dof_kmeans_expand_neighbors") c.bad_dof_kmeans_expand_neighbors = int(as_int(v, k));
                else if (k == "bad_dof_gmm_jump_threshold") c.bad_dof_gmm_jump_threshold = as_float(v, k);
                else if (k == "bad_dof_gmm_tol") c.bad_dof_gmm_tol = as_float(v, k);
                else if (k == "bad_dof_gmm_max_iterations") c.bad_dof_gmm_max_iterations = int(as_int(v, k));
                else if (k == "bad_dof_gmm_expand_neighbors") c.bad_dof_gmm_expand_neighbors = int(as_int(v, k));
                else if (k == "precision") c.precision = as_str(v, k);
                else if (k == "inner_spmv") c.inner_spmv = as_str(v, k);
                else if (k == "check_every") c.check_every = int(as_int(v, k));
                else if (k == "true_residual_period") c.true_residual_period = int(as_int(v, k));
                else if (k == "max_dense_component") c.max_dense_component = as_int(v, k);
                else if (k == "stiff_factorization") c.stiff_factorization = as_str(v, k);
                else if (k == "sparse_leaf_size") c.sparse_leaf_size = int(as_int(v, k));
                else if (k == "sparse_batch_rows") c.sparse_batch_rows = as_int(v, k);
                else if (k == "sparse_batch_bytes") c.sparse_batch_bytes = as_int(v, k);
                else if (k == "profile_kernels") c.profile_kernels = as_bool(v, k);
                else if (k == "partitioner") c.partitioner = as_str(v, k);
                else if (k == "max_coarse_levels") c.max_coarse_levels = int(as_int(v, k));
                else if (k == "partition_local_max") c.partition_local_max = int(as_int(v, k));
                else if (k == "bank_apply") c.bank_apply = as_str(v, k);
                else if (k == "mfma_flush") c.mfma_flush = int(as_int(v, k));
                else bad("unknown key MAS/" + k);
            }
        }
    }
    const json::Value *g = c.mas_mode() ? nullptr : doc.find("GPUHybrid");
    if (!g) return validate(c);
    if (g->kind != json::Value::Object) bad("\"GPUHybrid\" must be an object");
    for (auto &kv : g->obj) {
        const std::string &k = kv.first;
        const json::Value &v = kv.second;
        if (k == "max_iter") c.max_iter = int(as_int(v, k));
        else if (k == "relative_tolerance") c.relative_tolerance = as_float(v, k);
        else if (k == "absolute_tolerance") c.absolute_tolerance = as_float(v, k);
        else if (k == "block_size") c.block_size = int(as_int(v, k));
        else if (k == "theta") c.theta = as_float(v, k);
        else if (k == "do_hybrid_precond") c.do_hybrid_precond = as_bool(v, k);
        else if (k == "select_bad_dofs_from_l1_norm") c.select_bad_dofs_from_l1_norm = as_bool(v, k);
        else if (k == "conditioning_threshold") c.conditioning_threshold = as_float(v, k);
        else if (k == "gmm_jump_threshold") c.gmm_jump_threshold = as_float(v, k);
        else if (k == "gmm_tol") c.gmm_tol = as_float(v, k);
        else if (k == "max_gmm_iterations") c.max_gmm_iterations = int(as_int(v, k));
        else if (k == "decompose_subdomains") c.decompose_subdomains = as_bool(v, k);
        else if (k == "min_subdomain_size") c.min_subdomain_size = as_int(v, k);
        else if (k == "max_subdomain_size") c.max_subdomain_size = as_int(v, k);
        else if (k == "expand_subdomains") c.expand_subdomains = as_bool(v, k);
        else if (k == "max_dense_size") c.max_dense_size = as_int(v, k);
        else if (k == "use_float_on_subdomains") c.use_float_on_subdomains = as_bool(v, k);
        else if (k == "precision") c.precision = as_str(v, k);
        else if (k == "inner_spmv") c.inner_spmv = as_str(v, k);
        else if (k == "base_preconditioner") c.base_preconditioner = as_str(v, k);
        else if (k == "check_every") c.check_every = int(as_int(v, k));
        else if (k == "damping") c.damping = as_float(v, k);
        else if (k == "lanczos_steps") c.lanczos_steps = int(as_int(v, k));
        else if (k == "max_dense_component") c.max_dense_component = as_int(v, k);
        else if (k == "stiff_factorization") c.stiff_factorization = as_str(v, k);
        else if (k == "sparse_leaf_size") c.sparse_leaf_size = int(as_int(v, k));
        else if (k == "sparse_batch_rows") c.sparse_batch_rows = as_int(v, k);
        else if (k == "sparse_batch_bytes") c.sparse_batch_bytes = as_int(v, k);
        else if (k == "profile_kernels") c.profile_kernels = as_bool(v, k);
        else if (k == "true_residual_check") c.true_residual_check = as_bool(v, k);
        else if (k == "partitioner") c.partitioner = as_str(v, k);
        else if (k == "max_coarse_levels") c.max_coarse_levels = int(as_int(v, k));
        else if (k == "partition_local_max") c.partition_local_max = int(as_int(v, k));
        else if (k == "lazy_partitioning") c.lazy_partitioning = as_bool(v, k);
        else if (k == "bank_apply") c.bank_apply = as_str(v, k);
        else if (k == "mfma_flush") c.mfma_flush = int(as_int(v, k));
        else bad("unknown key GPUHybrid/" + k);
    }
    return validate(c);
}

Config validate(Config c)
{
    if (c.block_size < 1 || c.block_size > 3) bad("block_size must be 1, 2 or 3");
    if (c.max_iter < 0) bad("max_iter must be >= 0");
    if (c.precision != "fp64" && c.precision != "fp32" && c.precision != "bf16" && c.precision != "fp16")
        bad("precision must be fp64|fp32|bf16|fp16");
    if (c.inner_spmv != "auto" && c.inner_spmv != "fp64" && c.inner_spmv != "fp32" && c.inner_spmv != "bf16")
        bad("inner_spmv must be auto|fp64|fp32|bf16");
    if (c.base_preconditioner != "MAS" && c.base_preconditioner != "Jacobi" && c.base_preconditioner != "AMG")
        bad("base_preconditioner must be MAS|AMG|Jacobi");
    if (c.bank_apply != "vector" && c.bank_apply != "mfma") bad("bank_apply must be vector|mfma");
    if (c.stiff_factorization != "auto" && c.stiff_factorization != "dense" && c.stiff_factorization != "sparse")
        bad("stiff_factorization must be auto|dense|sparse");
    if (c.sparse_leaf_size < 1) bad("sparse_leaf_size must be >= 1");
    if (c.sparse_batch_rows < 0) bad("sparse_batch_rows must be >= 0");
    if (c.sparse_batch_bytes < 0) bad("sparse_batch_bytes must be >= 0");
    if (c.mfma_flush != 4 && c.mfma_flush != 8 && c.mfma_flush != 16 && c.mfma_flush != 32)
        bad("mfma_flush must be 4, 8, 16 or 32");
    if (c.check_every < 1) bad("check_every must be >= 1");
    if (c.true_residual_period < 0) bad("true_residual_period must be >= 0 (0: never)");
    if (c.partitioner != "device" && c.partitioner != "host") bad("partitioner must be device|host");
    if (c.max_coarse_levels < 0 || c.max_coarse_levels > 4) bad("max_coarse_levels must be in 0..4");
    if (c.partition_local_max < 0 || c.partition_local_max > 8192) bad("partition_local_max must be in 0..8192");
    if (!(c.damping > 0.0) || c.damping >= 4.0) bad("damping must be in (0, 4)");
    if (c.lanczos_steps < 1) bad("lanczos_steps must be >= 1");
    if (c.bad_dof_gmm_max_iterations < 0 || c.bad_dof_gmm_expand_neighbors < 0 ||
        c.bad_dof_kmeans_expand_neighbors < 0 || c.bad_dof_kmeans_max_iterations < 0)
        bad("bad_dof iteration counts and expand_neighbors must be >= 0");
    return c;
}

}  // namespace lss
