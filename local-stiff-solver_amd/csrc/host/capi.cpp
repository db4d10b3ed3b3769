// C ABI (include/lss.h) over HybridSolver.  Every entry point catches all
// exceptions and returns a status code; the message is kept per handle.
#include <cstring>
#include <exception>
#include <memory>
#include <string>

#include "lss.h"
#include "config.hpp"
#include "solver.hpp"
#include "sparse_stiff.hpp"
#include "../kernels/kernels.hpp"

struct lss_solver {
    std::unique_ptr<lss::HybridSolver> impl;
    std::string last_error;
};

namespace {

// message of the last failed handle-free call on this thread
thread_local std::string g_error;

template <class F>
int guarded_free(F &&f)
{
    try {
        f();
        g_error.clear();
        return LSS_OK;
    } catch (const lss::Error &e) {
        g_error = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_error = e.what();
        return LSS_ERR_HIP;
    }
}

template <class F>
int guarded(lss_solver *s, F &&f)
{
    if (!s) return LSS_ERR_ARG;
    try {
        f();
        s->last_error.clear();
        return LSS_OK;
    } catch (const lss::Error &e) {
        s->last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        s->last_error = "host allocation failed";
        return LSS_ERR_OOM;
    } catch (const std::exception &e) {
        s->last_error = e.what();
        return LSS_ERR_HIP;
    }
}

}  // namespace

extern "C" {

int lss_create(const char *json, lss_solver **out)
{
    if (!out) return LSS_ERR_ARG;
    *out = nullptr;
    auto s = new (std::nothrow) lss_solver();
    if (!s) return LSS_ERR_OOM;
    int rc = guarded(s, [&] {
        lss::Config cfg = lss::parse_config(json, lss::Config{});
        s->impl = std::make_unique<lss::HybridSolver>(cfg);
    });
    if (rc != LSS_OK) {
        // keep the handle so the caller can read the message, but mark it unusable;
        // the message is also the thread's handle-free error (lss_last_error(NULL))
        g_error = s->last_error;
        *out = s;
        return rc;
    }
    *out = s;
    return LSS_OK;
}

int lss_set_parameters(lss_solver *s, const char *json)
{
    return guarded(s, [&] {
        if (!s->impl) lss::fail(LSS_ERR_CONFIG, "solver was not created");
        s->impl->set_config(lss::parse_config(json, s->impl->config()));
    });
}

int lss_set_problematic_dofs(lss_solver *s, const int32_t *dofs, int64_t n)
{
    return guarded(s, [&] {
        if (!s->impl) lss::fail(LSS_ERR_CONFIG, "solver was not created");
        if (n < 0 || (n > 0 && !dofs)) lss::fail(LSS_ERR_ARG, "invalid DOF list");
        s->impl->set_problematic_dofs(dofs, n);
    });
}

int lss_analyze_pattern(lss_solver *s, int32_t, int64_t, const int32_t *, const int32_t *)
{
    return guarded(s, [&] {
        if (!s->impl) lss::fail(LSS_ERR_CONFIG, "solver was not created");
    });
}

int lss_factorize_csc(lss_solver *s, int32_t n, int64_t nnz, const int32_t *colptr, const int32_t *rowidx,
                      const double *vals)
{
    return guarded(s, [&] {
        if (!s->impl) lss::fail(LSS_ERR_CONFIG, "solver was not created");
        if (n <= 0 || nnz < 0 || !colptr || (nnz > 0 && (!rowidx || !vals))) lss::fail(LSS_ERR_ARG, "invalid matrix");
        if (int64_t(colptr[n]) != nnz) lss::fail(LSS_ERR_SIZE, "colptr[n] does not match nnz");
        s->impl->factorize_host(n, nnz, colptr, rowidx, vals);
    });
}

int lss_factorize_csc_device(lss_solver *s, int32_t n, int64_t nnz, const int32_t *d_colptr, const int32_t *d_rowidx,
                             const double *d_vals, void *stream)
{
    return guarded(s, [&] {
        if (!s->impl) lss::fail(LSS_ERR_CONFIG, "solver was not created");
        if (n <= 0 || nnz < 0 || !d_colptr || (nnz > 0 && (!d_rowidx || !d_vals)))
            lss::fail(LSS_ERR_ARG, "invalid matrix");
        s->impl->factorize_device(n, nnz, d_colptr, d_rowidx, d_vals, static_cast<hipStream_t>(stream));
    });
}

int lss_solve(lss_solver *s, const double *b, double *x)
{
    return guarded(s, [&] {
        if (!s->impl) lss::fail(LSS_ERR_CONFIG, "solver was not created");
        if (!b || !x) lss::fail(LSS_ERR_ARG, "null vector");
        s->impl->solve_host(b, x);
    });
}

int lss_solve_device(lss_solver *s, const double *d_b, double *d_x, void *stream)
{
    return guarded(s, [&] {
        if (!s->impl) lss::fail(LSS_ERR_CONFIG, "solver was not created");
        if (!d_b || !d_x) lss::fail(LSS_ERR_ARG, "null vector");
        s->impl->solve_device(d_b, d_x, static_cast<hipStream_t>(stream));
    });
}

int lss_solve_multi_device(lss_solver *s, int32_t k, const double *d_B, double *d_X, void *stream)
{
    return guarded(s, [&] {
        if (!s->impl) lss::fail(LSS_ERR_CONFIG, "solver was not created");
        if (k < 0 || (k > 0 && (!d_B || !d_X))) lss::fail(LSS_ERR_ARG, "invalid right-hand sides");
        s->impl->solve_multi_device(k, d_B, d_X, static_cast<hipStream_t>(stream));
    });
}

int lss_get_info(const lss_solver *s, char *buf, size_t cap, size_t *needed)
{
    if (!s || !s->impl) return LSS_ERR_ARG;
    try {
        const std::string j = s->impl->info_json();
        if (needed) *needed = j.size() + 1;
        if (buf && cap > 0) {
            const size_t m = std::min(cap - 1, j.size());
            std::memcpy(buf, j.data(), m);
            buf[m] = '\0';
        }
        return LSS_OK;
    } catch (...) {
        return LSS_ERR_HIP;
    }
}

const char *lss_name(const lss_solver *s)
{
    if (!s || !s->impl) return "";
    return s->impl->name();
}

int lss_get_stiff_dofs(const lss_solver *s, int32_t *dofs, int64_t cap, int64_t *n_dofs, int64_t *comp_off,
                       int64_t comp_cap, int64_t *n_comp)
{
    if (!s || !s->impl) return LSS_ERR_ARG;
    const auto &comps = s->impl->components();
    int64_t w = 0;
    if (comp_off && comp_cap > 0) comp_off[0] = 0;
    for (size_t c = 0; c < comps.size(); ++c) {
        for (int32_t d : comps[c]) {
            if (dofs && w < cap) dofs[w] = d;
            ++w;
        }
        if (comp_off && int64_t(c + 1) < comp_cap) comp_off[c + 1] = w;
    }
    if (n_dofs) *n_dofs = w;
    if (n_comp) *n_comp = int64_t(comps.size());
    return LSS_OK;
}

const char *lss_last_error(const lss_solver *s)
{
    if (!s) return g_error.empty() ? "null handle" : g_error.c_str();
    return s->last_error.c_str();
}

int lss_adjoint_pattern_grad(int32_t n, int64_t nnz, const int32_t *d_colptr, const int32_t *d_rowidx,
                             const double *d_lam, const double *d_x, double *d_grad, void *stream)
{
    return guarded_free([&] {
        if (n < 0 || nnz < 0 || (n > 0 && (!d_colptr || !d_lam || !d_x)) || (nnz > 0 && (!d_rowidx || !d_grad)))
            lss::fail(LSS_ERR_ARG, "invalid arguments");
        lss::k::pattern_grad(n, d_colptr, d_rowidx, d_lam, d_x, d_grad, static_cast<hipStream_t>(stream));
    });
}

int lss_adjoint_chain(int64_t nseg, const int64_t *d_seg_ptr, const int32_t *d_slot, const double *d_coef,
                      const double *d_grad_vals, double *d_out, void *stream)
{
    return guarded_free([&] {
        if (nseg < 0 || (nseg > 0 && (!d_seg_ptr || !d_slot || !d_coef || !d_grad_vals || !d_out)))
            lss::fail(LSS_ERR_ARG, "invalid arguments");
        lss::k::segment_dot(nseg, d_seg_ptr, d_slot, d_coef, d_grad_vals, d_out, static_cast<hipStream_t>(stream));
    });
}

int lss_residual_norm(int32_t n, int64_t nnz, const int32_t *d_colptr, const int32_t *d_rowidx, const double *d_vals,
                      const double *d_b, const double *d_x, double *d_out, void *stream)
{
    return guarded_free([&] {
        if (n < 0 || nnz < 0 || !d_out || (n > 0 && (!d_colptr || !d_b || !d_x)) ||
            (nnz > 0 && (!d_rowidx || !d_vals)))
            lss::fail(LSS_ERR_ARG, "invalid arguments");
        hipStream_t s = static_cast<hipStream_t>(stream);
        void *part = nullptr;
        LSS_HIP(hipMallocAsync(&part, size_t(4) * lss::k::residual_groups(n) * sizeof(double), s));
        try {
            lss::k::residual_norm_dd(n, d_colptr, d_rowidx, d_vals, d_b, d_x, static_cast<double *>(part), d_out, s);
        } catch (...) {
            (void)hipFreeAsync(part, s);
            throw;
        }
        LSS_HIP(hipFreeAsync(part, s));
    });
}

int lss_residual_norm_noalloc(int32_t n, int64_t nnz, const int32_t *d_colptr, const int32_t *d_rowidx,
                              const double *d_vals, const double *d_b, const double *d_x, double *d_out, void *stream)
{
    return guarded_free([&] {
        if (n < 0 || nnz < 0 || !d_out || (n > 0 && (!d_colptr || !d_b || !d_x)) ||
            (nnz > 0 && (!d_rowidx || !d_vals)))
            lss::fail(LSS_ERR_ARG, "invalid arguments");
        // the grid is capped, so every size fits the static partials array; the check
        // guards the invariant should the cap and the array ever disagree
        if (lss::k::residual_groups(n) > lss::k::kResidualGroups)
            lss::fail(LSS_ERR_SIZE, "residual needs more partial sums than the static array holds");
        lss::k::residual_norm_dd(n, d_colptr, d_rowidx, d_vals, d_b, d_x, lss::k::residual_scratch(), d_out,
                                 static_cast<hipStream_t>(stream));
    });
}

void lss_destroy(lss_solver *s) { delete s; }

int lss_debug_nested_dissection(int32_t m, const int32_t *adj_ptr, const int32_t *adj_idx, int32_t leaf,
                                int32_t *order, int32_t *nfronts, int64_t *nstruct, int32_t *front_p0,
                                int32_t *front_np, int32_t *front_parent, int64_t *struct_ptr, int32_t *struct_idx)
{
    return guarded_free([&] {
        if (m < 0 || !adj_ptr || (m > 0 && adj_ptr[m] > 0 && !adj_idx) || !nfronts || !nstruct)
            lss::fail(LSS_ERR_ARG, "invalid arguments");
        for (int32_t v = 0; v < m; ++v)
            if (adj_ptr[v + 1] < adj_ptr[v]) lss::fail(LSS_ERR_ARG, "adj_ptr must be non-decreasing");
        for (int32_t e = 0; e < (m > 0 ? adj_ptr[m] : 0); ++e)
            if (adj_idx[e] < 0 || adj_idx[e] >= m) lss::fail(LSS_ERR_ARG, "adj_idx out of range");
        const lss::FrontTree T = lss::nested_dissection(m, adj_ptr, adj_idx, leaf);
        *nfronts = T.fronts();
        *nstruct = int64_t(T.uidx.size());
        if (!order) return;
        if (!front_p0 || !front_np || !front_parent || !struct_ptr || (!T.uidx.empty() && !struct_idx))
            lss::fail(LSS_ERR_ARG, "output arrays missing");
        std::memcpy(order, T.order.data(), size_t(m) * sizeof(int32_t));
        std::memcpy(front_p0, T.p0.data(), T.p0.size() * sizeof(int32_t));
        std::memcpy(front_np, T.np.data(), T.np.size() * sizeof(int32_t));
        std::memcpy(front_parent, T.parent.data(), T.parent.size() * sizeof(int32_t));
        std::memcpy(struct_ptr, T.uptr.data(), T.uptr.size() * sizeof(int64_t));
        if (!T.uidx.empty()) std::memcpy(struct_idx, T.uidx.data(), T.uidx.size() * sizeof(int32_t));
    });
}

int lss_debug_row_norms(const lss_solver *s, double *out, int64_t n)
{
    return guarded(const_cast<lss_solver *>(s), [&] {
        if (!s->impl || n != s->impl->size()) lss::fail(LSS_ERR_SIZE, "size mismatch");
        s->impl->debug_row_norms(out);
    });
}

int lss_debug_gmm(const lss_solver *s, double *params, int32_t *iters, int32_t *jumped, int64_t *selected)
{
    return guarded(const_cast<lss_solver *>(s), [&] {
        if (!s->impl) lss::fail(LSS_ERR_CONFIG, "solver was not created");
        const auto &g = s->impl->gmm();
        const double v[8] = {g.global_mean, g.global_var, g.mean0, g.mean1, g.var0, g.var1, g.w0, g.w1};
        if (params) std::memcpy(params, v, sizeof(v));
        if (iters) *iters = g.iters;
        if (jumped) *jumped = g.jumped;
        if (selected) *selected = g.selected;
    });
}

int lss_debug_partition(const lss_solver *s, int32_t *perm, int64_t nb, int32_t *part_off, int64_t cap,
                        int32_t *nparts)
{
    return guarded(const_cast<lss_solver *>(s), [&] {
        if (!s->impl) lss::fail(LSS_ERR_CONFIG, "solver was not created");
        const auto &p = s->impl->perm();
        const auto &o = s->impl->part_off();
        if (nparts) *nparts = o.empty() ? 0 : int32_t(o.size()) - 1;
        if (perm) {
            if (nb != int64_t(p.size())) lss::fail(LSS_ERR_SIZE, "perm size mismatch");
            std::memcpy(perm, p.data(), p.size() * sizeof(int32_t));
        }
        if (part_off) {
            if (cap < int64_t(o.size())) lss::fail(LSS_ERR_SIZE, "part_off capacity too small");
            std::memcpy(part_off, o.data(), o.size() * sizeof(int32_t));
        }
    });
}

int lss_debug_amg(const lss_solver *s, int32_t level, int32_t *nlevels, int32_t *n, int64_t *nnz_a, int64_t *nnz_p,
                  int8_t *cf, int64_t *a_ptr, int32_t *a_col, double *a_val, int64_t *p_ptr, int32_t *p_col,
                  double *p_val)
{
    return guarded(const_cast<lss_solver *>(s), [&] {
        if (!s->impl) lss::fail(LSS_ERR_CONFIG, "solver was not created");
        s->impl->debug_amg(level, nlevels, n, nnz_a, nnz_p, cf, a_ptr, a_col, a_val, p_ptr, p_col, p_val);
    });
}

int lss_debug_spmv(lss_solver *s, const double *x, double *y)
{
    return guarded(s, [&] {
        if (!s->impl) lss::fail(LSS_ERR_CONFIG, "solver was not created");
        s->impl->debug_spmv(x, y);
    });
}

int lss_debug_base_apply(lss_solver *s, const double *r, double *z)
{
    return guarded(s, [&] {
        if (!s->impl) lss::fail(LSS_ERR_CONFIG, "solver was not created");
        s->impl->debug_base(r, z);
    });
}

int lss_debug_precond_apply(lss_solver *s, const double *r, double *z)
{
    return guarded(s, [&] {
        if (!s->impl) lss::fail(LSS_ERR_CONFIG, "solver was not created");
        s->impl->debug_precond(r, z);
    });
}

}  // extern "C"
