// The reference's standalone "MAS" solver (SRC/MASSolver.cu) on the same handle:
// PCG with B = MAS and, with use_bad_dof_precond, the additive stiff correction
// z = B r + S^T (S A S^T)^{-1} S r (MASSolver.cu:583-615) whose set S comes from one
// of three strategies on the partition-permuted operator:
//   elbow  (MAS/BadDOFPreconditioner.cu:61-100): |a_ii| sorted ascending, cut at
//          the largest deviation of log|a_ii| from the chord between its extremes;
//   kmean  (MAS/BadDofKMeanPreconditioner.cu:63-172): max |a_ij| (j != i), Lloyd's
//          two-means on the top fraction, ratio test, optional index windows;
//   gmm    (MAS/BadDofGMMPreconditioner.cu:86-258): row l1 norms, two-Gaussian EM
//          with a variance floor, the cluster with the larger |mean|, ratio test,
//          index windows of +-expand_neighbors positions.
// Positions count the non-padding scalar rows in permuted order.  The reference
// factorizes S A S^T as one sparse matrix (cuDSS); its connected components are
// independent blocks, so each is factorized densely here (same solution).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>

#include "solver.hpp"
#include "stiff_select.hpp"

namespace lss {
namespace {

using clk = std::chrono::steady_clock;
constexpr double kPi = 3.14159265358979323846;

double rel_change(double now, double before)
{
    // relative_change, BadDofGMMPreconditioner.cu:75-79
    return std::fabs(now - before) / std::max(std::fabs(before), 1e-300);
}

// union of the windows [p - e, p + e] around the given positions (BadDofGMMPreconditioner.cu:236-255)
std::vector<int64_t> expand_windows(const std::vector<int64_t> &pos, int64_t n, int e)
{
    if (e <= 0) return pos;
    std::vector<uint8_t> mark(size_t(n), 0);
    for (int64_t p : pos)
        for (int64_t j = std::max<int64_t>(0, p - e); j <= std::min<int64_t>(n - 1, p + e); ++j) mark[j] = 1;
    std::vector<int64_t> out;
    for (int64_t j = 0; j < n; ++j)
        if (mark[j]) out.push_back(j);
    return out;
}

template <class T>
std::vector<T> to_host(const T *d, size_t n, hipStream_t s)
{
    std::vector<T> h(n);
    if (n) LSS_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s));
    LSS_HIP(hipStreamSynchronize(s));
    return h;
}

double d2h(const double *d, hipStream_t s)
{
    double v = 0.0;
    LSS_HIP(hipMemcpyAsync(&v, d, sizeof(double), hipMemcpyDeviceToHost, s));
    LSS_HIP(hipStreamSynchronize(s));
    return v;
}

}  // namespace

// select_bad_dofs_from_l1_row_norm_gmm (BadDofGMMPreconditioner.cu:86-258) with the sums
// in the fixed reduction tree of the GPUHybrid selection.
std::vector<int64_t> HybridSolver::mas_gmm_positions(int64_t n, const double *x, hipStream_t s)
{
    MasSelection &st = mas_sel_;
    if (n < 2) return {};
    DBuf<double> g0(n), g1(n), part(4 * k::kCsumChunks), sums(8);
    k::canonical_sum(n, x, part.get(), sums.get(), s);
    k::minmax(n, x, part.get(), sums.get() + 1, s);
    double h[8];
    LSS_HIP(hipMemcpyAsync(h, sums.get(), 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    LSS_HIP(hipStreamSynchronize(s));
    st.global_mean = h[0] / double(n);
    k::gmm_var_global(n, x, st.global_mean, part.get(), s);
    k::csum_finish(1, part.get(), sums.get(), s);
    st.global_var = d2h(sums.get(), s) / double(n);
    double mean0 = h[1], mean1 = h[2];
    constexpr double var_reg = 1e-6;
    double var0 = std::max(st.global_var, var_reg), var1 = std::max(st.global_var, var_reg);
    double w0 = 0.5, w1 = 0.5;
    st.mean0 = mean0, st.mean1 = mean1, st.var0 = var0, st.var1 = var1;
    if (mean0 == mean1 || !std::isfinite(st.global_var)) return {};
    const int iterations = std::max(1, cfg_.bad_dof_gmm_max_iterations);
    int it = 0;
    for (; it < iterations; ++it) {
        k::GmmParams p;
        p.mean0 = mean0;
        p.mean1 = mean1;
        p.var0 = var0;
        p.var1 = var1;
        p.log_w0 = std::log(std::max(w0, std::numeric_limits<double>::min()));
        p.log_w1 = std::log(std::max(w1, std::numeric_limits<double>::min()));
        p.c0 = -0.5 * std::log(2.0 * kPi * var0);
        p.c1 = -0.5 * std::log(2.0 * kPi * var1);
        k::gmm_estep(n, x, p, g0.get(), g1.get(), part.get(), s);
        k::csum_finish(4, part.get(), sums.get(), s);
        LSS_HIP(hipMemcpyAsync(h, sums.get(), 4 * sizeof(double), hipMemcpyDeviceToHost, s));
        LSS_HIP(hipStreamSynchronize(s));
        const double sum_g0 = h[0], sum_g1 = h[1];
        if (sum_g0 <= 0.0 || sum_g1 <= 0.0) break;
        w0 = sum_g0 / double(n);
        w1 = sum_g1 / double(n);
        const double om0 = mean0, om1 = mean1, ov0 = var0, ov1 = var1;
        mean0 = h[2] / sum_g0;
        mean1 = h[3] / sum_g1;
        k::gmm_var(n, x, g0.get(), g1.get(), mean0, mean1, part.get(), s);
        k::csum_finish(2, part.get(), sums.get(), s);
        LSS_HIP(hipMemcpyAsync(h, sums.get(), 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        LSS_HIP(hipStreamSynchronize(s));
        var0 = h[0] / sum_g0 + var_reg;
        var1 = h[1] / sum_g1 + var_reg;
        st.mean0 = mean0, st.mean1 = mean1, st.var0 = var0, st.var1 = var1;
        const double tol = cfg_.bad_dof_gmm_tol;
        if (rel_change(mean0, om0) < tol && rel_change(mean1, om1) < tol && rel_change(var0, ov0) < tol &&
            rel_change(var1, ov1) < tol)
            break;
    }
    st.iters = it;
    const int high = std::fabs(mean0) > std::fabs(mean1) ? 0 : 1;
    const double high_mean = high == 0 ? mean0 : mean1, low_mean = high == 0 ? mean1 : mean0;
    st.ratio = std::fabs(high_mean) / std::fabs(low_mean);
    if (std::isnan(st.ratio) || st.ratio < cfg_.bad_dof_gmm_jump_threshold) return {};
    DBuf<uint8_t> fl(n);
    k::gmm_flags(n, g0.get(), g1.get(), fl.get(), s, high);
    const auto sel = k::flagged_indices(int32_t(n), fl.get(), s);
    const std::vector<int64_t> pos(sel.begin(), sel.end());
    st.selected = int64_t(pos.size());
    return expand_windows(pos, n, cfg_.bad_dof_gmm_expand_neighbors);
}

void HybridSolver::select_mas_bad(int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals,
                                  hipStream_t s)
{
    mas_sel_ = MasSelection{};
    mas_sel_.strategy = cfg_.bad_dof_strategy;
    const int64_t np = int64_t(nb_) * bs_;
    const int npad = int(np - n);
    const int32_t pblock = npad ? perm_[nb_ - 1] : -1;
    const int mode = cfg_.bad_dof_strategy == "gmm" ? 0 : (cfg_.bad_dof_strategy == "kmean" ? 2 : 1);
    DBuf<double> ind(np);
    k::mas_indicator(mode, nb_, bs_, d_rowptr_.get(), d_cols_.get(), d_bvals_.get(), pblock, npad, ind.get(), s);
    // positions: the non-padding permuted scalar rows in order
    std::vector<int64_t> pos2p;
    const int64_t m = n;
    DBuf<double> xs;
    const double *x = ind.get();
    if (npad) {
        auto h = to_host(ind.get(), size_t(np), s);
        std::vector<double> c;
        for (int64_t p = 0; p < np; ++p) {
            const int64_t I = p / bs_, a = p % bs_;
            if (I == pblock && a >= bs_ - npad) continue;
            pos2p.push_back(p);
            c.push_back(h[p]);
        }
        xs.upload(c.data(), c.size(), s);
        x = xs.get();
    }
    std::vector<int64_t> pos;
    if (mode == 0) {
        pos = mas_gmm_positions(m, x, s);
    } else if (m >= 1) {
        DBuf<double> keys(m);
        DBuf<int32_t> order(m);
        k::sort_positions(int32_t(m), x, mode == 2, keys.get(), order.get(), s);
        const auto hk = to_host(keys.get(), size_t(m), s);
        const auto ho = to_host(order.get(), size_t(m), s);
        if (mode == 1) {
            // select_bad_dofs_from_diagonal, BadDOFPreconditioner.cu:61-100
            int64_t cutoff = 0;
            if (m > 1) {
                std::vector<double> lg(static_cast<size_t>(m));
                for (int64_t i = 0; i < m; ++i) lg[i] = std::log(hk[i]);
                const double mn = lg.front(), mx = lg.back();
                double best = -1.0;
                for (int64_t i = 0; i < m; ++i) {
                    const double expected = (mx - mn) / (double(m) - 1.0) * double(i) + mn;
                    const double dev = std::fabs(expected - lg[i]);
                    if (dev > best) {
                        best = dev;
                        cutoff = i;
                    }
                }
            }
            mas_sel_.cutoff = cutoff;
            for (int64_t i = cutoff; i < m; ++i) pos.push_back(ho[i]);
            mas_sel_.selected = int64_t(pos.size());
        } else if (m >= 2) {
            // select_bad_dofs_from_diagonal_kmeans, BadDofKMeanPreconditioner.cu:63-172
            const int64_t sub = std::clamp<int64_t>(int64_t(double(m) * cfg_.bad_dof_kmeans_search_fraction), 2, m);
            double c[2] = {hk[0], hk[sub - 1]};
            int it = 0;
            for (; it < cfg_.bad_dof_kmeans_max_iterations; ++it) {
                double sum[2] = {0.0, 0.0};
                int64_t cnt[2] = {0, 0};
                for (int64_t i = 0; i < sub; ++i) {
                    const int cid = std::fabs(hk[i] - c[0]) < std::fabs(hk[i] - c[1]) ? 1 : 0;
                    cnt[cid]++;
                    sum[cid] += hk[i];
                }
                const double n0 = cnt[0] > 0 ? sum[0] / double(cnt[0]) : c[0];
                const double n1 = cnt[1] > 1 ? sum[1] / double(cnt[1]) : c[1];
                if (n0 == c[0] && n1 == c[1]) break;
                c[0] = n0;
                c[1] = n1;
            }
            if (c[0] < c[1]) std::swap(c[0], c[1]);
            mas_sel_.iters = it;
            mas_sel_.c0 = c[0];
            mas_sel_.c1 = c[1];
            mas_sel_.ratio = c[0] / c[1];
            if (!std::isnan(mas_sel_.ratio) && mas_sel_.ratio >= cfg_.bad_dof_kmeans_jump_threshold) {
                const double sep = 0.5 * (c[0] + c[1]);
                for (int64_t i = 0; i < sub; ++i)
                    if (hk[i] > sep) pos.push_back(ho[i]);
                mas_sel_.selected = int64_t(pos.size());
                std::sort(pos.begin(), pos.end());
                pos = expand_windows(pos, m, cfg_.bad_dof_kmeans_expand_neighbors);
            }
        }
    }
    mas_sel_.expanded = int64_t(pos.size());
    n_selected_ = mas_sel_.selected;
    n_filtered_ = n_expanded_ = int64_t(pos.size());
    if (pos.empty()) return;
    // positions -> permuted scalars -> original DOFs
    std::vector<int32_t> inv(nb_);
    for (int32_t I = 0; I < nb_; ++I) inv[perm_[I]] = I;
    std::vector<int32_t> S;
    S.reserve(pos.size());
    for (int64_t q : pos) {
        const int64_t p = npad ? pos2p[q] : q;
        S.push_back(int32_t(int64_t(inv[p / bs_]) * bs_ + p % bs_));
    }
    std::sort(S.begin(), S.end());
    RowSet rows = fetch_rows(S, ptr, idx, vals, false, s);
    comps_ = decompose_subdomains(n, rows, S, std::numeric_limits<int64_t>::max());
}

// MASSolverImpl::pcg_solve, MASSolver.cu:618-744: x0 = 0 (MASSolver.cu:500-504), the
// true residual every true_residual_period iterations, convergence tested every 10
// iterations on ||r||^2 (or r.z) against the initial value.
void HybridSolver::run_pcg_mas(hipStream_t s)
{
    const int64_t np = int64_t(nb_) * bs_;
    dev::PcgState *st = st_.get();
    DBuf<double> sc(1);
    std::memset(h_st_, 0, sizeof(dev::PcgState));
    LSS_HIP(hipMemcpyAsync(st, h_st_, sizeof(dev::PcgState), hipMemcpyHostToDevice, s));
    k::fill(np, xp_.get(), 0.0, s);
    k::copy(np, bp_.get(), r_.get(), s);  // r = b - A 0
    precond_apply(r_.get(), z_.get(), nullptr, nullptr, s);
    k::copy(np, z_.get(), p_.get(), s);
    k::dot(np, r_.get(), z_.get(), partials_.get(), &st->gamma, s);
    const double rz0 = d2h(&st->gamma, s);
    if (std::isnan(rz0) || !std::isfinite(rz0)) fail(LSS_ERR_NOT_SPD, "[MAS] Invalid initial residual.");
    double rr0 = rz0;
    if (!cfg_.use_preconditioned_residual_norm) {
        k::dot(np, r_.get(), r_.get(), partials_.get(), sc.get(), s);
        rr0 = d2h(sc.get(), s);
    }
    const double rtol2 = cfg_.relative_tolerance * cfg_.relative_tolerance;
    const double atol2 = cfg_.absolute_tolerance * cfg_.absolute_tolerance;
    mas_status_ = 0;
    mas_residual_ = 0.0;
    int iters = 0;
    for (int kk = 1; kk <= cfg_.max_iter; ++kk) {
        k::spmv_pq(sell_, p_.get(), q_.get(), st, partials_.get(), kk - 1, s, 1, 0, dev::kFinishPqMas);
        k::axpy_alpha(np, 1.0, p_.get(), xp_.get(), st, s);
        if (cfg_.true_residual_period > 0 && kk % cfg_.true_residual_period == 0)
            k::spmv_residual(sell_, xp_.get(), bp_.get(), r_.get(), nullptr, s);
        else
            k::axpy_alpha(np, -1.0, q_.get(), r_.get(), st, s);
        precond_apply(r_.get(), z_.get(), nullptr, nullptr, s);
        const int g = k::dot_partials(np, r_.get(), z_.get(), partials_.get(), s);
        k::finish(g, partials_.get(), dev::kFinishRzMas, st, nullptr, 0, s);
        iters = kk;
        bool converged = false;
        if (kk % 10 == 0) {
            double rr;
            if (cfg_.use_preconditioned_residual_norm) {
                rr = d2h(&st->gamma, s);
            } else {
                k::dot(np, r_.get(), r_.get(), partials_.get(), sc.get(), s);
                rr = d2h(sc.get(), s);
            }
            mas_residual_ = std::sqrt(rr);
            if (rr <= rtol2 * rr0 || rr <= atol2) {
                mas_status_ = rr <= atol2 ? 2 : 1;
                converged = true;
            }
        }
        if (converged) break;
        k::pcg_update_p(np, p_.get(), z_.get(), st, s);
    }
    if (iters == cfg_.max_iter) mas_status_ = 3;
    num_iterations_ = iters;
    stop_code_ = mas_status_ == 1 ? dev::kConvergedRel : (mas_status_ == 2 ? dev::kConvergedAbs : 0);
    col_iters_.assign(1, iters);
    col_stop_.assign(1, stop_code_);
    // true residual ||A x - b||
    k::spmv_residual(sell_, xp_.get(), bp_.get(), q_.get(), nullptr, s);
    k::dot(np, q_.get(), q_.get(), partials_.get(), sc.get(), s);
    final_res_norm_ = std::sqrt(d2h(sc.get(), s));
    col_res_.assign(1, final_res_norm_);
}

}  // namespace lss
