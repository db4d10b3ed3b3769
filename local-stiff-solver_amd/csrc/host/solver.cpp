#include "solver.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <future>
#include <numeric>
#include <sstream>
#include <thread>

#include "../kernels/util_kernels.hpp"
#include "topology.hpp"

namespace lss {

// damping of the base preconditioner after a breakdown (omega * lambda_max < 2 holds
// while the Lanczos estimate is within 25 % of lambda_max)
constexpr double kSafeDamping = 1.5;
namespace {

using clk = std::chrono::steady_clock;
double since(clk::time_point t0) { return std::chrono::duration<double>(clk::now() - t0).count(); }

constexpr double kPi = 3.14159265358979323846;
constexpr size_t kPartialsCap = size_t(1) << 20;

template <class T>
std::vector<T> to_host(const DBuf<T> &d, size_t n, hipStream_t s)
{
    std::vector<T> h(n);
    d.download(h.data(), n, s);
    LSS_HIP(hipStreamSynchronize(s));
    return h;
}

double d2h_scalar(const double *d, hipStream_t s)
{
    double v = 0;
    LSS_HIP(hipMemcpyAsync(&v, d, sizeof(double), hipMemcpyDeviceToHost, s));
    LSS_HIP(hipStreamSynchronize(s));
    return v;
}

}  // namespace

HybridSolver::HybridSolver(const Config &cfg) : cfg_(cfg)
{
    // a blocking stream: work on it is ordered with the legacy default stream, so
    // callers that pass NULL (e.g. torch's default stream) need no extra sync
    LSS_HIP(hipStreamCreate(&stream_));
    // second stream (default priority) for the MAS coarse levels, ordered by events
    LSS_HIP(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    LSS_HIP(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
    LSS_HIP(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming));
    // third stream (default priority) for the stiff selection and factorization
    LSS_HIP(hipStreamCreateWithFlags(&aux_, hipStreamNonBlocking));
    LSS_HIP(hipEventCreateWithFlags(&ev_aux_, hipEventDisableTiming));
    void *p = nullptr;
    LSS_HIP(hipHostMalloc(&p, sizeof(dev::PcgState) * kMaxCols, hipHostMallocDefault));
    h_st_ = static_cast<dev::PcgState *>(p);
    std::memset(h_st_, 0, sizeof(dev::PcgState) * kMaxCols);
    st_.resize(kMaxCols);
    st_.zero(stream_);
    partials_.resize(kPartialsCap);
    d_status_.resize(1);
    sync(stream_);
}

HybridSolver::~HybridSolver()
{
    release();
    for (auto e : ev_pool_) (void)hipEventDestroy(e);
    if (h_st_) (void)hipHostFree(h_st_);
    if (stream_) (void)hipStreamDestroy(stream_);
    if (ev_fork_) (void)hipEventDestroy(ev_fork_);
    if (ev_join_) (void)hipEventDestroy(ev_join_);
    if (side_) (void)hipStreamDestroy(side_);
    if (ev_aux_) (void)hipEventDestroy(ev_aux_);
    if (aux_) (void)hipStreamDestroy(aux_);
}

const char *HybridSolver::name() const
{
    if (cfg_.mas_mode()) return "MAS";  // MASSolver.cu:785-788
    // GPUHybrid.hpp:57-67
    if (cfg_.do_hybrid_precond) return cfg_.select_bad_dofs_from_l1_norm ? "GPUHybrid" : "GPUAMGF";
    return "GPUAMG";
}

k::Precision HybridSolver::prec() const
{
    if (cfg_.precision == "fp32") return k::Precision::FP32;
    if (cfg_.precision == "bf16") return k::Precision::BF16;
    if (cfg_.precision == "fp16") return k::Precision::FP16;
    return k::Precision::FP64;
}

void HybridSolver::set_problematic_dofs(const int32_t *dofs, int64_t n)
{
    user_set_.assign(dofs, dofs + n);
    std::sort(user_set_.begin(), user_set_.end());
    user_set_.erase(std::unique(user_set_.begin(), user_set_.end()), user_set_.end());
    user_set_given_ = true;
}

void HybridSolver::reset_factorization()
{
    // Device buffers of the operator, stiff blocks and vectors keep their allocations:
    // DBuf::resize keeps one whose size is unchanged, so a factorization of a matrix of
    // the same shape allocates nothing for them.  Each is rewritten before it is read.
    // The level-0 bank factors (most of the MAS bytes) are kept aside the same way.
    factorized_ = false;
    if (!mas_.empty()) spare_inv0_ = std::move(mas_[0].inv);
    mas_.clear();
    comps_.clear();
    ncomp_ = 0;
    ns_ = ns_dense_ = 0;
    sparse_.clear();
    sparse_off_.clear();
    omega_ = 1.0;
    d_unit_dense_.release();
    dense_bins_ = dense_dim_ = dense_dofs_ = 0;
    bin_sizes_.clear();
    amg_.clear();
}

void HybridSolver::release()
{
    factorized_ = false;
    amg_.clear();
    amg_work_ = k::AmgWork{};
    mas_.clear();
    spare_inv0_.release();
    comps_.clear();
    d_perm_.release();
    d_rowptr_.release();
    d_cols_.release();
    d_bvals_.release();
    d_slice_ptr_.release();
    d_scols_.release();
    d_svals_.release();
    d_svals32_.release();
    d_svals16_.release();
    d_dinv_.release();
    d_comp_off_.release();
    d_mat_off_.release();
    d_pdofs_.release();
    d_row_comp_.release();
    d_L_.release();
    d_swork_.release();
    d_unit_dense_.release();
    d_sparse_t_.release();
    sparse_.clear();
    sparse_off_.clear();
    xp_.release(), bp_.release(), r_.release(), p_.release(), q_.release(), z_.release(), u_.release();
    z32_.release();
    ncomp_ = 0;
    ns_ = ns_dense_ = 0;
    omega_ = 1.0;
}

// ------------------------------------------------------------------ factorize
void HybridSolver::factorize_host(int32_t n, int64_t nnz, const int32_t *colptr, const int32_t *rowidx,
                                  const double *vals)
{
    DBuf<int32_t> dp, di;
    DBuf<double> dv;
    dp.upload(colptr, size_t(n) + 1, stream_);
    di.upload(rowidx, size_t(nnz), stream_);
    dv.upload(vals, size_t(nnz), stream_);
    factorize_device(n, nnz, dp.get(), di.get(), dv.get(), stream_);
    sync(stream_);
}

void HybridSolver::factorize_device(int32_t n, int64_t nnz, const int32_t *ptr, const int32_t *idx, const double *vals,
                                    hipStream_t s)
{
    // GPUHybrid::factorize, GPUHybrid.cu:195-282
    if (!s) s = stream_;
    if (n <= 0) fail(LSS_ERR_SIZE, "matrix must have at least one row");
    auto t0 = clk::now();
    reset_factorization();
    n_ = n;
    bs_ = cfg_.block_size;
    nb_ = (n + bs_ - 1) / bs_;
    nnz_ = nnz;
    int32_t last = 0;
    LSS_HIP(hipMemcpyAsync(&last, ptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    sync(s);
    if (int64_t(last) != nnz) fail(LSS_ERR_SIZE, "colptr[n] does not match nnz");
    gmm_ = GmmResult{};
    filter_ = FilterStats{};
    n_selected_ = n_filtered_ = n_expanded_ = 0;
    t_select_ = t_stiff_ = 0.0;

    // The hybrid preconditioner's stiff selection (row norms, GMM, components: host
    // round trips and host graph work) needs neither the operator nor the MAS, and
    // its stiff factorization is not needed by the omega estimate.  Both run on a
    // worker thread and stream: the selection while this thread builds the operator,
    // partition and MAS, the factorization (after the MAS, which shares d_status_)
    // while it estimates omega.
    const bool overlap = !cfg_.mas_mode() && cfg_.do_hybrid_precond;
    std::thread worker;
    std::exception_ptr select_err, stiff_err;
    std::promise<void> selected_p, mas_done_p;
    std::future<void> selected = selected_p.get_future();
    std::shared_future<void> mas_done = mas_done_p.get_future().share();
    std::atomic<bool> abandon{false};
    bool mas_signalled = false;
    auto join = [&] {
        if (!worker.joinable()) return;
        if (!mas_signalled) {
            abandon = true;
            mas_done_p.set_value();
            mas_signalled = true;
        }
        worker.join();
    };
    if (overlap) {
        int dev = 0;
        LSS_HIP(hipGetDevice(&dev));
        LSS_HIP(hipEventRecord(ev_aux_, s));  // the input arrays are ready on s
        LSS_HIP(hipStreamWaitEvent(aux_, ev_aux_, 0));
        worker = std::thread([&, dev] {
            try {
                LSS_HIP(hipSetDevice(dev));
                const auto tw = clk::now();
                select_stiff(n, ptr, idx, vals, aux_);
                t_select_ = since(tw);
            } catch (...) {
                select_err = std::current_exception();
            }
            selected_p.set_value();
            if (select_err) return;
            mas_done.wait();
            if (abandon) return;
            try {
                const auto tw = clk::now();
                if (!comps_.empty()) build_stiff(n, ptr, idx, vals, aux_);
                sync(aux_);
                t_stiff_ = since(tw);
            } catch (...) {
                stiff_err = std::current_exception();
            }
        });
    }
    try {
        auto t = clk::now();
        build_operator(n, nnz, ptr, idx, vals, s);
        sync(s);
        t_operator_ = since(t) - t_partition_;

        t = clk::now();
        if (cfg_.base_preconditioner != "AMG") amg_work_ = k::AmgWork{};
        if (cfg_.base_preconditioner == "MAS") {
            build_mas(s);
        } else if (cfg_.base_preconditioner == "AMG") {
            build_amg(n, ptr, idx, vals, s);
        } else {
            d_dinv_.resize(size_t(nb_) * bs_);
            k::diag_inverse(nb_, bs_, d_rowptr_.get(), d_cols_.get(), d_bvals_.get(), d_dinv_.get(), s);
        }
        sync(s);
        t_mas_ = since(t);

        if (overlap) {
            mas_done_p.set_value();
            mas_signalled = true;
        } else {
            if (cfg_.mas_mode() && cfg_.use_bad_dof_precond) {
                // the MAS solver selects on the partition-permuted operator (MASSolver.cu:398-432)
                t = clk::now();
                select_mas_bad(n, ptr, idx, vals, s);
                t_select_ += since(t);
            }
            t = clk::now();
            if (!comps_.empty()) build_stiff(n, ptr, idx, vals, s);
            sync(s);
            t_stiff_ = since(t);
        }

        alloc_vectors();
        if (overlap) {
            selected.wait();  // estimate_omega needs to know whether the stiff set is empty
            if (select_err) std::rethrow_exception(select_err);
        }
        t = clk::now();
        estimate_omega(s);
        t_omega_ = since(t);
    } catch (...) {
        join();
        throw;
    }
    join();
    if (stiff_err) std::rethrow_exception(stiff_err);
    factorized_ = true;
    t_factorize_ = since(t0);
}

RowSet HybridSolver::fetch_rows(const std::vector<int32_t> &rows, const int32_t *ptr, const int32_t *idx,
                                const double *vals, bool with_vals, hipStream_t s)
{
    RowSet rs;
    rs.rows = rows;
    const int64_t m = int64_t(rows.size());
    rs.ptr.assign(m + 1, 0);
    if (m == 0) return rs;
    DBuf<int32_t> d_rows, d_len(m);
    d_rows.upload(rows.data(), m, s);
    // row lengths on the device: only m values cross to the host, not the column pointers
    k::row_lengths(m, d_rows.get(), ptr, d_len.get(), s);
    const auto len = to_host(d_len, m, s);
    for (int64_t k = 0; k < m; ++k) rs.ptr[k + 1] = rs.ptr[k] + len[k];
    const int64_t tot = rs.ptr[m];
    DBuf<int64_t> d_optr;
    d_optr.upload(rs.ptr.data(), m + 1, s);
    DBuf<int32_t> d_oidx(std::max<int64_t>(tot, 1));
    DBuf<double> d_ovals;
    if (with_vals) d_ovals.resize(std::max<int64_t>(tot, 1));
    k::gather_rows(m, d_rows.get(), ptr, idx, vals, d_optr.get(), d_oidx.get(), with_vals ? d_ovals.get() : nullptr,
                   s);
    rs.idx.resize(tot);
    d_oidx.download(rs.idx.data(), tot, s);
    if (with_vals) {
        rs.vals.resize(tot);
        d_ovals.download(rs.vals.data(), tot, s);
    }
    sync(s);
    return rs;
}

void HybridSolver::select_stiff(int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals, hipStream_t s)
{
    std::vector<int32_t> S;
    if (cfg_.select_bad_dofs_from_l1_norm) {
        // GPUHybrid::select_bad_dofs, GPUHybrid.cu:737-883
        d_norms_.resize(n);
        k::row_l1(n, ptr, vals, d_norms_.get(), s);
        DBuf<double> g0(n), g1(n), part(4 * k::kCsumChunks), sums(8);
        const double *x = d_norms_.get();
        k::canonical_sum(n, x, part.get(), sums.get(), s);
        k::minmax(n, x, part.get(), sums.get() + 1, s);
        double h[8];
        sums.download(h, 3, s);
        sync(s);
        const double global_mean = h[0] / n;
        k::gmm_var_global(n, x, global_mean, part.get(), s);
        k::csum_finish(1, part.get(), sums.get(), s);
        const double global_var = d2h_scalar(sums.get(), s) / n;
        double mean0 = h[1], mean1 = h[2], var0 = global_var, var1 = global_var, w0 = 0.5, w1 = 0.5;
        const double var_reg = 1e-6;
        int it;
        for (it = 0; it < cfg_.max_gmm_iterations; ++it) {
            k::GmmParams p;
            p.mean0 = mean0;
            p.mean1 = mean1;
            p.var0 = var0;
            p.var1 = var1;
            p.log_w0 = std::log(w0);
            p.log_w1 = std::log(w1);
            p.c0 = -0.5 * std::log(2.0 * kPi * var0);
            p.c1 = -0.5 * std::log(2.0 * kPi * var1);
            k::gmm_estep(n, x, p, g0.get(), g1.get(), part.get(), s);
            k::csum_finish(4, part.get(), sums.get(), s);
            sums.download(h, 4, s);
            sync(s);
            const double sum_g0 = h[0], sum_g1 = h[1];
            w0 = sum_g0 / n;
            w1 = sum_g1 / n;
            const double om0 = mean0, om1 = mean1, ov0 = var0, ov1 = var1;
            mean0 = h[2] / sum_g0;
            mean1 = h[3] / sum_g1;
            k::gmm_var(n, x, g0.get(), g1.get(), mean0, mean1, part.get(), s);
            k::csum_finish(2, part.get(), sums.get(), s);
            sums.download(h, 2, s);
            sync(s);
            var0 = h[0] / sum_g0 + var_reg;
            var1 = h[1] / sum_g1 + var_reg;
            if (std::fabs(mean0 - om0) / std::fabs(om0) < cfg_.gmm_tol &&
                std::fabs(mean1 - om1) / std::fabs(om1) < cfg_.gmm_tol &&
                std::fabs(var0 - ov0) / std::fabs(ov0) < cfg_.gmm_tol &&
                std::fabs(var1 - ov1) / std::fabs(ov1) < cfg_.gmm_tol)
                break;
        }
        gmm_.global_mean = global_mean;
        gmm_.global_var = global_var;
        gmm_.mean0 = mean0;
        gmm_.mean1 = mean1;
        gmm_.var0 = var0;
        gmm_.var1 = var1;
        gmm_.w0 = w0;
        gmm_.w1 = w1;
        gmm_.iters = it;
        gmm_.jumped = std::fabs(mean1) / std::fabs(mean0) > cfg_.gmm_jump_threshold;
        if (gmm_.jumped) {
            DBuf<uint8_t> fl(n);
            k::gmm_flags(n, g0.get(), g1.get(), fl.get(), s);
            S = k::flagged_indices(n, fl.get(), s);
        }
        gmm_.selected = int64_t(S.size());
    } else {
        for (int32_t d : user_set_)
            if (d >= 0 && d < n) S.push_back(d);
    }
    n_selected_ = int64_t(S.size());
    if (cfg_.decompose_subdomains && !S.empty()) {
        RowSet rows = fetch_rows(S, ptr, idx, vals, true, s);
        S = filter_subdomains(n, rows, S, cfg_.min_subdomain_size, cfg_.max_subdomain_size,
                              cfg_.conditioning_threshold, filter_);
    }
    n_filtered_ = int64_t(S.size());
    if (cfg_.expand_subdomains && !S.empty()) {
        RowSet rows = fetch_rows(S, ptr, idx, nullptr, false, s);
        S = expand_subdomains(n, rows, S);
    }
    n_expanded_ = int64_t(S.size());
    comps_.clear();
    if (cfg_.decompose_subdomains) {
        if (!S.empty()) {
            RowSet rows = fetch_rows(S, ptr, idx, nullptr, false, s);
            comps_ = decompose_subdomains(n, rows, S, cfg_.max_subdomain_size);
        }
    } else if (!S.empty()) {
        comps_.push_back(S);
    }
}

void HybridSolver::build_operator(int32_t n, int64_t nnz, const int32_t *ptr, const int32_t *idx, const double *vals,
                                  hipStream_t s)
{
    const int bb = bs_ * bs_;
    const bool mas = cfg_.base_preconditioner == "MAS";
    const bool dev_topo = mas && cfg_.partitioner == "device";
    // scalar rows -> block rows (original numbering)
    DBuf<int32_t> cnt(nb_), flag(1);
    flag.zero(s);
    k::bsr_count(n, bs_, ptr, idx, cnt.get(), flag.get(), s);
    // block row pointer: scanned on the device when no int32 prefix can overflow (every
    // block holds a stored entry or is a padded tail diagonal: nnzb <= nnz + nb)
    const bool dev_scan = dev_topo && nnz + nb_ < (int64_t(1) << 31);
    Pattern g;
    g.n = nb_;
    DBuf<int32_t> d_rp0(static_cast<size_t>(nb_) + 1), tsum(std::max(div_up(nb_, 1024), 1));
    std::vector<int32_t> hc;
    if (dev_scan) {
        k::excl_scan_i32(nb_, cnt.get(), d_rp0.get(), tsum.get(), d_rp0.get() + nb_, s);
        int32_t h[2] = {0, 0};
        flag.download(&h[0], 1, s);
        LSS_HIP(hipMemcpyAsync(&h[1], d_rp0.get() + nb_, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        sync(s);
        if (h[0]) fail(LSS_ERR_ARG, "row indices must be strictly ascending within each column");
        nnzb_ = h[1];
    } else {
        hc = to_host(cnt, nb_, s);
        int32_t hflag = 0;
        flag.download(&hflag, 1, s);
        sync(s);
        if (hflag) fail(LSS_ERR_ARG, "row indices must be strictly ascending within each column");
        g.rowptr.assign(nb_ + 1, 0);
        int64_t acc = 0;
        for (int32_t I = 0; I < nb_; ++I) {
            acc += hc[I];
            if (acc >= (int64_t(1) << 31)) fail(LSS_ERR_UNSUPPORTED, "more than 2^31 nonzero blocks");
            g.rowptr[I + 1] = int32_t(acc);
        }
        nnzb_ = acc;
        d_rp0.upload(g.rowptr.data(), nb_ + 1, s);
    }
    // The unpermuted BSR (columns c0, values v0) is scratch until bsr_permute.  It lives
    // in the sliced arrays, which are written only after it is consumed and hold at
    // least nnzb slots (slices add padding): a refactorization of the same shape
    // allocates nothing for it.
    if (d_scols_.size() < size_t(std::max<int64_t>(nnzb_, 1))) d_scols_.resize(std::max<int64_t>(nnzb_, 1));
    if (d_svals_.size() < size_t(std::max<int64_t>(nnzb_ * bb, 1))) d_svals_.resize(std::max<int64_t>(nnzb_ * bb, 1));
    int32_t *c0 = d_scols_.get();
    double *v0 = d_svals_.get();
    LSS_HIP(hipMemsetAsync(v0, 0, size_t(nnzb_) * bb * sizeof(double), s));
    k::bsr_fill(n, bs_, ptr, idx, vals, d_rp0.get(), c0, v0, s);
    if (!dev_topo) {  // the host partition and coarse spaces read the pattern
        g.cols.resize(nnzb_);
        LSS_HIP(hipMemcpyAsync(g.cols.data(), c0, size_t(nnzb_) * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    }
    sync(s);

    // partition (MAS banks) and permutation; with lazy_partitioning the first partition is
    // kept while the block count is unchanged (MASSolver.cu:389-405, which does not compare
    // patterns either: the banks are a grouping of block rows, valid for any pattern)
    auto tp = clk::now();
    DBuf<int32_t> d_off;
    const bool reuse = mas && cfg_.lazy_partitioning && perm_.size() == size_t(nb_) && part_off_.size() > 1 &&
                       d_perm_.size() == size_t(nb_);
    if (reuse) {
        if (dev_topo) d_off.upload(part_off_.data(), part_off_.size(), s);
    } else if (dev_topo) {
        d_perm_.resize(nb_);
        const int32_t nparts = k::device_partition(nb_, d_rp0.get(), c0, 32, d_perm_.get(), d_off, s,
                                                    cfg_.partition_local_max);
        perm_.resize(nb_);
        part_off_.resize(size_t(nparts) + 1);
        d_perm_.download(perm_.data(), nb_, s);
        d_off.download(part_off_.data(), part_off_.size(), s);
        sync(s);
    } else if (mas) {
        std::vector<int32_t> part_of;
        int32_t nparts = 0;
        partition_graph(g, 32, part_of, nparts);
        partition_permutation(part_of, nparts, perm_, part_off_);
    } else {
        perm_.resize(nb_);
        std::iota(perm_.begin(), perm_.end(), 0);
        part_off_.clear();
    }
    partition_reused_ = reuse;
    // row pointer of the permuted pattern
    t_partition_ = since(tp);
    if (!dev_topo) d_perm_.upload(perm_.data(), nb_, s);
    std::vector<int32_t> prow;
    if (dev_scan) {
        d_rowptr_.resize(size_t(nb_) + 1);
        DBuf<int32_t> plen(nb_);
        k::permute_lengths(nb_, d_perm_.get(), cnt.get(), plen.get(), s);
        k::excl_scan_i32(nb_, plen.get(), d_rowptr_.get(), tsum.get(), d_rowptr_.get() + nb_, s);
        sync(s);  // plen is released at scope end
    } else {
        prow.assign(size_t(nb_) + 1, 0);
        for (int32_t I = 0; I < nb_; ++I) prow[size_t(perm_[I]) + 1] = hc[I];
        for (int32_t I = 0; I < nb_; ++I) prow[size_t(I) + 1] += prow[I];
        d_rowptr_.upload(prow.data(), nb_ + 1, s);
    }
    d_cols_.resize(std::max<int64_t>(nnzb_, 1));
    d_bvals_.resize(std::max<int64_t>(nnzb_ * bb, 1));
    k::bsr_permute(nb_, bs_, d_rp0.get(), c0, v0, d_perm_.get(), d_rowptr_.get(), d_cols_.get(),
                   d_bvals_.get(), s);

    // sliced block-ELL copy for SpMV
    const int H = 64 / bs_;
    const int32_t nsl = (nb_ + H - 1) / H;
    std::vector<int64_t> sp(nsl + 1, 0);
    {
        DBuf<int32_t> wd(std::max(nsl, 1));
        k::sell_width(nb_, bs_, d_rowptr_.get(), wd.get(), s);
        const auto w = to_host(wd, size_t(nsl), s);
        for (int32_t sl = 0; sl < nsl; ++sl) sp[sl + 1] = sp[sl] + int64_t(w[sl]) * H;
    }
    sell_slots_ = sp[nsl];
    d_slice_ptr_.upload(sp.data(), nsl + 1, s);
    sync(s);  // bsr_permute has read c0 and v0 before the sliced arrays are resized
    // sell_fill writes every slot (padding: column -1, zero values)
    d_scols_.resize(std::max<int64_t>(sell_slots_, 1));
    const int64_t nsv = k::sell_value_count(sell_slots_, bs_);
    d_svals_.resize(std::max<int64_t>(nsv, 1));
    const bool v32 = cfg_.fp32_inner() && !cfg_.bf16_inner(), v16 = cfg_.bf16_inner();
    if (v32) d_svals32_.resize(std::max<int64_t>(nsv, 1));
    else d_svals32_.release();
    if (v16) d_svals16_.resize(std::max<int64_t>(nsv, 1));
    else d_svals16_.release();
    k::sell_fill(nb_, bs_, d_rowptr_.get(), d_cols_.get(), d_bvals_.get(), d_slice_ptr_.get(), d_scols_.get(),
                 d_svals_.get(), v32 ? d_svals32_.get() : nullptr, v16 ? d_svals16_.get() : nullptr, s);
    sell_.nb = nb_;
    sell_.bs = bs_;
    sell_.nslices = nsl;
    sell_.slice_ptr = d_slice_ptr_.get();
    sell_.cols = d_scols_.get();
    sell_.vals = d_svals_.get();
    sell_.vals32 = v32 ? d_svals32_.get() : nullptr;
    sell_.vals16 = v16 ? d_svals16_.get() : nullptr;
    sync(s);
    level_nodes_.clear();
    mas_.clear();
    if (!mas) return;
    if (dev_topo && build_levels_device(std::move(d_off), int32_t(part_off_.size()) - 1, s)) return;
    // host coarse spaces (host partition, or a bank too dense for the device sort)
    Pattern gp;
    if (dev_topo) {
        gp.n = nb_;
        if (prow.empty()) prow = to_host(d_rowptr_, size_t(nb_) + 1, s);
        gp.rowptr = prow;
        gp.cols.resize(nnzb_);
        d_cols_.download(gp.cols.data(), nnzb_, s);
        sync(s);
    } else {
        gp = permute_pattern(g, perm_);
    }
    build_levels_host(gp, s);
}

bool HybridSolver::build_levels_device(DBuf<int32_t> &&part_off, int32_t nparts, hipStream_t s)
{
    const int bb = bs_ * bs_;
    mas_.clear();
    level_nodes_.clear();
    mas_.emplace_back();
    mas_[0].nodes = nb_;
    mas_[0].nbanks = nparts;
    mas_[0].bank_off = std::move(part_off);
    level_nodes_.push_back(nb_);
    for (int lv = 0; lv < cfg_.max_coarse_levels; ++lv) {
        const size_t l = mas_.size() - 1;
        const int32_t *rp = l == 0 ? d_rowptr_.get() : mas_[l].rowptr.get();
        const int32_t *cl = l == 0 ? d_cols_.get() : mas_[l].cols.get();
        const int64_t fine_nnz = l == 0 ? nnzb_ : int64_t(mas_[l].cols.size());
        DBuf<int32_t> parent, child, base, count;
        DBuf<int64_t> child_ptr;
        const int32_t coarse_n = k::device_components(mas_[l].nodes, rp, cl, mas_[l].nbanks, mas_[l].bank_off.get(),
                                                      parent, child_ptr, child, base, count, s);
        // stop when nothing merges (MASPreconditioner.cu:449-454); level 0 always coarsens
        if (lv > 0 && coarse_n == mas_[l].nodes) break;
        DevLevel nxt;
        nxt.nodes = coarse_n;
        DBuf<int64_t> sptr;
        DBuf<int32_t> sblk;
        int64_t nslots = 0;
        if (!k::device_galerkin_pattern(fine_nnz, rp, cl, mas_[l].nbanks, mas_[l].bank_off.get(), base.get(),
                                        count.get(), parent.get(), coarse_n, nxt.rowptr, nxt.cols, sptr, sblk, nslots,
                                        s)) {
            sync(s);
            mas_.clear();
            level_nodes_.clear();
            return false;
        }
        nxt.vals.resize(std::max<int64_t>(nslots * bb, 1));
        const double *fine_vals = l == 0 ? d_bvals_.get() : mas_[l].vals.get();
        k::mas_galerkin(nslots, bs_, sptr.get(), sblk.get(), fine_vals, nxt.vals.get(), s);
        nxt.r.resize(size_t(coarse_n) * bs_ * kMaxCols);
        nxt.w.resize(size_t(coarse_n) * bs_ * kMaxCols);
        nxt.nbanks = k::uniform_banks(coarse_n, nxt.bank_off, s);
        mas_[l].parent = std::move(parent);
        sync(s);  // the source lists are released at scope end
        mas_.push_back(std::move(nxt));
        level_nodes_.push_back(coarse_n);
    }
    sync(s);
    return true;
}

void HybridSolver::build_levels_host(const Pattern &gp, hipStream_t s)
{
    const int bb = bs_ * bs_;
    std::vector<MasLevel> levels;
    build_mas_levels(gp, part_off_, cfg_.max_coarse_levels, levels);
    mas_.resize(levels.size());
    for (size_t l = 0; l < levels.size(); ++l) {
        const MasLevel &h = levels[l];
        DevLevel &d = mas_[l];
        d.nodes = h.nodes;
        d.nbanks = int32_t(h.bank_off.size()) - 1;
        d.bank_off.upload(h.bank_off.data(), h.bank_off.size(), s);
        if (l >= 1) {
            d.rowptr.upload(h.op.rowptr.data(), h.op.rowptr.size(), s);
            d.cols.upload(h.op.cols.data(), std::max<size_t>(h.op.cols.size(), 1), s);
            const int64_t nslots = int64_t(h.op.cols.size());
            d.vals.resize(std::max<int64_t>(nslots * bb, 1));
            DBuf<int64_t> sptr;
            DBuf<int32_t> sblk;
            sptr.upload(h.src_ptr.data(), h.src_ptr.size(), s);
            sblk.upload(h.src_blk.data(), std::max<size_t>(h.src_blk.size(), 1), s);
            const double *fine_vals = (l == 1) ? d_bvals_.get() : mas_[l - 1].vals.get();
            k::mas_galerkin(nslots, bs_, sptr.get(), sblk.get(), fine_vals, d.vals.get(), s);
            d.r.resize(size_t(h.nodes) * bs_ * kMaxCols);
            d.w.resize(size_t(h.nodes) * bs_ * kMaxCols);
            sync(s);  // sptr / sblk are released at scope end
        }
        if (!h.parent.empty()) d.parent.upload(h.parent.data(), h.parent.size(), s);
        level_nodes_.push_back(h.nodes);
    }
    sync(s);
}

void HybridSolver::build_mas(hipStream_t s)
{
    const size_t eb = k::precision_bytes(prec());
    d_status_.zero(s);
    mas_banks_ = 0;
    mas_bytes_ = 0;
    for (size_t l = 0; l < mas_.size(); ++l) {
        DevLevel &d = mas_[l];
        const int64_t S = k::mas_bank_entries(prec(), bs_, l == 0);
        const size_t bytes = std::max<size_t>(size_t(d.nbanks) * S * eb, 16);
        if (l == 0) {
            if (spare_inv0_.size() == bytes) d.inv = std::move(spare_inv0_);
            spare_inv0_.release();
        }
        d.inv.resize(bytes);
        const int32_t *rp = l == 0 ? d_rowptr_.get() : d.rowptr.get();
        const int32_t *cl = l == 0 ? d_cols_.get() : d.cols.get();
        const double *vl = l == 0 ? d_bvals_.get() : d.vals.get();
        k::mas_bank_invert(prec(), bs_, d.nbanks, d.bank_off.get(), rp, cl, vl, d.inv.get(), d_status_.get(), s,
                           l == 0);
        mas_banks_ += d.nbanks;
        mas_bytes_ += double(d.nbanks) * S * eb;
    }
    int st = 0;
    d_status_.download(&st, 1, s);
    sync(s);
    // level operators of the coarse levels are only needed to build the banks
    for (size_t l = 1; l < mas_.size(); ++l) {
        mas_[l].rowptr.release();
        mas_[l].cols.release();
        mas_[l].vals.release();
    }
    mas_levels_ = int(mas_.size());
    if (st) fail(LSS_ERR_NOT_SPD, "MAS bank inverse failed (zero pivot or non-finite value)");
}

void HybridSolver::build_amg(int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals, hipStream_t s)
{
    // BoomerAMG setup (GPUHybrid.cu:357-376 with the options of 285-339), on the device:
    // per level the l1 norms, strength, PMIS splitting, extended+i interpolation
    // (Pmax 4), R = P^T and the Galerkin operator R (A P), down to at most 5 points or
    // 25 levels, or until a splitting gives no C point or only C points.  Level 0 is
    // the input itself (CSC column j = operator row j, as hypre is given it).
    constexpr int kMaxLevels = 25, kMinCoarse = 5, kPmax = 4;
    constexpr double kMaxRowSum = 0.9;
    const double theta = bs_ > 1 ? cfg_.theta : 0.25;
    const int nf = bs_;
    const int64_t np = int64_t(nb_) * bs_;
    amg_times_ = AmgSetupTimes{};
    auto t0 = clk::now();
    auto lap = [&](double &acc) {
        sync(s);
        acc += since(t0);
        t0 = clk::now();
    };
    DBuf<int64_t> p0(static_cast<size_t>(n) + 1);
    k::amg_widen_ptr(int64_t(n) + 1, ptr, p0.get(), s);
    const k::AmgCsr A0{n, p0.get(), idx, vals, nnz_};
    DBuf<int32_t> func;
    if (nf > 1) {
        std::vector<int32_t> f0(static_cast<size_t>(n));
        for (int32_t i = 0; i < n; ++i) f0[i] = i % nf;
        func.upload(f0.data(), f0.size(), s);
    }
    amg_.clear();
    amg_cf_.clear();
    amg_sizes_.clear();
    amg_nnz_.clear();
    amg_.reserve(kMaxLevels);
    amg_.emplace_back();
    for (int l = 0;; ++l) {
        const k::AmgCsr A = l == 0 ? A0 : amg_[l].A.view();
        AmgLevel &L = amg_[l];
        const int32_t m = A.n;
        L.n = m;
        amg_sizes_.push_back(m);
        amg_nnz_.push_back(A.nnz);
        // level 0 works on the padded vectors of the sliced operator: its tail rows
        // have a unit diagonal, l1 norm 1
        const size_t len = l == 0 ? size_t(np) : size_t(std::max(m, 1));
        L.dinv.resize(len);
        k::amg_l1_inverse(A, L.dinv.get(), s);
        if (l == 0 && np > n) k::fill(np - n, L.dinv.get() + n, 1.0, s);
        if (l > 0) {
            L.b.resize(len);
            L.x.resize(len);
        }
        L.t.resize(len);
        lap(amg_times_.norms);
        amg_cf_.emplace_back();
        if (l + 1 >= kMaxLevels || m <= kMinCoarse) break;
        k::AmgCsrBuf S;
        k::amg_strength_device(A, nf > 1 ? func.get() : nullptr, nf, theta, kMaxRowSum, S, amg_work_, s);
        lap(amg_times_.strength);
        DBuf<int8_t> cf(static_cast<size_t>(std::max(m, 1)));
        const int32_t nc = k::amg_pmis_device(S.view(), cf.get(), amg_work_, s);
        lap(amg_times_.pmis);
        if (nc == 0 || nc == m) break;
        amg_cf_.back().resize(size_t(m));
        cf.download(amg_cf_.back().data(), size_t(m), s);
        DBuf<int32_t> cfunc(static_cast<size_t>(nc));
        k::amg_interp_device(A, S.view(), cf.get(), nf > 1 ? func.get() : nullptr, nf, kPmax, L.P,
                             nf > 1 ? cfunc.get() : nullptr, amg_work_, s);
        lap(amg_times_.interp);
        k::amg_transpose_device(L.P.view(), nc, L.R, amg_work_, s);
        lap(amg_times_.transpose);
        k::AmgCsrBuf AP;
        k::amg_spgemm_device(A, L.P.view(), nc, AP, amg_work_, s);
        amg_.emplace_back();  // L is not used past this point (the vector may have moved)
        k::amg_spgemm_device(amg_[l].R.view(), AP.view(), nc, amg_[l + 1].A, amg_work_, s);
        lap(amg_times_.rap);
        func = std::move(cfunc);
    }
    amg_z_.resize(size_t(np));
    sync(s);
}

void HybridSolver::amg_vcycle(const double *b0, double *x0, hipStream_t s)
{
    // one V-cycle from x = 0 (GPUHybrid.cu:677-689: HYPRE_BoomerAMGSolve, one
    // iteration): l1-Jacobi pre-smoothing, residual, restriction R = P^T down to the
    // coarsest level (one l1-Jacobi sweep there), then prolongation and l1-Jacobi
    // post-smoothing up.  Level 0 uses the sliced block operator for its products.
    const int nl = int(amg_.size());
    const int64_t np = int64_t(nb_) * bs_;
    auto bvec = [&](int l) { return l == 0 ? b0 : amg_[l].b.get(); };
    auto xvec = [&](int l) { return l == 0 ? x0 : amg_[l].x.get(); };
    for (int l = 0; l + 1 < nl; ++l) {
        AmgLevel &L = amg_[l];
        const int64_t m = l == 0 ? np : L.n;
        k::amg_level_smooth_pre(int32_t(m), L.dinv.get(), bvec(l), xvec(l), s);
        if (l == 0) k::spmv_residual(sell_, xvec(0), bvec(0), L.t.get(), nullptr, s);
        else k::amg_residual(L.A.view(), xvec(l), bvec(l), L.t.get(), s);
        k::amg_spmv(L.R.view(), L.t.get(), amg_[l + 1].b.get(), s);
    }
    {
        AmgLevel &C = amg_[nl - 1];
        const int64_t m = nl == 1 ? np : C.n;
        k::amg_level_smooth_pre(int32_t(m), C.dinv.get(), bvec(nl - 1), xvec(nl - 1), s);
    }
    // the post-smoothed vector of a coarse level is written to its t (the smoother reads
    // x while it writes), and the prolongation to the next finer level reads it there
    const double *xc = xvec(nl - 1);
    for (int l = nl - 2; l >= 0; --l) {
        AmgLevel &L = amg_[l];
        k::amg_spmv_add(L.P.view(), xc, xvec(l), s);
        if (l == 0) {
            k::spmv_residual(sell_, xvec(0), bvec(0), L.t.get(), nullptr, s);
            k::amg_add_diag_scaled(np, L.dinv.get(), L.t.get(), xvec(0), s);
        } else {
            k::amg_smooth_post(L.A.view(), L.dinv.get(), xvec(l), bvec(l), L.t.get(), s);
            xc = L.t.get();
        }
    }
}

void HybridSolver::build_stiff(int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals, hipStream_t s)
{
    // factorize_submatrix, GPUHybrid.cu:993-1411, with a dense FP64 Cholesky per unit.
    // Units: the components, except that with max_dense_size > 0 the components of at
    // most max_dense_size DOFs go to the dense branch (GPUHybrid.cu:1253-1406): first-fit
    // packed, in component order, into bins of the largest such size, each bin one unit
    // (its DOFs the concatenation of its components) whose explicit inverse is applied
    // as one product.  The larger components come first, then the bins (the order of
    // d_all_bad_dofs, GPUHybrid.cu:244-259).  Components are decoupled (no stored entry
    // joins two of them), so a bin's S A S^T is block diagonal and its identity padding
    // in the reference (a bin of fewer DOFs than the largest) changes nothing.
    std::vector<std::vector<int32_t>> units;
    std::vector<uint8_t> dense;
    dense_bins_ = 0;
    dense_dim_ = 0;
    dense_dofs_ = 0;
    if (cfg_.max_dense_size > 0) {
        for (auto &c : comps_)
            if (int64_t(c.size()) <= cfg_.max_dense_size) dense_dim_ = std::max<int64_t>(dense_dim_, c.size());
        std::vector<std::vector<int32_t>> bins;
        for (auto &c : comps_) {
            if (int64_t(c.size()) > cfg_.max_dense_size) {
                units.push_back(c);
                dense.push_back(0);
                continue;
            }
            dense_dofs_ += int64_t(c.size());
            bool placed = false;
            for (auto &b : bins)
                if (int64_t(b.size() + c.size()) <= dense_dim_) {
                    b.insert(b.end(), c.begin(), c.end());
                    placed = true;
                    break;
                }
            if (!placed) bins.push_back(c);
        }
        dense_bins_ = int64_t(bins.size());
        bin_sizes_.clear();
        for (auto &b : bins) {
            bin_sizes_.push_back(int64_t(b.size()));
            units.push_back(std::move(b));
            dense.push_back(1);
        }
    } else {
        units = comps_;
        dense.assign(units.size(), 0);
        bin_sizes_.clear();
    }
    // Units above max_dense_component take the multifrontal path ("stiff_factorization"
    // "auto"; "sparse": every unit outside the dense branch; "dense": such units are
    // rejected).  The dense units keep the order of the list, the sparse ones follow.
    std::vector<std::vector<int32_t>> sparse_units;
    {
        std::vector<std::vector<int32_t>> du;
        std::vector<uint8_t> dd;
        max_comp_ = 0;
        for (size_t c = 0; c < units.size(); ++c) {
            const int64_t m = int64_t(units[c].size());
            max_comp_ = std::max(max_comp_, m);
            const bool sp = !dense[c] && (cfg_.stiff_factorization == "sparse" ||
                                          (cfg_.stiff_factorization == "auto" && m > cfg_.max_dense_component));
            if (!sp && m > cfg_.max_dense_component)
                fail(LSS_ERR_UNSUPPORTED, "stiff component of " + std::to_string(m) +
                                              " DOFs exceeds max_dense_component (" +
                                              std::to_string(cfg_.max_dense_component) + ")");
            if (sp) {
                sparse_units.push_back(std::move(units[c]));
            } else {
                du.push_back(std::move(units[c]));
                dd.push_back(dense[c]);
            }
        }
        units = std::move(du);
        dense = std::move(dd);
    }
    ncomp_ = int32_t(units.size());
    std::vector<int64_t> off(ncomp_ + 1, 0), moff(ncomp_ + 1, 0);
    for (int32_t c = 0; c < ncomp_; ++c) {
        const int64_t m = int64_t(units[c].size());
        off[c + 1] = off[c] + m;
        moff[c + 1] = moff[c] + m * m;
    }
    ns_dense_ = off[ncomp_];
    sparse_off_.assign(1, 0);
    for (auto &u : sparse_units) sparse_off_.push_back(sparse_off_.back() + int64_t(u.size()));
    ns_ = ns_dense_ + sparse_off_.back();
    std::vector<int32_t> dofs(ns_dense_), pd(ns_);
    for (int32_t c = 0; c < ncomp_; ++c)
        for (size_t r = 0; r < units[c].size(); ++r) {
            const int32_t d = units[c][r];
            dofs[off[c] + r] = d;
            pd[off[c] + r] = perm_[d / bs_] * bs_ + d % bs_;
        }
    for (size_t u = 0; u < sparse_units.size(); ++u)
        for (size_t r = 0; r < sparse_units[u].size(); ++r) {
            const int32_t d = sparse_units[u][r];
            pd[ns_dense_ + sparse_off_[u] + r] = perm_[d / bs_] * bs_ + d % bs_;
        }
    d_pdofs_.upload(pd.data(), ns_, s);
    sparse_.clear();
    sparse_.resize(sparse_units.size());
    if (!sparse_units.empty()) {
        std::vector<int32_t> h_ptr(size_t(n) + 1);
        LSS_HIP(hipMemcpyAsync(h_ptr.data(), ptr, (size_t(n) + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        sync(s);
        int64_t mmax = 0;
        d_status_.zero(s);
        for (size_t u = 0; u < sparse_units.size(); ++u) {
            sparse_[u].factorize(n, sparse_units[u], h_ptr, ptr, idx, vals, cfg_.sparse_leaf_size,
                                 cfg_.sparse_batch_rows, cfg_.sparse_batch_bytes, kMaxCols, dense_,
                                 d_status_.get(), s);
            mmax = std::max<int64_t>(mmax, int64_t(sparse_units[u].size()));
        }
        int st = 0;
        d_status_.download(&st, 1, s);
        sync(s);
        if (st) fail(LSS_ERR_NOT_SPD, "stiff block Cholesky failed: a selected sub-block is not SPD");
        d_sparse_t_.resize(size_t(kMaxCols) * mmax);
    } else {
        d_sparse_t_.release();
    }
    if (ncomp_ == 0) {
        d_unit_dense_.release();
        d_swork_.resize(1);
        return;
    }
    d_comp_off_.upload(off.data(), ncomp_ + 1, s);
    d_mat_off_.upload(moff.data(), ncomp_ + 1, s);
    std::vector<int32_t> rc(ns_dense_);
    for (int32_t c = 0; c < ncomp_; ++c)
        for (int64_t k = off[c]; k < off[c + 1]; ++k) rc[k] = c;
    d_row_comp_.upload(rc.data(), ns_dense_, s);
    DBuf<int32_t> d_dofs, g2l(n), comp_of(n);
    d_dofs.upload(dofs.data(), ns_dense_, s);
    k::stiff_maps(ncomp_, n, d_comp_off_.get(), d_dofs.get(), g2l.get(), comp_of.get(), s);
    d_L_.resize(moff[ncomp_]);
    d_L_.zero(s);
    k::stiff_extract(ncomp_, d_comp_off_.get(), d_mat_off_.get(), d_dofs.get(), g2l.get(), comp_of.get(), ptr, idx,
                     vals, d_L_.get(), s);
    d_status_.zero(s);
    DBuf<double> G(moff[ncomp_]);
    // up to stiff_small_max() DOFs: one workgroup per component; larger: blocked path
    k::stiff_cholesky(ncomp_, d_comp_off_.get(), d_mat_off_.get(), d_L_.get(), d_status_.get(), s);
    int64_t max_small = 0, work = 0;
    for (int32_t c = 0; c < ncomp_; ++c) {
        const int64_t m = off[c + 1] - off[c];
        if (m <= k::stiff_small_max()) max_small = std::max(max_small, m);
        else work = std::max(work, DenseBlas::work_size(m));
    }
    if (work > 0) {
        DBuf<double> w(work);
        for (int32_t c = 0; c < ncomp_; ++c) {
            const int64_t m = off[c + 1] - off[c];
            if (m > k::stiff_small_max())
                dense_.factor_inverse(m, d_L_.get() + moff[c], G.get() + moff[c], w.get(), d_status_.get(), s);
        }
        sync(s);  // w is released at scope end
    }
    int st = 0;
    d_status_.download(&st, 1, s);
    sync(s);
    if (st) fail(LSS_ERR_NOT_SPD, "stiff block Cholesky failed: a selected sub-block is not SPD");
    k::stiff_inverse_factor(ncomp_, int32_t(max_small), d_comp_off_.get(), d_mat_off_.get(), d_L_.get(), G.get(), s);
    sync(s);
    d_L_ = std::move(G);  // only the inverse factor is needed by the apply
    // dense branch: the bins' storage takes their explicit inverses M = G^T G
    std::vector<int32_t> dl;
    int64_t dmax = 0;
    for (int32_t c = 0; c < ncomp_; ++c)
        if (dense[c]) {
            dl.push_back(c);
            dmax = std::max(dmax, off[c + 1] - off[c]);
        }
    if (!dl.empty()) {
        DBuf<int32_t> d_dl;
        d_dl.upload(dl.data(), dl.size(), s);
        DBuf<double> M(moff[ncomp_]);
        k::explicit_inverse(int32_t(dl.size()), d_dl.get(), int32_t(dmax), d_comp_off_.get(), d_mat_off_.get(),
                            d_L_.get(), M.get(), s);
        for (int32_t c : dl)
            LSS_HIP(hipMemcpyAsync(d_L_.get() + moff[c], M.get() + moff[c], size_t(moff[c + 1] - moff[c]) * sizeof(double),
                                   hipMemcpyDeviceToDevice, s));
        sync(s);  // M and the list are released at scope end
        d_unit_dense_.upload(dense.data(), dense.size(), s);
    } else {
        d_unit_dense_.release();
    }
    d_swork_.resize(std::max<int64_t>(2 * kMaxCols * ns_dense_, 1));
}

void HybridSolver::alloc_vectors()
{
    const size_t np = size_t(nb_) * bs_;
    xp_.resize(np), bp_.resize(np), r_.resize(np), p_.resize(np), q_.resize(np), z_.resize(np), u_.resize(np);
    if (cfg_.fp32_inner()) z32_.resize(np);
    kcap_ = 1;
    // partials of one reduction for kMaxCols columns
    int64_t per = std::max<int64_t>(k::partials_per_column(int64_t(np)), k::spmv_partials_per_column(sell_));
    if (!mas_.empty()) per = std::max<int64_t>(per, mas_[0].nbanks);
    partials_.resize(std::max<size_t>(kPartialsCap, size_t(per) * kMaxCols));
}

void HybridSolver::ensure_columns(int k)
{
    if (k <= kcap_) return;
    const size_t m = size_t(nb_) * bs_ * k;
    xp_.resize(m), bp_.resize(m), r_.resize(m), p_.resize(m), q_.resize(m), z_.resize(m), u_.resize(m);
    if (cfg_.fp32_inner()) z32_.resize(m);
    kcap_ = k;
}

namespace {

// Largest eigenvalue of the symmetric tridiagonal (d, e) by Sturm-count bisection.
double tridiag_max_eig(const std::vector<double> &d, const std::vector<double> &e)
{
    const int m = int(d.size());
    double lo = 0.0, hi = 0.0;
    for (int i = 0; i < m; ++i) {
        const double rad = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(e[i]) : 0.0);
        hi = std::max(hi, d[i] + rad);
        lo = std::min(lo, d[i] - rad);
    }
    auto count_below = [&](double x) {
        int cnt = 0;
        double q = d[0] - x;
        if (q < 0) ++cnt;
        for (int i = 1; i < m; ++i) {
            if (q == 0.0) q = 1e-300;
            q = d[i] - x - e[i - 1] * e[i - 1] / q;
            if (q < 0) ++cnt;
        }
        return cnt;
    };
    for (int it = 0; it < 200 && hi - lo > 1e-13 * std::max(1.0, std::fabs(hi)); ++it) {
        const double mid = 0.5 * (lo + hi);
        if (count_below(mid) >= m) hi = mid;
        else lo = mid;
    }
    return hi;
}

}  // namespace

void HybridSolver::estimate_omega(hipStream_t s)
{
    omega_ = 1.0;
    // comps_ (not ncomp_, which the stiff factorization may be writing) tells whether
    // there is a stiff set
    if (cfg_.mas_mode() || !cfg_.do_hybrid_precond || comps_.empty()) return;
    // a V-cycle of a convergent multigrid has lambda(B A) in (0, 2): the reference
    // composes it undamped (GPUHybrid.cu:521-542)
    if (cfg_.base_preconditioner == "AMG") return;
    // The additive MAS can have lambda_max(B A) > 2, which makes the multiplicative
    // sweep (I - BA)(I - PA)(I - BA) non-contractive and the composed preconditioner
    // indefinite.  omega = damping / theta, theta the largest Ritz value of B A after
    // lanczos_steps steps of B-preconditioned CG (CG-Lanczos) on A x = v, v hashed;
    // theta <= lambda_max(B A) and converges to it quickly.
    const int64_t np = int64_t(nb_) * bs_;
    double *r = r_.get(), *z = z_.get(), *p = p_.get(), *q = q_.get();
    k::hash_vector(n_, u_.get(), s);
    k::fill(np, r, 0.0, s);
    k::permute_in(nb_, bs_, n_, d_perm_.get(), u_.get(), r, s);
    DBuf<double> sc(1);
    base_apply(r, 1.0, nullptr, z, nullptr, nullptr, s);
    k::copy(np, z, p, s);
    k::dot(np, r, z, partials_.get(), sc.get(), s);
    double rz = d2h_scalar(sc.get(), s);
    std::vector<double> alpha, beta;
    for (int j = 0; j < cfg_.lanczos_steps && rz > 0.0; ++j) {
        k::spmv(sell_, p, q, nullptr, s);
        k::dot(np, p, q, partials_.get(), sc.get(), s);
        const double pq = d2h_scalar(sc.get(), s);
        if (!(pq > 0.0)) break;
        const double a = rz / pq;
        k::axpy(np, -a, q, r, s);
        base_apply(r, 1.0, nullptr, z, nullptr, nullptr, s);
        k::dot(np, r, z, partials_.get(), sc.get(), s);
        const double rz2 = d2h_scalar(sc.get(), s);
        const double b = rz2 / rz;
        alpha.push_back(a);
        beta.push_back(b);
        if (!(rz2 > 0.0)) break;
        rz = rz2;
        k::xpby(np, z, b, p, s);
    }
    const int m = int(alpha.size());
    if (m == 0) fail(LSS_ERR_NOT_SPD, "lambda_max(BA) estimate failed: B or A is not positive definite");
    std::vector<double> d(m), e(m > 0 ? m - 1 : 0);
    for (int j = 0; j < m; ++j) {
        d[j] = 1.0 / alpha[j] + (j > 0 ? beta[j - 1] / alpha[j - 1] : 0.0);
        if (j + 1 < m) e[j] = std::sqrt(beta[j]) / alpha[j];
    }
    const double theta = tridiag_max_eig(d, e);
    if (!(theta > 0.0) || !std::isfinite(theta))
        fail(LSS_ERR_NOT_SPD, "lambda_max(BA) estimate is not positive: " + std::to_string(theta));
    lambda_max_ = theta;
    omega_ = cfg_.damping / theta;
}

// ------------------------------------------------------------------ preconditioner
void HybridSolver::base_apply(const double *r, double scale, const double *add, double *z, dev::PcgState *st,
                              const int *stop, hipStream_t s, bool first, int k, float *z32)
{
    const int64_t np = int64_t(nb_) * bs_;
    const int op = first ? dev::kFinishRzFirst : dev::kFinishRz;
    if (cfg_.base_preconditioner == "AMG") {
        // z = scale * V(r) + add per column, then r . z for the PCG.  The V-cycle has no
        // stop test: after a column stops, the PCG updates leave its x and r alone.
        for (int c = 0; c < k; ++c) {
            const double *rc = r + c * np;
            double *zc = z + c * np;
            amg_vcycle(rc, amg_z_.get(), s);
            if (add) {
                if (add + c * np != zc) k::copy(np, add + c * np, zc, s);
                k::axpy(np, scale, amg_z_.get(), zc, s);
            } else {
                k::copy(np, amg_z_.get(), zc, s);
                if (scale != 1.0) k::scale(np, zc, scale, s);
            }
        }
        if (st) {
            for (int c = 0; c < k; ++c)
                k::dot_partials(np, r_.get() + c * np, z + c * np, partials_.get() + c * k::partials_per_column(np), s);
            k::finish(k::partials_per_column(np), partials_.get(), op, st, nullptr, 0, s, k);
        }
        if (z32) k::to_f32(int64_t(k) * np, z, z32, s);
        return;
    }
    if (cfg_.base_preconditioner != "MAS") {
        const int g = k::jacobi_apply(np, d_dinv_.get(), r, scale, add, z, st ? r_.get() : nullptr, st,
                                      partials_.get(), stop, s, k, np);
        if (st) k::finish(g, partials_.get(), op, st, nullptr, 0, s, k);
        if (z32) k::to_f32(int64_t(k) * np, z, z32, s);
        return;
    }
    const int L = int(mas_.size()) - 1;
    auto ld = [&](int l) { return l == 0 ? np : int64_t(mas_[l].nodes) * bs_; };
    const bool prof = cfg_.profile_kernels && st && k == 1;
    if (L == 0) {
        if (prof) prof_begin(2, s);
        k::mas_level0_apply(prec(), bs_, mas_[0].nbanks, mas_[0].bank_off.get(), mas_[0].inv.get(), r, nullptr,
                            nullptr, scale, add, z, st ? r_.get() : nullptr, st, partials_.get(), stop, s, k, np, 0,
                            cfg_.mfma_flush_steps());
        if (prof) prof_end(s);
        if (st) k::finish(mas_[0].nbanks, partials_.get(), op, st, nullptr, 0, s, k);
        if (z32) k::to_f32(int64_t(k) * np, z, z32, s);
        return;
    }
    // The coarse levels (restriction chain, then the coarse banks from the top down) are a
    // chain of small latency-bound launches; they run on a second stream while the level-0
    // banks solve on s, and the prolongated level-1 correction is added after the join.
    LSS_HIP(hipEventRecord(ev_fork_, s));
    LSS_HIP(hipStreamWaitEvent(side_, ev_fork_, 0));
    const double *src = r;
    for (int l = 1; l <= L; ++l) {
        k::mas_restrict(bs_, mas_[l - 1].nbanks, mas_[l - 1].bank_off.get(), mas_[l - 1].parent.get(), src,
                        mas_[l].r.get(), stop, side_, k, ld(l - 1), ld(l));
        src = mas_[l].r.get();
    }
    for (int l = L; l >= 1; --l) {
        const bool top = (l == L);
        k::mas_coarse_apply(prec(), bs_, mas_[l].nodes, mas_[l].inv.get(), mas_[l].r.get(),
                            top ? nullptr : mas_[l].parent.get(), top ? nullptr : mas_[l + 1].w.get(),
                            mas_[l].w.get(), stop, side_, k, ld(l), top ? 0 : ld(l + 1));
    }
    LSS_HIP(hipEventRecord(ev_join_, side_));
    if (prof) prof_begin(2, s);
    k::mas_level0_apply(prec(), bs_, mas_[0].nbanks, mas_[0].bank_off.get(), mas_[0].inv.get(), r, nullptr, nullptr,
                        scale, add, z, nullptr, nullptr, partials_.get(), stop, s, k, np, 0, cfg_.mfma_flush_steps());
    if (prof) prof_end(s);
    LSS_HIP(hipStreamWaitEvent(s, ev_join_, 0));
    const int g = k::mas_prolong_add(np, bs_, mas_[0].parent.get(), mas_[1].w.get(), scale, z,
                                     st ? r_.get() : nullptr, partials_.get(), stop, s, k, np, ld(1), z32);
    if (st) k::finish(g, partials_.get(), op, st, nullptr, 0, s, k);
}

void HybridSolver::sparse_correct(const double *r, double *z, const int *stop, hipStream_t s, int k, int64_t ld,
                                  bool additive, float *z32)
{
    // per sparse unit: t = (r - A z)[unit rows] (additive: r[unit rows]), z[unit] += (S A S^T)^{-1} t
    for (size_t u = 0; u < sparse_.size(); ++u) {
        const int32_t *pd = d_pdofs_.get() + ns_dense_ + sparse_off_[u];
        k::stiff_residual_rows(sparse_[u].dofs(), bs_, pd, d_rowptr_.get(), d_cols_.get(), d_bvals_.get(), r,
                               additive ? nullptr : z, d_sparse_t_.get(), stop, s, k, ld);
        sparse_[u].solve_add(d_sparse_t_.get(), pd, z, z32, ld, k, stop, s);
    }
}

void HybridSolver::precond_apply(const double *r, double *z, dev::PcgState *st, const int *stop, hipStream_t s,
                                 bool first, int k)
{
    const int64_t np = int64_t(nb_) * bs_;
    if (cfg_.mas_mode()) {
        // MASSolverImpl::apply_preconditioner, MASSolver.cu:583-615: z = B r (+ S^T D^{-1} S r)
        base_apply(r, 1.0, nullptr, z, st, stop, s, first, k);
        if (ncomp_)
            k::stiff_correct(ncomp_, ns_dense_, bs_, d_comp_off_.get(), d_mat_off_.get(), d_row_comp_.get(),
                             d_pdofs_.get(), d_L_.get(), d_rowptr_.get(), d_cols_.get(), d_bvals_.get(), r, z,
                             d_swork_.get(), d_swork_.get() + kMaxCols * ns_dense_, stop, s, k, np, true, nullptr,
                             d_unit_dense_.get());
        sparse_correct(r, z, stop, s, k, np, true, nullptr);
        return;
    }
    // custom_mixed_precond_iter, GPUHybrid.cu:521-542
    if (!cfg_.do_hybrid_precond || !has_stiff()) {
        base_apply(r, 1.0, nullptr, z, st, stop, s, first, k);
        return;
    }
    // with FP32 values the inner SpMV also gathers an FP32 copy of z2 (half the gathered
    // bytes), written by the producers of z1 and of its stiff entries
    const bool f32 = cfg_.fp32_inner();
    float *z32 = f32 ? z32_.get() : nullptr;
    base_apply(r, omega_, nullptr, z, nullptr, stop, s, false, k, z32);  // z1 = B r
    // z2 = z1 + P (r - A z1)  (dss_precond_iter, GPUHybrid.cu:544-675)
    k::stiff_correct(ncomp_, ns_dense_, bs_, d_comp_off_.get(), d_mat_off_.get(), d_row_comp_.get(),
                     d_pdofs_.get(), d_L_.get(), d_rowptr_.get(), d_cols_.get(), d_bvals_.get(), r, z, d_swork_.get(),
                     d_swork_.get() + kMaxCols * ns_dense_, stop, s, k, np, false, z32, d_unit_dense_.get());
    // units are decoupled (no stored entry joins two of them), so the dense units'
    // corrections leave the sparse units' residual rows unchanged
    sparse_correct(r, z, stop, s, k, np, false, z32);
    // u = r - A z2
    const bool prof = cfg_.profile_kernels && st && k == 1;
    if (prof) prof_begin(1, s);
    k::spmv_residual(sell_, z, r, u_.get(), stop, s, f32, k, np, z32);
    if (prof) prof_end(s);
    if (cfg_.fp32_inner())
        k::stiff_rows_exact(ns_, bs_, d_pdofs_.get(), d_rowptr_.get(), d_cols_.get(), d_bvals_.get(), r, z, u_.get(),
                            stop, s, k, np);
    // z = z2 + B u
    base_apply(u_.get(), omega_, z, z, st, stop, s, first, k);
}

// ------------------------------------------------------------------ PCG
void HybridSolver::prof_begin(int kind, hipStream_t s)
{
    int *cnt = kind == 0 ? &prof_spmv_outer_n_ : (kind == 1 ? &prof_spmv_inner_n_ : &prof_mas0_n_);
    if (*cnt >= 200) {
        pending_skip_ = true;
        return;
    }
    pending_skip_ = false;
    ++*cnt;
    if (ev_used_ + 2 > ev_pool_.size()) {
        for (int q = 0; q < 64; ++q) {
            hipEvent_t e;
            LSS_HIP(hipEventCreate(&e));
            ev_pool_.push_back(e);
        }
    }
    LSS_HIP(hipEventRecord(ev_pool_[ev_used_], s));
    ev_marks_.emplace_back(kind, int(ev_used_));
    ev_used_ += 2;
}

void HybridSolver::prof_end(hipStream_t s)
{
    // closes the pair opened by the matching prof_begin (kinds never nest)
    if (pending_skip_ || ev_marks_.empty()) return;
    LSS_HIP(hipEventRecord(ev_pool_[ev_marks_.back().second + 1], s));
}

hipEvent_t HybridSolver::prof_end_event()
{
    if (pending_skip_ || ev_marks_.empty()) return nullptr;
    return ev_pool_[ev_marks_.back().second + 1];
}

void HybridSolver::prof_collect()
{
    double sum[3] = {0, 0, 0};
    int cnt[3] = {0, 0, 0};
    for (auto &m : ev_marks_) {
        float ms = 0;
        LSS_HIP(hipEventElapsedTime(&ms, ev_pool_[m.second], ev_pool_[m.second + 1]));
        sum[m.first] += ms;
        cnt[m.first]++;
    }
    prof_spmv_outer_ms_ = cnt[0] ? sum[0] / cnt[0] : 0;
    prof_spmv_inner_ms_ = cnt[1] ? sum[1] / cnt[1] : 0;
    prof_mas0_ms_ = cnt[2] ? sum[2] / cnt[2] : 0;
    prof_spmv_outer_n_ = cnt[0];
    prof_spmv_inner_n_ = cnt[1];
    prof_mas0_n_ = cnt[2];
}

void HybridSolver::run_pcg(hipStream_t s, int k)
{
    // GPUHybrid::pcg_solve, GPUHybrid.cu:1413-1512, scalars on the device.  k > 1:
    // k independent PCG recurrences in lockstep; every launch serves all columns,
    // a column that stops keeps its x and r, and the loop ends when all stopped.
    const int64_t np = int64_t(nb_) * bs_;
    dev::PcgState *st = st_.get();
    DBuf<double> sc(k);
    double bb[kMaxCols];
    k::dot(np, bp_.get(), bp_.get(), partials_.get(), sc.get(), s, k, np);
    sc.download(bb, k, s);
    sync(s);
    std::memset(h_st_, 0, sizeof(dev::PcgState) * k);
    bool any = false;
    for (int c = 0; c < k; ++c) {
        h_st_[c].bb = bb[c];
        h_st_[c].rel_eps = cfg_.relative_tolerance * cfg_.relative_tolerance;
        h_st_[c].abs_eps = cfg_.absolute_tolerance * cfg_.absolute_tolerance;
        if (!(bb[c] > 0.0)) {  // zero right-hand side: x = 0 without iterating
            h_st_[c].stop = dev::kZeroRhs;
            k::fill(np, xp_.get() + c * np, 0.0, s);
        } else {
            any = true;
        }
    }
    col_iters_.assign(k, 0);
    col_res_.assign(k, 0.0);
    col_stop_.assign(k, dev::kZeroRhs);
    restarts_ = 0;
    if (!any) {
        num_iterations_ = 0;
        final_res_norm_ = 0.0;
        stop_code_ = dev::kZeroRhs;
        return;
    }
    LSS_HIP(hipMemcpyAsync(st, h_st_, sizeof(dev::PcgState) * k, hipMemcpyHostToDevice, s));
    ev_marks_.clear();
    ev_used_ = 0;
    prof_spmv_outer_n_ = prof_spmv_inner_n_ = prof_mas0_n_ = 0;

    auto poll = [&]() {
        LSS_HIP(hipMemcpyAsync(h_st_, st, sizeof(dev::PcgState) * k, hipMemcpyDeviceToHost, s));
        sync(s);
        for (int c = 0; c < k; ++c)
            if (!h_st_[c].stop) return false;
        return true;
    };
    k::spmv_residual(sell_, xp_.get(), bp_.get(), r_.get(), nullptr, s, false, k, np);  // r = b - A x
    precond_apply(r_.get(), z_.get(), st, nullptr, s, true, k);                        // z = M r, gamma = r.z
    k::copy(np * k, z_.get(), p_.get(), s);
    int it = 0;
    while (true) {
        bool stopped = false;
        while (it < cfg_.max_iter && !stopped) {
            const int batch = std::min(cfg_.check_every, cfg_.max_iter - it);
            for (int b = 0; b < batch; ++b, ++it) {
                const bool prof = cfg_.profile_kernels && k == 1;
                // the timed pair closes after the SpMV kernel, before its finisher
                if (prof) prof_begin(0, s);
                k::spmv_pq(sell_, p_.get(), q_.get(), st, partials_.get(), it, s, k, np, dev::kFinishPq,
                           prof ? prof_end_event() : nullptr);
                k::pcg_update_xr(np, xp_.get(), r_.get(), p_.get(), q_.get(), st, partials_.get(), s, k, np);
                precond_apply(r_.get(), z_.get(), st, &st->stop, s, false, k);
                k::pcg_update_p(np, p_.get(), z_.get(), st, s, k, np);
            }
            stopped = poll();
        }
        const int code = h_st_[0].stop;
        if (k > 1 || !cfg_.true_residual_check || (code != dev::kConvergedRel && code != dev::kConvergedAbs)) break;
        // The recurrence residual drifts from b - A x on badly conditioned systems;
        // accept only a true residual inside the tolerance, else restart from it.
        k::spmv_residual(sell_, xp_.get(), bp_.get(), r_.get(), nullptr, s);
        k::dot(np, r_.get(), r_.get(), partials_.get(), sc.get(), s);
        const double rr = d2h_scalar(sc.get(), s);
        if ((h_st_->rel_eps > 0 && rr / bb[0] < h_st_->rel_eps) || (h_st_->abs_eps > 0 && rr < h_st_->abs_eps))
            break;
        if (it >= cfg_.max_iter || restarts_ >= 50) break;
        ++restarts_;
        h_st_->stop = 0;
        LSS_HIP(hipMemcpyAsync(&st->stop, &h_st_->stop, sizeof(int), hipMemcpyHostToDevice, s));
        precond_apply(r_.get(), z_.get(), st, nullptr, s, true);
        k::copy(np, z_.get(), p_.get(), s);
    }
    poll();
    // true residual ||A x - b|| per column (GPUHybrid.cu:383-386)
    k::spmv_residual(sell_, xp_.get(), bp_.get(), q_.get(), nullptr, s, false, k, np);
    k::dot(np, q_.get(), q_.get(), partials_.get(), sc.get(), s, k, np);
    double rr[kMaxCols];
    sc.download(rr, k, s);
    sync(s);
    num_iterations_ = 0;
    final_res_norm_ = 0.0;
    stop_code_ = -1;
    for (int c = 0; c < k; ++c) {
        col_iters_[c] = h_st_[c].iters;
        col_stop_[c] = h_st_[c].stop;
        col_res_[c] = std::sqrt(rr[c]);
        num_iterations_ = std::max(num_iterations_, col_iters_[c]);
        final_res_norm_ = std::max(final_res_norm_, col_res_[c]);
        // the solve reports the first column that did not converge, else column 0
        const bool conv = col_stop_[c] == dev::kConvergedRel || col_stop_[c] == dev::kConvergedAbs ||
                          col_stop_[c] == dev::kZeroRhs;
        if (stop_code_ < 0 && !conv) stop_code_ = col_stop_[c];
    }
    if (stop_code_ < 0) stop_code_ = col_stop_[0];
    if (cfg_.profile_kernels && k == 1) prof_collect();
}

void HybridSolver::solve_device(const double *b, double *x, hipStream_t s)
{
    if (!factorized_) fail(LSS_ERR_NOT_FACTORIZED, "solve called before factorize");
    if (!s) s = stream_;
    auto t0 = clk::now();
    const int64_t np = int64_t(nb_) * bs_;
    if (np != n_) {
        k::fill(np, bp_.get(), 0.0, s);
        k::fill(np, xp_.get(), 0.0, s);
    }
    k::permute_in(nb_, bs_, n_, d_perm_.get(), b, bp_.get(), s);
    k::permute_in(nb_, bs_, n_, d_perm_.get(), x, xp_.get(), s);
    // omega = damping / theta keeps omega * lambda_max(B A) < 2 while theta is within
    // damping / 2 of lambda_max; should the sweep still lose definiteness (a PCG
    // breakdown), solve again from the same start with the safe damping
    const bool damped = !cfg_.mas_mode() && cfg_.do_hybrid_precond && has_stiff() && cfg_.damping > kSafeDamping &&
                        cfg_.base_preconditioner != "AMG";
    if (damped) {
        x0p_.resize(size_t(np));
        k::copy(np, xp_.get(), x0p_.get(), s);
    }
    damping_retry_ = false;
    if (cfg_.mas_mode()) run_pcg_mas(s);
    else run_pcg(s);
    if (damped && dev::is_breakdown(stop_code_)) {
        damping_retry_ = true;
        const double omega = omega_;
        omega_ = kSafeDamping / lambda_max_;
        k::copy(np, x0p_.get(), xp_.get(), s);
        run_pcg(s);
        omega_used_ = omega_;
        omega_ = omega;
    } else {
        omega_used_ = omega_;
    }
    k::permute_out(nb_, bs_, n_, d_perm_.get(), xp_.get(), x, s);
    sync(s);
    last_k_ = 1;
    t_solve_ = since(t0);
}

void HybridSolver::solve_host(const double *b, double *x)
{
    DBuf<double> db, dx;
    db.upload(b, n_, stream_);
    dx.upload(x, n_, stream_);
    solve_device(db.get(), dx.get(), stream_);
    dx.download(x, n_, stream_);
    sync(stream_);
}

void HybridSolver::solve_multi_device(int32_t nrhs, const double *B, double *X, hipStream_t s)
{
    if (!factorized_) fail(LSS_ERR_NOT_FACTORIZED, "solve called before factorize");
    if (nrhs <= 0) fail(LSS_ERR_ARG, "number of right-hand sides must be positive");
    if (!s) s = stream_;
    // the restart logic of true_residual_check is per solve, and the MAS solver's loop
    // polls the host every 10 iterations: keep those one by one
    if (nrhs == 1 || cfg_.true_residual_check || cfg_.mas_mode()) {
        for (int32_t c = 0; c < nrhs; ++c) solve_device(B + int64_t(c) * n_, X + int64_t(c) * n_, s);
        return;
    }
    auto t0 = clk::now();
    const int64_t np = int64_t(nb_) * bs_;
    const bool damped = cfg_.do_hybrid_precond && has_stiff() && cfg_.damping > kSafeDamping &&
                        cfg_.base_preconditioner != "AMG";
    std::vector<int> iters(nrhs), stops(nrhs);
    std::vector<double> res(nrhs);
    bool retried = false;
    for (int32_t c0 = 0; c0 < nrhs;) {
        const int rem = nrhs - c0;
        const int k = rem >= kMaxCols ? kMaxCols : (rem >= 2 ? 2 : 1);
        ensure_columns(k);
        if (np != n_) {
            k::fill(np * k, bp_.get(), 0.0, s);
            k::fill(np * k, xp_.get(), 0.0, s);
        }
        for (int c = 0; c < k; ++c) {
            k::permute_in(nb_, bs_, n_, d_perm_.get(), B + int64_t(c0 + c) * n_, bp_.get() + c * np, s);
            k::permute_in(nb_, bs_, n_, d_perm_.get(), X + int64_t(c0 + c) * n_, xp_.get() + c * np, s);
        }
        if (damped) {
            x0p_.resize(size_t(np) * k);
            k::copy(np * k, xp_.get(), x0p_.get(), s);
        }
        run_pcg(s, k);
        std::vector<int> redo;
        for (int c = 0; c < k; ++c) {
            iters[c0 + c] = col_iters_[c];
            stops[c0 + c] = col_stop_[c];
            res[c0 + c] = col_res_[c];
            if (damped && dev::is_breakdown(col_stop_[c])) redo.push_back(c);
            else k::permute_out(nb_, bs_, n_, d_perm_.get(), xp_.get() + c * np, X + int64_t(c0 + c) * n_, s);
        }
        // a column whose sweep lost definiteness is solved again alone with the safe damping
        for (int c : redo) {
            retried = true;
            k::copy(np, bp_.get() + c * np, bp_.get(), s);
            k::copy(np, x0p_.get() + c * np, xp_.get(), s);
            const double omega = omega_;
            omega_ = kSafeDamping / lambda_max_;
            run_pcg(s, 1);
            omega_ = omega;
            iters[c0 + c] = num_iterations_;
            stops[c0 + c] = stop_code_;
            res[c0 + c] = final_res_norm_;
            k::permute_out(nb_, bs_, n_, d_perm_.get(), xp_.get(), X + int64_t(c0 + c) * n_, s);
        }
        c0 += k;
    }
    sync(s);
    col_iters_ = iters;
    col_stop_ = stops;
    col_res_ = res;
    num_iterations_ = *std::max_element(iters.begin(), iters.end());
    final_res_norm_ = *std::max_element(res.begin(), res.end());
    stop_code_ = stops[0];
    for (int32_t c = 0; c < nrhs; ++c)
        if (stops[c] != dev::kConvergedRel && stops[c] != dev::kConvergedAbs && stops[c] != dev::kZeroRhs) {
            stop_code_ = stops[c];
            break;
        }
    damping_retry_ = retried;
    omega_used_ = omega_;
    last_k_ = nrhs;
    t_solve_ = since(t0);
}

// ------------------------------------------------------------------ inspection
void HybridSolver::debug_row_norms(double *out) const
{
    if (!d_norms_) fail(LSS_ERR_NOT_FACTORIZED, "row norms are computed by factorize when stiff DOFs are selected");
    d_norms_.download(out, n_, stream_);
    LSS_HIP(hipStreamSynchronize(stream_));
}

void HybridSolver::debug_amg(int32_t level, int32_t *nlevels, int32_t *n, int64_t *nnz_a, int64_t *nnz_p,
                             int8_t *cf, int64_t *a_ptr, int32_t *a_col, double *a_val, int64_t *p_ptr,
                             int32_t *p_col, double *p_val)
{
    if (!factorized_ || amg_.empty()) fail(LSS_ERR_NOT_FACTORIZED, "no AMG hierarchy (base_preconditioner AMG)");
    if (nlevels) *nlevels = int32_t(amg_.size());
    if (level < 0 || level >= int32_t(amg_.size())) {
        if (n || nnz_a || nnz_p || cf || a_ptr || p_ptr) fail(LSS_ERR_ARG, "AMG level out of range");
        return;
    }
    const AmgLevel &L = amg_[level];
    if (n) *n = L.n;
    const bool has_p = L.P.ptr.get() != nullptr;
    if (nnz_a) *nnz_a = level == 0 ? nnz_ : L.A.nnz;
    if (nnz_p) *nnz_p = has_p ? L.P.nnz : 0;
    if (cf && !amg_cf_[level].empty()) std::memcpy(cf, amg_cf_[level].data(), amg_cf_[level].size());
    if (level > 0 && a_ptr) {
        L.A.ptr.download(a_ptr, size_t(L.n) + 1, stream_);
        L.A.col.download(a_col, size_t(L.A.nnz), stream_);
        L.A.val.download(a_val, size_t(L.A.nnz), stream_);
    }
    if (has_p && p_ptr) {
        L.P.ptr.download(p_ptr, size_t(L.n) + 1, stream_);
        L.P.col.download(p_col, size_t(L.P.nnz), stream_);
        L.P.val.download(p_val, size_t(L.P.nnz), stream_);
    }
    sync(stream_);
}

void HybridSolver::debug_spmv(const double *x, double *y)
{
    if (!factorized_) fail(LSS_ERR_NOT_FACTORIZED, "not factorized");
    DBuf<double> dx;
    dx.upload(x, n_, stream_);
    k::fill(int64_t(nb_) * bs_, xp_.get(), 0.0, stream_);
    k::permute_in(nb_, bs_, n_, d_perm_.get(), dx.get(), xp_.get(), stream_);
    k::spmv(sell_, xp_.get(), q_.get(), nullptr, stream_);
    k::permute_out(nb_, bs_, n_, d_perm_.get(), q_.get(), dx.get(), stream_);
    dx.download(y, n_, stream_);
    sync(stream_);
}

void HybridSolver::debug_base(const double *r, double *z)
{
    if (!factorized_) fail(LSS_ERR_NOT_FACTORIZED, "not factorized");
    DBuf<double> dr;
    dr.upload(r, n_, stream_);
    k::fill(int64_t(nb_) * bs_, bp_.get(), 0.0, stream_);
    k::permute_in(nb_, bs_, n_, d_perm_.get(), dr.get(), bp_.get(), stream_);
    base_apply(bp_.get(), 1.0, nullptr, z_.get(), nullptr, nullptr, stream_);
    k::permute_out(nb_, bs_, n_, d_perm_.get(), z_.get(), dr.get(), stream_);
    dr.download(z, n_, stream_);
    sync(stream_);
}

void HybridSolver::debug_precond(const double *r, double *z)
{
    if (!factorized_) fail(LSS_ERR_NOT_FACTORIZED, "not factorized");
    DBuf<double> dr;
    dr.upload(r, n_, stream_);
    k::fill(int64_t(nb_) * bs_, r_.get(), 0.0, stream_);
    k::permute_in(nb_, bs_, n_, d_perm_.get(), dr.get(), r_.get(), stream_);
    precond_apply(r_.get(), z_.get(), nullptr, nullptr, stream_);
    k::permute_out(nb_, bs_, n_, d_perm_.get(), z_.get(), dr.get(), stream_);
    dr.download(z, n_, stream_);
    sync(stream_);
}

std::string HybridSolver::info_json() const
{
    static const char *stops[] = {"max_iter",           "converged_rel",       "converged_abs", "err_zero_sdotp",
                                  "err_negative_alpha", "err_subnormal_alpha", "zero_rhs"};
    auto stop_name = [&](int c) { return stops[(c >= 0 && c < 7) ? c : 0]; };
    int64_t nst = 0;
    for (auto &c : comps_) nst += int64_t(c.size());
    std::ostringstream o;
    o.precision(17);
    o << "{\"num_iterations\":" << num_iterations_ << ",\"final_res_norm\":" << final_res_norm_
      << ",\"restarts\":" << restarts_ << ",\"damping_retry\":" << (damping_retry_ ? "true" : "false")
      << ",\"omega_used\":" << omega_used_ << ",\"solver\":\""
      << name() << "\",\"stop_reason\":\"" << stop_name(stop_code_)
      << "\",\"n\":" << n_ << ",\"nnz\":" << nnz_ << ",\"block_size\":" << bs_ << ",\"nnz_blocks\":" << nnzb_
      << ",\"sell_slots\":" << sell_slots_ << ",\"partition_reused\":" << (partition_reused_ ? "true" : "false") << ",\"precision\":\"" << cfg_.precision
      << "\",\"bank_apply\":\"" << (cfg_.mfma_flush_steps() ? "mfma" : "vector") << "\",\"mfma_flush\":"
      << cfg_.mfma_flush_steps() << ",\"inner_spmv_fp32\":"
      << (cfg_.fp32_inner() ? "true" : "false") << ",\"inner_spmv_values\":\""
      << (cfg_.bf16_inner() ? "bf16" : (cfg_.fp32_inner() ? "fp32" : "fp64")) << "\"" << ",\"omega\":" << omega_ << ",\"lambda_max_BA\":" << lambda_max_ << ",\"timings\":{\"factorize\":"
      << t_factorize_ << ",\"select\":" << t_select_ << ",\"operator\":" << t_operator_ << ",\"partition\":"
      << t_partition_ << ",\"mas\":" << t_mas_ << ",\"stiff\":" << t_stiff_ << ",\"omega\":" << t_omega_
      << ",\"solve\":" << t_solve_ << "},\"stiff\":{\"gmm\":{\"global_mean\":" << gmm_.global_mean
      << ",\"global_var\":" << gmm_.global_var << ",\"mean0\":" << gmm_.mean0 << ",\"mean1\":" << gmm_.mean1
      << ",\"var0\":" << gmm_.var0 << ",\"var1\":" << gmm_.var1 << ",\"w0\":" << gmm_.w0 << ",\"w1\":" << gmm_.w1
      << ",\"iters\":" << gmm_.iters << ",\"jumped\":" << gmm_.jumped << ",\"selected\":" << gmm_.selected
      << "},\"n_selected\":" << n_selected_ << ",\"n_filtered\":" << n_filtered_ << ",\"n_expanded\":" << n_expanded_
      << ",\"filter\":{\"too_small\":" << filter_.too_small << ",\"too_large\":" << filter_.too_large
      << ",\"not_poorly_conditioned\":" << filter_.not_poorly_conditioned << "},\"n_components\":" << comps_.size()
      << ",\"n_stiff_dofs\":" << nst << ",\"max_component\":" << max_comp_ << ",\"dense_branch\":{\"bins\":"
      << dense_bins_ << ",\"dim\":" << dense_dim_ << ",\"dofs\":" << dense_dofs_ << ",\"bin_sizes\":[";
    for (size_t b = 0; b < bin_sizes_.size(); ++b) o << (b ? "," : "") << bin_sizes_[b];
    o << "]},\"sparse\":{\"units\":" << sparse_.size() << ",\"dofs\":[";
    for (size_t u = 0; u < sparse_.size(); ++u) o << (u ? "," : "") << sparse_[u].dofs();
    o << "],\"fronts\":[";
    for (size_t u = 0; u < sparse_.size(); ++u) o << (u ? "," : "") << sparse_[u].fronts();
    o << "],\"levels\":[";
    for (size_t u = 0; u < sparse_.size(); ++u) o << (u ? "," : "") << sparse_[u].levels();
    o << "],\"max_front\":[";
    for (size_t u = 0; u < sparse_.size(); ++u) o << (u ? "," : "") << sparse_[u].max_front();
    o << "],\"factor_entries\":[";
    for (size_t u = 0; u < sparse_.size(); ++u) o << (u ? "," : "") << sparse_[u].factor_entries();
    o << "],\"update_bytes\":[";
    for (size_t u = 0; u < sparse_.size(); ++u) o << (u ? "," : "") << sparse_[u].stack_bytes();
    o << "],\"batched_fronts\":[";
    for (size_t u = 0; u < sparse_.size(); ++u) o << (u ? "," : "") << sparse_[u].batched_fronts();
    o << "],\"front_groups\":[";
    for (size_t u = 0; u < sparse_.size(); ++u) o << (u ? "," : "") << sparse_[u].front_groups();
    o << "],\"t_symbolic\":[";
    for (size_t u = 0; u < sparse_.size(); ++u) o << (u ? "," : "") << sparse_[u].t_symbolic;
    o << "],\"t_numeric\":[";
    for (size_t u = 0; u < sparse_.size(); ++u) o << (u ? "," : "") << sparse_[u].t_numeric;
    o << "]}},\"mas\":{\"levels\":" << mas_levels_
      << ",\"banks\":" << mas_banks_ << ",\"inverse_bytes\":" << mas_bytes_ << ",\"parts\":"
      << (part_off_.empty() ? 0 : part_off_.size() - 1) << ",\"level_nodes\":[";
    for (size_t l = 0; l < level_nodes_.size(); ++l) o << (l ? "," : "") << level_nodes_[l];
    o << "]},\"amg\":{\"levels\":" << amg_.size() << ",\"rows\":[";
    for (size_t l = 0; l < amg_sizes_.size(); ++l) o << (l ? "," : "") << amg_sizes_[l];
    o << "],\"nnz\":[";
    for (size_t l = 0; l < amg_nnz_.size(); ++l) o << (l ? "," : "") << amg_nnz_[l];
    o << "],\"setup_s\":{\"norms\":" << amg_times_.norms << ",\"strength\":" << amg_times_.strength
      << ",\"pmis\":" << amg_times_.pmis << ",\"interp\":" << amg_times_.interp << ",\"transpose\":"
      << amg_times_.transpose << ",\"rap\":" << amg_times_.rap << "}},\"kernel_ms\":{\"spmv_outer\":" << prof_spmv_outer_ms_ << ",\"spmv_outer_n\":" << prof_spmv_outer_n_
      << ",\"spmv_inner\":" << prof_spmv_inner_ms_ << ",\"spmv_inner_n\":" << prof_spmv_inner_n_
      << ",\"mas_level0\":" << prof_mas0_ms_ << ",\"mas_level0_n\":" << prof_mas0_n_ << "}";
    if (cfg_.mas_mode()) {
        // MASSolver::get_info, MASSolver.cu:311-316
        static const char *status[] = {"Running", "Reach relative tolerance", "Reach absolute tolerance",
                                       "Reach max iterations"};
        const MasSelection &m = mas_sel_;
        o << ",\"solver_iter\":" << num_iterations_ << ",\"solver_error\":" << mas_residual_
          << ",\"solver_status\":\"" << status[(mas_status_ >= 0 && mas_status_ < 4) ? mas_status_ : 0]
          << "\",\"bad_dof\":{\"enabled\":" << (cfg_.use_bad_dof_precond ? "true" : "false") << ",\"strategy\":\""
          << cfg_.bad_dof_strategy << "\",\"selected\":" << m.selected << ",\"expanded\":" << m.expanded
          << ",\"mean0\":" << m.mean0 << ",\"mean1\":" << m.mean1 << ",\"var0\":" << m.var0 << ",\"var1\":" << m.var1
          << ",\"global_mean\":" << m.global_mean << ",\"global_var\":" << m.global_var << ",\"iters\":" << m.iters
          << ",\"ratio\":" << m.ratio << ",\"c0\":" << m.c0 << ",\"c1\":" << m.c1 << ",\"cutoff\":" << m.cutoff
          << "}";
    }
    if (last_k_ > 1) {
        // per right-hand side of the last multi solve
        o << ",\"columns\":[";
        for (size_t c = 0; c < col_iters_.size(); ++c)
            o << (c ? "," : "") << "{\"num_iterations\":" << col_iters_[c] << ",\"final_res_norm\":" << col_res_[c]
              << ",\"stop_reason\":\"" << stop_name(col_stop_[c]) << "\"}";
        o << "]";
    }
    o << "}";
    return o.str();
}

}  // namespace lss
