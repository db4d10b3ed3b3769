// HybridSolver: the MI355X implementation of GPUHybrid factorize()/solve()
// (src/polysolve/linear/GPUHybrid.hpp:27-192) with B = MAS.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <string>
#include <vector>

#include "../common/runtime.hpp"
#include "../kernels/kernels.hpp"
#include "../kernels/mas_kernels.hpp"
#include "config.hpp"
#include "dense.hpp"
#include "sparse_stiff.hpp"
#include "stiff_select.hpp"
#include "topology.hpp"

namespace lss {

struct GmmResult {
    double global_mean = 0, global_var = 0, mean0 = 0, mean1 = 0, var0 = 0, var1 = 0, w0 = 0, w1 = 0;
    int iters = 0;
    int jumped = 0;
    int64_t selected = 0;
};

class HybridSolver {
public:
    explicit HybridSolver(const Config &cfg);
    ~HybridSolver();
    HybridSolver(const HybridSolver &) = delete;
    HybridSolver &operator=(const HybridSolver &) = delete;

    // A change of any key that shapes the factorization drops it (buffers and layouts
    // were built for the old keys): solve then reports LSS_ERR_NOT_FACTORIZED.
    void set_config(const Config &cfg)
    {
        if (factorized_ && !cfg.same_factorization(cfg_)) reset_factorization();
        cfg_ = cfg;
    }
    const Config &config() const { return cfg_; }
    void set_problematic_dofs(const int32_t *dofs, int64_t n);

    void factorize_host(int32_t n, int64_t nnz, const int32_t *colptr, const int32_t *rowidx, const double *vals);
    void factorize_device(int32_t n, int64_t nnz, const int32_t *ptr, const int32_t *idx, const double *vals,
                          hipStream_t s);
    void solve_host(const double *b, double *x);
    void solve_device(const double *b, double *x, hipStream_t s);
    // nrhs right-hand sides (column c at B + c * n), solved in fused groups of up
    // to kMaxCols columns that share every matrix and bank read
    void solve_multi_device(int32_t nrhs, const double *B, double *X, hipStream_t s);
    static constexpr int kMaxCols = 4;

    std::string info_json() const;
    const char *name() const;
    bool factorized() const { return factorized_; }
    int32_t size() const { return n_; }

    // inspection
    void debug_row_norms(double *out) const;
    const GmmResult &gmm() const { return gmm_; }
    const std::vector<int32_t> &perm() const { return perm_; }
    const std::vector<int32_t> &part_off() const { return part_off_; }
    const std::vector<std::vector<int32_t>> &components() const { return comps_; }
    void debug_spmv(const double *x, double *y);
    void debug_amg(int32_t level, int32_t *nlevels, int32_t *n, int64_t *nnz_a, int64_t *nnz_p, int8_t *cf,
                   int64_t *a_ptr, int32_t *a_col, double *a_val, int64_t *p_ptr, int32_t *p_col, double *p_val);
    void debug_base(const double *r, double *z);
    void debug_precond(const double *r, double *z);

private:
    struct DevLevel {
        int32_t nodes = 0, nbanks = 0;
        DBuf<int32_t> bank_off;
        DBuf<uint8_t> inv;         // bank factors, precision-dependent element type and layout
        DBuf<int32_t> rowptr, cols;  // level operator (levels >= 1)
        DBuf<double> vals;
        DBuf<int32_t> parent;      // to the next level (restriction: children of a coarse node lie in one bank)
        DBuf<double> r, w;         // restricted residual / prolongated correction
    };

    void release();              // frees every device buffer
    void reset_factorization();  // clears the factorization; buffers stay allocated for reuse
    void select_stiff(int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals, hipStream_t s);
    // MAS solver: stiff set of the elbow / k-means / GMM strategy on the permuted operator
    void select_mas_bad(int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals, hipStream_t s);
    std::vector<int64_t> mas_gmm_positions(int64_t n, const double *x, hipStream_t s);
    void run_pcg_mas(hipStream_t s);
    RowSet fetch_rows(const std::vector<int32_t> &rows, const int32_t *ptr, const int32_t *idx, const double *vals,
                      bool with_vals, hipStream_t s);
    void build_operator(int32_t n, int64_t nnz, const int32_t *ptr, const int32_t *idx, const double *vals,
                        hipStream_t s);
    // MAS coarse spaces: device kernels (false if a bank is too dense for them) or host
    bool build_levels_device(DBuf<int32_t> &&part_off, int32_t nparts, hipStream_t s);
    void build_levels_host(const Pattern &gp, hipStream_t s);
    void build_mas(hipStream_t s);
    // AMG base preconditioner: hierarchy set up on the host, V-cycle on the device
    void build_amg(int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals, hipStream_t s);
    void amg_vcycle(const double *b, double *x, hipStream_t s);
    void build_stiff(int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals, hipStream_t s);
    void estimate_omega(hipStream_t s);
    void alloc_vectors();
    void ensure_columns(int k);  // PCG vectors for k columns (stride nb * bs)

    // preconditioner pieces on permuted vectors
    // st != nullptr: gamma = r . z is reduced into the PCG state (first: before the loop)
    // k columns at stride nb * bs, states st[0..k), stop = &st[0].stop
    // z32 != nullptr: also an FP32 copy of z (k columns at stride n)
    void base_apply(const double *r, double scale, const double *add, double *z, dev::PcgState *st, const int *stop,
                    hipStream_t s, bool first = false, int k = 1, float *z32 = nullptr);
    void precond_apply(const double *r, double *z, dev::PcgState *st, const int *stop, hipStream_t s,
                       bool first = false, int k = 1);
    void run_pcg(hipStream_t s, int k = 1);

    k::Precision prec() const;
    void sync(hipStream_t s) const { LSS_HIP(hipStreamSynchronize(s)); }

    Config cfg_;
    hipStream_t stream_ = nullptr;
    hipStream_t side_ = nullptr;  // MAS coarse levels, concurrent with the level-0 banks
    hipEvent_t ev_fork_ = nullptr, ev_join_ = nullptr;
    // stiff selection and factorization on a worker thread during factorize
    hipStream_t aux_ = nullptr;
    hipEvent_t ev_aux_ = nullptr;
    bool factorized_ = false;
    int32_t n_ = 0, nb_ = 0;
    int bs_ = 1;
    int64_t nnz_ = 0, nnzb_ = 0;

    // stiff selection
    std::vector<int32_t> user_set_;
    bool user_set_given_ = false;
    GmmResult gmm_;
    FilterStats filter_;
    int64_t n_selected_ = 0, n_filtered_ = 0, n_expanded_ = 0;
    std::vector<std::vector<int32_t>> comps_;
    DBuf<double> d_norms_;

    // permutation / operator
    std::vector<int32_t> perm_, part_off_;
    DBuf<int32_t> d_perm_;
    DBuf<int32_t> d_rowptr_, d_cols_;
    DBuf<double> d_bvals_;
    DBuf<int64_t> d_slice_ptr_;
    DBuf<int32_t> d_scols_;
    DBuf<double> d_svals_;
    DBuf<float> d_svals32_;
    DBuf<uint16_t> d_svals16_;  // bfloat16 copy (inner_spmv "bf16")
    k::SellView sell_;
    int64_t sell_slots_ = 0;

    // base preconditioner
    std::vector<DevLevel> mas_;
    DBuf<uint8_t> spare_inv0_;  // the previous factorization's level-0 factor buffer, for reuse
    std::vector<int32_t> level_nodes_;
    DBuf<double> d_dinv_;
    struct AmgLevel {
        int32_t n = 0;
        k::AmgCsrBuf A, P, R;  // A: levels >= 1 (level 0 uses the sliced operator)
        DBuf<double> dinv, b, x, t;
    };
    std::vector<AmgLevel> amg_;
    k::AmgWork amg_work_;  // setup scratch, kept for refactorizations
    bool partition_reused_ = false;
    std::vector<std::vector<int8_t>> amg_cf_;  // C/F splitting per level (host, inspection)
    std::vector<int32_t> amg_sizes_;
    std::vector<int64_t> amg_nnz_;
    struct AmgSetupTimes {
        double norms = 0, strength = 0, pmis = 0, interp = 0, transpose = 0, rap = 0;
    } amg_times_;
    DBuf<double> amg_z_;  // V-cycle output before scaling

    // stiff blocks: ncomp_ dense units (their rows first in d_pdofs_), then the units
    // factored by the multifrontal path (sparse_, rows from ns_dense_ on); ns_ = all rows
    int32_t ncomp_ = 0;
    int64_t ns_ = 0, ns_dense_ = 0;
    std::vector<SparseFactor> sparse_;
    std::vector<int64_t> sparse_off_;  // first row of each sparse unit after ns_dense_
    DBuf<double> d_sparse_t_;          // residual of a sparse unit's rows, kMaxCols columns
    bool has_stiff() const { return ncomp_ > 0 || !sparse_.empty(); }
    void sparse_correct(const double *r, double *z, const int *stop, hipStream_t s, int k, int64_t ld, bool additive,
                        float *z32);
    int64_t max_comp_ = 0;
    DBuf<int64_t> d_comp_off_, d_mat_off_;
    DBuf<int32_t> d_pdofs_, d_row_comp_;
    DBuf<double> d_L_, d_swork_;
    DBuf<uint8_t> d_unit_dense_;  // per unit: 1 = dense-branch bin holding its explicit inverse
    int64_t dense_bins_ = 0, dense_dim_ = 0, dense_dofs_ = 0;
    std::vector<int64_t> bin_sizes_;
    DenseBlas dense_;  // blocked factorization of large components
    double omega_ = 1.0;
    double omega_used_ = 1.0;  // omega of the last solve (after a damping retry)
    bool damping_retry_ = false;
    DBuf<double> x0p_;         // start vector kept for a damping retry
    double lambda_max_ = 0.0;

    // PCG
    DBuf<double> xp_, bp_, r_, p_, q_, z_, u_;
    DBuf<float> z32_;  // FP32 copy of z2 for the FP32 preconditioner-internal SpMV
    DBuf<double> partials_;
    DBuf<dev::PcgState> st_;         // kMaxCols states
    dev::PcgState *h_st_ = nullptr;  // pinned mirror
    int kcap_ = 1;                   // columns the PCG vectors hold
    DBuf<int> d_status_;

    // info
    int num_iterations_ = 0;
    double final_res_norm_ = 0.0;
    int stop_code_ = 0;
    int restarts_ = 0;
    int last_k_ = 1;  // right-hand sides of the last solve
    // MAS solver (MASSolver.cu): status, selection statistics
    int mas_status_ = 0;  // 0 running, 1 relative tolerance, 2 absolute tolerance, 3 max iterations
    double mas_residual_ = 0.0;
    struct MasSelection {
        std::string strategy;
        double mean0 = 0, mean1 = 0, var0 = 0, var1 = 0, global_mean = 0, global_var = 0, ratio = 0;
        double c0 = 0, c1 = 0;  // k-means centres
        int iters = 0;
        int64_t cutoff = -1;  // elbow position
        int64_t selected = 0, expanded = 0;
    } mas_sel_;
    std::vector<int> col_iters_, col_stop_;
    std::vector<double> col_res_;
    double t_factorize_ = 0, t_select_ = 0, t_operator_ = 0, t_partition_ = 0, t_mas_ = 0, t_stiff_ = 0,
           t_omega_ = 0, t_solve_ = 0;
    int mas_levels_ = 0;
    int64_t mas_banks_ = 0;
    double mas_bytes_ = 0;
    // kernel timing (profile_kernels)
    double prof_spmv_outer_ms_ = 0, prof_spmv_inner_ms_ = 0, prof_mas0_ms_ = 0;
    int prof_spmv_outer_n_ = 0, prof_spmv_inner_n_ = 0, prof_mas0_n_ = 0;
    bool pending_skip_ = false;
    std::vector<hipEvent_t> ev_pool_;
    std::vector<std::pair<int, int>> ev_marks_;  // (kind, event index)
    size_t ev_used_ = 0;
    void prof_begin(int kind, hipStream_t s);
    void prof_end(hipStream_t s);
    hipEvent_t prof_end_event();  // the closing event of the open pair (nullptr when skipped)
    void prof_collect();
};

}  // namespace lss
