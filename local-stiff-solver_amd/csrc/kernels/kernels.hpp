// Launchers of the hand-written gfx950 kernels.  All take an explicit stream.
#pragma once
#include <vector>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_utils.hpp"
#include "../common/runtime.hpp"

namespace lss::k {

// ---------------------------------------------------------------- formats
// Scalar CSR (operator rows) -> block rows: number of distinct block columns per block row.
void bsr_count(int32_t n, int bs, const int32_t *ptr, const int32_t *idx, int32_t *cnt, int32_t *unsorted_flag,
               hipStream_t s);
// Fill block rows (block-major, row-major inside a block).  Missing tail scalars of
// the last block row get a unit diagonal (reference BSRMatrix.cu:123-134, 189-191).
void bsr_fill(int32_t n, int bs, const int32_t *ptr, const int32_t *idx, const double *vals, const int32_t *rowptr,
              int32_t *cols, double *bvals, hipStream_t s);
// Relabel block rows/cols: new row perm[I] <- old row I, columns mapped and sorted.
void bsr_permute(int32_t nb, int bs, const int32_t *rowptr, const int32_t *cols, const double *vals,
                 const int32_t *perm, const int32_t *new_rowptr, int32_t *new_cols, double *new_vals, hipStream_t s);
// plen[perm[I]] = len[I]: row lengths in the permuted numbering.
void permute_lengths(int32_t nb, const int32_t *perm, const int32_t *len, int32_t *plen, hipStream_t s);
// Sliced block-ELL layout (one wave per slice of 64/bs block rows).
void sell_width(int32_t nb, int bs, const int32_t *rowptr, int32_t *slice_width, hipStream_t s);
void sell_fill(int32_t nb, int bs, const int32_t *rowptr, const int32_t *cols, const double *vals,
               const int64_t *slice_ptr, int32_t *scols, double *svals, float *svals32, uint16_t *svals16,
               hipStream_t s);

// ---------------------------------------------------------------- SpMV
struct SellView {
    int32_t nb = 0;
    int bs = 1;
    int32_t nslices = 0;
    const int64_t *slice_ptr = nullptr;  // nslices + 1, in slots (slot = one block)
    const int32_t *cols = nullptr;       // per slot, -1 = padding
    const double *vals = nullptr;        // per slot bs*bs
    const float *vals32 = nullptr;       // optional FP32 copy
    const uint16_t *vals16 = nullptr;    // optional bfloat16 copy (taken before vals32)
};

// y = A x (sub == nullptr) or y = sub - A x.  spmv_pq also reduces p . A p and
// updates alpha / the breakdown flags of the PCG state.  With stop != nullptr the
// kernels do nothing once the stop words of all k columns are set.  k = 1, 2 or 4
// right-hand sides share one pass over the matrix (column c at v + c * ld).
void spmv(const SellView &A, const double *x, double *y, const double *sub, hipStream_t s);
// `after`, when given, is recorded on s between the SpMV and the finisher (kernel timing).
void spmv_pq(const SellView &A, const double *p, double *q, dev::PcgState *st, double *partials, int iter,
             hipStream_t s, int k = 1, int64_t ld = 0, int finish_op = dev::kFinishPq, hipEvent_t after = nullptr);
// fp32: the FP32 (or, when present, bfloat16) values with x32, an FP32 copy of x at the
// same stride (x unused)
void spmv_residual(const SellView &A, const double *x, const double *b, double *r, const int *stop, hipStream_t s,
                   bool fp32 = false, int k = 1, int64_t ld = 0, const float *x32 = nullptr);
// y = float(x) over n entries
void to_f32(int64_t n, const double *x, float *y, hipStream_t s);
// Partials one column of spmv_pq writes.
int spmv_partials_per_column(const SellView &A);
// Values of the sliced layout for `slots` block slots (counted per block row).
int64_t sell_value_count(int64_t slots, int bs);

// ---------------------------------------------------------------- BLAS-1 / PCG
constexpr int kCsumChunks = 1024;  // canonical reduction tree (shared with the oracle)
constexpr int kCsumLanes = 256;
// Canonical-order sums used by the stiff DOF selection.
void canonical_sum(int64_t n, const double *v, double *partials, double *out, hipStream_t s);

// Several right-hand sides (k > 1): column c of a vector v is v + c * ld, its
// partials start at partials + c * count and its PCG state is st[c].

// Sum count partials in index order and apply the scalar update `op` (dev::FinishOp),
// one workgroup per column.
void finish(int count, const double *partials, int op, dev::PcgState *st, double *out, int iter, hipStream_t s,
            int k = 1);
// Partials one column of dot / pcg_update_xr writes for vectors of length n.
int partials_per_column(int64_t n);
// per-workgroup partials of a . b (one column); returns their count (reduce with finish)
int dot_partials(int64_t n, const double *a, const double *b, double *partials, hipStream_t s);
// out[c] = a_c . b_c
void dot(int64_t n, const double *a, const double *b, double *partials, double *out, hipStream_t s, int k = 1,
         int64_t ld = 0);
// x += alpha p; r -= alpha q; rr = r.r; convergence test (columns that stopped are left alone).
void pcg_update_xr(int64_t n, double *x, double *r, const double *p, const double *q, dev::PcgState *st,
                   double *partials, hipStream_t s, int k = 1, int64_t ld = 0);
// p = beta p + z.
void pcg_update_p(int64_t n, double *p, const double *z, const dev::PcgState *st, hipStream_t s, int k = 1,
                  int64_t ld = 0);
// y = sign * alpha * x + y with alpha from the PCG state (MASSolver.cu:651-666 axpby)
void axpy_alpha(int64_t n, double sign, const double *x, double *y, const dev::PcgState *st, hipStream_t s);
void copy(int64_t n, const double *a, double *b, hipStream_t s);
void fill(int64_t n, double *a, double v, hipStream_t s);
void scale(int64_t n, double *a, double v, hipStream_t s);
void axpy(int64_t n, double a, const double *x, double *y, hipStream_t s);  // y += a x
void xpby(int64_t n, const double *x, double b, double *y, hipStream_t s);  // y = x + b y
// permuted[i] = src[perm-mapped], for node permutation with block size bs
void permute_in(int32_t nb, int bs, int64_t n, const int32_t *perm, const double *src, double *dst, hipStream_t s);
void permute_out(int32_t nb, int bs, int64_t n, const int32_t *perm, const double *src, double *dst, hipStream_t s);

// ---------------------------------------------------------------- stiff DOF selection
void row_l1(int32_t n, const int32_t *ptr, const double *vals, double *out, hipStream_t s);
// min and max of v into out2[0..1]; part holds 2 * kMinmaxBlocks partials
constexpr int kMinmaxBlocks = 512;
void minmax(int64_t n, const double *v, double *part, double *out2, hipStream_t s);
struct GmmParams {
    double mean0, mean1, var0, var1, log_w0, log_w1, c0, c1;
};
// E-step: gamma0/gamma1 + canonical partial sums of {g0, g1, g0 x, g1 x}.
void gmm_estep(int64_t n, const double *x, GmmParams p, double *g0, double *g1, double *partials4, hipStream_t s);
// canonical partial sums of {g0 (x-m0)^2, g1 (x-m1)^2}
void gmm_var(int64_t n, const double *x, const double *g0, const double *g1, double m0, double m1,
             double *partials2, hipStream_t s);
void gmm_var_global(int64_t n, const double *x, double mean, double *partials1, hipStream_t s);
void csum_finish(int nsums, const double *partials, double *out, hipStream_t s);
// flags[i] = g1[i] > g0[i] (high = 1) or g0[i] > g1[i] (high = 0)
void gmm_flags(int64_t n, const double *g0, const double *g1, uint8_t *flags, hipStream_t s, int high = 1);
// Ascending indices of the nonzero flags (compacted on the device; synchronizes s).
std::vector<int32_t> flagged_indices(int32_t n, const uint8_t *flags, hipStream_t s);
// len[k] = ptr[rows[k] + 1] - ptr[rows[k]] for k < m.
void row_lengths(int64_t m, const int32_t *rows, const int32_t *ptr, int32_t *len, hipStream_t s);

// ---------------------------------------------------------------- MAS solver selection
// Per permuted scalar row of the block operator: mode 0 sum |a_ij|, 1 |a_ii|,
// 2 max |a_ij| (j != i), each in the reference's loop order; padding rows get 0.
void mas_indicator(int mode, int32_t nb, int bs, const int32_t *rowptr, const int32_t *cols, const double *vals,
                   int32_t pblock, int npad, double *out, hipStream_t s);
// Stable sort of n keys (ascending or descending) with their positions 0..n-1; synchronizes s.
void sort_positions(int32_t n, const double *keys, bool descending, double *keys_out, int32_t *pos_out,
                    hipStream_t s);

// ---------------------------------------------------------------- partition
// Recursive breadth-first bisection of the block graph (rowptr/cols, rows
// ascending) into parts of at most max_part nodes; writes perm[old] = new and
// part_off (nparts + 1 entries, resized); returns nparts.  Synchronizes s.
// Subsets of at most local_max nodes (capped at 8192; 0: none) finish their
// recursion in one workgroup each, with the same result.
int32_t device_partition(int32_t nb, const int32_t *rowptr, const int32_t *cols, int max_part, int32_t *perm,
                         DBuf<int32_t> &part_off, hipStream_t s, int32_t local_max = 8192);

// Exclusive scan of n int32 values (out may not alias a); tsum holds
// ceil(n / 1024) values; the total goes to *total when it is not null.
void excl_scan_i32(int32_t n, const int32_t *a, int32_t *out, int32_t *tsum, int32_t *total, hipStream_t s);

// ---------------------------------------------------------------- MAS coarse spaces
// Components of each bank (at most 32 nodes): parent[fine_n], child_ptr[coarse_n + 1],
// child[fine_n] (children ascending), base/count per bank; returns coarse_n.  Synchronizes s.
int32_t device_components(int32_t fine_n, const int32_t *rowptr, const int32_t *cols, int32_t nbanks,
                          const int32_t *bank_off, DBuf<int32_t> &parent, DBuf<int64_t> &child_ptr,
                          DBuf<int32_t> &child, DBuf<int32_t> &base, DBuf<int32_t> &count, hipStream_t s);
// Pattern of R A R^T and the fine entries each coarse entry sums (ascending).
// Returns false (nothing built) when a bank holds more entries than one workgroup sorts.
bool device_galerkin_pattern(int64_t fine_nnz, const int32_t *rowptr, const int32_t *cols, int32_t nbanks,
                             const int32_t *bank_off, const int32_t *base, const int32_t *count,
                             const int32_t *parent, int32_t coarse_n, DBuf<int32_t> &crowptr, DBuf<int32_t> &ccols,
                             DBuf<int64_t> &src_ptr, DBuf<int32_t> &src_blk, int64_t &nslots, hipStream_t s);
// bank_off = {0, 32, 64, ..., n}; returns the bank count.
int32_t uniform_banks(int32_t n, DBuf<int32_t> &bank_off, hipStream_t s);

// ---------------------------------------------------------------- AMG V-cycle
// A scalar CSR matrix in device memory (n rows).
struct AmgCsr {
    int32_t n = 0;
    const int64_t *ptr = nullptr;
    const int32_t *col = nullptr;
    const double *val = nullptr;
    int64_t nnz = 0;
};
// Owning scalar CSR on the device (n x m, nnz entries, columns ascending per row).
struct AmgCsrBuf {
    int32_t n = 0, m = 0;
    int64_t nnz = 0;
    DBuf<int64_t> ptr;
    DBuf<int32_t> col;
    DBuf<double> val;
    AmgCsr view() const { return AmgCsr{n, ptr.get(), col.get(), val.get(), nnz}; }
};
// Scratch of the AMG setup: grows to the largest product of a setup and is kept
// (a refactorization reuses it without new allocations).
struct AmgWork {
    DBuf<uint64_t> key, key2;
    DBuf<double> val, val2;
    DBuf<uint8_t> tmp;
    DBuf<int32_t> head, uidx, row;
    DBuf<int64_t> cnt, off;
};
// AMG setup on the device (amg_setup.hip).  All of them wait for their work.
// p64[0..count) = p32[0..count)
void amg_widen_ptr(int64_t count, const int32_t *p32, int64_t *p64, hipStream_t s);
// dinv_i = 1 / sum_j |a_ij| (0 for an empty row)
void amg_l1_inverse(const AmgCsr &A, double *dinv, hipStream_t s);
// classical strength of connection (pattern only) with the "unknown" approach for nf > 1
void amg_strength_device(const AmgCsr &A, const int32_t *func, int nf, double theta, double max_row_sum,
                         AmgCsrBuf &S, AmgWork &w, hipStream_t s);
// PMIS splitting: cf = 1 (C), -1 (F), -3 (no strong connection); returns the C count
int32_t amg_pmis_device(const AmgCsr &S, int8_t *cf, AmgWork &w, hipStream_t s);
// extended+i interpolation truncated to pmax entries (n x C count); cfunc gets the
// functions of the C points in coarse order
void amg_interp_device(const AmgCsr &A, const AmgCsr &S, const int8_t *cf, const int32_t *func, int nf, int pmax,
                       AmgCsrBuf &P, int32_t *cfunc, AmgWork &w, hipStream_t s);
// R = P^T (P has m columns); a row of R lists the rows of P in ascending order
void amg_transpose_device(const AmgCsr &P, int32_t m, AmgCsrBuf &R, AmgWork &w, hipStream_t s);
// Z = X Y (Y has m columns): each entry the sum from 0.0 of its products in the
// order of X's row and Y's rows
void amg_spgemm_device(const AmgCsr &X, const AmgCsr &Y, int32_t m, AmgCsrBuf &Z, AmgWork &w, hipStream_t s);
// x = dinv .* b (one l1-Jacobi sweep from zero)
void amg_level_smooth_pre(int32_t n, const double *dinv, const double *b, double *x, hipStream_t s);
// x += dinv .* t
void amg_add_diag_scaled(int64_t n, const double *dinv, const double *t, double *x, hipStream_t s);
// r = b - A x
void amg_residual(const AmgCsr &A, const double *x, const double *b, double *r, hipStream_t s);
// xout = x + dinv .* (b - A x) (xout must not alias x)
void amg_smooth_post(const AmgCsr &A, const double *dinv, const double *x, const double *b, double *xout,
                     hipStream_t s);
// y = M x; y += M x
void amg_spmv(const AmgCsr &M, const double *x, double *y, hipStream_t s);
void amg_spmv_add(const AmgCsr &M, const double *x, double *y, hipStream_t s);

// ---------------------------------------------------------------- true residual
// out[0] = ||b - A x||_2, out[1] = ||b||_2 (device, two doubles) in double-double
// arithmetic; the CSC arrays are read as operator rows.  partials holds
// 4 * residual_groups(n) doubles (at most 4 * kResidualGroups).
constexpr int kResidualGroups = 2048;
int residual_groups(int32_t n);
void residual_norm_dd(int32_t n, const int32_t *colptr, const int32_t *rowidx, const double *vals, const double *b,
                      const double *x, double *partials, double *out, hipStream_t s);
// Device address of the current GPU's static partials array (4 * kResidualGroups doubles),
// looked up once per device; nothing is allocated.
double *residual_scratch();

// ---------------------------------------------------------------- adjoint
// grad[e] = -lam[j] * x[rowidx[e]] for every entry e of CSC column j (operator entry (j, rowidx[e])).
void pattern_grad(int32_t n, const int32_t *colptr, const int32_t *rowidx, const double *lam, const double *x,
                  double *grad, hipStream_t s);
// out[s] = sum over k in [ptr[s], ptr[s+1]) of coef[k] * g[slot[k]].
void segment_dot(int64_t nseg, const int64_t *ptr, const int32_t *slot, const double *coef, const double *g,
                 double *out, hipStream_t s);

}  // namespace lss::k
