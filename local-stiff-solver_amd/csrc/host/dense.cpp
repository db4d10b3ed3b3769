// Blocked dense Cholesky and inverse factor of one large stiff component.
//
// The reference factorizes each stiff component with cuDSS (sparse LDL^T,
// GPUHybrid.cu:1231-1249) and applies it by triangular solves.  Components up
// to a few hundred DOFs are factored and inverted by one workgroup each
// (stiff.hip); larger ones (contact clusters merged by the one-ring expansion
// reach 16 000 DOFs on a 160^3 grid) take this blocked right-looking path:
//   for each diagonal block k (128 rows):
//     L_kk = chol(A_kk), G_kk = L_kk^{-1}            own kernel, in LDS
//     L_ik = A_ik G_kk^T                                GEMM
//     A_22 -= L_i L_i^T  (lower triangle)               SYRK
// then G = L^{-1} by recursive halving, G_21 = -G_22 (L_21 G_11), two GEMMs per
// split, the diagonal blocks being the G_kk above.  The GEMM/SYRK updates are
// plain dense BLAS-3 calls to rocBLAS; everything runs on the caller's stream.
// Matrices are row-major; a row-major X is the column-major X^T, so a
// row-major product C = A op(B) is the column-major C^T = op(B)^T A^T.
#include "dense.hpp"

#include <rocblas/rocblas.h>

#include <algorithm>
#include <string>

#include "../common/runtime.hpp"
#include "../kernels/kernels.hpp"
#include "../kernels/mas_kernels.hpp"

namespace lss {
namespace {

void rb_check(rocblas_status st, const char *what)
{
    if (st != rocblas_status_success)
        fail(LSS_ERR_HIP, std::string("rocBLAS ") + what + ": " + rocblas_status_to_string(st));
}

// row-major C (r x c, ldc) = alpha A (r x k, lda) op(B) + beta C; op(B) = B (k x c, ldb) or B^T (B: c x k)
void gemm_rm(rocblas_handle h, bool trans_b, int64_t r, int64_t c, int64_t k, double alpha, const double *A,
             int64_t lda, const double *B, int64_t ldb, double beta, double *C, int64_t ldc)
{
    rb_check(rocblas_dgemm(h, trans_b ? rocblas_operation_transpose : rocblas_operation_none,
                           rocblas_operation_none, rocblas_int(c), rocblas_int(r), rocblas_int(k), &alpha, B,
                           rocblas_int(ldb), A, rocblas_int(lda), &beta, C, rocblas_int(ldc)),
             "dgemm");
}

}  // namespace

DenseBlas::~DenseBlas()
{
    if (handle_) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(handle_));
    if (handle2_) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(handle2_));
    if (ev_col_) (void)hipEventDestroy(ev_col_);
    if (ev_pan_) (void)hipEventDestroy(ev_pan_);
    if (s2_) (void)hipStreamDestroy(s2_);
}

void DenseBlas::init()
{
    if (handle_) return;
    rocblas_handle h = nullptr, h2 = nullptr;
    rb_check(rocblas_create_handle(&h), "create_handle");
    handle_ = h;
    rb_check(rocblas_create_handle(&h2), "create_handle");
    handle2_ = h2;
    LSS_HIP(hipStreamCreateWithFlags(&s2_, hipStreamNonBlocking));
    LSS_HIP(hipEventCreateWithFlags(&ev_col_, hipEventDisableTiming));
    LSS_HIP(hipEventCreateWithFlags(&ev_pan_, hipEventDisableTiming));
}

int64_t DenseBlas::work_size(int64_t m)
{
    const int64_t nb = k::dense_block_size();
    const int64_t half = m / 2 + nb;
    return std::max(2 * m * nb, half * half);  // two panels in flight; the inverse's products
}

void DenseBlas::factor_inverse(int64_t m, double *A, double *G, double *work, int *status, hipStream_t s)
{
    partial_factor_inverse(m, m, A, G, work, status, s);
    k::dense_mirror(m, G, s);
}

void DenseBlas::partial_factor_inverse(int64_t m, int64_t p, double *A, double *G, double *work, int *status,
                                       hipStream_t s)
{
    init();
    rocblas_handle h = static_cast<rocblas_handle>(handle_), h2 = static_cast<rocblas_handle>(handle2_);
    rb_check(rocblas_set_stream(h, s), "set_stream");
    rb_check(rocblas_set_stream(h2, s2_), "set_stream");
    const int64_t nb = k::dense_block_size();
    // the GEMMs read whole blocks of G (the diagonal blocks' and G11's upper parts included):
    // those entries must be zero, whatever the buffer held before
    k::fill(p * m, G, 0.0, s);
    // panels alternate between two workspaces: one is read by the trailing update while the
    // next is written
    double *W[2] = {work, work + m * nb};
    // diagonal block k0 (L_kk, G_kk) and the panel L_ik = A_ik G_kk^T into Wb (r x kb), then into A
    auto panel = [&](rocblas_handle hh, hipStream_t ss, int64_t k0, double *Wb) {
        const int64_t kb = std::min(nb, p - k0), r = m - k0 - kb;
        k::dense_diag_block(m, int(k0), int(kb), A, G, status, ss);
        if (r == 0) return;
        double *Aik = A + (k0 + kb) * m + k0;
        gemm_rm(hh, true, r, kb, kb, 1.0, Aik, m, G + k0 * m + k0, m, 0.0, Wb, kb);
        LSS_HIP(hipMemcpy2DAsync(Aik, size_t(m) * sizeof(double), Wb, size_t(kb) * sizeof(double),
                                 size_t(kb) * sizeof(double), size_t(r), hipMemcpyDeviceToDevice, ss));
    };
    // trailing lower triangle from row/column c0: A(c0:, c0:) -= L L^T, L = rows of Wb from c0 (n x kb).
    // Column-major view: the upper triangle of A^T, and L L^T = V^T V with V = the rows (kb x n).
    auto syrk = [&](int64_t c0, int64_t n, int64_t kb, const double *L) {
        const double alpha = -1.0, beta = 1.0;
        rb_check(rocblas_dsyrk(h, rocblas_fill_upper, rocblas_operation_transpose, rocblas_int(n), rocblas_int(kb),
                               &alpha, L, rocblas_int(kb), &beta, A + c0 * m + c0, rocblas_int(m)),
                 "dsyrk");
    };
    panel(h, s, 0, W[0]);
    int b = 0;
    for (int64_t k0 = 0; k0 < p; k0 += nb, b ^= 1) {
        const int64_t kb = std::min(nb, p - k0), r = m - k0 - kb, k1 = k0 + kb;
        if (r == 0) break;
        const double *Wk = W[b];
        if (k1 >= p) {  // last pivot block: the update rows' Schur complement
            syrk(k1, r, kb, Wk);
            break;
        }
        // look-ahead: block column k+1 first (rows k1 .. m-1), then its diagonal block and
        // panel on the second stream while the rest of this step's update runs here
        const int64_t kb1 = std::min(nb, p - k1), k2 = k1 + kb1;
        gemm_rm(h, true, r, kb1, kb, -1.0, Wk, kb, Wk, kb, 1.0, A + k1 * m + k1, m);
        LSS_HIP(hipEventRecord(ev_col_, s));
        LSS_HIP(hipStreamWaitEvent(s2_, ev_col_, 0));
        panel(h2, s2_, k1, W[b ^ 1]);
        LSS_HIP(hipEventRecord(ev_pan_, s2_));
        if (m > k2) syrk(k2, m - k2, kb, Wk + kb1 * kb);
        LSS_HIP(hipStreamWaitEvent(s, ev_pan_, 0));
    }
    inverse(0, p, m, A, G, work, s);
}

// G(a:b, a:b) = L(a:b, a:b)^{-1}; a is a multiple of the block size, b = a + a
// multiple of it or m.  Ranges of one block are the diagonal inverses already in G.
void DenseBlas::inverse(int64_t a, int64_t b, int64_t m, const double *L, double *G, double *work, hipStream_t s)
{
    const int64_t nb = k::dense_block_size();
    const int64_t n = b - a;
    if (n <= nb) return;
    const int64_t blocks = (n + nb - 1) / nb;
    const int64_t sp = a + (blocks / 2) * nb;
    inverse(a, sp, m, L, G, work, s);
    inverse(sp, b, m, L, G, work, s);
    rocblas_handle h = static_cast<rocblas_handle>(handle_);
    const int64_t r2 = b - sp, c1 = sp - a;
    // work = L_21 G_11 (r2 x c1); G_21 = -G_22 work
    gemm_rm(h, false, r2, c1, c1, 1.0, L + sp * m + a, m, G + a * m + a, m, 0.0, work, c1);
    gemm_rm(h, false, r2, c1, r2, -1.0, G + sp * m + sp, m, work, c1, 0.0, G + sp * m + a, m);
}

}  // namespace lss
