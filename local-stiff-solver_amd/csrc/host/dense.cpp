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
}

int64_t DenseBlas::work_size(int64_t m)
{
    const int64_t nb = k::dense_block_size();
    const int64_t half = m / 2 + nb;
    return std::max(m * nb, half * half);
}

void DenseBlas::factor_inverse(int64_t m, double *A, double *G, double *work, int *status, hipStream_t s)
{
    partial_factor_inverse(m, m, A, G, work, status, s);
    k::dense_mirror(m, G, s);
}

void DenseBlas::partial_factor_inverse(int64_t m, int64_t p, double *A, double *G, double *work, int *status,
                                       hipStream_t s)
{
    if (!handle_) {
        rocblas_handle h = nullptr;
        rb_check(rocblas_create_handle(&h), "create_handle");
        handle_ = h;
    }
    rocblas_handle h = static_cast<rocblas_handle>(handle_);
    rb_check(rocblas_set_stream(h, s), "set_stream");
    const int64_t nb = k::dense_block_size();
    // the GEMMs read whole blocks of G (the diagonal blocks' and G11's upper parts included):
    // those entries must be zero, whatever the buffer held before
    k::fill(p * m, G, 0.0, s);
    for (int64_t k0 = 0; k0 < p; k0 += nb) {
        const int64_t kb = std::min(nb, p - k0), r = m - k0 - kb;
        k::dense_diag_block(m, int(k0), int(kb), A, G, status, s);
        if (r == 0) break;
        // panel L_ik = A_ik G_kk^T into work (r x kb), then back into A
        double *Aik = A + (k0 + kb) * m + k0;
        gemm_rm(h, true, r, kb, kb, 1.0, Aik, m, G + k0 * m + k0, m, 0.0, work, kb);
        LSS_HIP(hipMemcpy2DAsync(Aik, size_t(m) * sizeof(double), work, size_t(kb) * sizeof(double),
                                 size_t(kb) * sizeof(double), size_t(r), hipMemcpyDeviceToDevice, s));
        // trailing lower triangle: A_22 -= L_i L_i^T.  Column-major view: the
        // upper triangle of A_22^T, and L_i L_i^T = W^T W with W = work (kb x r).
        const double alpha = -1.0, beta = 1.0;
        rb_check(rocblas_dsyrk(h, rocblas_fill_upper, rocblas_operation_transpose, rocblas_int(r), rocblas_int(kb),
                               &alpha, work, rocblas_int(kb), &beta, A + (k0 + kb) * m + k0 + kb, rocblas_int(m)),
                 "dsyrk");
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
