// V-cycle kernels of the AMG base preconditioner: scalar CSR products with a
// group of lanes per row (the level operators, interpolation and restriction of
// a classical AMG hierarchy have tens of entries per row), fused with the
// l1-Jacobi smoothing steps.  Every output row is written by one lane group in a
// fixed order, so an application is bit-reproducible.
#include "kernels.hpp"
#include "../common/runtime.hpp"

namespace lss::k {
namespace {

constexpr int kAmgThreads = 256;
constexpr int kLanesPerRow = 8;
constexpr int kRowsPerGroup = kAmgThreads / kLanesPerRow;

// mode 0: y = A x; 1: y = b - A x; 2: y = xo + dinv (b - A x); 3: y += A x
template <int MODE>
__global__ __launch_bounds__(kAmgThreads) void k_csr(int32_t n, const int64_t *__restrict__ ptr,
                                                     const int32_t *__restrict__ col, const double *__restrict__ val,
                                                     const double *__restrict__ x, const double *__restrict__ b,
                                                     const double *__restrict__ xo, const double *__restrict__ dinv,
                                                     double *__restrict__ y)
{
    const int sub = threadIdx.x % kLanesPerRow;
    const int64_t i = int64_t(blockIdx.x) * kRowsPerGroup + threadIdx.x / kLanesPerRow;
    double acc = 0.0;
    if (i < n)
        for (int64_t e = ptr[i] + sub; e < ptr[i + 1]; e += kLanesPerRow) acc += val[e] * x[col[e]];
#pragma unroll
    for (int off = kLanesPerRow / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kLanesPerRow);
    if (i >= n || sub != 0) return;
    if constexpr (MODE == 0) y[i] = acc;
    else if constexpr (MODE == 1) y[i] = b[i] - acc;
    else if constexpr (MODE == 2) y[i] = xo[i] + dinv[i] * (b[i] - acc);
    else y[i] += acc;
}

__global__ __launch_bounds__(256) void k_jacobi_zero(int32_t n, const double *__restrict__ dinv,
                                                     const double *__restrict__ b, double *__restrict__ x)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) x[i] = dinv[i] * b[i];
}

__global__ __launch_bounds__(256) void k_add_diag_scaled(int64_t n, const double *__restrict__ dinv,
                                                         const double *__restrict__ t, double *__restrict__ x)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) x[i] += dinv[i] * t[i];
}

template <int MODE>
void launch_csr(int32_t n, const int64_t *ptr, const int32_t *col, const double *val, const double *x, const double *b,
                const double *xo, const double *dinv, double *y, hipStream_t s)
{
    if (n <= 0) return;
    k_csr<MODE><<<div_up(n, kRowsPerGroup), kAmgThreads, 0, s>>>(n, ptr, col, val, x, b, xo, dinv, y);
    LSS_LAUNCH_CHECK();
}

}  // namespace

void amg_level_smooth_pre(int32_t n, const double *dinv, const double *b, double *x, hipStream_t s)
{
    if (n <= 0) return;
    k_jacobi_zero<<<div_up(n, 256), 256, 0, s>>>(n, dinv, b, x);
    LSS_LAUNCH_CHECK();
}

void amg_add_diag_scaled(int64_t n, const double *dinv, const double *t, double *x, hipStream_t s)
{
    if (n <= 0) return;
    k_add_diag_scaled<<<div_up(n, 256), 256, 0, s>>>(n, dinv, t, x);
    LSS_LAUNCH_CHECK();
}

void amg_residual(const AmgCsr &A, const double *x, const double *b, double *r, hipStream_t s)
{
    launch_csr<1>(A.n, A.ptr, A.col, A.val, x, b, nullptr, nullptr, r, s);
}

void amg_smooth_post(const AmgCsr &A, const double *dinv, const double *x, const double *b, double *xout,
                     hipStream_t s)
{
    launch_csr<2>(A.n, A.ptr, A.col, A.val, x, b, x, dinv, xout, s);
}

void amg_spmv(const AmgCsr &M, const double *x, double *y, hipStream_t s)
{
    launch_csr<0>(M.n, M.ptr, M.col, M.val, x, nullptr, nullptr, nullptr, y, s);
}

void amg_spmv_add(const AmgCsr &M, const double *x, double *y, hipStream_t s)
{
    launch_csr<3>(M.n, M.ptr, M.col, M.val, x, nullptr, nullptr, nullptr, y, s);
}

}  // namespace lss::k
