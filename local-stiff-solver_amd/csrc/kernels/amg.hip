// V-cycle kernels of the AMG base preconditioner: scalar CSR products with a
// group of lanes per row sized to the matrix (interpolation rows hold at most 4
// entries, coarse operators of 3D elasticity 80-350), fused with the l1-Jacobi
// smoothing steps.  Every output row is written by one lane group in a
// fixed order, so an application is bit-reproducible.
#include "kernels.hpp"
#include "../common/runtime.hpp"

namespace lss::k {
namespace {

constexpr int kAmgThreads = 256;

// mode 0: y = A x; 1: y = b - A x; 2: y = xo + dinv (b - A x); 3: y += A x.
// L lanes per row (2..64, chosen from the matrix's mean row length), each lane
// taking every L-th entry of the row, two entries per step with both loads issued
// first; matrix entries are streamed with non-temporal loads so that x stays in
// the caches.  The lane sums are combined by a fixed butterfly.
template <int MODE, int L>
__global__ __launch_bounds__(kAmgThreads) void k_csr(int32_t n, const int64_t *__restrict__ ptr,
                                                     const int32_t *__restrict__ col, const double *__restrict__ val,
                                                     const double *__restrict__ x, const double *__restrict__ b,
                                                     const double *__restrict__ xo, const double *__restrict__ dinv,
                                                     double *__restrict__ y)
{
    constexpr int R = kAmgThreads / L;
    const int sub = threadIdx.x % L;
    const int64_t i = int64_t(blockIdx.x) * R + threadIdx.x / L;
    double a0 = 0.0, a1 = 0.0;
    if (i < n) {
        const int64_t end = ptr[i + 1];
        int64_t e = ptr[i] + sub;
        for (; e + L < end; e += 2 * L) {
            const int32_t c0 = __builtin_nontemporal_load(col + e), c1 = __builtin_nontemporal_load(col + e + L);
            const double v0 = __builtin_nontemporal_load(val + e), v1 = __builtin_nontemporal_load(val + e + L);
            a0 += v0 * x[c0];
            a1 += v1 * x[c1];
        }
        if (e < end) a0 += __builtin_nontemporal_load(val + e) * x[__builtin_nontemporal_load(col + e)];
    }
    double acc = a0 + a1;
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, L);
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

template <int MODE, int L>
void launch_csr_l(const AmgCsr &M, const double *x, const double *b, const double *xo, const double *dinv, double *y,
                  hipStream_t s)
{
    k_csr<MODE, L><<<div_up(M.n, kAmgThreads / L), kAmgThreads, 0, s>>>(M.n, M.ptr, M.col, M.val, x, b, xo, dinv, y);
    LSS_LAUNCH_CHECK();
}

// lanes per row: the smallest power of two >= a third of the mean row length, 2..64
int csr_lanes(const AmgCsr &M)
{
    const double mean = M.n > 0 ? double(M.nnz) / M.n : 0.0;
    int l = 2;
    while (l < 64 && l * 3 < mean) l *= 2;
    return l;
}

template <int MODE>
void launch_csr(const AmgCsr &M, const double *x, const double *b, const double *xo, const double *dinv, double *y,
                hipStream_t s)
{
    if (M.n <= 0) return;
    switch (csr_lanes(M)) {
        case 2: launch_csr_l<MODE, 2>(M, x, b, xo, dinv, y, s); break;
        case 4: launch_csr_l<MODE, 4>(M, x, b, xo, dinv, y, s); break;
        case 8: launch_csr_l<MODE, 8>(M, x, b, xo, dinv, y, s); break;
        case 16: launch_csr_l<MODE, 16>(M, x, b, xo, dinv, y, s); break;
        case 32: launch_csr_l<MODE, 32>(M, x, b, xo, dinv, y, s); break;
        default: launch_csr_l<MODE, 64>(M, x, b, xo, dinv, y, s); break;
    }
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
    launch_csr<1>(A, x, b, nullptr, nullptr, r, s);
}

void amg_smooth_post(const AmgCsr &A, const double *dinv, const double *x, const double *b, double *xout,
                     hipStream_t s)
{
    launch_csr<2>(A, x, b, x, dinv, xout, s);
}

void amg_spmv(const AmgCsr &M, const double *x, double *y, hipStream_t s)
{
    launch_csr<0>(M, x, nullptr, nullptr, nullptr, y, s);
}

void amg_spmv_add(const AmgCsr &M, const double *x, double *y, hipStream_t s)
{
    launch_csr<3>(M, x, nullptr, nullptr, nullptr, y, s);
}

}  // namespace lss::k
