// True residual norm ||b - A x||_2 (and ||b||_2) in double-double arithmetic.
//
// The reference reports the FP64 residual ||A x - b|| of the solution
// (GPUHybrid.cu:383-386).  For stiff systems that FP64 evaluation has an error
// floor of about eps * || |A| |x| ||, far above a 1e-8 relative residual when
// entries span 1e8.  Here every product a_ij x_j is split exactly into two
// doubles (fma), the row sum b_i - sum_j a_ij x_j is carried as an unevaluated
// double-double (about 106 significant bits), and the squares are summed in
// double-double too, so the computed norm is accurate to a few units of the
// last place of the residual itself.
//
// Layout: CSC arrays read as operator rows (column j is row j, GPUHybrid.cu:401-433).
// A grid of at most kResidualGroups workgroups walks the rows; each workgroup
// stores four doubles (r.r and b.b as hi/lo pairs) into a partials buffer that a
// one-workgroup finisher sums in index order, so the result is deterministic.
#include <algorithm>
#include <mutex>

#include "kernels.hpp"
#include "../common/runtime.hpp"

// No contraction of a product into a neighbouring sum anywhere below: TwoSum must see
// the rounded product p = a * b, or its error term is wrong (the only fused
// multiply-adds are the explicit ones of two_prod).
#pragma clang fp contract(off)

namespace lss::k {
namespace {

constexpr int kResThreads = 256;
constexpr int kLanesPerRow = 16;
constexpr int kRowsPerGroup = kResThreads / kLanesPerRow;

struct DD {
    double hi, lo;
};

// Knuth's TwoSum: s + e == a + b exactly.
__device__ __forceinline__ DD two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}

// |a| >= |b| (or a == 0): s + e == a + b exactly.
__device__ __forceinline__ DD quick_two_sum(double a, double b)
{
    const double s = a + b;
    return {s, b - (s - a)};
}

// Exact product with the fused multiply-add: p + e == a * b.
__device__ __forceinline__ DD two_prod(double a, double b)
{
    const double p = a * b;
    return {p, fma(a, b, -p)};
}

__device__ __forceinline__ DD dd_add(DD a, DD b)
{
    DD s = two_sum(a.hi, b.hi);
    const DD t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}

__device__ __forceinline__ DD dd_sqr(DD a)
{
    DD p = two_prod(a.hi, a.hi);
    p.lo += 2.0 * a.hi * a.lo;
    return quick_two_sum(p.hi, p.lo);
}

__device__ __forceinline__ DD shfl_xor_dd(DD v, int off)
{
    return {__shfl_xor(v.hi, off, dev::kWave), __shfl_xor(v.lo, off, dev::kWave)};
}

// partials[4 * g + {0,1,2,3}] = (r.r hi, r.r lo, b.b hi, b.b lo) of workgroup g.
__global__ __launch_bounds__(kResThreads) void k_residual_dd(int32_t n, const int32_t *__restrict__ colptr,
                                                            const int32_t *__restrict__ rowidx,
                                                            const double *__restrict__ vals,
                                                            const double *__restrict__ b,
                                                            const double *__restrict__ x, double *__restrict__ partials)
{
    __shared__ double sh[4][kResThreads / dev::kWave];
    const int sub = threadIdx.x % kLanesPerRow;
    DD rr{0.0, 0.0}, bb{0.0, 0.0};
    for (int64_t j = int64_t(blockIdx.x) * kRowsPerGroup + threadIdx.x / kLanesPerRow; j < n;
         j += int64_t(gridDim.x) * kRowsPerGroup) {
        DD acc{0.0, 0.0};
        const int32_t lo = colptr[j], hi = colptr[j + 1];
        for (int32_t e = lo + sub; e < hi; e += kLanesPerRow) acc = dd_add(acc, two_prod(vals[e], x[rowidx[e]]));
#pragma unroll
        for (int off = kLanesPerRow / 2; off > 0; off >>= 1) acc = dd_add(acc, shfl_xor_dd(acc, off));
        if (sub == 0) {
            const double bj = b[j];
            const DD r = dd_add(DD{bj, 0.0}, DD{-acc.hi, -acc.lo});
            rr = dd_add(rr, dd_sqr(r));
            bb = dd_add(bb, two_prod(bj, bj));
        }
    }
    // lanes that led no row hold zeros; a fixed butterfly over the wave, then the waves in order
#pragma unroll
    for (int off = dev::kWave / 2; off > 0; off >>= 1) {
        rr = dd_add(rr, shfl_xor_dd(rr, off));
        bb = dd_add(bb, shfl_xor_dd(bb, off));
    }
    const int lane = threadIdx.x & (dev::kWave - 1), wid = threadIdx.x / dev::kWave;
    if (lane == 0) {
        sh[0][wid] = rr.hi, sh[1][wid] = rr.lo, sh[2][wid] = bb.hi, sh[3][wid] = bb.lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        DD tr{sh[0][0], sh[1][0]}, tb{sh[2][0], sh[3][0]};
        for (int w = 1; w < kResThreads / dev::kWave; ++w) {
            tr = dd_add(tr, DD{sh[0][w], sh[1][w]});
            tb = dd_add(tb, DD{sh[2][w], sh[3][w]});
        }
        double *p = partials + 4 * int64_t(blockIdx.x);
        p[0] = tr.hi, p[1] = tr.lo, p[2] = tb.hi, p[3] = tb.lo;
    }
}

__device__ __forceinline__ double dd_sqrt(DD a)
{
    if (!(a.hi > 0.0)) return 0.0;
    const double s = sqrt(a.hi);
    // one Newton step on the double-double value: s + (a - s^2) / (2 s)
    const DD s2 = two_prod(s, s);
    const double d = ((a.hi - s2.hi) - s2.lo) + a.lo;
    return s + d / (2.0 * s);
}

// out[0] = ||b - A x||, out[1] = ||b||; ngroups partials in order.
__global__ __launch_bounds__(kResThreads) void k_residual_finish(int ngroups, const double *__restrict__ partials,
                                                                double *__restrict__ out)
{
    __shared__ double sh[4][kResThreads];
    DD rr{0.0, 0.0}, bb{0.0, 0.0};
    for (int g = threadIdx.x; g < ngroups; g += kResThreads) {
        rr = dd_add(rr, DD{partials[4 * g], partials[4 * g + 1]});
        bb = dd_add(bb, DD{partials[4 * g + 2], partials[4 * g + 3]});
    }
    sh[0][threadIdx.x] = rr.hi, sh[1][threadIdx.x] = rr.lo, sh[2][threadIdx.x] = bb.hi, sh[3][threadIdx.x] = bb.lo;
    __syncthreads();
    for (int s = kResThreads / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) {
            const int t = threadIdx.x, u = t + s;
            const DD a = dd_add(DD{sh[0][t], sh[1][t]}, DD{sh[0][u], sh[1][u]});
            const DD c = dd_add(DD{sh[2][t], sh[3][t]}, DD{sh[2][u], sh[3][u]});
            sh[0][t] = a.hi, sh[1][t] = a.lo, sh[2][t] = c.hi, sh[3][t] = c.lo;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = dd_sqrt(DD{sh[0][0], sh[1][0]});
        out[1] = dd_sqrt(DD{sh[2][0], sh[3][0]});
    }
}

// Partial sums of the allocation-free entry point: 4 doubles per workgroup of the
// capped grid (64 KiB).  Each GPU the library runs on holds its own copy until the
// process ends; host code reaches it only through the address hipGetSymbolAddress
// returns for the current device (the array's name is a host-side stand-in).
static __device__ __align__(256) double scratch[4 * kResidualGroups];

}  // namespace

double *residual_scratch()
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static double *addr[kMaxDevices] = {};
    int dev = 0;
    LSS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) fail(LSS_ERR_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        LSS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(LSS_ERR_HIP, "no device address for the residual scratch array");
        addr[dev] = static_cast<double *>(p);
    }
    return addr[dev];
}

int residual_groups(int32_t n)
{
    return int(std::max<int64_t>(1, std::min<int64_t>(kResidualGroups, div_up(int64_t(n), kRowsPerGroup))));
}

void residual_norm_dd(int32_t n, const int32_t *colptr, const int32_t *rowidx, const double *vals, const double *b,
                      const double *x, double *partials, double *out, hipStream_t s)
{
    const int g = residual_groups(n);
    k_residual_dd<<<g, kResThreads, 0, s>>>(n, colptr, rowidx, vals, b, x, partials);
    LSS_LAUNCH_CHECK();
    k_residual_finish<<<1, kResThreads, 0, s>>>(g, partials, out);
    LSS_LAUNCH_CHECK();
}

}  // namespace lss::k
