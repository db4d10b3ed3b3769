// Stiff DOF selection on the device: row l1 norms and the 1-D two-component
// Gaussian mixture EM of GPUHybrid::select_bad_dofs (GPUHybrid.cu:737-883).
//
// Every sum uses one fixed reduction tree (1024 contiguous chunks x 256 lanes,
// each lane summing its strided elements in index order, then pairwise trees
// over lanes and chunks), so the result is bit-stable run to run and equal to
// the CPU restatement's canonical sum.  The reference reduces with thrust in an
// unspecified order.  Expressions keep the reference's operation order and this
// file is compiled without FP contraction.
#include <vector>

#include "kernels.hpp"
#include "../common/runtime.hpp"

#pragma clang fp contract(off)

namespace lss::k {
namespace {

constexpr int C = kCsumChunks;
constexpr int L = kCsumLanes;

inline int64_t chunk_len(int64_t n)
{
    int64_t c = (n + C - 1) / C;
    return (c + L - 1) / L * L;
}

// Lane tree over L values in shared memory: stride L/2 .. 1, result in sh[0].
__device__ __forceinline__ double lane_tree(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int s = L / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// GPUHybrid.cu:753-765: each row summed sequentially in storage order.  A
// block of L rows stages the contiguous value range of its rows through LDS
// in coalesced tiles; every thread adds the values of its own row in order.
constexpr int kL1Tile = 2048;
__global__ __launch_bounds__(L) void k_row_l1(int32_t n, const int32_t *__restrict__ ptr,
                                              const double *__restrict__ vals, double *__restrict__ out)
{
    __shared__ double tile[kL1Tile];
    const int32_t r0 = int32_t(blockIdx.x) * L;
    const int32_t i = r0 + threadIdx.x;
    const int32_t lo = ptr[r0], hi = ptr[min(r0 + L, n)];
    const int32_t a = i < n ? ptr[i] : 0, e = i < n ? ptr[i + 1] : 0;
    double s = 0.0;
    for (int32_t c0 = lo; c0 < hi; c0 += kL1Tile) {  // uniform over the block
        for (int q = threadIdx.x; q < kL1Tile; q += L)
            if (c0 + q < hi) tile[q] = fabs(vals[c0 + q]);
        __syncthreads();
        const int32_t from = max(a, c0), to = min(e, c0 + kL1Tile);
        for (int32_t j = from; j < to; ++j) s += tile[j - c0];
        __syncthreads();
    }
    if (i < n) out[i] = s;
}

__global__ __launch_bounds__(L) void k_csum(int64_t n, int64_t clen, const double *__restrict__ v,
                                            double *__restrict__ partials)
{
    __shared__ double sh[L];
    const int64_t lo = int64_t(blockIdx.x) * clen, hi = min(lo + clen, n);
    double a = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += L) a += v[i];
    const double r = lane_tree(a, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

__global__ __launch_bounds__(L) void k_var_global(int64_t n, int64_t clen, const double *__restrict__ v, double mean,
                                                  double *__restrict__ partials)
{
    __shared__ double sh[L];
    const int64_t lo = int64_t(blockIdx.x) * clen, hi = min(lo + clen, n);
    double a = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += L) {
        const double d = v[i] - mean;
        a += d * d;
    }
    const double r = lane_tree(a, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

__global__ __launch_bounds__(L) void k_estep(int64_t n, int64_t clen, const double *__restrict__ x, GmmParams p,
                                             double *__restrict__ g0, double *__restrict__ g1,
                                             double *__restrict__ partials)
{
    // GPUHybrid.cu:797-826 (E-step) fused with the sums of 830-840
    __shared__ double sh[L];
    const int64_t lo = int64_t(blockIdx.x) * clen, hi = min(lo + clen, n);
    double s0 = 0.0, s1 = 0.0, sx0 = 0.0, sx1 = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += L) {
        const double xi = x[i];
        const double log_N0 = p.c0 - 0.5 * (xi - p.mean0) * (xi - p.mean0) / p.var0;
        const double log_N1 = p.c1 - 0.5 * (xi - p.mean1) * (xi - p.mean1) / p.var1;
        const double log_g0 = p.log_w0 + log_N0;
        const double log_g1 = p.log_w1 + log_N1;
        const double mx = fmax(log_g0, log_g1);
        const double log_total = mx + log(exp(log_g0 - mx) + exp(log_g1 - mx));
        const double a0 = exp(log_g0 - log_total);
        const double a1 = exp(log_g1 - log_total);
        g0[i] = a0;
        g1[i] = a1;
        s0 += a0;
        s1 += a1;
        sx0 += a0 * xi;
        sx1 += a1 * xi;
    }
    double r = lane_tree(s0, sh);
    if (threadIdx.x == 0) partials[0 * C + blockIdx.x] = r;
    r = lane_tree(s1, sh);
    if (threadIdx.x == 0) partials[1 * C + blockIdx.x] = r;
    r = lane_tree(sx0, sh);
    if (threadIdx.x == 0) partials[2 * C + blockIdx.x] = r;
    r = lane_tree(sx1, sh);
    if (threadIdx.x == 0) partials[3 * C + blockIdx.x] = r;
}

__global__ __launch_bounds__(L) void k_var(int64_t n, int64_t clen, const double *__restrict__ x,
                                           const double *__restrict__ g0, const double *__restrict__ g1, double m0,
                                           double m1, double *__restrict__ partials)
{
    // GPUHybrid.cu:836-848
    __shared__ double sh[L];
    const int64_t lo = int64_t(blockIdx.x) * clen, hi = min(lo + clen, n);
    double a0 = 0.0, a1 = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += L) {
        const double xi = x[i];
        a0 += g0[i] * (xi - m0) * (xi - m0);
        a1 += g1[i] * (xi - m1) * (xi - m1);
    }
    double r = lane_tree(a0, sh);
    if (threadIdx.x == 0) partials[0 * C + blockIdx.x] = r;
    r = lane_tree(a1, sh);
    if (threadIdx.x == 0) partials[1 * C + blockIdx.x] = r;
}

// Chunk tree: stride C/2 .. 1 over the C partials of sum blockIdx.x.
__global__ __launch_bounds__(C / 2) void k_csum_finish(const double *__restrict__ partials, double *__restrict__ out)
{
    __shared__ double sh[C / 2];
    const double *p = partials + int64_t(blockIdx.x) * C;
    sh[threadIdx.x] = p[threadIdx.x] + p[threadIdx.x + C / 2];
    __syncthreads();
    for (int s = C / 4; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_minmax_finish(const double *__restrict__ part, double *__restrict__ out)
{
    __shared__ double smin[256], smax[256];
    double lo = INFINITY, hi = -INFINITY;
    for (int b = threadIdx.x; b < kMinmaxBlocks; b += 256) {
        lo = fmin(lo, part[2 * b]);
        hi = fmax(hi, part[2 * b + 1]);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = smin[0];
        out[1] = smax[0];
    }
}

// min / max are exact in any order: per-block partials, then one block (k_minmax_finish)
__global__ __launch_bounds__(256) void k_minmax(int64_t n, const double *__restrict__ v, double *__restrict__ part)
{
    __shared__ double smin[256], smax[256];
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        lo = fmin(lo, v[i]);
        hi = fmax(hi, v[i]);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = smin[0];
        part[2 * blockIdx.x + 1] = smax[0];
    }
}

__global__ void k_flags(int64_t n, const double *__restrict__ g0, const double *__restrict__ g1,
                        uint8_t *__restrict__ flags, int high)
{
    // GPUHybrid.cu:862-868 (high = 1); BadDofGMMPreconditioner.cu:224-235 picks the cluster
    // with the larger |mean|
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) flags[i] = (high ? g0[i] < g1[i] : g1[i] < g0[i]) ? 1 : 0;
}

}  // namespace

void row_l1(int32_t n, const int32_t *ptr, const double *vals, double *out, hipStream_t s)
{
    if (n == 0) return;
    k_row_l1<<<div_up(n, L), L, 0, s>>>(n, ptr, vals, out);
    LSS_LAUNCH_CHECK();
}

void canonical_sum(int64_t n, const double *v, double *partials, double *out, hipStream_t s)
{
    k_csum<<<C, L, 0, s>>>(n, chunk_len(n), v, partials);
    LSS_LAUNCH_CHECK();
    k_csum_finish<<<1, C / 2, 0, s>>>(partials, out);
    LSS_LAUNCH_CHECK();
}

void gmm_var_global(int64_t n, const double *x, double mean, double *partials1, hipStream_t s)
{
    k_var_global<<<C, L, 0, s>>>(n, chunk_len(n), x, mean, partials1);
    LSS_LAUNCH_CHECK();
}

void minmax(int64_t n, const double *v, double *part, double *out2, hipStream_t s)
{
    k_minmax<<<kMinmaxBlocks, 256, 0, s>>>(n, v, part);
    k_minmax_finish<<<1, 256, 0, s>>>(part, out2);
    LSS_LAUNCH_CHECK();
}

void gmm_estep(int64_t n, const double *x, GmmParams p, double *g0, double *g1, double *partials4, hipStream_t s)
{
    k_estep<<<C, L, 0, s>>>(n, chunk_len(n), x, p, g0, g1, partials4);
    LSS_LAUNCH_CHECK();
}

void gmm_var(int64_t n, const double *x, const double *g0, const double *g1, double m0, double m1,
             double *partials2, hipStream_t s)
{
    k_var<<<C, L, 0, s>>>(n, chunk_len(n), x, g0, g1, m0, m1, partials2);
    LSS_LAUNCH_CHECK();
}

void csum_finish(int nsums, const double *partials, double *out, hipStream_t s)
{
    k_csum_finish<<<nsums, C / 2, 0, s>>>(partials, out);
    LSS_LAUNCH_CHECK();
}

void gmm_flags(int64_t n, const double *g0, const double *g1, uint8_t *flags, hipStream_t s, int high)
{
    if (n == 0) return;
    k_flags<<<div_up(n, 256), 256, 0, s>>>(n, g0, g1, flags, high);
    LSS_LAUNCH_CHECK();
}

namespace {

__global__ void k_flags_i32(int32_t n, const uint8_t *__restrict__ f, int32_t *__restrict__ out)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f[i] ? 1 : 0;
}

__global__ void k_scatter_flagged(int32_t n, const uint8_t *__restrict__ f, const int32_t *__restrict__ rank,
                                  int32_t *__restrict__ out)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && f[i]) out[rank[i]] = int32_t(i);
}

__global__ void k_row_lengths(int64_t m, const int32_t *__restrict__ rows, const int32_t *__restrict__ ptr,
                              int32_t *__restrict__ len)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < m) len[k] = ptr[rows[k] + 1] - ptr[rows[k]];
}

}  // namespace

std::vector<int32_t> flagged_indices(int32_t n, const uint8_t *flags, hipStream_t s)
{
    std::vector<int32_t> h;
    if (n <= 0) return h;
    DBuf<int32_t> f(n), rank(n), tsum(std::max(div_up(n, 1024), 1)), total(1);
    k_flags_i32<<<div_up(n, 256), 256, 0, s>>>(n, flags, f.get());
    excl_scan_i32(n, f.get(), rank.get(), tsum.get(), total.get(), s);
    int32_t cnt = 0;
    LSS_HIP(hipMemcpyAsync(&cnt, total.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
    LSS_HIP(hipStreamSynchronize(s));
    if (cnt == 0) return h;
    DBuf<int32_t> out(cnt);
    k_scatter_flagged<<<div_up(n, 256), 256, 0, s>>>(n, flags, rank.get(), out.get());
    LSS_LAUNCH_CHECK();
    h.resize(size_t(cnt));
    LSS_HIP(hipMemcpyAsync(h.data(), out.get(), sizeof(int32_t) * size_t(cnt), hipMemcpyDeviceToHost, s));
    LSS_HIP(hipStreamSynchronize(s));
    return h;
}

void row_lengths(int64_t m, const int32_t *rows, const int32_t *ptr, int32_t *len, hipStream_t s)
{
    if (m == 0) return;
    k_row_lengths<<<div_up(m, 256), 256, 0, s>>>(m, rows, ptr, len);
    LSS_LAUNCH_CHECK();
}

}  // namespace lss::k
