// MAS coarse spaces on the device (MAS/MASPreconditioner.cu:128-481): the
// nodes of level l+1 are the connected components of each bank of level l
// (one wave per bank, neighbour bit masks closed by Warshall's rule across the
// lanes), numbered bank by bank and inside a bank by smallest member; the
// coarse operator R A R^T gets its pattern and, per coarse entry, the list of
// fine entries it sums in ascending order (one workgroup per bank sorts the
// bank's (coarse row, coarse column, fine entry) keys in LDS).  Same result
// as the host build_mas_levels.
#include <hip/hip_runtime.h>

#include <climits>

#include "../common/runtime.hpp"
#include "kernels.hpp"

namespace lss::k {
namespace {

constexpr int kLT = 256;
constexpr int kWavesLT = kLT / dev::kWave;
constexpr int kSortCap = 16384;  // stored entries of one bank sorted in LDS (128 KiB of keys)
constexpr int kKeyBits = 26;     // entry offset inside its bank

__device__ __forceinline__ unsigned long long lanemask_lt(int lane)
{
    return (1ull << lane) - 1ull;
}

// components of each bank: local rank -> parent[v] (made global by k_children), count per bank
__global__ __launch_bounds__(kLT) void k_bank_cc(int32_t nbanks, const int32_t *__restrict__ bank_off,
                                                  const int32_t *__restrict__ rowptr, const int32_t *__restrict__ cols,
                                                  int32_t *__restrict__ rank, int32_t *__restrict__ count,
                                                  int32_t *__restrict__ flags)
{
    const int lane = threadIdx.x & (dev::kWave - 1);
    const int32_t b = int32_t(blockIdx.x) * kWavesLT + threadIdx.x / dev::kWave;
    if (b >= nbanks) return;
    const int32_t lo = bank_off[b], hi = bank_off[b + 1], m = hi - lo;
    if (m > 32) {
        if (lane == 0) flags[0] = 1;
        return;
    }
    unsigned nbr = 0;
    if (lane < m) {
        const int32_t v = lo + lane;
        nbr = 1u << lane;
        for (int32_t e = rowptr[v]; e < rowptr[v + 1]; ++e) {
            const int32_t c = cols[e];
            if (c >= lo && c < hi) nbr |= 1u << (c - lo);
        }
    }
    unsigned conn = nbr;
    for (int k = 0; k < m; ++k) {  // Warshall: closure over intermediate node k
        const unsigned ck = __shfl(conn, k, dev::kWave);
        if ((conn >> k) & 1u) conn |= ck;
    }
    const int lead = lane < m ? __builtin_ctz(conn) : 63;
    const unsigned long long leads = __ballot(lane < m && lead == lane);
    if (lane < m) rank[lo + lane] = __popcll(leads & lanemask_lt(lead));
    if (lane == 0) count[b] = __popcll(leads);
}

// parent = bank base + rank; children of each coarse node ascending, coarse
// nodes bank by bank and by rank
__global__ __launch_bounds__(kLT) void k_children(int32_t nbanks, const int32_t *__restrict__ bank_off,
                                                   const int32_t *__restrict__ base, int32_t *__restrict__ parent,
                                                   int64_t *__restrict__ child_ptr, int32_t *__restrict__ child)
{
    const int lane = threadIdx.x & (dev::kWave - 1);
    const int32_t b = int32_t(blockIdx.x) * kWavesLT + threadIdx.x / dev::kWave;
    if (b >= nbanks) return;
    const int32_t lo = bank_off[b], hi = bank_off[b + 1], m = hi - lo;
    const bool on = lane < m;
    const int32_t r = on ? parent[lo + lane] : 0;  // local rank from k_bank_cc
    // nodes of lower rank come first; equal rank in ascending lane order
    int32_t before = 0, same_before = 0;
    for (int q = 0; q < m; ++q) {
        const int32_t rq = __shfl(r, q, dev::kWave);
        before += rq < r;
        same_before += (q < lane) && rq == r;
    }
    if (on) {
        parent[lo + lane] = base[b] + r;
        child[lo + before + same_before] = lo + lane;
        if (same_before == 0) child_ptr[base[b] + r] = lo + before;
    }
}

__global__ void k_bank_max_entries(int32_t nbanks, const int32_t *__restrict__ bank_off,
                                   const int32_t *__restrict__ rowptr, int32_t *__restrict__ out)
{
    const int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nbanks) atomicMax(out, rowptr[bank_off[b + 1]] - rowptr[bank_off[b]]);
}

__global__ void k_set_last_i64(int64_t *__restrict__ a, int32_t at, int64_t v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) a[at] = v;
}

__global__ void k_uniform_banks(int32_t n, int32_t nbanks, int32_t *__restrict__ bank_off)
{
    const int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b <= nbanks) bank_off[b] = min(b * 32, n);
}

// LDS bitonic sort of p (power of two) keys
__device__ void bitonic_sort(unsigned long long *key, int p)
{
    for (int k = 2; k <= p; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < p; i += kLT) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const unsigned long long a = key[i], c = key[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > c) == up) {
                        key[i] = c;
                        key[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// keys (coarse row inside the bank, coarse column, entry offset) of one bank, sorted
__device__ int load_sorted_keys(unsigned long long *key, int32_t lo, int32_t hi, int32_t base,
                                const int32_t *__restrict__ rowptr, const int32_t *__restrict__ cols,
                                const int32_t *__restrict__ parent)
{
    const int32_t sl = rowptr[lo], ns = rowptr[hi] - sl;
    int p = 1;
    while (p < ns) p <<= 1;
    for (int j = threadIdx.x; j < p; j += kLT) {
        unsigned long long k = ~0ull;
        if (j < ns) {
            const int32_t e = sl + j;
            int32_t r0 = lo, r1 = hi;  // row of entry e: rowptr[r] <= e < rowptr[r + 1]
            while (r1 - r0 > 1) {
                const int32_t mid = (r0 + r1) >> 1;
                if (rowptr[mid] <= e) r0 = mid;
                else r1 = mid;
            }
            const unsigned long long pa = unsigned(parent[r0] - base);
            const unsigned long long pb = unsigned(parent[cols[e]]);
            k = (pa << 58) | (pb << kKeyBits) | unsigned(j);
        }
        key[j] = k;
    }
    __syncthreads();
    bitonic_sort(key, p);
    return ns;
}

__device__ __forceinline__ bool run_start(const unsigned long long *key, int j)
{
    return j == 0 || (key[j] >> kKeyBits) != (key[j - 1] >> kKeyBits);
}

__global__ __launch_bounds__(kLT) void k_coarse_count(int32_t nbanks, const int32_t *__restrict__ bank_off,
                                                       const int32_t *__restrict__ base,
                                                       const int32_t *__restrict__ count,
                                                       const int32_t *__restrict__ rowptr,
                                                       const int32_t *__restrict__ cols,
                                                       const int32_t *__restrict__ parent,
                                                       int32_t *__restrict__ ccount, int32_t *__restrict__ flags)
{
    extern __shared__ unsigned long long key[];
    __shared__ int32_t per_row[32];
    const int32_t b = blockIdx.x;
    const int32_t lo = bank_off[b], hi = bank_off[b + 1];
    const int32_t ns = rowptr[hi] - rowptr[lo];
    if (ns > kSortCap) {
        if (threadIdx.x == 0) flags[1] = 1;
        return;
    }
    if (threadIdx.x < 32) per_row[threadIdx.x] = 0;
    load_sorted_keys(key, lo, hi, base[b], rowptr, cols, parent);
    for (int j = threadIdx.x; j < ns; j += kLT)
        if (run_start(key, j)) atomicAdd(&per_row[int(key[j] >> 58)], 1);
    __syncthreads();
    if (threadIdx.x < count[b]) ccount[base[b] + threadIdx.x] = per_row[threadIdx.x];
}

__global__ __launch_bounds__(kLT) void k_coarse_fill(int32_t nbanks, const int32_t *__restrict__ bank_off,
                                                      const int32_t *__restrict__ base,
                                                      const int32_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ cols,
                                                      const int32_t *__restrict__ parent,
                                                      const int32_t *__restrict__ crowptr, int32_t *__restrict__ ccols,
                                                      int64_t *__restrict__ src_ptr, int32_t *__restrict__ src_blk)
{
    extern __shared__ unsigned long long key[];
    __shared__ int32_t sh[kLT];
    const int32_t b = blockIdx.x;
    const int32_t lo = bank_off[b], hi = bank_off[b + 1];
    const int32_t sl = rowptr[lo];
    const int32_t ns = rowptr[hi] - sl;
    if (ns > kSortCap) return;
    load_sorted_keys(key, lo, hi, base[b], rowptr, cols, parent);
    // coarse entries of this bank are consecutive: the first is crowptr[base]
    int32_t run = crowptr[base[b]];
    for (int32_t j0 = 0; j0 < ns; j0 += kLT) {
        const int32_t j = j0 + threadIdx.x;
        const int32_t st = (j < ns && run_start(key, j)) ? 1 : 0;
        // inclusive count of run starts up to j
        sh[threadIdx.x] = st;
        __syncthreads();
        for (int d = 1; d < kLT; d <<= 1) {
            const int32_t a = threadIdx.x >= d ? sh[threadIdx.x - d] : 0;
            __syncthreads();
            sh[threadIdx.x] += a;
            __syncthreads();
        }
        const int32_t incl = sh[threadIdx.x], tot = sh[kLT - 1];
        __syncthreads();
        if (j < ns) {
            const int32_t t = run + incl - 1;  // coarse entry of this key
            if (st) {
                ccols[t] = int32_t((key[j] >> kKeyBits) & ((1ull << 32) - 1ull));
                src_ptr[t] = int64_t(sl) + j;
            }
            src_blk[sl + j] = sl + int32_t(key[j] & ((1ull << kKeyBits) - 1ull));
        }
        run += tot;
    }
}

}  // namespace

int32_t device_components(int32_t fine_n, const int32_t *rowptr, const int32_t *cols, int32_t nbanks,
                          const int32_t *bank_off, DBuf<int32_t> &parent, DBuf<int64_t> &child_ptr,
                          DBuf<int32_t> &child, DBuf<int32_t> &base, DBuf<int32_t> &count, hipStream_t s)
{
    DBuf<int32_t> flags(4);
    flags.zero(s);
    count.resize(std::max(nbanks, 1));
    parent.resize(std::max(fine_n, 1));
    child.resize(std::max(fine_n, 1));
    base.resize(std::max(nbanks, 1));
    k_bank_cc<<<div_up(nbanks, kWavesLT), kLT, 0, s>>>(nbanks, bank_off, rowptr, cols, parent.get(), count.get(),
                                                        flags.get());
    LSS_LAUNCH_CHECK();
    DBuf<int32_t> tsum(std::max(div_up(nbanks, 1024), 1));
    excl_scan_i32(nbanks, count.get(), base.get(), tsum.get(), flags.get() + 2, s);
    int32_t h[3] = {0, 0, 0};
    LSS_HIP(hipMemcpyAsync(h, flags.get(), 3 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    LSS_HIP(hipStreamSynchronize(s));
    if (h[0]) fail(LSS_ERR_HIP, "MAS bank larger than 32 nodes");
    const int32_t coarse_n = h[2];
    child_ptr.resize(size_t(coarse_n) + 1);
    k_children<<<div_up(nbanks, kWavesLT), kLT, 0, s>>>(nbanks, bank_off, base.get(), parent.get(), child_ptr.get(),
                                                         child.get());
    k_set_last_i64<<<1, 1, 0, s>>>(child_ptr.get(), coarse_n, fine_n);
    LSS_LAUNCH_CHECK();
    return coarse_n;
}

bool device_galerkin_pattern(int64_t fine_nnz, const int32_t *rowptr, const int32_t *cols, int32_t nbanks,
                             const int32_t *bank_off, const int32_t *base, const int32_t *count,
                             const int32_t *parent, int32_t coarse_n, DBuf<int32_t> &crowptr, DBuf<int32_t> &ccols,
                             DBuf<int64_t> &src_ptr, DBuf<int32_t> &src_blk, int64_t &nslots, hipStream_t s)
{
    DBuf<int32_t> ccount(size_t(coarse_n) + 1), flags(4);
    flags.zero(s);
    k_bank_max_entries<<<div_up(nbanks, 256), 256, 0, s>>>(nbanks, bank_off, rowptr, flags.get() + 2);
    LSS_LAUNCH_CHECK();
    int32_t maxe = 0;
    LSS_HIP(hipMemcpyAsync(&maxe, flags.get() + 2, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    LSS_HIP(hipStreamSynchronize(s));
    if (maxe > kSortCap) return false;  // a bank with more stored entries than one workgroup sorts
    int p = 1;
    while (p < maxe) p <<= 1;
    const size_t lds = size_t(p) * sizeof(unsigned long long);
    static bool attr = false;
    if (!attr) {
        const int cap = int(size_t(kSortCap) * sizeof(unsigned long long));
        LSS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_coarse_count),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, cap));
        LSS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_coarse_fill),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, cap));
        attr = true;
    }
    k_coarse_count<<<nbanks, kLT, lds, s>>>(nbanks, bank_off, base, count, rowptr, cols, parent, ccount.get(),
                                            flags.get());
    LSS_HIP(hipMemsetAsync(ccount.get() + coarse_n, 0, sizeof(int32_t), s));
    LSS_LAUNCH_CHECK();
    crowptr.resize(size_t(coarse_n) + 1);
    DBuf<int32_t> tsum(std::max(div_up(coarse_n + 1, 1024), 1));
    excl_scan_i32(coarse_n + 1, ccount.get(), crowptr.get(), tsum.get(), flags.get() + 3, s);
    int32_t h[4] = {0, 0, 0, 0};
    LSS_HIP(hipMemcpyAsync(h, flags.get(), 4 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    LSS_HIP(hipStreamSynchronize(s));
    if (h[1]) return false;
    nslots = h[3];
    ccols.resize(size_t(std::max<int64_t>(nslots, 1)));
    src_ptr.resize(size_t(nslots) + 1);
    src_blk.resize(size_t(std::max<int64_t>(fine_nnz, 1)));
    k_coarse_fill<<<nbanks, kLT, lds, s>>>(nbanks, bank_off, base, rowptr, cols, parent, crowptr.get(), ccols.get(),
                                           src_ptr.get(), src_blk.get());
    k_set_last_i64<<<1, 1, 0, s>>>(src_ptr.get(), int32_t(nslots), fine_nnz);
    LSS_LAUNCH_CHECK();
    return true;
}

int32_t uniform_banks(int32_t n, DBuf<int32_t> &bank_off, hipStream_t s)
{
    const int32_t nbanks = (n + 31) / 32;
    bank_off.resize(size_t(nbanks) + 1);
    k_uniform_banks<<<div_up(nbanks + 1, 256), 256, 0, s>>>(n, nbanks, bank_off.get());
    LSS_LAUNCH_CHECK();
    return nbanks;
}

}  // namespace lss::k
