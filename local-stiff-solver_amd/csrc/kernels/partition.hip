// Recursive breadth-first bisection of the block graph on the device.
//
// Same result as the host partition_graph (and the oracle's orc_partition):
// nodes without off-diagonal entries go to trailing parts; every subset of
// more than max_part nodes is searched breadth-first from its smallest member,
// again from the last node reached (a pseudo-peripheral node), and the first
// ceil(k/2)*max_part nodes of the second search go left; parts are numbered in
// recursion order.  Replaces KaMinPar (MAS/GraphPartition.cpp:176-289).
//
// All subsets of one recursion depth are searched together, one level of
// every search per step: a frontier node claims each unvisited neighbour with
// an atomic minimum of its frontier index, so the next level lists every node
// under the earliest frontier node that reaches it, in that node's adjacency
// order -- exactly the queue order of a serial search.  A subset with an empty
// frontier and unvisited members restarts from its smallest unvisited member,
// as the serial search does.  Subsets keep their members in one array, each
// in a contiguous ascending range; after the split the left members take the
// front of the range, so at the end the array lists the nodes part by part and
// is the inverse of the partition permutation.
#include <hip/hip_runtime.h>

#include <climits>
#include <vector>

#include "../common/runtime.hpp"
#include "kernels.hpp"

namespace lss::k {
namespace {

constexpr int kPT = 256;  // threads per block
constexpr int kTile = 4 * kPT;

struct Sub {
    int32_t off, m, base;      // member range [off, off + m) and first part id
    int32_t head, tail;        // current frontier = order[head, tail)
    int32_t cursor;            // restart scan position in the member range
    int32_t phase;             // 0 / 1 = first / second search, 2 = finished
    int32_t pad;
};

__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t *sh, int32_t &total)
{
    // exclusive scan over the block (kPT threads), fixed order
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int d = 1; d < kPT; d <<= 1) {
        const int32_t a = t >= d ? sh[t - d] : 0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    total = sh[kPT - 1];
    const int32_t incl = sh[t];
    __syncthreads();
    return incl - v;
}

__device__ __forceinline__ int32_t block_sum_i(int32_t v, int32_t *sh)
{
    int32_t tot;
    (void)block_excl_scan(v, sh, tot);
    return tot;
}

// ---------------------------------------------------------------- set-up
__global__ __launch_bounds__(kPT) void k_coupled(int32_t nb, const int32_t *__restrict__ rowptr,
                                                  const int32_t *__restrict__ cols, int32_t *__restrict__ flag)
{
    const int64_t v = int64_t(blockIdx.x) * kPT + threadIdx.x;
    if (v >= nb) return;
    int32_t c = 0;
    for (int32_t e = rowptr[v]; e < rowptr[v + 1]; ++e) c |= cols[e] != v;
    flag[v] = c;
}

__global__ __launch_bounds__(kPT) void k_tile_count(int32_t n, const int32_t *__restrict__ flag,
                                                     int32_t *__restrict__ tcnt)
{
    __shared__ int32_t sh[kPT];
    const int64_t base = int64_t(blockIdx.x) * kTile;
    int32_t c = 0;
    for (int k = 0; k < 4; ++k) {
        const int64_t i = base + int64_t(k) * kPT + threadIdx.x;
        if (i < n) c += flag[i];
    }
    c = block_sum_i(c, sh);
    if (threadIdx.x == 0) tcnt[blockIdx.x] = c;
}

// exclusive scan of ntiles counts in one block; total -> *total
__global__ __launch_bounds__(kPT) void k_tile_scan(int32_t ntiles, int32_t *__restrict__ tcnt,
                                                    int32_t *__restrict__ total)
{
    __shared__ int32_t sh[kPT];
    int32_t run = 0;
    for (int32_t b = 0; b < ntiles; b += kPT) {
        const int32_t i = b + threadIdx.x;
        const int32_t v = i < ntiles ? tcnt[i] : 0;
        int32_t tot;
        const int32_t ex = block_excl_scan(v, sh, tot);
        if (i < ntiles) tcnt[i] = run + ex;
        run += tot;
    }
    if (threadIdx.x == 0) *total = run;
}

// coupled nodes first (ascending), then the others (ascending)
__global__ __launch_bounds__(kPT) void k_tile_scatter(int32_t n, const int32_t *__restrict__ flag,
                                                       const int32_t *__restrict__ toff,
                                                       const int32_t *__restrict__ total, int32_t *__restrict__ mem,
                                                       int32_t *__restrict__ sub)
{
    __shared__ int32_t sh[kPT];
    const int64_t base = int64_t(blockIdx.x) * kTile;
    const int32_t real_n = *total;
    int32_t run = toff[blockIdx.x];
    for (int k = 0; k < 4; ++k) {
        const int64_t i = base + int64_t(k) * kPT + threadIdx.x;
        const int32_t f = i < n ? flag[i] : 0;
        int32_t tot;
        const int32_t ex = block_excl_scan(f, sh, tot);
        if (i < n) {
            if (f) {
                mem[run + ex] = int32_t(i);
                sub[i] = 0;
            } else {
                mem[real_n + (int32_t(i) - (run + ex))] = int32_t(i);
                sub[i] = -1;
            }
        }
        run += tot;
    }
}

__global__ void k_fill_i32(int64_t n, int32_t *__restrict__ a, int32_t v)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------------------------------------------------------- one depth
// Leaves (m <= max_part) take their part id; larger subsets start the first search.
__global__ __launch_bounds__(64) void k_begin(int32_t nsub, Sub *__restrict__ S, const int32_t *__restrict__ mem,
                                              int32_t *__restrict__ order, int32_t *__restrict__ vis,
                                              int32_t *__restrict__ sub, int32_t *__restrict__ part_off,
                                              int32_t max_part, int32_t pid, int32_t *__restrict__ flags)
{
    const int32_t s = blockIdx.x;
    if (s >= nsub) return;
    Sub st = S[s];
    if (st.m == 0) {
        if (threadIdx.x == 0) S[s].phase = 2;
        return;
    }
    if (st.m <= max_part) {
        for (int32_t q = threadIdx.x; q < st.m; q += blockDim.x) sub[mem[st.off + q]] = -1;
        if (threadIdx.x == 0) {
            part_off[st.base] = st.off;
            S[s].phase = 2;
        }
        return;
    }
    if (threadIdx.x == 0) {
        const int32_t v0 = mem[st.off];
        order[st.off] = v0;
        vis[v0] = pid;
        st.head = st.off;
        st.tail = st.off + 1;
        st.cursor = st.off;
        st.phase = 0;
        S[s] = st;
        flags[0] = 1;
    }
}

struct Range {
    int32_t lo, hi;  // frontier indices of this block
};

__device__ __forceinline__ Range block_range(const Sub &st, int32_t b, int32_t bx)
{
    const int32_t f = st.tail - st.head;
    int32_t chunk = (f + bx - 1) / bx;
    chunk = chunk < kPT ? kPT : ((chunk + kPT - 1) / kPT) * kPT;
    Range r;
    r.lo = min(f, b * chunk);
    r.hi = min(f, r.lo + chunk);
    return r;
}

__global__ __launch_bounds__(kPT) void k_claim(int32_t bx, const Sub *__restrict__ S, const int32_t *__restrict__ rowptr,
                                                const int32_t *__restrict__ cols, const int32_t *__restrict__ order,
                                                const int32_t *__restrict__ vis, const int32_t *__restrict__ sub,
                                                int32_t *__restrict__ claim, int32_t pid0)
{
    const int32_t s = blockIdx.x / bx, b = blockIdx.x % bx;
    const Sub st = S[s];
    if (st.phase == 2) return;
    const int32_t pid = pid0 + st.phase;
    const Range r = block_range(st, b, bx);
    for (int32_t i = r.lo + threadIdx.x; i < r.hi; i += kPT) {
        const int32_t u = order[st.head + i];
        for (int32_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
            const int32_t w = cols[e];
            if (sub[w] == s && vis[w] != pid) atomicMin(&claim[w], i);
        }
    }
}

__device__ __forceinline__ int32_t owned_children(int32_t i, int32_t u, int32_t s, int32_t pid,
                                                  const int32_t *__restrict__ rowptr, const int32_t *__restrict__ cols,
                                                  const int32_t *__restrict__ vis, const int32_t *__restrict__ sub,
                                                  const int32_t *__restrict__ claim)
{
    int32_t c = 0;
    for (int32_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
        const int32_t w = cols[e];
        c += (sub[w] == s && vis[w] != pid && claim[w] == i);
    }
    return c;
}

__global__ __launch_bounds__(kPT) void k_count(int32_t bx, const Sub *__restrict__ S, const int32_t *__restrict__ rowptr,
                                                const int32_t *__restrict__ cols, const int32_t *__restrict__ order,
                                                const int32_t *__restrict__ vis, const int32_t *__restrict__ sub,
                                                const int32_t *__restrict__ claim, int32_t *__restrict__ bsum,
                                                int32_t pid0)
{
    __shared__ int32_t sh[kPT];
    const int32_t s = blockIdx.x / bx, b = blockIdx.x % bx;
    const Sub st = S[s];
    if (st.phase == 2) return;
    const int32_t pid = pid0 + st.phase;
    const Range r = block_range(st, b, bx);
    int32_t c = 0;
    for (int32_t i = r.lo + threadIdx.x; i < r.hi; i += kPT)
        c += owned_children(i, order[st.head + i], s, pid, rowptr, cols, vis, sub, claim);
    c = block_sum_i(c, sh);
    if (threadIdx.x == 0) bsum[blockIdx.x] = c;
}

__global__ __launch_bounds__(kPT) void k_write(int32_t bx, const Sub *__restrict__ S, const int32_t *__restrict__ rowptr,
                                                const int32_t *__restrict__ cols, int32_t *__restrict__ order,
                                                int32_t *__restrict__ vis, const int32_t *__restrict__ sub,
                                                int32_t *__restrict__ claim, const int32_t *__restrict__ bsum,
                                                int32_t pid0)
{
    __shared__ int32_t sh[kPT];
    const int32_t s = blockIdx.x / bx, b = blockIdx.x % bx;
    const Sub st = S[s];
    if (st.phase == 2) return;
    const int32_t pid = pid0 + st.phase;
    const Range r = block_range(st, b, bx);
    if (r.lo >= r.hi) return;  // uniform over the block
    // children of the earlier blocks of this subset come first
    int32_t before = 0;
    for (int32_t q = threadIdx.x; q < b; q += kPT) before += bsum[int64_t(s) * bx + q];
    int32_t at = st.tail + block_sum_i(before, sh);
    for (int32_t t0 = r.lo; t0 < r.hi; t0 += kPT) {
        const int32_t i = t0 + threadIdx.x;
        const int32_t u = i < r.hi ? order[st.head + i] : 0;
        const int32_t c = i < r.hi ? owned_children(i, u, s, pid, rowptr, cols, vis, sub, claim) : 0;
        int32_t tot;
        int32_t w_at = at + block_excl_scan(c, sh, tot);
        if (c > 0) {
            for (int32_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
                const int32_t w = cols[e];
                if (sub[w] == s && vis[w] != pid && claim[w] == i) {
                    order[w_at++] = w;
                    vis[w] = pid;
                    claim[w] = INT_MAX;
                }
            }
        }
        at += tot;
    }
}

// Close the level: advance the frontier, restart from the smallest unvisited
// member, or move to the next search / finish.
__global__ __launch_bounds__(kPT) void k_advance(int32_t bx, Sub *__restrict__ S, const int32_t *__restrict__ bsum,
                                                  const int32_t *__restrict__ mem, int32_t *__restrict__ order,
                                                  int32_t *__restrict__ vis, int32_t pid0, int32_t count_active,
                                                  int32_t *__restrict__ flags)
{
    __shared__ int32_t sh[kPT];
    __shared__ int32_t found;
    const int32_t s = blockIdx.x;
    Sub st = S[s];
    if (st.phase == 2) return;
    int32_t add = 0;
    for (int32_t q = threadIdx.x; q < bx; q += kPT) add += bsum[int64_t(s) * bx + q];
    add = block_sum_i(add, sh);
    st.head = st.tail;
    st.tail += add;
    int32_t pid = pid0 + st.phase;
    if (st.head == st.tail) {
        if (st.tail - st.off < st.m) {
            // smallest unvisited member at or after the cursor
            if (threadIdx.x == 0) found = INT_MAX;
            __syncthreads();
            for (int32_t q0 = st.cursor; q0 < st.off + st.m; q0 += kPT) {
                const int32_t q = q0 + threadIdx.x;
                if (q < st.off + st.m && vis[mem[q]] != pid) atomicMin(&found, q);
                __syncthreads();
                if (found != INT_MAX) break;
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                if (found == INT_MAX) {  // cannot happen: every visited member is queued once
                    st.phase = 2;
                    flags[3] = 1;
                } else {
                    const int32_t v = mem[found];
                    order[st.tail] = v;
                    vis[v] = pid;
                    st.tail += 1;
                    st.cursor = found + 1;
                }
            }
        } else if (st.phase == 0) {
            if (threadIdx.x == 0) {
                const int32_t start = order[st.off + st.m - 1];
                st.phase = 1;
                pid = pid0 + 1;
                order[st.off] = start;
                vis[start] = pid;
                st.head = st.off;
                st.tail = st.off + 1;
                st.cursor = st.off;
            }
        } else {
            st.phase = 2;
        }
    }
    if (threadIdx.x == 0) {
        S[s] = st;
        if (count_active && st.phase != 2) atomicAdd(&flags[1], 1);
    }
}

// Split every searched subset: the first nl nodes of its second search go
// left; members stay ascending on both sides.  Leaves keep their members and
// pass them on (k_begin finalizes a leaf again at every depth, idempotently).
__global__ __launch_bounds__(kPT) void k_split(int32_t nsub, const Sub *__restrict__ S, Sub *__restrict__ N,
                                                const int32_t *__restrict__ order, const int32_t *__restrict__ mem,
                                                int32_t *__restrict__ mem2, int32_t *__restrict__ sub,
                                                int32_t *__restrict__ left, int32_t max_part)
{
    __shared__ int32_t sh[kPT];
    const int32_t s = blockIdx.x;
    const Sub st = S[s];
    Sub a{}, c{};
    a.off = st.off;
    a.base = st.base;
    c.off = st.off + st.m;
    c.base = st.base;
    if (st.m <= max_part) {
        // a leaf is carried to the next depth unchanged (first child), so each
        // depth rewrites the whole coupled range of the other member buffer
        for (int32_t q = threadIdx.x; q < st.m; q += kPT) mem2[st.off + q] = mem[st.off + q];
        a.m = st.m;
    } else {
        const int32_t k = (st.m + max_part - 1) / max_part;
        const int32_t kl = (k + 1) / 2;
        const int32_t nl = kl * max_part;
        for (int32_t q = threadIdx.x; q < st.m; q += kPT) left[order[st.off + q]] = q < nl;
        __syncthreads();
        int32_t run = 0;  // left members placed so far
        for (int32_t q0 = 0; q0 < st.m; q0 += kPT) {
            const int32_t q = q0 + threadIdx.x;
            const int32_t v = q < st.m ? mem[st.off + q] : 0;
            const int32_t l = q < st.m ? left[v] : 0;
            int32_t tot;
            const int32_t ex = block_excl_scan(l, sh, tot);
            if (q < st.m) {
                if (l) {
                    mem2[st.off + run + ex] = v;
                    sub[v] = 2 * s;
                } else {
                    mem2[st.off + nl + (q - run - ex)] = v;
                    sub[v] = 2 * s + 1;
                }
            }
            run += tot;
        }
        a.m = nl;
        c.off = st.off + nl;
        c.m = st.m - nl;
        c.base = st.base + kl;
    }
    if (threadIdx.x == 0) {
        a.phase = c.phase = 2;
        N[2 * s] = a;
        N[2 * s + 1] = c;
    }
}

__global__ void k_inverse_perm(int32_t nb, const int32_t *__restrict__ mem, int32_t *__restrict__ perm)
{
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q < nb) perm[mem[q]] = int32_t(q);
}

__global__ void k_tail_part_off(int32_t real_parts, int32_t iso_parts, int32_t real_n, int32_t nb, int32_t max_part,
                                int32_t *__restrict__ part_off)
{
    const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < iso_parts) part_off[real_parts + j] = real_n + j * max_part;
    if (j == 0) part_off[real_parts + iso_parts] = nb;
}

}  // namespace

int32_t device_partition(int32_t nb, const int32_t *rowptr, const int32_t *cols, int max_part, int32_t *perm,
                         DBuf<int32_t> &part_off, hipStream_t s)
{
    DBuf<int32_t> flag(nb), mem(nb), mem2(nb), sub(nb), vis(nb), claim(nb), order(nb);
    const int32_t ntiles = div_up(nb, kTile);
    DBuf<int32_t> tcnt(std::max(ntiles, 1)), dflags(4);
    dflags.zero(s);
    k_coupled<<<div_up(nb, kPT), kPT, 0, s>>>(nb, rowptr, cols, flag.get());
    k_tile_count<<<ntiles, kPT, 0, s>>>(nb, flag.get(), tcnt.get());
    k_tile_scan<<<1, kPT, 0, s>>>(ntiles, tcnt.get(), dflags.get() + 2);
    k_tile_scatter<<<ntiles, kPT, 0, s>>>(nb, flag.get(), tcnt.get(), dflags.get() + 2, mem.get(), sub.get());
    LSS_LAUNCH_CHECK();
    // the splits rewrite only the coupled range [0, real_n) of the second member
    // buffer; both buffers must hold the trailing uncoupled nodes
    LSS_HIP(hipMemcpyAsync(mem2.get(), mem.get(), size_t(nb) * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    int32_t real_n = 0;
    LSS_HIP(hipMemcpyAsync(&real_n, dflags.get() + 2, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    LSS_HIP(hipStreamSynchronize(s));
    const int32_t real_parts = (real_n + max_part - 1) / max_part;
    const int32_t iso_parts = (nb - real_n + max_part - 1) / max_part;
    const int32_t nparts = real_parts + iso_parts;
    part_off.resize(size_t(nparts) + 1);
    k_tail_part_off<<<div_up(std::max(iso_parts, 1), 256), 256, 0, s>>>(real_parts, iso_parts, real_n, nb, max_part,
                                                                       part_off.get());
    k_fill_i32<<<div_up(nb, 256), 256, 0, s>>>(nb, vis.get(), 0);
    k_fill_i32<<<div_up(nb, 256), 256, 0, s>>>(nb, claim.get(), INT_MAX);
    LSS_LAUNCH_CHECK();

    std::vector<Sub> h0(1);
    h0[0] = Sub{0, real_n, 0, 0, 0, 0, 2, 0};
    DBuf<Sub> S, N;
    S.upload(h0.data(), 1, s);
    DBuf<int32_t> bsum;
    int32_t nsub = 1;
    int32_t hflags[2] = {0, 0};
    for (int depth = 0; real_n > 0; ++depth) {
        const int32_t pid0 = 2 * depth + 1;
        LSS_HIP(hipMemsetAsync(dflags.get(), 0, 2 * sizeof(int32_t), s));
        k_begin<<<nsub, 64, 0, s>>>(nsub, S.get(), mem.get(), order.get(), vis.get(), sub.get(), part_off.get(),
                                    max_part, pid0, dflags.get());
        LSS_LAUNCH_CHECK();
        LSS_HIP(hipMemcpyAsync(hflags, dflags.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
        LSS_HIP(hipStreamSynchronize(s));
        if (!hflags[0]) break;  // only leaves at this depth: the member array is final
        const int32_t bx = std::max(1, 1024 >> depth);
        bsum.resize(size_t(nsub) * bx);
        const int64_t grid = int64_t(nsub) * bx;
        // every step visits a node or restarts a search: two searches per
        // subset bound the steps by 4 nb
        const int64_t max_batches = (4 * int64_t(nb)) / 8 + 16;
        for (int64_t batch = 0;; ++batch) {
            if (batch >= max_batches) fail(LSS_ERR_HIP, "device partition did not terminate");
            constexpr int kBatch = 8;
            LSS_HIP(hipMemsetAsync(dflags.get() + 1, 0, sizeof(int32_t), s));
            for (int j = 0; j < kBatch; ++j) {
                k_claim<<<grid, kPT, 0, s>>>(bx, S.get(), rowptr, cols, order.get(), vis.get(), sub.get(), claim.get(),
                                             pid0);
                k_count<<<grid, kPT, 0, s>>>(bx, S.get(), rowptr, cols, order.get(), vis.get(), sub.get(), claim.get(),
                                             bsum.get(), pid0);
                k_write<<<grid, kPT, 0, s>>>(bx, S.get(), rowptr, cols, order.get(), vis.get(), sub.get(), claim.get(),
                                             bsum.get(), pid0);
                k_advance<<<nsub, kPT, 0, s>>>(bx, S.get(), bsum.get(), mem.get(), order.get(), vis.get(), pid0,
                                               j == kBatch - 1, dflags.get());
            }
            LSS_LAUNCH_CHECK();
            LSS_HIP(hipMemcpyAsync(hflags + 1, dflags.get() + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
            LSS_HIP(hipStreamSynchronize(s));
            if (hflags[1] == 0) break;
        }
        N.resize(size_t(nsub) * 2);
        k_split<<<nsub, kPT, 0, s>>>(nsub, S.get(), N.get(), order.get(), mem.get(), mem2.get(), sub.get(), flag.get(),
                                     max_part);
        LSS_LAUNCH_CHECK();
        std::swap(mem, mem2);
        std::swap(S, N);
        nsub *= 2;
        // the next depth's subsets start as "finished"; k_begin activates them
    }
    int32_t bad = 0;
    LSS_HIP(hipMemcpyAsync(&bad, dflags.get() + 3, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    LSS_HIP(hipStreamSynchronize(s));
    if (bad) fail(LSS_ERR_HIP, "device partition lost track of a subset");
    k_inverse_perm<<<div_up(nb, 256), 256, 0, s>>>(nb, mem.get(), perm);
    LSS_LAUNCH_CHECK();
    LSS_HIP(hipStreamSynchronize(s));
    return nparts;
}

}  // namespace lss::k
