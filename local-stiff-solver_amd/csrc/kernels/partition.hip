// Recursive breadth-first bisection of the block graph on the device.
//
// Same result as the host partition_graph (and the oracle's orc_partition):
// nodes without off-diagonal entries go to trailing parts; every subset of
// more than max_part nodes is ordered by (breadth-first distance, node id)
// from its smallest member, again from the last node of that order (farthest,
// then largest id: a pseudo-peripheral node), and the first ceil(k/2)*max_part
// nodes of the second order go left; parts are numbered in recursion order.
// A search that runs out restarts from the smallest unreached member, one
// level beyond the deepest level so far.  Replaces KaMinPar
// (MAS/GraphPartition.cpp:176-289).
//
// All subsets of one recursion depth are searched together, one level of
// every search per step (two launches: expand, advance).  Only distances
// matter, so a level's nodes are appended in any order: 16 lanes share a
// frontier node's row, a wave appends its new nodes with one atomic.  The
// split then needs the cut level and, inside it, the id rank: two segmented
// scans over the member array, which keeps each subset in a contiguous
// ascending range; the left members take the front of the range, so at the
// end the array lists the nodes part by part and is the inverse of the
// partition permutation.
#include <hip/hip_runtime.h>

#include <climits>
#include <vector>

#include "../common/runtime.hpp"
#include "kernels.hpp"

namespace lss::k {
namespace {

constexpr int kPT = 256;   // threads per block
constexpr int kTile = 4 * kPT;
constexpr int kRowLanes = 16;  // lanes per frontier row in k_expand

struct Sub {
    int32_t off, m, base, phase;      // members mem[off, off+m); first part id; 0/1 searching, 2 finished
    int32_t head, tail, qend, level;  // frontier order[head, tail) at distance `level`; qend = append cursor
    int32_t cursor, lh, lt, cut;      // restart scan position; last non-empty frontier; cut level (-1: none)
    int32_t need, carried, pad0, pad1;  // cut-level nodes going left; 1 = leaf carried unchanged
};

__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t *sh, int32_t &total)
{
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

__host__ __device__ __forceinline__ int32_t left_count(int32_t m, int32_t max_part)
{
    const int32_t k = (m + max_part - 1) / max_part;
    return ((k + 1) / 2) * max_part;
}

// ---------------------------------------------------------------- exclusive scan of int32 flags
__global__ __launch_bounds__(kPT) void k_tile_sum(int32_t n, const int32_t *__restrict__ a, int32_t *__restrict__ tsum)
{
    __shared__ int32_t sh[kPT];
    const int64_t base = int64_t(blockIdx.x) * kTile;
    int32_t c = 0;
    for (int k = 0; k < 4; ++k) {
        const int64_t i = base + int64_t(k) * kPT + threadIdx.x;
        if (i < n) c += a[i];
    }
    c = block_sum_i(c, sh);
    if (threadIdx.x == 0) tsum[blockIdx.x] = c;
}

// in-place exclusive scan of ntiles values in one block; total -> *total
__global__ __launch_bounds__(kPT) void k_tile_scan(int32_t ntiles, int32_t *__restrict__ tsum,
                                                    int32_t *__restrict__ total)
{
    __shared__ int32_t sh[kPT];
    int32_t run = 0;
    for (int32_t b = 0; b < ntiles; b += kPT) {
        const int32_t i = b + threadIdx.x;
        const int32_t v = i < ntiles ? tsum[i] : 0;
        int32_t tot;
        const int32_t ex = block_excl_scan(v, sh, tot);
        if (i < ntiles) tsum[i] = run + ex;
        run += tot;
    }
    if (threadIdx.x == 0 && total) *total = run;
}

__global__ __launch_bounds__(kPT) void k_tile_apply(int32_t n, const int32_t *__restrict__ a,
                                                     const int32_t *__restrict__ toff, int32_t *__restrict__ out)
{
    __shared__ int32_t sh[kPT];
    const int64_t base = int64_t(blockIdx.x) * kTile;
    int32_t run = toff[blockIdx.x];
    for (int k = 0; k < 4; ++k) {
        const int64_t i = base + int64_t(k) * kPT + threadIdx.x;
        const int32_t v = i < n ? a[i] : 0;
        int32_t tot;
        const int32_t ex = block_excl_scan(v, sh, tot);
        if (i < n) out[i] = run + ex;
        run += tot;
    }
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

// coupled nodes first (ascending), then the others (ascending)
__global__ void k_place(int32_t nb, const int32_t *__restrict__ flag, const int32_t *__restrict__ rank,
                        const int32_t *__restrict__ total, int32_t *__restrict__ mem, int32_t *__restrict__ sub)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    if (flag[i]) {
        mem[rank[i]] = int32_t(i);
        sub[i] = 0;
    } else {
        mem[*total + (int32_t(i) - rank[i])] = int32_t(i);
        sub[i] = -1;
    }
}

__global__ void k_fill_i32(int64_t n, int32_t *__restrict__ a, int32_t v)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------------------------------------------------------- one depth
// Leaves (m <= max_part) take their part id once; larger subsets start the
// first search from their smallest member.
__global__ void k_begin(int32_t nsub, Sub *__restrict__ S, const int32_t *__restrict__ mem,
                        int32_t *__restrict__ order, int32_t *__restrict__ vis, int32_t *__restrict__ dist,
                        int32_t *__restrict__ part_off, int32_t max_part, int32_t pid, int32_t *__restrict__ flags)
{
    const int32_t s = int32_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= nsub) return;
    Sub st = S[s];
    st.phase = 2;
    if (st.m > 0 && !st.carried) {
        if (st.m <= max_part) {
            part_off[st.base] = st.off;
            st.carried = 1;
        } else {
            const int32_t v0 = mem[st.off];
            order[st.off] = v0;
            vis[v0] = pid;
            dist[v0] = 0;
            st.phase = 0;
            st.head = st.off;
            st.tail = st.qend = st.off + 1;
            st.level = 0;
            st.cursor = st.off;
            st.lh = st.off;
            st.lt = st.off + 1;
            st.cut = -1;
            st.need = 0;
            flags[0] = 1;
        }
    }
    S[s] = st;
}

// One level of every search: 16 lanes per frontier row; a node is taken by
// the lane whose compare-and-swap marks it and collected in LDS, and the
// workgroup appends its new nodes to the subset's queue with one atomic per
// flush (a device-scope atomic on the subset's queue end is the contended
// resource when one subset has a frontier of ~10^4 rows).
constexpr int kExpBuf = 2048;

__global__ __launch_bounds__(kPT) void k_expand(int32_t bx, Sub *__restrict__ S, const int32_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ cols, int32_t *__restrict__ order,
                                                 int32_t *__restrict__ vis, int32_t *__restrict__ dist,
                                                 const int32_t *__restrict__ sub, int32_t pid0)
{
    __shared__ int32_t buf[kExpBuf];
    __shared__ int32_t cnt, base, wlen[kPT / dev::kWave];
    const int32_t s = blockIdx.x / bx, b = blockIdx.x % bx;
    const Sub st = S[s];
    if (st.phase == 2) return;
    const int32_t pid = pid0 + st.phase;
    const int32_t f = st.tail - st.head;
    constexpr int kRowsPerPass = kPT / kRowLanes;
    int32_t chunk = (f + bx - 1) / bx;
    chunk = ((chunk + kRowsPerPass - 1) / kRowsPerPass) * kRowsPerPass;
    const int32_t lo = min(f, b * chunk), hi = min(f, lo + chunk);
    if (lo >= hi) return;  // uniform over the block
    const int t = threadIdx.x, lane = t & (dev::kWave - 1);
    const int g = t / kRowLanes, gl = t % kRowLanes;
    const int32_t dnew = st.level + 1;
    auto flush = [&]() {  // every thread; cnt was last changed before a barrier
        if (t == 0) base = atomicAdd(&S[s].qend, cnt);
        __syncthreads();
        for (int32_t i = t; i < cnt; i += kPT) order[base + i] = buf[i];
        __syncthreads();
        if (t == 0) cnt = 0;
        __syncthreads();
    };
    if (t == 0) cnt = 0;
    for (int32_t i0 = lo; i0 < hi; i0 += kRowsPerPass) {
        const int32_t i = i0 + g;
        int32_t r0 = 0, r1 = 0;
        if (i < hi) {
            const int32_t u = order[st.head + i];
            r0 = rowptr[u];
            r1 = rowptr[u + 1];
        }
        // longest row of the block: every lane runs the same number of rounds
        int32_t len = r1 - r0;
        for (int o = 32; o > 0; o >>= 1) len = max(len, __shfl_xor(len, o, dev::kWave));
        __syncthreads();
        if (lane == 0) wlen[t / dev::kWave] = len;
        __syncthreads();
        for (int w = 0; w < kPT / dev::kWave; ++w) len = max(len, wlen[w]);
        for (int32_t k = 0; k < len; k += kRowLanes) {
            const int32_t e = r0 + k + gl;
            if (e < r1) {
                const int32_t c = cols[e];
                if (sub[c] == s) {
                    const int32_t old = vis[c];
                    if (old != pid && atomicCAS(&vis[c], old, pid) == old) {
                        dist[c] = dnew;
                        buf[atomicAdd(&cnt, 1)] = c;
                    }
                }
            }
            __syncthreads();
            if (cnt > kExpBuf - kPT) flush();  // room for one more round
        }
    }
    __syncthreads();
    if (cnt > 0) flush();
}

// Close the level of every search: new frontier, cut level of the second
// search, restart from the smallest unreached member, or switch searches.
__global__ __launch_bounds__(kPT) void k_advance(Sub *__restrict__ S, const int32_t *__restrict__ mem,
                                                  int32_t *__restrict__ order, int32_t *__restrict__ vis,
                                                  int32_t *__restrict__ dist, int32_t pid0, int32_t max_part,
                                                  int32_t count_active, int32_t *__restrict__ active,
                                                  int32_t *__restrict__ flags)
{
    __shared__ int32_t sh[kPT];
    __shared__ int32_t found;
    const int32_t s = blockIdx.x;
    Sub st = S[s];
    if (st.phase == 2) return;
    const int32_t nl = left_count(st.m, max_part);
    st.level += 1;
    st.head = st.tail;
    st.tail = st.qend;
    if (st.tail > st.head) {
        st.lh = st.head;
        st.lt = st.tail;
    } else if (st.tail - st.off < st.m) {
        // smallest unreached member at or after the cursor
        const int32_t pid = pid0 + st.phase;
        if (threadIdx.x == 0) found = INT_MAX;
        __syncthreads();
        for (int32_t q0 = st.cursor; q0 < st.off + st.m; q0 += kPT) {
            const int32_t q = q0 + threadIdx.x;
            if (q < st.off + st.m && vis[mem[q]] != pid) atomicMin(&found, q);
            __syncthreads();
            if (found != INT_MAX) break;
            __syncthreads();
        }
        const int32_t fq = found;
        if (fq == INT_MAX) {  // cannot happen: every reached member is queued once
            st.phase = 2;
            if (threadIdx.x == 0) flags[3] = 1;
        } else {
            if (threadIdx.x == 0) {
                const int32_t v = mem[fq];
                order[st.tail] = v;
                vis[v] = pid;
                dist[v] = st.level;
            }
            st.tail += 1;
            st.qend = st.tail;
            st.cursor = fq + 1;
            st.lh = st.head;
            st.lt = st.tail;
        }
    }
    if (st.phase == 1 && st.cut < 0 && st.tail - st.off >= nl) {
        st.cut = st.level;
        st.need = nl - (st.head - st.off);
    }
    if (st.phase != 2 && st.head == st.tail) {  // every member reached
        if (st.phase == 0) {
            // farthest level's largest id starts the second search
            int32_t best = -1;
            for (int32_t q = st.lh + threadIdx.x; q < st.lt; q += kPT) best = max(best, order[q]);
            sh[threadIdx.x] = best;
            __syncthreads();
            for (int d = kPT / 2; d > 0; d >>= 1) {
                if (threadIdx.x < d) sh[threadIdx.x] = max(sh[threadIdx.x], sh[threadIdx.x + d]);
                __syncthreads();
            }
            const int32_t r2 = sh[0];
            if (threadIdx.x == 0) {
                order[st.off] = r2;
                vis[r2] = pid0 + 1;
                dist[r2] = 0;
            }
            st.phase = 1;
            st.head = st.off;
            st.tail = st.qend = st.off + 1;
            st.level = 0;
            st.cursor = st.off;
            st.lh = st.off;
            st.lt = st.off + 1;
        } else {
            st.phase = 2;
        }
    }
    if (threadIdx.x == 0) {
        S[s] = st;
        if (count_active && st.phase != 2) atomicAdd(active, 1);
    }
}

// Split: cut-level flags, then left flags (ranks come from two scans).
__global__ void k_cut_flags(int32_t real_n, const Sub *__restrict__ S, const int32_t *__restrict__ mem,
                            const int32_t *__restrict__ sub, const int32_t *__restrict__ dist, int32_t max_part,
                            int32_t *__restrict__ incut)
{
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= real_n) return;
    const int32_t v = mem[q];
    const Sub &st = S[sub[v]];
    incut[q] = (!st.carried && st.m > max_part && dist[v] == st.cut) ? 1 : 0;
}

__global__ void k_left_flags(int32_t real_n, const Sub *__restrict__ S, const int32_t *__restrict__ mem,
                             const int32_t *__restrict__ sub, const int32_t *__restrict__ dist,
                             const int32_t *__restrict__ cutrank, int32_t max_part, int32_t *__restrict__ left)
{
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= real_n) return;
    const int32_t v = mem[q];
    const Sub &st = S[sub[v]];
    int32_t l = 0;
    if (!st.carried && st.m > max_part) {
        const int32_t d = dist[v];
        l = d < st.cut || (d == st.cut && cutrank[q] - cutrank[st.off] < st.need);
    }
    left[q] = l;
}

__global__ void k_scatter(int32_t real_n, const Sub *__restrict__ S, const int32_t *__restrict__ mem,
                          int32_t *__restrict__ mem2, int32_t *__restrict__ sub, const int32_t *__restrict__ left,
                          const int32_t *__restrict__ lrank, int32_t max_part)
{
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= real_n) return;
    const int32_t v = mem[q];
    const int32_t s = sub[v];
    const Sub &st = S[s];
    if (st.carried || st.m <= max_part) {
        mem2[q] = v;
        sub[v] = 2 * s;
        return;
    }
    const int32_t lr = lrank[q] - lrank[st.off];
    if (left[q]) {
        mem2[st.off + lr] = v;
        sub[v] = 2 * s;
    } else {
        mem2[st.off + left_count(st.m, max_part) + (int32_t(q) - st.off - lr)] = v;
        sub[v] = 2 * s + 1;
    }
}

__global__ void k_children(int32_t nsub, const Sub *__restrict__ S, Sub *__restrict__ N, int32_t max_part)
{
    const int32_t s = int32_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= nsub) return;
    const Sub st = S[s];
    Sub a{}, c{};
    a.off = st.off;
    a.base = st.base;
    a.phase = c.phase = 2;
    c.off = st.off + st.m;
    c.base = st.base;
    if (st.carried || st.m <= max_part) {
        a.m = st.m;
        a.carried = st.carried;
    } else {
        const int32_t nl = left_count(st.m, max_part);
        a.m = nl;
        c.off = st.off + nl;
        c.m = st.m - nl;
        c.base = st.base + nl / max_part;
    }
    N[2 * s] = a;
    N[2 * s + 1] = c;
}

// ---------------------------------------------------------------- small subsets in one workgroup
// Every subset of at most T nodes finishes its whole recursion in one
// workgroup: the searches, cuts and splits above, over all parts of one
// recursion depth at once, with the per-node state in LDS.  A node is named
// by its rank u in the subset (rank order = id order, as the member array is
// ascending); key[v] = position of v in the member array, so a neighbour c
// belongs to the subset iff key[c] - off is in [0, m).  vd[u] packs the
// search stamp (high 16 bits) and the distance; a level is expanded by
// marking nodes, not by a queue: nodes at the same distance are only ever
// ordered by id.  The subset's adjacency is first rewritten as ranks (int16,
// -1 outside the subset) into adj, rows in member order from roff[off], so
// a search step waits on one global load instead of the row / column / key
// chain.
constexpr int kLT = 1024;      // threads per workgroup
constexpr int kLocalCap = 8192;  // most nodes per subset (8 positions per thread)

struct LocalLds {
    int32_t *vd, *aoff, *desc[2], *ph, *lv, *rc, *ct, *nd, *cn, *cm, *cf, *rq, *misc;
    uint16_t *mem[2], *lsub, *scan;
};

__host__ __device__ inline size_t local_lds_bytes(int32_t T, int32_t idcap)
{
    return size_t(2 * T + 1) * sizeof(int32_t) + size_t(idcap) * 17 * sizeof(int32_t) + 32 * sizeof(int32_t) +
           size_t(T) * 4 * sizeof(uint16_t);
}

__device__ inline LocalLds local_layout(unsigned char *lds, int32_t T, int32_t idcap)
{
    LocalLds L;
    L.vd = reinterpret_cast<int32_t *>(lds);
    L.aoff = L.vd + T;            // T + 1 row offsets into adj
    L.desc[0] = L.aoff + T + 1;  // off, m, base, carried (idcap each)
    L.desc[1] = L.desc[0] + 4 * idcap;
    L.ph = L.desc[1] + 4 * idcap;  // search state: phase (2 = none), level, reached, cut, need
    L.lv = L.ph + idcap;
    L.rc = L.lv + idcap;
    L.ct = L.rc + idcap;
    L.nd = L.ct + idcap;
    L.cn = L.nd + idcap;  // per step: new nodes, largest frontier rank, restart rank, restart request
    L.cm = L.cn + idcap;
    L.cf = L.cm + idcap;
    L.rq = L.cf + idcap;
    L.misc = L.rq + idcap;  // [0,16) wave sums, [16,18) active, [18] error, [19] any search, [20,22) restart
    L.mem[0] = reinterpret_cast<uint16_t *>(L.misc + 32);
    L.mem[1] = L.mem[0] + T;
    L.lsub = L.mem[1] + T;
    L.scan = L.lsub + T;
    return L;
}

// exclusive scan over the kLT threads; wsum holds 16 ints
__device__ __forceinline__ int32_t wg_excl_scan(int32_t v, int32_t *wsum, int32_t &total)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int32_t x = v;
    for (int d = 1; d < 64; d <<= 1) {
        const int32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (t < 64) {
        int32_t a = t < kLT / 64 ? wsum[t] : 0;
        for (int d = 1; d < kLT / 64; d <<= 1) {
            const int32_t y = __shfl_up(a, d, 64);
            if (lane >= d) a += y;
        }
        if (t < kLT / 64) wsum[t] = a;
    }
    __syncthreads();
    total = wsum[kLT / 64 - 1];
    const int32_t r = (w ? wsum[w - 1] : 0) + x - v;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kLT) void k_local_partition(const Sub *__restrict__ S, const int32_t *__restrict__ rowptr,
                                                        const int32_t *__restrict__ cols, int32_t *__restrict__ mem,
                                                        const int32_t *__restrict__ key,
                                                        const int32_t *__restrict__ roff, int16_t *__restrict__ adj,
                                                        int32_t *__restrict__ part_off, int32_t max_part, int32_t T,
                                                        int32_t idcap, int32_t *__restrict__ flags)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const Sub top = S[blockIdx.x];
    const int t = threadIdx.x;
    if (top.m == 0 || top.carried) return;
    if (top.m <= max_part) {  // a leaf seen for the first time
        if (t == 0) part_off[top.base] = top.off;
        return;
    }
    if (top.m > T) {
        if (t == 0) flags[3] = 1;
        return;
    }
    const int32_t O = top.off, M = top.m;
    LocalLds L = local_layout(lds, T, idcap);
    for (int32_t u = t; u < M; u += kLT) {
        L.vd[u] = 0;
        L.mem[0][u] = uint16_t(u);
        L.lsub[u] = 0;
    }
    for (int32_t j = t; j < idcap; j += kLT) {
        L.ph[j] = 2;
        L.cn[j] = 0;
        L.cm[j] = -1;
        L.cf[j] = INT_MAX;
        L.rq[j] = 0;
    }
    // ranks of the neighbours inside the subset, row by row in member order
    const int32_t A0 = roff[O];
    for (int32_t u = t; u < M; u += kLT) {
        const int32_t g = mem[O + u];
        const int32_t r0 = rowptr[g], r1 = rowptr[g + 1], a = roff[O + u];
        L.aoff[u] = a - A0;
        if (u == M - 1) L.aoff[M] = a - A0 + (r1 - r0);
        for (int32_t e0 = r0; e0 < r1; e0 += 16) {
            int32_t c[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) c[i] = e0 + i < r1 ? cols[e0 + i] : -1;
#pragma unroll
            for (int i = 0; i < 16; ++i)
                if (c[i] >= 0) {
                    const int32_t q = key[c[i]] - O;
                    adj[a + (e0 - r0) + i] = int16_t(uint32_t(q) < uint32_t(M) ? q : -1);
                }
        }
    }
    if (t == 0) {
        L.desc[0][0] = O;
        L.desc[0][idcap] = M;
        L.desc[0][2 * idcap] = top.base;
        L.desc[0][3 * idcap] = 0;
        L.misc[18] = 0;
    }
    __syncthreads();
    const int32_t P = (M + kLT - 1) / kLT;  // member positions per thread in the splits
    const int32_t q0 = min(M, t * P), q1 = min(M, q0 + P);
    int cb = 0;  // current descriptor / member buffer
    int32_t ncur = 1;
    for (int depth = 0;; ++depth) {
        int32_t *D = cb ? L.desc[1] : L.desc[0];
        const uint16_t *mc = cb ? L.mem[1] : L.mem[0];
        const int32_t st0 = 2 * depth + 1;  // stamp of the first search; the second uses st0 + 1
        if (t == 0) {
            L.misc[16] = L.misc[17] = L.misc[20] = L.misc[21] = 0;
            L.misc[19] = 0;
        }
        __syncthreads();
        // leaves take their part id once; larger parts start the first search from their smallest member
        for (int32_t j = t; j < ncur; j += kLT) {
            const int32_t m = D[idcap + j];
            L.ph[j] = 2;
            if (m > 0 && !D[3 * idcap + j]) {
                if (m <= max_part) {
                    part_off[D[2 * idcap + j]] = D[j];
                    D[3 * idcap + j] = 1;
                } else {
                    L.vd[mc[D[j] - O]] = st0 << 16;
                    L.ph[j] = 0;
                    L.lv[j] = 0;
                    L.rc[j] = 1;
                    L.ct[j] = -1;
                    L.nd[j] = 0;
                    L.misc[19] = 1;
                }
            }
        }
        __syncthreads();
        if (!L.misc[19]) break;
        for (int32_t step = 0;; ++step) {
            if (step > 4 * M + 16) {  // each step reaches a node, restarts or switches searches
                if (t == 0) L.misc[18] = 1;
                break;
            }
            // expand: mark the unvisited neighbours of every frontier node
            for (int32_t u = t; u < M; u += kLT) {
                const int j = L.lsub[u];
                const int p = L.ph[j];
                if (p == 2) continue;
                const int32_t cv = ((st0 + p) << 16) | L.lv[j];
                if (L.vd[u] != cv) continue;
                const int32_t a1 = L.aoff[u + 1];
                for (int32_t a0 = L.aoff[u]; a0 < a1; a0 += 16) {
                    int32_t q[16];
#pragma unroll
                    for (int i = 0; i < 16; ++i) q[i] = a0 + i < a1 ? int32_t(adj[A0 + a0 + i]) : -1;
#pragma unroll
                    for (int i = 0; i < 16; ++i)
                        if (q[i] >= 0 && L.lsub[q[i]] == j && (L.vd[q[i]] >> 16) != st0 + p) L.vd[q[i]] = cv + 1;
                }
            }
            __syncthreads();
            // count the new level; largest rank of the current level
            for (int32_t u = t; u < M; u += kLT) {
                const int j = L.lsub[u];
                const int p = L.ph[j];
                if (p == 2) continue;
                const int32_t cv = ((st0 + p) << 16) | L.lv[j];
                const int32_t v = L.vd[u];
                if (v == cv + 1) atomicAdd(&L.cn[j], 1);
                else if (v == cv) atomicMax(&L.cm[j], u);
            }
            __syncthreads();
            const int slot = step & 1;
            if (t == 0) {
                L.misc[16 + (slot ^ 1)] = 0;
                L.misc[20 + (slot ^ 1)] = 0;
            }
            for (int32_t j = t; j < ncur; j += kLT) {
                const int p = L.ph[j];
                if (p == 2) continue;
                const int32_t m = D[idcap + j];
                const int32_t before = L.rc[j], lvl = L.lv[j] + 1, nw = L.cn[j], far = L.cm[j];
                L.cn[j] = 0;
                L.cm[j] = -1;
                if (nw > 0) {
                    L.rc[j] = before + nw;
                    L.lv[j] = lvl;
                    const int32_t nl = left_count(m, max_part);
                    if (p == 1 && L.ct[j] < 0 && before + nw >= nl) {
                        L.ct[j] = lvl;
                        L.nd[j] = nl - before;
                    }
                    L.misc[16 + slot] = 1;
                } else if (before < m) {  // restart from the smallest unreached member
                    L.rq[j] = 1;
                    L.misc[20 + slot] = 1;
                    L.misc[16 + slot] = 1;
                } else if (p == 0) {  // farthest level's largest id starts the second search
                    if (far < 0) {
                        L.misc[18] = 1;
                        L.ph[j] = 2;
                        continue;
                    }
                    L.vd[far] = (st0 + 1) << 16;
                    L.ph[j] = 1;
                    L.lv[j] = 0;
                    L.rc[j] = 1;
                    L.misc[16 + slot] = 1;
                } else {
                    L.ph[j] = 2;
                }
            }
            __syncthreads();
            if (L.misc[20 + slot]) {
                for (int32_t u = t; u < M; u += kLT) {
                    const int j = L.lsub[u];
                    if (L.rq[j] && (L.vd[u] >> 16) != st0 + L.ph[j]) atomicMin(&L.cf[j], u);
                }
                __syncthreads();
                for (int32_t j = t; j < ncur; j += kLT) {
                    if (!L.rq[j]) continue;
                    L.rq[j] = 0;
                    const int32_t v = L.cf[j];
                    L.cf[j] = INT_MAX;
                    const int p = L.ph[j];
                    if (v == INT_MAX) {
                        L.misc[18] = 1;
                        L.ph[j] = 2;
                        continue;
                    }
                    const int32_t lvl = L.lv[j] + 1, before = L.rc[j];
                    L.vd[v] = ((st0 + p) << 16) | lvl;
                    L.lv[j] = lvl;
                    L.rc[j] = before + 1;
                    const int32_t nl = left_count(D[idcap + j], max_part);
                    if (p == 1 && L.ct[j] < 0 && before + 1 >= nl) {
                        L.ct[j] = lvl;
                        L.nd[j] = nl - before;
                    }
                }
                __syncthreads();
            }
            if (!L.misc[16 + slot]) break;
        }
        __syncthreads();
        if (L.misc[18]) break;
        if (2 * ncur > idcap) {
            if (t == 0) L.misc[18] = 1;
            break;
        }
        // split every searched part: rank inside the cut level, then left rank, in member order
        auto searched = [&](int j) { return D[idcap + j] > max_part && !D[3 * idcap + j]; };
        int32_t c = 0;
        for (int32_t q = q0; q < q1; ++q) {
            const int u = mc[q], j = L.lsub[u];
            c += searched(j) && (L.vd[u] & 0xFFFF) == L.ct[j];
        }
        int32_t tot;
        int32_t run = wg_excl_scan(c, L.misc, tot);
        for (int32_t q = q0; q < q1; ++q) {
            const int u = mc[q], j = L.lsub[u];
            L.scan[q] = uint16_t(run);
            run += searched(j) && (L.vd[u] & 0xFFFF) == L.ct[j];
        }
        __syncthreads();
        uint32_t bits = 0;
        c = 0;
        for (int32_t q = q0; q < q1; ++q) {
            const int u = mc[q], j = L.lsub[u];
            if (!searched(j)) continue;
            const int32_t d = L.vd[u] & 0xFFFF, cut = L.ct[j];
            const bool left = d < cut || (d == cut && int32_t(L.scan[q]) - int32_t(L.scan[D[j] - O]) < L.nd[j]);
            bits |= uint32_t(left) << (q - q0);
            c += left;
        }
        __syncthreads();
        run = wg_excl_scan(c, L.misc, tot);
        for (int32_t q = q0; q < q1; ++q) {
            L.scan[q] = uint16_t(run);
            run += (bits >> (q - q0)) & 1u;
        }
        __syncthreads();
        uint16_t *mn = cb ? L.mem[0] : L.mem[1];
        for (int32_t q = q0; q < q1; ++q) {
            const int u = mc[q], j = L.lsub[u];
            if (!searched(j)) {
                mn[q] = uint16_t(u);
                L.lsub[u] = uint16_t(2 * j);
                continue;
            }
            const int32_t o = D[j] - O, lr = int32_t(L.scan[q]) - int32_t(L.scan[o]);
            if ((bits >> (q - q0)) & 1u) {
                mn[o + lr] = uint16_t(u);
                L.lsub[u] = uint16_t(2 * j);
            } else {
                mn[o + left_count(D[idcap + j], max_part) + (q - o - lr)] = uint16_t(u);
                L.lsub[u] = uint16_t(2 * j + 1);
            }
        }
        int32_t *N = cb ? L.desc[0] : L.desc[1];
        for (int32_t j = t; j < ncur; j += kLT) {
            const int32_t off = D[j], m = D[idcap + j], base = D[2 * idcap + j], car = D[3 * idcap + j];
            int32_t am = m, acar = car, coff = off + m, cm = 0, cbase = base;
            if (!car && m > max_part) {
                const int32_t nl = left_count(m, max_part);
                am = nl;
                coff = off + nl;
                cm = m - nl;
                cbase = base + nl / max_part;
            }
            N[2 * j] = off;
            N[idcap + 2 * j] = am;
            N[2 * idcap + 2 * j] = base;
            N[3 * idcap + 2 * j] = acar;
            N[2 * j + 1] = coff;
            N[idcap + 2 * j + 1] = cm;
            N[2 * idcap + 2 * j + 1] = cbase;
            N[3 * idcap + 2 * j + 1] = 0;
        }
        __syncthreads();
        cb ^= 1;
        ncur *= 2;
    }
    __syncthreads();
    if (L.misc[18]) {
        if (t == 0) flags[3] = 1;
        return;
    }
    // the member array of the subset, part by part
    int32_t g[kLocalCap / kLT];
    const uint16_t *mc = cb ? L.mem[1] : L.mem[0];
#pragma unroll
    for (int i = 0; i < kLocalCap / kLT; ++i)
        if (q0 + i < q1) g[i] = mem[O + mc[q0 + i]];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kLocalCap / kLT; ++i)
        if (q0 + i < q1) mem[O + q0 + i] = g[i];
}

__global__ void k_member_row_len(int32_t n, const int32_t *__restrict__ mem, const int32_t *__restrict__ rowptr,
                                 int32_t *__restrict__ len)
{
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q < n) len[q] = rowptr[mem[q] + 1] - rowptr[mem[q]];
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

void excl_scan_i32(int32_t n, const int32_t *a, int32_t *out, int32_t *tsum, int32_t *total, hipStream_t s)
{
    const int32_t ntiles = div_up(n, kTile);
    if (ntiles == 0) return;
    k_tile_sum<<<ntiles, kPT, 0, s>>>(n, a, tsum);
    k_tile_scan<<<1, kPT, 0, s>>>(ntiles, tsum, total);
    k_tile_apply<<<ntiles, kPT, 0, s>>>(n, a, tsum, out);
}

// recursion depths until every part of a subset of m nodes has at most max_part
static int split_depths(int32_t m, int32_t max_part)
{
    int d = 0;
    for (; m > max_part; ++d) m = left_count(m, max_part);  // the left part is the larger one
    return d;
}

int32_t device_partition(int32_t nb, const int32_t *rowptr, const int32_t *cols, int max_part, int32_t *perm,
                         DBuf<int32_t> &part_off, hipStream_t s, int32_t local_max)
{
    DBuf<int32_t> flag(nb), rank(nb), mem(nb), mem2(nb), sub(nb), vis(nb), dist(nb), order(nb), aux(nb);
    DBuf<int32_t> tsum(std::max(div_up(nb, kTile), 1)), dflags(8);
    dflags.zero(s);
    k_coupled<<<div_up(nb, kPT), kPT, 0, s>>>(nb, rowptr, cols, flag.get());
    excl_scan_i32(nb, flag.get(), rank.get(), tsum.get(), dflags.get() + 2, s);
    k_place<<<div_up(nb, 256), 256, 0, s>>>(nb, flag.get(), rank.get(), dflags.get() + 2, mem.get(), sub.get());
    LSS_LAUNCH_CHECK();
    // the splits rewrite only the coupled range [0, real_n) of the second
    // member buffer; both buffers hold the trailing uncoupled nodes
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
    LSS_LAUNCH_CHECK();

    std::vector<Sub> h0(1);
    h0[0] = Sub{};
    h0[0].m = real_n;
    h0[0].phase = 2;
    DBuf<Sub> S, N;
    S.upload(h0.data(), 1, s);
    int32_t nsub = 1;
    int32_t hflags[2] = {0, 0};
    // two pinned active counts and their events
    struct Pinned {
        int32_t *p = nullptr;
        hipEvent_t ev[2] = {nullptr, nullptr};
        ~Pinned()
        {
            if (p) (void)hipHostFree(p);
            for (hipEvent_t e : ev)
                if (e) (void)hipEventDestroy(e);
        }
    } hact;
    LSS_HIP(hipHostMalloc(reinterpret_cast<void **>(&hact.p), 2 * sizeof(int32_t), hipHostMallocDefault));
    for (hipEvent_t &e : hact.ev) LSS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    int32_t maxm = real_n;  // largest subset of the current depth
    local_max = std::min<int32_t>(local_max, kLocalCap);
    for (int depth = 0; real_n > 0; ++depth) {
        if (maxm <= local_max && maxm > max_part) {
            // every subset of this depth finishes in one workgroup
            const int32_t T = maxm;
            const int32_t idcap = std::max(2, 1 << split_depths(T, max_part));
            const size_t lds = local_lds_bytes(T, idcap);
            int dev = 0, cap = 0;
            LSS_HIP(hipGetDevice(&dev));
            LSS_HIP(hipDeviceGetAttribute(&cap, hipDeviceAttributeSharedMemPerBlockOptin, dev));
            if (lds <= size_t(cap)) {
                LSS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_local_partition),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
                k_inverse_perm<<<div_up(nb, 256), 256, 0, s>>>(nb, mem.get(), aux.get());
                // row offsets of the rank adjacency: scan of the row lengths in member order
                k_member_row_len<<<div_up(real_n, 256), 256, 0, s>>>(real_n, mem.get(), rowptr, flag.get());
                excl_scan_i32(real_n, flag.get(), rank.get(), tsum.get(), dflags.get() + 4, s);
                int32_t nadj = 0;
                LSS_HIP(hipMemcpyAsync(&nadj, dflags.get() + 4, sizeof(int32_t), hipMemcpyDeviceToHost, s));
                LSS_HIP(hipStreamSynchronize(s));
                DBuf<int16_t> adj(std::max(nadj, 1));
                k_local_partition<<<nsub, kLT, lds, s>>>(S.get(), rowptr, cols, mem.get(), aux.get(), rank.get(),
                                                         adj.get(), part_off.get(), max_part, T, idcap, dflags.get());
                LSS_LAUNCH_CHECK();
                LSS_HIP(hipStreamSynchronize(s));  // adj is freed at scope end
                break;
            }
        }
        maxm = maxm > max_part ? left_count(maxm, max_part) : maxm;
        const int32_t pid0 = 2 * depth + 1;
        LSS_HIP(hipMemsetAsync(dflags.get(), 0, 2 * sizeof(int32_t), s));
        k_begin<<<div_up(nsub, 256), 256, 0, s>>>(nsub, S.get(), mem.get(), order.get(), vis.get(), dist.get(),
                                                  part_off.get(), max_part, pid0, dflags.get());
        LSS_LAUNCH_CHECK();
        LSS_HIP(hipMemcpyAsync(hflags, dflags.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
        LSS_HIP(hipStreamSynchronize(s));
        if (!hflags[0]) break;  // only leaves at this depth: the member array is final
        // about 2048 blocks per level step: a frontier of up to 32 768 rows takes one pass
        // of 16 rows per block, a larger one several; more blocks would only add
        // dispatch rounds of blocks that find no rows (most frontiers are far smaller)
        const int32_t bx = std::max(1, 2048 >> depth);
        const int64_t grid = int64_t(nsub) * bx;
        // each step reaches a node, restarts or switches searches: 4 nb bounds the steps
        constexpr int kBatch = 8;
        const int64_t max_batches = (4 * int64_t(nb)) / kBatch + 16;
        // batches of level steps, each ending with a count of the searches still active
        // (device slot 5 + parity, copied to pinned host memory); the next batch is queued
        // before the host waits for the count of the current one
        auto enqueue = [&](int slot) {
            int32_t *act = dflags.get() + 5 + slot;
            LSS_HIP(hipMemsetAsync(act, 0, sizeof(int32_t), s));
            for (int j = 0; j < kBatch; ++j) {
                k_expand<<<grid, kPT, 0, s>>>(bx, S.get(), rowptr, cols, order.get(), vis.get(), dist.get(),
                                              sub.get(), pid0);
                k_advance<<<nsub, kPT, 0, s>>>(S.get(), mem.get(), order.get(), vis.get(), dist.get(), pid0, max_part,
                                               j == kBatch - 1, act, dflags.get());
            }
            LSS_LAUNCH_CHECK();
            LSS_HIP(hipMemcpyAsync(hact.p + slot, act, sizeof(int32_t), hipMemcpyDeviceToHost, s));
            LSS_HIP(hipEventRecord(hact.ev[slot], s));
        };
        enqueue(0);
        for (int64_t batch = 0;; ++batch) {
            if (batch >= max_batches) {
                LSS_HIP(hipStreamSynchronize(s));
                fail(LSS_ERR_HIP, "device partition did not terminate");
            }
            enqueue(int((batch + 1) & 1));
            LSS_HIP(hipEventSynchronize(hact.ev[batch & 1]));
            if (hact.p[batch & 1] == 0) break;
        }
        // split every searched subset (flag, rank and aux are free here)
        const int gq = div_up(real_n, 256);
        k_cut_flags<<<gq, 256, 0, s>>>(real_n, S.get(), mem.get(), sub.get(), dist.get(), max_part, flag.get());
        excl_scan_i32(real_n, flag.get(), rank.get(), tsum.get(), nullptr, s);
        k_left_flags<<<gq, 256, 0, s>>>(real_n, S.get(), mem.get(), sub.get(), dist.get(), rank.get(), max_part,
                                        aux.get());
        excl_scan_i32(real_n, aux.get(), rank.get(), tsum.get(), nullptr, s);
        k_scatter<<<gq, 256, 0, s>>>(real_n, S.get(), mem.get(), mem2.get(), sub.get(), aux.get(), rank.get(),
                                     max_part);
        N.resize(size_t(nsub) * 2);
        k_children<<<div_up(nsub, 256), 256, 0, s>>>(nsub, S.get(), N.get(), max_part);
        LSS_LAUNCH_CHECK();
        std::swap(mem, mem2);
        std::swap(S, N);
        nsub *= 2;
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
