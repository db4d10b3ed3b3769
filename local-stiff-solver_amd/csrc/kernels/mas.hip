// Multilevel additive Schwarz (MAS) preconditioner on gfx950.
//
// Semantics follow MAS/MASPreconditioner.cu:33-1164: level-0 banks are the
// graph-partition parts (<= 32 nodes); coarse level l+1 nodes are the bank-local
// connected components of level l; every bank matrix (dim 32*bs, zero diagonal
// entries set to 1) is inverted; the apply is
//     z = sum_l R_l^T  A_{l,bank}^{-1}  R_l r.
// The reference stores A_b^{-1} packed-upper from a Gauss-Jordan sweep
// (799-936); here each bank keeps its inverse Cholesky factor G = L^{-1}
// (A_b = L L^T), packed lower, the same number of values, and the apply computes
// G^T (G x).  The bank operator is then SPD by construction in every storage
// precision, where an explicit inverse of a bank with condition number 1e10-1e11
// (contact stiffness) can lose definiteness already in FP64.
// Layout and schedule are this build's own:
//  * parts are contiguous ranges of the partition-permuted numbering, so the
//    level-0 gather/scatter of the reference (padded copy, 593-659 / 745-777)
//    disappears: a bank reads and writes one contiguous slice of r and z;
//  * coarse operators are the Galerkin products A_{l+1} = R A_l R^T, summed in
//    a fixed source order (no FP atomics, reference 494-591 uses atomicAdd);
//  * one workgroup assembles AND inverts a bank in LDS (reference: separate
//    fill, pad and invert kernels, 939-979);
//  * the apply runs restriction level by level, the coarse banks from the
//    coarsest level down (each adding its parent's prolongated correction), and
//    finally the level-0 banks, whose epilogue adds the coarse correction,
//    scales, optionally adds a vector and reduces r.z;
//  * a bank apply splits both triangular products into 32 x 32 tiles so that
//    every lane of three waves runs 32 steps; 1, 2 or 4 right-hand sides share
//    each staged factor.
#include "mas_kernels.hpp"
#include "../common/runtime.hpp"

namespace lss::k {
namespace {

__device__ __forceinline__ int packed_index(int D, int i, int j)
{
    // row-major packed upper triangle, i <= j (MASPreconditioner.cu:483-491)
    return i * D - (i * (i + 1)) / 2 + j;
}

template <class T>
__device__ __forceinline__ T to_store(double v);
template <>
__device__ __forceinline__ double to_store<double>(double v) { return v; }
template <>
__device__ __forceinline__ float to_store<float>(double v) { return static_cast<float>(v); }
template <>
__device__ __forceinline__ __hip_bfloat16 to_store<__hip_bfloat16>(double v)
{
    return __float2bfloat16(static_cast<float>(v));
}

template <>
__device__ __forceinline__ _Float16 to_store<_Float16>(double v)
{
    return static_cast<_Float16>(static_cast<float>(v));
}

template <class T>
__device__ __forceinline__ double to_lds_any(T v) { return double(v); }
template <>
__device__ __forceinline__ double to_lds_any<__hip_bfloat16>(__hip_bfloat16 v) { return double(__bfloat162float(v)); }

template <class T>
__device__ __forceinline__ float to_lds(T v);
template <>
__device__ __forceinline__ float to_lds<float>(float v) { return v; }
template <>
__device__ __forceinline__ float to_lds<__hip_bfloat16>(__hip_bfloat16 v) { return __bfloat162float(v); }
template <>
__device__ __forceinline__ float to_lds<_Float16>(_Float16 v) { return static_cast<float>(v); }

constexpr int kInvThreads = 256;

__host__ __device__ constexpr int packed_lower(int i, int j)
{
    // row-major packed lower triangle, j <= i
    return (i * (i + 1)) / 2 + j;
}

// Tiled layout of the FP32 level-0 bank factors, in HBM and in LDS.  The lower
// triangle of a D x D factor (D = 32 BS) is cut into BS (BS + 1) / 2 tiles of
// 32 x 32, tile (b, s), s <= b, holding rows 32b.. and columns 32s.., numbered
// tile_index(b, s).
//  * HBM: the tiles in that order; an off-diagonal tile as 32 rows of 32
//    floats, row r of a diagonal tile as its r + 1 entries followed by zeros up
//    to a multiple of 4 (576 floats per diagonal tile; 4 800 per bank for bs = 3
//    against 4 656 packed).  Every 16-byte vector then lies in one tile row.
//  * LDS: every tile as 32 rows of kTileLd floats (16-byte aligned rows, so a
//    row segment is one 16-byte read), tiles 32 LDS banks apart; the entries
//    above the diagonal of a diagonal tile are zeroed once per workgroup and
//    never written again, so neither triangular product masks its loads.
constexpr int kTileLd = 36;
constexpr int kTileStride = 32 * kTileLd + 32;
constexpr int kDiagTileFloats = 576;

__host__ __device__ constexpr int tile_index(int b, int s) { return b * (b + 1) / 2 + s; }
__host__ __device__ constexpr int diag_row_len(int r) { return 4 * ((r + 4) / 4); }
__host__ __device__ constexpr int tiled_floats(int bs) { return bs * kDiagTileFloats + bs * (bs - 1) / 2 * 1024; }

// src[q]: packed-lower index of tiled HBM entry q (-1: zero padding);
// stage[v]: LDS byte offset of the 16-byte vector v of a tiled factor
template <int BS>
struct TiledLayout {
    static constexpr int ST = tiled_floats(BS);
    int16_t src[ST];
    uint16_t stage[ST / 4];
    constexpr TiledLayout() : src{}, stage{}
    {
        int q = 0;
        for (int b = 0; b < BS; ++b)
            for (int s = 0; s <= b; ++s)
                for (int r = 0; r < 32; ++r) {
                    const int len = s < b ? 32 : diag_row_len(r);
                    for (int c = 0; c < len; ++c, ++q) {
                        const int i = 32 * b + r, j = 32 * s + c;
                        src[q] = j <= i ? int16_t(packed_lower(i, j)) : int16_t(-1);
                        if (c % 4 == 0)
                            stage[q / 4] = uint16_t(4 * (tile_index(b, s) * kTileStride + r * kTileLd + c));
                    }
                }
    }
};
static_assert(diag_row_len(0) == 4 && diag_row_len(3) == 4 && diag_row_len(4) == 8 && diag_row_len(31) == 32,
              "rows padded to whole 16-byte vectors");
static_assert(tiled_floats(3) == 4800 && 4 * (tile_index(2, 2) * kTileStride + 31 * kTileLd + 31) < 65536,
              "tiled sizes and 16-bit LDS offsets");
__constant__ TiledLayout<1> kTiled1 = TiledLayout<1>();
__constant__ TiledLayout<2> kTiled2 = TiledLayout<2>();
__constant__ TiledLayout<3> kTiled3 = TiledLayout<3>();

template <int BS>
__device__ __forceinline__ const TiledLayout<BS> &tiled_layout()
{
    if constexpr (BS == 1) return kTiled1;
    else if constexpr (BS == 2) return kTiled2;
    else return kTiled3;
}

// Assemble the bank matrices of one level from its block operator, factor
// A_b = L L^T and invert L; store G = L^{-1} packed lower.  The matrix lives
// in registers: thread (a, b) of a 16 x 16 grid owns the entries
// (a + 16u, b + 16v), u, v < D / 16, so a rank-1 step of the right-looking
// Cholesky (and of the forward elimination L X = I that gives G) costs each
// thread 2 D / 16 LDS reads of the broadcast column / row and (D / 16)^2
// FMAs, with two barriers per step.  The packed lower triangle in LDS holds
// the assembled matrix, then L (read by the elimination), then G for the
// coalesced store.
template <int BS, class TS, bool TILED>
__global__ __launch_bounds__(kInvThreads) void k_bank_factor(int32_t nbanks, const int32_t *__restrict__ bank_off,
                                                             const int32_t *__restrict__ rowptr,
                                                             const int32_t *__restrict__ cols,
                                                             const double *__restrict__ vals, TS *__restrict__ out,
                                                             int *__restrict__ status)
{
    constexpr int D = 32 * BS;
    constexpr int R = D / 16;
    constexpr int S = D * (D + 1) / 2;
    constexpr int BB = BS * BS;
    constexpr int NW = kInvThreads / 64;
    static_assert(kInvThreads == 256, "16 x 16 thread grid");
    extern __shared__ double lds[];
    double *P = lds;       // packed lower triangle, S doubles
    double *vec = P + S;   // D doubles: column k of L, or row k of X
    double *dg = vec + D;  // the next diagonal entry
    __shared__ int bad;
    const int32_t bank = blockIdx.x;
    if (bank >= nbanks) return;
    const int32_t lo = bank_off[bank], hi = bank_off[bank + 1];
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int ta = t >> 4, tb = t & 15;
    if (t == 0) bad = 0;
    for (int q = t; q < S + D; q += kInvThreads) P[q] = 0.0;  // and vec
    __syncthreads();
    // assemble the lower triangle: block (i, J), J <= i, both nodes in the bank
    for (int32_t i = lo + wid; i < hi; i += NW) {
        for (int32_t k = rowptr[i] + lane; k < rowptr[i + 1]; k += 64) {
            const int32_t J = cols[k];
            if (J < lo || J > i) continue;
            const int r0 = (i - lo) * BS, c0 = (J - lo) * BS;
#pragma unroll
            for (int a = 0; a < BS; ++a)
#pragma unroll
                for (int b = 0; b < BS; ++b)
                    if (c0 + b <= r0 + a) P[packed_lower(r0 + a, c0 + b)] = vals[int64_t(k) * BB + a * BS + b];
        }
    }
    __syncthreads();
    // slots v <= u only: (a + 16u, b + 16v) with v > u lies above the diagonal
    double m[R][R];
#pragma unroll
    for (int u = 0; u < R; ++u)
#pragma unroll
        for (int v = 0; v <= u; ++v) {
            const int i = ta + 16 * u, j = tb + 16 * v;
            double x = j <= i ? P[packed_lower(i, j)] : 0.0;
            if (i == j && x == 0.0) x = 1.0;  // zero diagonal -> 1 (MASPreconditioner.cu:779-796): padding, empty rows
            m[u][v] = x;
        }
    if (t == 0) *dg = m[0][0];
    __syncthreads();
    // Cholesky, right-looking: scale column k, then the rank-1 update of the trailing lower triangle
    for (int k = 0; k < D; ++k) {
        const double d = *dg;
        if (t == 0 && (!(d > 0.0) || !isfinite(d))) bad = 1;
        const double rdk = 1.0 / sqrt(d);
        if (tb == (k & 15)) {
#pragma unroll
            for (int u = 0; u < R; ++u)
#pragma unroll
                for (int v = 0; v <= u; ++v) {
                    if (v != (k >> 4)) continue;
                    const int i = ta + 16 * u;
                    if (i > k) {
                        const double x = m[u][v] * rdk;
                        m[u][v] = x;
                        vec[i] = x;
                    } else if (i == k) {
                        m[u][v] = sqrt(d);
                    }
                }
        }
        __syncthreads();
        // slot columns with every column <= k are finished (a uniform test); inside the
        // others a zero factor leaves finished rows / columns unchanged
        double ci[R], cj[R];
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const double a = vec[ta + 16 * u], c = vec[tb + 16 * u];
            ci[u] = ta + 16 * u > k ? a : 0.0;
            cj[u] = tb + 16 * u > k ? c : 0.0;
        }
#pragma unroll
        for (int u = 0; u < R; ++u)
#pragma unroll
            for (int v = 0; v <= u; ++v)
                if (16 * v + 15 > k) m[u][v] -= ci[u] * cj[v];
        if (k + 1 < D && ta == ((k + 1) & 15) && tb == ta) {
#pragma unroll
            for (int u = 0; u < R; ++u)
                if (u == ((k + 1) >> 4)) *dg = m[u][u];
        }
        __syncthreads();
    }
    // L to LDS; X = L^{-1} by forward elimination of the identity, row k final at step k
#pragma unroll
    for (int u = 0; u < R; ++u)
#pragma unroll
        for (int v = 0; v <= u; ++v) {
            const int i = ta + 16 * u, j = tb + 16 * v;
            if (j <= i) P[packed_lower(i, j)] = m[u][v];
        }
    double x[R][R];
#pragma unroll
    for (int u = 0; u < R; ++u)
#pragma unroll
        for (int v = 0; v <= u; ++v) x[u][v] = (u == v && ta == tb) ? 1.0 : 0.0;
    __syncthreads();
    for (int k = 0; k < D; ++k) {
        if (ta == (k & 15)) {
            const double lkk = P[packed_lower(k, k)];
#pragma unroll
            for (int u = 0; u < R; ++u) {
                if (u != (k >> 4)) continue;
#pragma unroll
                for (int v = 0; v <= u; ++v) {
                    const int c = tb + 16 * v;
                    if (c <= k) {
                        const double y = x[u][v] / lkk;
                        x[u][v] = y;
                        vec[c] = y;
                    }
                }
            }
        }
        __syncthreads();
        double li[R], rc[R];
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int i = ta + 16 * u;
            const double a = P[packed_lower(i > k ? i : k, k)], c = vec[tb + 16 * u];
            li[u] = i > k ? a : 0.0;
            rc[u] = tb + 16 * u <= k ? c : 0.0;
        }
#pragma unroll
        for (int u = 0; u < R; ++u)
#pragma unroll
            for (int v = 0; v <= u; ++v)
                if (16 * u + 15 > k && 16 * v <= k) x[u][v] -= li[u] * rc[v];
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < R; ++u)
#pragma unroll
        for (int v = 0; v <= u; ++v) {
            const int i = ta + 16 * u, j = tb + 16 * v;
            if (j <= i) P[packed_lower(i, j)] = x[u][v];
        }
    __syncthreads();
    bool ok = true;
    if constexpr (TILED) {
        constexpr int ST = tiled_floats(BS);
        const int16_t *src = tiled_layout<BS>().src;
        TS *dst = out + int64_t(bank) * ST;
        for (int q = t; q < ST; q += kInvThreads) {
            const int p = src[q];
            const double v = p >= 0 ? P[p] : 0.0;
            if (!isfinite(v)) ok = false;
            dst[q] = to_store<TS>(v);
        }
    } else {
        TS *dst = out + int64_t(bank) * S;
        for (int q = t; q < S; q += kInvThreads) {
            const double v = P[q];
            if (!isfinite(v)) ok = false;
            dst[q] = to_store<TS>(v);
        }
    }
    if (!ok || (t == 0 && bad)) atomicOr(status, 1);
}

// Galerkin coarse operator values: target slot t sums its source blocks in order.
__global__ void k_galerkin(int64_t nslots, int bb, const int64_t *__restrict__ src_ptr,
                           const int32_t *__restrict__ src_blk, const double *__restrict__ fine_vals,
                           double *__restrict__ coarse_vals)
{
    const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= nslots * bb) return;
    const int64_t t = g / bb;
    const int e = int(g - t * bb);
    double s = 0.0;
    for (int64_t q = src_ptr[t]; q < src_ptr[t + 1]; ++q) s += fine_vals[int64_t(src_blk[q]) * bb + e];
    coarse_vals[g] = s;
}

// Restriction r_l = R r_{l-1}: a coarse node sums its children, which are the
// nodes of one component of one fine bank (at most 32 nodes).  32 lanes per
// fine bank (lane j = node lo + j).  A bank whose nodes share one parent (the
// usual case at level 1) is summed by a 32-lane butterfly; otherwise the bank's
// parents and values are staged in LDS and the first child of each parent sums
// that parent's children in index order.  Both orders are fixed.
template <int BS, int K>
__global__ __launch_bounds__(256) void k_restrict(int32_t nbanks, const int32_t *__restrict__ bank_off,
                                                  const int32_t *__restrict__ parent, const double *__restrict__ fine,
                                                  int64_t ld_fine, double *__restrict__ coarse, int64_t ld_coarse,
                                                  const int *__restrict__ stop)
{
    if (dev::all_stopped(stop, K)) return;
    __shared__ int32_t spar[8][32];
    __shared__ double sval[8][32 * BS * K];
    const int g = threadIdx.x >> 5, j = threadIdx.x & 31;  // bank group, lane in the group
    const int32_t bank = int32_t(blockIdx.x) * 8 + g;
    const int32_t lo = bank < nbanks ? bank_off[bank] : 0;
    const bool own = bank < nbanks && j < bank_off[bank + 1] - lo;
    const int32_t node = lo + j;
    const int32_t p = own ? parent[node] : -1 - j;  // never matches a real parent
    double v[K][BS];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < BS; ++c) v[k][c] = own ? fine[k * ld_fine + int64_t(node) * BS + c] : 0.0;
    const int32_t p0 = __shfl(p, 0, 32);
    const unsigned long long same = __ballot(!own || p == p0);
    const bool single = ((same >> (threadIdx.x & 32)) & 0xffffffffull) == 0xffffffffull;  // per half-wave
    if (single) {
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int c = 0; c < BS; ++c) {
                double a = v[k][c];
#pragma unroll
                for (int off = 16; off > 0; off >>= 1) a += __shfl_xor(a, off, 32);
                v[k][c] = a;
            }
        if (j == 0 && bank < nbanks)
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int c = 0; c < BS; ++c) coarse[k * ld_coarse + int64_t(p0) * BS + c] = v[k][c];
        return;
    }
    // several parents: both halves of a wave may come here; LDS is per half-wave
    spar[g][j] = p;
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < BS; ++c) sval[g][(j * BS + c) * K + k] = v[k][c];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (!own) return;
    bool first = true;
    for (int i = 0; i < 32; ++i) first &= !(i < j && spar[g][i] == p);
    if (!first) return;
    double acc[K][BS];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < BS; ++c) acc[k][c] = 0.0;
    for (int i = j; i < 32; ++i) {
        if (spar[g][i] != p) continue;
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int c = 0; c < BS; ++c) acc[k][c] += sval[g][(i * BS + c) * K + k];
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < BS; ++c) coarse[k * ld_coarse + int64_t(p) * BS + c] = acc[k][c];
}

// y = G^T (G x) per bank, G packed lower in LDS, for K columns.
//
// The matrix is cut into 32 x 32 tiles.  Phase 1 (t = G x): a half-wave task
// (b, s) multiplies rows 32b..32b+31 of G by entries 32s..32s+31 of x (s <= b,
// the diagonal tile masked to its lower triangle); phase 2 (y = G^T t): task
// (b, s) multiplies columns 32b..32b+31 by entries 32s..32s+31 of t (s >= b).
// Each lane owns one row (phase 1) or column (phase 2) of its tile and walks
// the 32 entries of the segment, so every lane runs the same 32 steps: for
// bs = 3 the six tasks of a phase fill three waves (96 wave-steps per phase,
// against 160 for one thread per row, whose longest row sets the pace).  The
// 32 lanes of a half-wave read the same vector entry (an LDS broadcast); the
// vectors are stored with one pad entry per 32 (entry e at e + e / 32) so the
// two halves of a wave read different LDS banks.  Each LDS load of G serves
// all K columns.  Partial sums of a row (column) over its tiles are added in
// tile order.
struct ApplyTask {
    int b, s;
};

template <int BS>
__device__ __forceinline__ ApplyTask apply_task(int phase, int w, int h)
{
    if constexpr (BS == 1) {
        return h == 0 ? ApplyTask{0, 0} : ApplyTask{-1, -1};
    } else if constexpr (BS == 2) {
        if (phase == 1) return w == 0 ? (h == 0 ? ApplyTask{0, 0} : ApplyTask{1, 0}) : (h == 0 ? ApplyTask{1, 1} : ApplyTask{-1, -1});
        return w == 0 ? (h == 0 ? ApplyTask{0, 1} : ApplyTask{1, 1}) : (h == 0 ? ApplyTask{0, 0} : ApplyTask{-1, -1});
    } else {
        if (phase == 1) {
            if (w == 0) return h == 0 ? ApplyTask{0, 0} : ApplyTask{1, 0};
            if (w == 1) return h == 0 ? ApplyTask{1, 1} : ApplyTask{2, 1};
            return h == 0 ? ApplyTask{2, 0} : ApplyTask{2, 2};
        }
        if (w == 0) return h == 0 ? ApplyTask{0, 1} : ApplyTask{1, 1};
        if (w == 1) return h == 0 ? ApplyTask{0, 2} : ApplyTask{1, 2};
        return h == 0 ? ApplyTask{0, 0} : ApplyTask{2, 2};
    }
}

// padded position of vector entry e (one pad entry per 32)
__device__ __forceinline__ int vpad(int e) { return e + (e >> 5); }

// xs: K padded bank vectors of stride XS = D + BS (zero past the bank's rows).
// ts, P, Q: LDS scratch (K * XS, K * BS * D, K * BS * D doubles).  After the
// final barrier, y_k(j) = sum over s = j / 32 .. BS - 1 of Q[(k * BS + s) * D + j].
template <int BS, class TL, int K>
// P (phase-1 partials) and Q (phase-2 partials) may be one buffer: P is last read
// before the barrier that precedes the first write of Q.
__device__ __forceinline__ void bank_solve(const TL *__restrict__ G, const double *__restrict__ xs,
                                           double *__restrict__ ts, double *P, double *Q)
{
    constexpr int D = 32 * BS;
    constexpr int XS = D + BS;
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63, h = lane >> 5, l = lane & 31;
    if (w < BS) {  // phase 1: P(s, row) = G(row, 32s..32s+31) x(32s..32s+31)
        const ApplyTask tB0 = apply_task<BS>(1, w, 1);
        const bool hasB = tB0.b >= 0;
        const ApplyTask t = (h && hasB) ? tB0 : apply_task<BS>(1, w, 0);  // a missing half repeats half A
        const int row = t.b * 32 + l;
        const int rbase = packed_lower(row, 0);
        const int xbase = vpad(t.s * 32);
        double a0[K], a1[K];
#pragma unroll
        for (int k = 0; k < K; ++k) a0[k] = a1[k] = 0.0;
#pragma unroll
        for (int u0 = 0; u0 < 32; u0 += 8) {
            double g[8], xv[K][8];
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const int j = t.s * 32 + u0 + v;  // in bounds for every lane; masked above the diagonal
                const double gv = double(G[rbase + j]);
                g[v] = j <= row ? gv : 0.0;
#pragma unroll
                for (int k = 0; k < K; ++k) xv[k][v] = xs[k * XS + xbase + u0 + v];
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                a0[k] += g[0] * xv[k][0] + g[2] * xv[k][2] + g[4] * xv[k][4] + g[6] * xv[k][6];
                a1[k] += g[1] * xv[k][1] + g[3] * xv[k][3] + g[5] * xv[k][5] + g[7] * xv[k][7];
            }
        }
        if (h == 0 || hasB)
#pragma unroll
            for (int k = 0; k < K; ++k) P[(k * BS + t.s) * D + row] = a0[k] + a1[k];
    }
    __syncthreads();
    // t = G x, padded like x
    for (int e = threadIdx.x; e < K * D; e += blockDim.x) {
        const int k = e / D, i = e - k * D;
        double v = P[(k * BS) * D + i];
        for (int q = 1; q <= i / 32; ++q) v += P[(k * BS + q) * D + i];
        ts[k * XS + vpad(i)] = v;
    }
    __syncthreads();
    if (w < BS) {  // phase 2: Q(s, col) = G(32s..32s+31, col)^T t(32s..32s+31)
        const ApplyTask tB0 = apply_task<BS>(2, w, 1);
        const bool hasB = tB0.b >= 0;
        const ApplyTask t = (h && hasB) ? tB0 : apply_task<BS>(2, w, 0);
        const int col = t.b * 32 + l;
        const int tbase = vpad(t.s * 32);
        double a0[K], a1[K];
#pragma unroll
        for (int k = 0; k < K; ++k) a0[k] = a1[k] = 0.0;
#pragma unroll
        for (int u0 = 0; u0 < 32; u0 += 8) {
            double g[8], tv[K][8];
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const int i = t.s * 32 + u0 + v;
                const double gv = double(G[packed_lower(i, 0) + col]);  // in bounds; masked below
                g[v] = i >= col ? gv : 0.0;
#pragma unroll
                for (int k = 0; k < K; ++k) tv[k][v] = ts[k * XS + tbase + u0 + v];
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                a0[k] += g[0] * tv[k][0] + g[2] * tv[k][2] + g[4] * tv[k][4] + g[6] * tv[k][6];
                a1[k] += g[1] * tv[k][1] + g[3] * tv[k][3] + g[5] * tv[k][5] + g[7] * tv[k][7];
            }
        }
        if (h == 0 || hasB)
#pragma unroll
            for (int k = 0; k < K; ++k) Q[(k * BS + t.s) * D + col] = a0[k] + a1[k];
    }
    __syncthreads();
}

// y_k(j) from the phase-2 partials
template <int BS>
__device__ __forceinline__ double bank_result(const double *__restrict__ Q, int k, int j)
{
    constexpr int D = 32 * BS;
    double v = Q[(k * BS + j / 32) * D + j];
    for (int q = j / 32 + 1; q < BS; ++q) v += Q[(k * BS + q) * D + j];
    return v;
}

template <class TS>
struct LdsType {
    using type = double;
};
template <>
struct LdsType<float> {
    using type = float;
};
template <>
struct LdsType<__hip_bfloat16> {
    using type = float;
};
template <>
struct LdsType<_Float16> {
    using type = float;
};

// Stage one bank's packed factor into LDS: every thread issues all of its
// 16-byte global loads (non-temporal: each factor is read once per apply) before
// the first LDS store, so a workgroup keeps the whole bank in flight.
template <class TS, class TL, int S, int NT>
__device__ __forceinline__ void stage_matrix(const TS *__restrict__ src, TL *__restrict__ dst)
{
    constexpr int VEC = 16 / sizeof(TS);           // elements per 16-byte load
    constexpr int NV = S / VEC;                    // full vectors
    constexpr int PER = (NV + NT - 1) / NT;
    typedef unsigned int V __attribute__((ext_vector_type(4)));
    const V *sv = reinterpret_cast<const V *>(src);
    V reg[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int t = threadIdx.x + q * NT;
        if (t < NV) reg[q] = __builtin_nontemporal_load(&sv[t]);
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int t = threadIdx.x + q * NT;
        if (t < NV) {
            const TS *e = reinterpret_cast<const TS *>(&reg[q]);
#pragma unroll
            for (int u = 0; u < VEC; ++u) dst[t * VEC + u] = static_cast<TL>(to_lds_any<TS>(e[u]));
        }
    }
    for (int t = NV * VEC + threadIdx.x; t < S; t += NT) dst[t] = static_cast<TL>(to_lds_any<TS>(src[t]));
}

constexpr int kApplyThreads = 256;
static_assert(kApplyThreads >= 96, "one thread per bank row in the epilogue");

// stage_matrix in two steps: the 16-byte loads of a bank into registers, and
// (after they land) the conversion into LDS, so that a workgroup can issue the
// loads of its next bank before it solves the current one
template <class TS, int S, int NT>
struct MatrixRegs {
    static constexpr int VEC = 16 / sizeof(TS);
    static constexpr int NV = S / VEC;
    static constexpr int PER = (NV + NT - 1) / NT;
    typedef unsigned int V __attribute__((ext_vector_type(4)));
    V reg[PER];
    __device__ __forceinline__ void issue(const TS *__restrict__ src)
    {
        const V *sv = reinterpret_cast<const V *>(src);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int t = threadIdx.x + q * NT;
            if (t < NV) reg[q] = __builtin_nontemporal_load(&sv[t]);
        }
    }
    template <class TL>
    __device__ __forceinline__ void store(const TS *__restrict__ src, TL *__restrict__ dst) const
    {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int t = threadIdx.x + q * NT;
            if (t < NV) {
                const TS *e = reinterpret_cast<const TS *>(&reg[q]);
#pragma unroll
                for (int u = 0; u < VEC; ++u) dst[t * VEC + u] = static_cast<TL>(to_lds_any<TS>(e[u]));
            }
        }
        for (int t = NV * VEC + threadIdx.x; t < S; t += NT) dst[t] = static_cast<TL>(to_lds_any<TS>(src[t]));
    }
};

// bank vector of each column into LDS (padded positions, zero past n entries)
template <int BS, int K>
__device__ __forceinline__ void load_bank_vector(const double *__restrict__ src, int64_t ld, int n,
                                                 double *__restrict__ xs)
{
    constexpr int D = 32 * BS;
    constexpr int XS = D + BS;
    for (int e = threadIdx.x; e < K * D; e += blockDim.x) {
        const int k = e / D, i = e - k * D;
        xs[k * XS + vpad(i)] = i < n ? src[k * ld + i] : 0.0;
    }
}

// Coarse level: y = A^{-1} r per bank of 32 nodes; w = y + w_parent[parent].
// K columns: r and w at stride ld, w_parent at stride ld_parent.  The epilogue's
// gathers are issued before the bank is staged, so their latency overlaps it.
template <int BS, class TS, int K>
__global__ __launch_bounds__(kApplyThreads) void k_coarse_apply(int32_t nodes, const TS *__restrict__ inv,
                                                                const double *__restrict__ r, int64_t ld,
                                                                const int32_t *__restrict__ parent,
                                                                const double *__restrict__ w_parent,
                                                                int64_t ld_parent, double *__restrict__ w,
                                                                const int *__restrict__ stop)
{
    constexpr int D = 32 * BS;
    constexpr int S = D * (D + 1) / 2;
    using TL = typename LdsType<TS>::type;
    if (dev::all_stopped(stop, K)) return;
    __shared__ __attribute__((aligned(16))) TL A[S + 4];
    __shared__ double xs[K * (D + BS)], ts[K * (D + BS)], PQ[K * BS * D];  // P, then Q (bank_solve)
    const int32_t bank = blockIdx.x;
    const int32_t lo = bank * 32;
    const int n = min(32, nodes - lo) * BS;
    const int i = threadIdx.x;
    double wp[K];
    if (i < n && w_parent) {
        const int64_t pi = int64_t(parent[lo + i / BS]) * BS + (i % BS);
#pragma unroll
        for (int k = 0; k < K; ++k) wp[k] = w_parent[k * ld_parent + pi];
    }
    load_bank_vector<BS, K>(r + int64_t(lo) * BS, ld, n, xs);
    stage_matrix<TS, TL, S, kApplyThreads>(inv + int64_t(bank) * S, A);
    __syncthreads();
    // rows beyond the bank's nodes are identity padding with zero right-hand side
    bank_solve<BS, TL, K>(A, xs, ts, PQ, PQ);
    if (i < n) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double y = bank_result<BS>(PQ, k, i);
            if (w_parent) y += wp[k];
            w[k * ld + int64_t(lo) * BS + i] = y;
        }
    }
}

// Level 0: z = scale * (A_bank^{-1} r_bank + w1[parent]) (+ add); optional r.z reduction.
// K columns: r, add, z, dotv at stride ld, w1 at stride ld1; the partials of
// column k start at partials + k * nbanks.
template <int BS, class TS, bool DOT, int K>
__global__ __launch_bounds__(kApplyThreads) void k_level0_apply(int32_t nbanks, const int32_t *__restrict__ part_off,
                                                                const TS *__restrict__ inv,
                                                                const double *__restrict__ r, int64_t ld,
                                                                const int32_t *__restrict__ parent,
                                                                const double *__restrict__ w1, int64_t ld1,
                                                                double scale, const double *__restrict__ add,
                                                                double *__restrict__ z,
                                                                const double *__restrict__ dotv, double *partials,
                                                                const int *__restrict__ stop)
{
    constexpr int D = 32 * BS;
    constexpr int S = D * (D + 1) / 2;
    using TL = typename LdsType<TS>::type;
    if (dev::all_stopped(stop, K)) return;
    __shared__ __attribute__((aligned(16))) TL A[S + 4];
    __shared__ double xs[K * (D + BS)], ts[K * (D + BS)], PQ[K * BS * D];  // P, then Q (bank_solve)
    const int32_t bank = blockIdx.x;
    const int32_t lo = part_off[bank];
    const int n = (part_off[bank + 1] - lo) * BS;
    const int i = threadIdx.x;
    const int64_t gi = int64_t(lo) * BS + i;
    // epilogue operands first: their latency overlaps the staging of the bank
    double cw[K], ca[K], cd[K];
    if (i < n) {
        const int64_t pi = w1 ? int64_t(parent[lo + i / BS]) * BS + (i % BS) : 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            cw[k] = w1 ? w1[k * ld1 + pi] : 0.0;
            ca[k] = add ? add[k * ld + gi] : 0.0;
            if constexpr (DOT) cd[k] = dotv[k * ld + gi];
        }
    }
    load_bank_vector<BS, K>(r + int64_t(lo) * BS, ld, n, xs);
    stage_matrix<TS, TL, S, kApplyThreads>(inv + int64_t(bank) * S, A);
    __syncthreads();
    bank_solve<BS, TL, K>(A, xs, ts, PQ, PQ);
    double part[K];
#pragma unroll
    for (int k = 0; k < K; ++k) part[k] = 0.0;
    if (i < n) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double y = bank_result<BS>(PQ, k, i);
            if (w1) y += cw[k];
            y *= scale;
            if (add) y += ca[k];
            z[k * ld + gi] = y;
            if constexpr (DOT) part[k] = cd[k] * y;
        }
    }
    if constexpr (DOT) {
        __shared__ double sh[kApplyThreads / 64];
#pragma unroll
        for (int k = 0; k < K; ++k)
            dev::store_partial(dev::block_sum<kApplyThreads>(part[k], sh), partials + int64_t(k) * nbanks);
    }
}

// Level 0 for one right-hand side, two banks per workgroup (k_level0_apply's
// steps per bank); the FP64 banks (37 KB each) gain most from the overlap.
template <int BS, class TS, bool DOT, int K>
__global__ __launch_bounds__(kApplyThreads) void k_level0_apply2(int32_t nbanks, const int32_t *__restrict__ part_off,
                                                                const TS *__restrict__ inv,
                                                                const double *__restrict__ r, int64_t ld,
                                                                const int32_t *__restrict__ parent,
                                                                const double *__restrict__ w1, int64_t ld1,
                                                                double scale, const double *__restrict__ add,
                                                                double *__restrict__ z,
                                                                const double *__restrict__ dotv, double *partials,
                                                                const int *__restrict__ stop)
{
    constexpr int D = 32 * BS;
    constexpr int S = D * (D + 1) / 2;
    using TL = typename LdsType<TS>::type;
    if (dev::all_stopped(stop, K)) return;
    __shared__ __attribute__((aligned(16))) TL A[S + 4];
    __shared__ double xs[K * (D + BS)], ts[K * (D + BS)], PQ[K * BS * D];  // P, then Q (bank_solve)
    const int i = threadIdx.x;
    // two consecutive banks per workgroup; the matrix loads of the second are in
    // flight while the first is solved
    const int32_t first = int32_t(blockIdx.x) * 2;
    const int32_t last = min(nbanks, first + 2);
    MatrixRegs<TS, S, kApplyThreads> mr;
    mr.issue(inv + int64_t(first) * S);
    for (int32_t bank = first; bank < last; ++bank) {
        const int32_t lo = part_off[bank];
        const int n = (part_off[bank + 1] - lo) * BS;
        const int64_t gi = int64_t(lo) * BS + i;
        // epilogue operands first: their latency overlaps the staging of the bank
        double cw[K], ca[K], cd[K];
        if (i < n) {
            const int64_t pi = w1 ? int64_t(parent[lo + i / BS]) * BS + (i % BS) : 0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                cw[k] = w1 ? w1[k * ld1 + pi] : 0.0;
                ca[k] = add ? add[k * ld + gi] : 0.0;
                if constexpr (DOT) cd[k] = dotv[k * ld + gi];
            }
        }
        load_bank_vector<BS, K>(r + int64_t(lo) * BS, ld, n, xs);
        mr.store(inv + int64_t(bank) * S, A);
        __syncthreads();
        if (bank + 1 < last) mr.issue(inv + int64_t(bank + 1) * S);
        bank_solve<BS, TL, K>(A, xs, ts, PQ, PQ);
        double part[K];
#pragma unroll
        for (int k = 0; k < K; ++k) part[k] = 0.0;
        if (i < n) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                double y = bank_result<BS>(PQ, k, i);
                if (w1) y += cw[k];
                y *= scale;
                if (add) y += ca[k];
                z[k * ld + gi] = y;
                if constexpr (DOT) part[k] = cd[k] * y;
            }
        }
        if constexpr (DOT) {
            __shared__ double sh[kApplyThreads / 64];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const double tot = dev::block_sum<kApplyThreads>(part[k], sh);
                if (threadIdx.x == 0) partials[int64_t(k) * nbanks + bank] = tot;
            }
        }
        __syncthreads();  // A, xs and Q are rewritten for the next bank
    }
}

// bank_solve on the tiled LDS layout: the same tasks and the same products in
// the same order (identical results), with the diagonal tiles' zero upper
// halves in place of the masks and 16-byte row reads in phase 1.
template <int BS, int K>
__device__ __forceinline__ void tile_solve(const float *__restrict__ Gt, const double *__restrict__ xs,
                                           double *__restrict__ ts, double *P, double *Q)
{
    constexpr int D = 32 * BS;
    constexpr int XS = D + BS;
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63, h = lane >> 5, l = lane & 31;
    typedef float F4 __attribute__((ext_vector_type(4)));
    if (w < BS) {  // phase 1: P(s, row) = G(row, 32s..32s+31) x(32s..32s+31)
        const ApplyTask tB0 = apply_task<BS>(1, w, 1);
        const bool hasB = tB0.b >= 0;
        const ApplyTask t = (h && hasB) ? tB0 : apply_task<BS>(1, w, 0);
        const int row = t.b * 32 + l;
        const float *G = Gt + tile_index(t.b, t.s) * kTileStride + l * kTileLd;
        const int xbase = vpad(t.s * 32);
        double a0[K], a1[K];
#pragma unroll
        for (int k = 0; k < K; ++k) a0[k] = a1[k] = 0.0;
#pragma unroll
        for (int u0 = 0; u0 < 32; u0 += 8) {
            const F4 ga = *reinterpret_cast<const F4 *>(G + u0);
            const F4 gb = *reinterpret_cast<const F4 *>(G + u0 + 4);
            const double g[8] = {double(ga.x), double(ga.y), double(ga.z), double(ga.w),
                                 double(gb.x), double(gb.y), double(gb.z), double(gb.w)};
            double xv[K][8];
#pragma unroll
            for (int v = 0; v < 8; ++v)
#pragma unroll
                for (int k = 0; k < K; ++k) xv[k][v] = xs[k * XS + xbase + u0 + v];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                a0[k] += g[0] * xv[k][0] + g[2] * xv[k][2] + g[4] * xv[k][4] + g[6] * xv[k][6];
                a1[k] += g[1] * xv[k][1] + g[3] * xv[k][3] + g[5] * xv[k][5] + g[7] * xv[k][7];
            }
        }
        if (h == 0 || hasB)
#pragma unroll
            for (int k = 0; k < K; ++k) P[(k * BS + t.s) * D + row] = a0[k] + a1[k];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < K * D; e += blockDim.x) {
        const int k = e / D, i = e - k * D;
        double v = P[(k * BS) * D + i];
        for (int q = 1; q <= i / 32; ++q) v += P[(k * BS + q) * D + i];
        ts[k * XS + vpad(i)] = v;
    }
    __syncthreads();
    if (w < BS) {  // phase 2: Q(s, col) = G(32s..32s+31, col)^T t(32s..32s+31)
        const ApplyTask tB0 = apply_task<BS>(2, w, 1);
        const bool hasB = tB0.b >= 0;
        const ApplyTask t = (h && hasB) ? tB0 : apply_task<BS>(2, w, 0);
        const int col = t.b * 32 + l;
        const float *G = Gt + tile_index(t.s, t.b) * kTileStride + l;
        const int tbase = vpad(t.s * 32);
        double a0[K], a1[K];
#pragma unroll
        for (int k = 0; k < K; ++k) a0[k] = a1[k] = 0.0;
#pragma unroll
        for (int u0 = 0; u0 < 32; u0 += 8) {
            double g[8], tv[K][8];
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                g[v] = double(G[(u0 + v) * kTileLd]);
#pragma unroll
                for (int k = 0; k < K; ++k) tv[k][v] = ts[k * XS + tbase + u0 + v];
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                a0[k] += g[0] * tv[k][0] + g[2] * tv[k][2] + g[4] * tv[k][4] + g[6] * tv[k][6];
                a1[k] += g[1] * tv[k][1] + g[3] * tv[k][3] + g[5] * tv[k][5] + g[7] * tv[k][7];
            }
        }
        if (h == 0 || hasB)
#pragma unroll
            for (int k = 0; k < K; ++k) Q[(k * BS + t.s) * D + col] = a0[k] + a1[k];
    }
    __syncthreads();
}

// Waves per SIMD asked of the compiler, as many workgroups as the LDS of a CU
// holds for bs = 3: 5 of 32 KB for one column (the variant without the dot
// product fits the 96 registers that takes; with it, or for bs = 2, the kernel
// would spill and keeps the compiler's own choice), 4 of 36 KB for two columns,
// 3 of 44 KB for four (again not with the dot product, which would spill).
// banks per workgroup of the tiled level-0 apply (the grid is nbanks / this)
__host__ __device__ constexpr int tiled_banks_per_group(int k) { return k == 1 ? 4 : (k == 2 ? 2 : 1); }

template <int BS, bool DOT, int K>
constexpr int kTiledWaves = BS != 3 ? 1 : (K == 1 ? (DOT ? 1 : 5) : (K == 2 ? 4 : (DOT ? 1 : 3)));

// Level 0 with FP32 factors in the tiled layout: one workgroup solves two
// consecutive banks for 1 or 2 right-hand sides (the second bank's factor loads
// in flight while the first solves), one bank for 4 columns.  Each 16-byte
// vector of a factor goes to LDS with one 16-byte store at its offset from the
// constant layout table.  Epilogue as in k_level0_apply.
template <int BS, bool DOT, int K>
__global__ __launch_bounds__(kApplyThreads, (kTiledWaves<BS, DOT, K>)) void k_level0_apply_tiled(int32_t nbanks,
                                                                     const int32_t *__restrict__ part_off,
                                                                     const float *__restrict__ inv,
                                                                     const double *__restrict__ r, int64_t ld,
                                                                     const int32_t *__restrict__ parent,
                                                                     const double *__restrict__ w1, int64_t ld1,
                                                                     double scale, const double *__restrict__ add,
                                                                     double *__restrict__ z,
                                                                     const double *__restrict__ dotv,
                                                                     double *partials, const int *__restrict__ stop)
{
    constexpr int D = 32 * BS;
    constexpr int ST = tiled_floats(BS);
    constexpr int NTILE = BS * (BS + 1) / 2;
    constexpr int NB = tiled_banks_per_group(K);
    using Regs = MatrixRegs<float, ST, kApplyThreads>;
    constexpr int NV = Regs::NV, PER = Regs::PER;
    static_assert(ST % 4 == 0, "whole 16-byte vectors");
    if (dev::all_stopped(stop, K)) return;
    __shared__ __attribute__((aligned(16))) float Gt[NTILE * kTileStride];
    __shared__ double xs[K * (D + BS)], ts[K * (D + BS)], PQ[K * BS * D];  // P, then Q (tile_solve)
    const int i = threadIdx.x;
    const int32_t first = int32_t(blockIdx.x) * NB;
    if (first >= nbanks) return;
    const int32_t last = min(nbanks, first + NB);
    Regs mr;
    mr.issue(inv + int64_t(first) * ST);
    typedef float F4 __attribute__((ext_vector_type(4)));
    for (int e = i; e < NTILE * kTileStride / 4; e += kApplyThreads)
        reinterpret_cast<F4 *>(Gt)[e] = F4{0.0f, 0.0f, 0.0f, 0.0f};
    __syncthreads();  // the zeros land before any factor entry
    const uint16_t *stage = tiled_layout<BS>().stage;
    char *Gb = reinterpret_cast<char *>(Gt);
    for (int32_t bank = first; bank < last; ++bank) {
        const int32_t lo = part_off[bank];
        const int n = (part_off[bank + 1] - lo) * BS;
        const int64_t gi = int64_t(lo) * BS + i;
        int off[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int t = i + q * kApplyThreads;
            off[q] = t < NV ? stage[t] : 0;
        }
        // epilogue operands first: their latency overlaps the staging of the bank
        double cw[K], ca[K], cd[K];
        if (i < n) {
            const int64_t pi = w1 ? int64_t(parent[lo + i / BS]) * BS + (i % BS) : 0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                cw[k] = w1 ? w1[k * ld1 + pi] : 0.0;
                ca[k] = add ? add[k * ld + gi] : 0.0;
                if constexpr (DOT) cd[k] = dotv[k * ld + gi];
            }
        }
        load_bank_vector<BS, K>(r + int64_t(lo) * BS, ld, n, xs);
#pragma unroll
        for (int q = 0; q < PER; ++q)
            if (i + q * kApplyThreads < NV) *reinterpret_cast<typename Regs::V *>(Gb + off[q]) = mr.reg[q];
        __syncthreads();
        if (bank + 1 < last) mr.issue(inv + int64_t(bank + 1) * ST);
        tile_solve<BS, K>(Gt, xs, ts, PQ, PQ);
        double part[K];
#pragma unroll
        for (int k = 0; k < K; ++k) part[k] = 0.0;
        if (i < n) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                double y = bank_result<BS>(PQ, k, i);
                if (w1) y += cw[k];
                y *= scale;
                if (add) y += ca[k];
                z[k * ld + gi] = y;
                if constexpr (DOT) part[k] = cd[k] * y;
            }
        }
        if constexpr (DOT) {
            __shared__ double sh[kApplyThreads / 64];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const double tot = dev::block_sum<kApplyThreads>(part[k], sh);
                if (threadIdx.x == 0) partials[int64_t(k) * nbanks + bank] = tot;
            }
        }
        __syncthreads();  // Gt (tile rows), xs and Q are rewritten for the next bank
    }
}

// ------------------------------------------------------------ MFMA bank solve
// The same G^T (G x) on the tiled LDS factor with v_mfma_f32_4x4x1f32, whose 16
// blocks each multiply a 4-entry column of A by a 4-entry row of B per step:
// lane 4b + i holds A_b[i], lane 4b + j holds B_b[j], and lane 4b + j, register
// v, accumulates D_b[v][j] = RN32(D_b[v][j] + A_b[v] B_b[j]) (exact product,
// one rounding per step).  An item is 4 rows (phase 1) or 4 columns (phase 2) of
// one 32 x 32 tile and runs 32 steps; a wave runs 16 items at once, one per block.
// B carries each right-hand side as two FP32 columns, hi = RN32(x) and
// lo = RN32(x - hi), so x enters with 48 bits; every FL steps the FP32
// accumulators are added into FP64 ones, which bounds the FP32 part of the sum to
// FL products.  Partials of a row (column) over its tiles, hi + lo first, are
// summed in tile order as in tile_solve.
template <int BS>
constexpr int mfma_vs() { return 32 * BS + 4; }  // floats per LDS vector variant (16-byte multiple)

// tile (b, s) of tile_index t
__device__ __forceinline__ void tile_of(int t, int &b, int &s)
{
    b = t >= 3 ? (t >= 6 ? 3 : 2) : (t >= 1 ? 1 : 0);
    s = t - b * (b + 1) / 2;
}

// the K bank vectors of src as hi/lo FP32 pairs: variant 2m + h of column m
template <int BS, int K>
__device__ __forceinline__ void split_bank_vector(const double *__restrict__ src, int64_t ld, int n,
                                                  float *__restrict__ xv)
{
    constexpr int D = 32 * BS, VS = mfma_vs<BS>();
    for (int e = threadIdx.x; e < K * D; e += blockDim.x) {
        const int k = e / D, i = e - k * D;
        const double v = i < n ? src[k * ld + i] : 0.0;
        const float hi = float(v);
        xv[(2 * k) * VS + i] = hi;
        xv[(2 * k + 1) * VS + i] = float(v - double(hi));
    }
}

// one 32-step item on 16 blocks: A from a (strided by astep floats per step),
// B from the variant vectors at bv[m] (contiguous); acc[m][v] gets the sum
template <int K, int FL>
__device__ __forceinline__ void mfma_item(const float *__restrict__ a, int astep, const float *const *bv,
                                          double (&acc)[K == 4 ? 2 : 1][4])
{
    constexpr int NM = K == 4 ? 2 : 1;
    typedef float F4 __attribute__((ext_vector_type(4)));
    F4 c[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        c[m] = F4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[m][v] = 0.0;
    }
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        float av[4];
        if (astep == 1) {
            const F4 t = *reinterpret_cast<const F4 *>(a + 4 * g);
            av[0] = t.x; av[1] = t.y; av[2] = t.z; av[3] = t.w;
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) av[u] = a[(4 * g + u) * astep];
        }
        F4 b[NM];
#pragma unroll
        for (int m = 0; m < NM; ++m) b[m] = *reinterpret_cast<const F4 *>(bv[m] + 4 * g);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int m = 0; m < NM; ++m) c[m] = __builtin_amdgcn_mfma_f32_4x4x1f32(av[u], b[m][u], c[m], 0, 0, 0);
        if ((4 * g + 4) % FL == 0) {
#pragma unroll
            for (int m = 0; m < NM; ++m) {
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[m][v] += double(c[m][v]);
                c[m] = F4{0.0f, 0.0f, 0.0f, 0.0f};
            }
        }
    }
}

// G^T (G x) for K columns: xv holds x (split_bank_vector), tv receives t = G x the
// same way; on return (after the final barrier) y_k(j) = bank_result<BS>(Q, k, j).
// P and Q may be one buffer (K * BS * D doubles).
template <int BS, int K, int FL>
__device__ __forceinline__ void mfma_solve(const float *__restrict__ Gt, const float *__restrict__ xv,
                                           float *__restrict__ tv, double *P, double *Q)
{
    constexpr int D = 32 * BS, VS = mfma_vs<BS>();
    constexpr int NITEM = 4 * BS * (BS + 1);  // 8 row (column) groups per tile
    constexpr int NM = K == 4 ? 2 : 1;
    constexpr int NV = 2 * K;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, blk = lane >> 2, q = lane & 3;
    // variant of B column q in MFMA m (K = 1: columns 2, 3 repeat hi, lo; never stored)
    const int var0 = q % (NV < 4 ? NV : 4);
    const bool owner = q < NV;  // lane holds a real variant
    const int item_raw = w * 16 + blk;
    const bool live = item_raw < NITEM;
    const int item = live ? item_raw : item_raw % NITEM;  // dead blocks of a live wave repeat live items
    const int t = item >> 3, grp = item & 7;
    int tb, ts;
    tile_of(t, tb, ts);
    const float *tile = Gt + t * kTileStride;
    double acc[NM][4];
    if (w * 16 < NITEM) {  // phase 1: rows 32 tb + 4 grp + v of G times x over tile column ts
        const float *bv[NM];
#pragma unroll
        for (int m = 0; m < NM; ++m) bv[m] = xv + (var0 + 4 * m) * VS + 32 * ts;
        mfma_item<K, FL>(tile + (4 * grp + q) * kTileLd, 1, bv, acc);
#pragma unroll
        for (int m = 0; m < NM; ++m)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[m][v] += __shfl_xor(acc[m][v], 1);  // hi + lo
        if (live && owner && (q & 1) == 0)
#pragma unroll
            for (int m = 0; m < NM; ++m)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    P[((var0 / 2 + 2 * m) * BS + ts) * D + 32 * tb + 4 * grp + v] = acc[m][v];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < K * D; e += blockDim.x) {
        const int k = e / D, i = e - k * D;
        double v = P[(k * BS) * D + i];
        for (int s = 1; s <= i / 32; ++s) v += P[(k * BS + s) * D + i];
        const float hi = float(v);
        tv[(2 * k) * VS + i] = hi;
        tv[(2 * k + 1) * VS + i] = float(v - double(hi));
    }
    __syncthreads();
    if (w * 16 < NITEM) {  // phase 2: columns 32 ts + 4 grp + v of G^T t over tile row tb
        const float *bv[NM];
#pragma unroll
        for (int m = 0; m < NM; ++m) bv[m] = tv + (var0 + 4 * m) * VS + 32 * tb;
        mfma_item<K, FL>(tile + 4 * grp + q, kTileLd, bv, acc);
#pragma unroll
        for (int m = 0; m < NM; ++m)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[m][v] += __shfl_xor(acc[m][v], 1);
        if (live && owner && (q & 1) == 0)
#pragma unroll
            for (int m = 0; m < NM; ++m)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    Q[((var0 / 2 + 2 * m) * BS + tb) * D + 32 * ts + 4 * grp + v] = acc[m][v];
    }
    __syncthreads();
}

// k_level0_apply_tiled with the MFMA bank solve (FP32 factors, tiled layout)
template <int BS, bool DOT, int K, int FL>
__global__ __launch_bounds__(kApplyThreads, (kTiledWaves<BS, DOT, K>)) void k_level0_apply_mfma(int32_t nbanks,
                                                                     const int32_t *__restrict__ part_off,
                                                                     const float *__restrict__ inv,
                                                                     const double *__restrict__ r, int64_t ld,
                                                                     const int32_t *__restrict__ parent,
                                                                     const double *__restrict__ w1, int64_t ld1,
                                                                     double scale, const double *__restrict__ add,
                                                                     double *__restrict__ z,
                                                                     const double *__restrict__ dotv,
                                                                     double *partials, const int *__restrict__ stop)
{
    constexpr int D = 32 * BS, VS = mfma_vs<BS>();
    constexpr int ST = tiled_floats(BS);
    constexpr int NTILE = BS * (BS + 1) / 2;
    constexpr int NB = tiled_banks_per_group(K);
    using Regs = MatrixRegs<float, ST, kApplyThreads>;
    constexpr int NV = Regs::NV, PER = Regs::PER;
    static_assert(ST % 4 == 0 && VS % 4 == 0 && BS <= 3, "whole 16-byte vectors, tile_of covers 6 tiles");
    static_assert(32 % FL == 0 && FL % 4 == 0, "flush period divides the 32 steps of an item");
    if (dev::all_stopped(stop, K)) return;
    __shared__ __attribute__((aligned(16))) float Gt[NTILE * kTileStride];
    __shared__ __attribute__((aligned(16))) float xv[2 * K * VS], tv[2 * K * VS];
    __shared__ double PQ[K * BS * D];
    const int i = threadIdx.x;
    const int32_t first = int32_t(blockIdx.x) * NB;
    if (first >= nbanks) return;
    const int32_t last = min(nbanks, first + NB);
    Regs mr;
    mr.issue(inv + int64_t(first) * ST);
    typedef float F4 __attribute__((ext_vector_type(4)));
    for (int e = i; e < NTILE * kTileStride / 4; e += kApplyThreads)
        reinterpret_cast<F4 *>(Gt)[e] = F4{0.0f, 0.0f, 0.0f, 0.0f};
    __syncthreads();  // the zeros land before any factor entry
    const uint16_t *stage = tiled_layout<BS>().stage;
    char *Gb = reinterpret_cast<char *>(Gt);
    for (int32_t bank = first; bank < last; ++bank) {
        const int32_t lo = part_off[bank];
        const int n = (part_off[bank + 1] - lo) * BS;
        const int64_t gi = int64_t(lo) * BS + i;
        int off[PER];
#pragma unroll
        for (int qq = 0; qq < PER; ++qq) {
            const int t = i + qq * kApplyThreads;
            off[qq] = t < NV ? stage[t] : 0;
        }
        double cw[K], ca[K], cd[K];
        if (i < n) {
            const int64_t pi = w1 ? int64_t(parent[lo + i / BS]) * BS + (i % BS) : 0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                cw[k] = w1 ? w1[k * ld1 + pi] : 0.0;
                ca[k] = add ? add[k * ld + gi] : 0.0;
                if constexpr (DOT) cd[k] = dotv[k * ld + gi];
            }
        }
        split_bank_vector<BS, K>(r + int64_t(lo) * BS, ld, n, xv);
#pragma unroll
        for (int qq = 0; qq < PER; ++qq)
            if (i + qq * kApplyThreads < NV) *reinterpret_cast<typename Regs::V *>(Gb + off[qq]) = mr.reg[qq];
        __syncthreads();
        if (bank + 1 < last) mr.issue(inv + int64_t(bank + 1) * ST);
        mfma_solve<BS, K, FL>(Gt, xv, tv, PQ, PQ);
        double part[K];
#pragma unroll
        for (int k = 0; k < K; ++k) part[k] = 0.0;
        if (i < n) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                double y = bank_result<BS>(PQ, k, i);
                if (w1) y += cw[k];
                y *= scale;
                if (add) y += ca[k];
                z[k * ld + gi] = y;
                if constexpr (DOT) part[k] = cd[k] * y;
            }
        }
        if constexpr (DOT) {
            __shared__ double sh[kApplyThreads / 64];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const double tot = dev::block_sum<kApplyThreads>(part[k], sh);
                if (threadIdx.x == 0) partials[int64_t(k) * nbanks + bank] = tot;
            }
        }
        __syncthreads();  // Gt, xv and Q are rewritten for the next bank
    }
}

// z += scale * w1[parent] (the level-1 correction prolongated to level 0) with an
// optional per-workgroup partial of dotv . z and an optional FP32 copy z32 of the
// result (same stride); grid row y = column y.
template <bool DOT>
__global__ __launch_bounds__(256) void k_prolong_add(int64_t np, int bs, const int32_t *__restrict__ parent,
                                                     const double *__restrict__ w1, int64_t ld1, double scale,
                                                     double *__restrict__ z, int64_t ld,
                                                     const double *__restrict__ dotv, double *partials,
                                                     const int *__restrict__ stop, float *__restrict__ z32)
{
    if (dev::all_stopped(stop, gridDim.y)) return;
    const int64_t o = blockIdx.y * ld;
    w1 += blockIdx.y * ld1;
    z += o;
    if (z32) z32 += o;
    double part = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < np; i += int64_t(gridDim.x) * 256) {
        const int64_t node = i / bs;
        const double v = z[i] + scale * w1[int64_t(parent[node]) * bs + (i - node * bs)];
        z[i] = v;
        if (z32) z32[i] = float(v);
        if constexpr (DOT) part += dotv[o + i] * v;
    }
    if constexpr (DOT) {
        __shared__ double sh[4];
        dev::store_partial(dev::block_sum<256>(part, sh), partials + int64_t(blockIdx.y) * gridDim.x);
    }
}

template <class TS, bool TILED = false>
void launch_bank_invert(int bs, int32_t nbanks, const int32_t *bank_off, const int32_t *rowptr, const int32_t *cols,
                        const double *vals, TS *out, int *status, hipStream_t s)
{
    if (nbanks == 0) return;
    const int D = 32 * bs;
    const size_t lds = size_t(D * (D + 1) / 2 + D + 2) * sizeof(double);
    switch (bs) {
        case 1:
            k_bank_factor<1, TS, TILED><<<nbanks, kInvThreads, lds, s>>>(nbanks, bank_off, rowptr, cols, vals, out,
                                                                         status);
            break;
        case 2:
            k_bank_factor<2, TS, TILED><<<nbanks, kInvThreads, lds, s>>>(nbanks, bank_off, rowptr, cols, vals, out,
                                                                         status);
            break;
        case 3:
            k_bank_factor<3, TS, TILED><<<nbanks, kInvThreads, lds, s>>>(nbanks, bank_off, rowptr, cols, vals, out,
                                                                         status);
            break;
        default: fail(LSS_ERR_UNSUPPORTED, "block_size must be 1, 2 or 3");
    }
    LSS_LAUNCH_CHECK();
}

}  // namespace

void mas_set_lds_limits()
{
    static bool done = false;
    if (done) return;
    const size_t lds3 = size_t(96 * 97 / 2 + 96 + 2) * sizeof(double);
    LSS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bank_factor<3, double, false>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds3)));
    LSS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bank_factor<3, float, false>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds3)));
    LSS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bank_factor<3, float, true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds3)));
    LSS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bank_factor<3, __hip_bfloat16, false>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds3)));
    LSS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bank_factor<3, _Float16, false>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds3)));
    done = true;
}

int64_t mas_bank_entries(Precision prec, int bs, bool level0)
{
    if (level0 && prec == Precision::FP32) return tiled_floats(bs);
    const int64_t D = 32 * bs;
    return D * (D + 1) / 2;
}

void mas_bank_invert(Precision prec, int bs, int32_t nbanks, const int32_t *bank_off, const int32_t *rowptr,
                     const int32_t *cols, const double *vals, void *out, int *status, hipStream_t s, bool level0)
{
    mas_set_lds_limits();
    switch (prec) {
        case Precision::FP64:
            launch_bank_invert<double>(bs, nbanks, bank_off, rowptr, cols, vals, static_cast<double *>(out), status, s);
            break;
        case Precision::FP32:
            if (level0)
                launch_bank_invert<float, true>(bs, nbanks, bank_off, rowptr, cols, vals, static_cast<float *>(out),
                                                status, s);
            else
                launch_bank_invert<float>(bs, nbanks, bank_off, rowptr, cols, vals, static_cast<float *>(out), status,
                                          s);
            break;
        case Precision::BF16:
            launch_bank_invert<__hip_bfloat16>(bs, nbanks, bank_off, rowptr, cols, vals,
                                               static_cast<__hip_bfloat16 *>(out), status, s);
            break;
        case Precision::FP16:
            launch_bank_invert<_Float16>(bs, nbanks, bank_off, rowptr, cols, vals, static_cast<_Float16 *>(out),
                                         status, s);
            break;
    }
}

void mas_galerkin(int64_t nslots, int bs, const int64_t *src_ptr, const int32_t *src_blk, const double *fine_vals,
                  double *coarse_vals, hipStream_t s)
{
    const int64_t work = nslots * bs * bs;
    if (work == 0) return;
    k_galerkin<<<div_up(work, 256), 256, 0, s>>>(nslots, bs * bs, src_ptr, src_blk, fine_vals, coarse_vals);
    LSS_LAUNCH_CHECK();
}

void mas_restrict(int bs, int32_t fine_nbanks, const int32_t *fine_bank_off, const int32_t *fine_parent,
                  const double *fine, double *coarse, const int *stop, hipStream_t s, int k, int64_t ld_fine,
                  int64_t ld_coarse)
{
    if (fine_nbanks == 0) return;
    const int grid = div_up(fine_nbanks, 8);
#define LSS_RS(BS, K)                                                                                         \
    k_restrict<BS, K><<<grid, 256, 0, s>>>(fine_nbanks, fine_bank_off, fine_parent, fine, ld_fine, coarse,    \
                                           ld_coarse, stop)
#define LSS_RS_K(BS)                                                                                          \
    switch (k) {                                                                                              \
        case 1: LSS_RS(BS, 1); break;                                                                         \
        case 2: LSS_RS(BS, 2); break;                                                                         \
        case 4: LSS_RS(BS, 4); break;                                                                         \
        default: fail(LSS_ERR_ARG, "a fused MAS apply takes 1, 2 or 4 right-hand sides");                     \
    }
    switch (bs) {
        case 1: LSS_RS_K(1); break;
        case 2: LSS_RS_K(2); break;
        case 3: LSS_RS_K(3); break;
        default: fail(LSS_ERR_UNSUPPORTED, "block_size must be 1, 2 or 3");
    }
#undef LSS_RS_K
#undef LSS_RS
    LSS_LAUNCH_CHECK();
}

namespace {

template <int K>
void coarse_apply_k(Precision prec, int bs, int32_t nodes, const void *inv, const double *r, int64_t ld,
                    const int32_t *parent, const double *w_parent, int64_t ld_parent, double *w, const int *stop,
                    hipStream_t s)
{
    const int32_t nbanks = div_up(nodes, 32);
#define LSS_CA(BS, TS)                                                                                         \
    k_coarse_apply<BS, TS, K><<<nbanks, kApplyThreads, 0, s>>>(nodes, static_cast<const TS *>(inv), r, ld,      \
                                                               parent, w_parent, ld_parent, w, stop)
#define LSS_CA_BS(TS)                                                                                          \
    switch (bs) {                                                                                              \
        case 1: LSS_CA(1, TS); break;                                                                          \
        case 2: LSS_CA(2, TS); break;                                                                          \
        case 3: LSS_CA(3, TS); break;                                                                          \
        default: fail(LSS_ERR_UNSUPPORTED, "block_size must be 1, 2 or 3");                                    \
    }
    switch (prec) {
        case Precision::FP64: LSS_CA_BS(double); break;
        case Precision::FP32: LSS_CA_BS(float); break;
        case Precision::BF16: LSS_CA_BS(__hip_bfloat16); break;
        case Precision::FP16: LSS_CA_BS(_Float16); break;
    }
#undef LSS_CA_BS
#undef LSS_CA
    LSS_LAUNCH_CHECK();
}

template <int K>
void level0_apply_k(Precision prec, int bs, int32_t nbanks, const int32_t *part_off, const void *inv,
                    const double *r, int64_t ld, const int32_t *parent, const double *w1, int64_t ld1, double scale,
                    const double *add, double *z, const double *dotv, bool dot, double *partials, const int *stop,
                    hipStream_t s, int mfma_flush)
{
#define LSS_L0_MFMA(BS, DOT, FL)                                                                               \
    k_level0_apply_mfma<BS, DOT, K, FL><<<div_up(nbanks, tiled_banks_per_group(K)), kApplyThreads, 0, s>>>(    \
        nbanks, part_off, static_cast<const float *>(inv), r, ld, parent, w1, ld1, scale, add, z, dotv, partials, \
        stop)
#define LSS_L0(BS, TS, DOT)                                                                                    \
    if constexpr (std::is_same_v<TS, float>) {                                                                 \
        switch (mfma_flush) {                                                                                  \
            case 0:                                                                                            \
                k_level0_apply_tiled<BS, DOT, K><<<div_up(nbanks, tiled_banks_per_group(K)), kApplyThreads, 0, \
                                                   s>>>(nbanks, part_off, static_cast<const float *>(inv), r, ld, \
                                                        parent, w1, ld1, scale, add, z, dotv, partials, stop); \
                break;                                                                                         \
            case 4: LSS_L0_MFMA(BS, DOT, 4); break;                                                            \
            case 8: LSS_L0_MFMA(BS, DOT, 8); break;                                                            \
            case 16: LSS_L0_MFMA(BS, DOT, 16); break;                                                          \
            case 32: LSS_L0_MFMA(BS, DOT, 32); break;                                                          \
            default: fail(LSS_ERR_ARG, "mfma_flush must be 0, 4, 8, 16 or 32");                                \
        }                                                                                                      \
    } else if constexpr (K == 1)                                                                                 \
        k_level0_apply2<BS, TS, DOT, K><<<div_up(nbanks, 2), kApplyThreads, 0, s>>>(                           \
            nbanks, part_off, static_cast<const TS *>(inv), r, ld, parent, w1, ld1, scale, add, z, dotv,        \
            partials, stop);                                                                                   \
    else                                                                                                       \
        k_level0_apply<BS, TS, DOT, K><<<nbanks, kApplyThreads, 0, s>>>(nbanks, part_off,                      \
                                                                        static_cast<const TS *>(inv), r, ld,   \
                                                                        parent, w1, ld1, scale, add, z, dotv,  \
                                                                        partials, stop)
#define LSS_L0_BS(TS)                                                                                          \
    switch (bs) {                                                                                              \
        case 1: if (dot) LSS_L0(1, TS, true); else LSS_L0(1, TS, false); break;                                \
        case 2: if (dot) LSS_L0(2, TS, true); else LSS_L0(2, TS, false); break;                                \
        case 3: if (dot) LSS_L0(3, TS, true); else LSS_L0(3, TS, false); break;                                \
        default: fail(LSS_ERR_UNSUPPORTED, "block_size must be 1, 2 or 3");                                    \
    }
    switch (prec) {
        case Precision::FP64: LSS_L0_BS(double); break;
        case Precision::FP32: LSS_L0_BS(float); break;
        case Precision::BF16: LSS_L0_BS(__hip_bfloat16); break;
        case Precision::FP16: LSS_L0_BS(_Float16); break;
    }
#undef LSS_L0_BS
#undef LSS_L0
#undef LSS_L0_MFMA
    LSS_LAUNCH_CHECK();
}

}  // namespace

int mas_prolong_add(int64_t np, int bs, const int32_t *parent, const double *w1, double scale, double *z,
                    const double *dotv, double *partials, const int *stop, hipStream_t s, int k, int64_t ld,
                    int64_t ld1, float *z32)
{
    if (np == 0) return 0;
    const int grid = int(std::min<int64_t>(std::max<int64_t>(div_up(np, 256 * 4), 1), 4096));
    if (dotv)
        k_prolong_add<true><<<dim3(grid, k), 256, 0, s>>>(np, bs, parent, w1, ld1, scale, z, ld, dotv, partials,
                                                          stop, z32);
    else
        k_prolong_add<false><<<dim3(grid, k), 256, 0, s>>>(np, bs, parent, w1, ld1, scale, z, ld, nullptr,
                                                           nullptr, stop, z32);
    LSS_LAUNCH_CHECK();
    return grid;
}

void mas_coarse_apply(Precision prec, int bs, int32_t nodes, const void *inv, const double *r, const int32_t *parent,
                      const double *w_parent, double *w, const int *stop, hipStream_t s, int k, int64_t ld,
                      int64_t ld_parent)
{
    if (nodes == 0) return;
    switch (k) {
        case 1: coarse_apply_k<1>(prec, bs, nodes, inv, r, ld, parent, w_parent, ld_parent, w, stop, s); break;
        case 2: coarse_apply_k<2>(prec, bs, nodes, inv, r, ld, parent, w_parent, ld_parent, w, stop, s); break;
        case 4: coarse_apply_k<4>(prec, bs, nodes, inv, r, ld, parent, w_parent, ld_parent, w, stop, s); break;
        default: fail(LSS_ERR_ARG, "a fused MAS apply takes 1, 2 or 4 right-hand sides");
    }
}

void mas_level0_apply(Precision prec, int bs, int32_t nbanks, const int32_t *part_off, const void *inv,
                      const double *r, const int32_t *parent, const double *w1, double scale, const double *add,
                      double *z, const double *dotv, dev::PcgState *st, double *partials, const int *stop,
                      hipStream_t s, int k, int64_t ld, int64_t ld1, int mfma_flush)
{
    if (nbanks == 0) return;
    const bool dot = st != nullptr;
    switch (k) {
        case 1:
            level0_apply_k<1>(prec, bs, nbanks, part_off, inv, r, ld, parent, w1, ld1, scale, add, z, dotv, dot,
                              partials, stop, s, mfma_flush);
            break;
        case 2:
            level0_apply_k<2>(prec, bs, nbanks, part_off, inv, r, ld, parent, w1, ld1, scale, add, z, dotv, dot,
                              partials, stop, s, mfma_flush);
            break;
        case 4:
            level0_apply_k<4>(prec, bs, nbanks, part_off, inv, r, ld, parent, w1, ld1, scale, add, z, dotv, dot,
                              partials, stop, s, mfma_flush);
            break;
        default: fail(LSS_ERR_ARG, "a fused MAS apply takes 1, 2 or 4 right-hand sides");
    }
}

}  // namespace lss::k
