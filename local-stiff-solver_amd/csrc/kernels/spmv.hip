// Block sparse formats and SpMV for gfx950.
//
// Matrix layout in HBM (sliced block-ELL, "SELL-H/bs"): block rows are grouped
// into slices of H = 64/bs rows, one wave64 per slice.  Lane l of the
// wave owns column r = l % bs of the blocks of block row h = l / bs, so a wave
// covers 64 (bs = 1), 64 (bs = 2) or 63 (bs = 3) scalar columns.  Slot j of slice
// s holds block j of each of the H rows; its H*bs*bs values are stored
// lane-major: entry (c, r) of the block of lane l = h*bs + r sits at c*H*bs + l,
// so each of the bs value loads of a wave reads one contiguous span of 8*H*bs
// bytes (504 B for bs = 3), and a lane gathers one x entry per slot.  The bs
// lanes of a block row add their partial rows once per slice (lane shuffles).
// Rows shorter than the slice width are padded with column -1 and zero values.
//
// Reference counterparts: BSR layout MAS/BSRMatrix.hpp:10-57 (row-major blocks,
// tail block diagonal padded with 1, BSRMatrix.cu:123-134, 189-191); SpMV
// cuSPARSE BSR (MASSolver.cu:356-375) / hypre ParCSR matvec (GPUHybrid.cu:454-466).
#include "kernels.hpp"
#include "../common/runtime.hpp"

namespace lss::k {
namespace {

constexpr int kSpmvThreads = 256;  // 4 waves = 4 slices per workgroup

// CSR -> BSR, one wave per block row.  The bs scalar rows of block row I are
// the contiguous entry range [ptr[I*bs], ptr[I*bs + rows]).  An entry starts
// a block column of the row's union when it is the first of its block column
// in its scalar row and no earlier scalar row of the block row holds that
// block column (binary search; rows are ascending).  The union's block
// columns are collected in LDS and ranked by counting; a block row with more
// than kUnionCap distinct block columns is done by lane 0 alone.
constexpr int kUnionCap = 256;
// A block row of at most kStageCap entries is first copied to LDS, so that the
// union tests search LDS instead of global memory.
constexpr int kStageCap = 512;

struct RowRanges {
    int32_t p[3], e[3];
    int rows;
};

__device__ __forceinline__ RowRanges block_rows(int32_t n, int bs, const int32_t *__restrict__ ptr, int32_t I)
{
    RowRanges R;
    R.rows = min(bs, n - I * bs);
    for (int r = 0; r < 3; ++r) {
        R.p[r] = r < R.rows ? ptr[I * bs + r] : 0;
        R.e[r] = r < R.rows ? ptr[I * bs + r + 1] : 0;
    }
    return R;
}

// does scalar row [p, e) (ascending) hold an entry of block column c?
__device__ __forceinline__ bool row_has_block(const int32_t *__restrict__ idx, int32_t p, int32_t e, int bs, int32_t c)
{
    int32_t lo = p, hi = e;  // first entry >= c * bs
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (idx[mid] < c * bs) lo = mid + 1;
        else hi = mid;
    }
    return lo < e && idx[lo] < (c + 1) * bs;
}

// union-first test of entry k (row r of the block row)
__device__ __forceinline__ bool union_first(const int32_t *__restrict__ idx, const RowRanges &R, int r, int32_t k,
                                            int bs)
{
    const int32_t c = idx[k] / bs;
    if (k > R.p[r] && idx[k - 1] / bs == c) return false;
    for (int q = 0; q < r; ++q)
        if (row_has_block(idx, R.p[q], R.e[q], bs, c)) return false;
    return true;
}

__device__ __forceinline__ int row_of(const RowRanges &R, int32_t k)
{
    return k < R.e[0] ? 0 : (k < R.e[1] ? 1 : 2);
}

__global__ __launch_bounds__(256) void k_bsr_count(int32_t n, int bs, const int32_t *__restrict__ ptr,
                                                    const int32_t *__restrict__ idx, int32_t *__restrict__ cnt,
                                                    int32_t *__restrict__ unsorted)
{
    const int lane = threadIdx.x & 63;
    const int32_t nb = (n + bs - 1) / bs;
    const int32_t I = int32_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (I >= nb) return;
    __shared__ int32_t stage[4][kStageCap];
    RowRanges R = block_rows(n, bs, ptr, I);
    const int32_t lo = R.p[0], hi = R.e[R.rows - 1];
    int32_t c = 0, bad = 0, diag = 0;
    if (hi - lo <= kStageCap) {
        int32_t *ix = stage[threadIdx.x >> 6];
        for (int32_t k = lane; k < hi - lo; k += 64) ix[k] = idx[lo + k];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int r = 0; r < R.rows; ++r) {
            R.p[r] -= lo;
            R.e[r] -= lo;
        }
        for (int32_t k = lane; k < hi - lo; k += 64) {
            const int r = row_of(R, k);
            if (k > R.p[r] && ix[k] <= ix[k - 1]) bad = 1;
            const bool first = union_first(ix, R, r, k, bs);
            c += first;
            diag |= first && ix[k] / bs == I;
        }
    } else {
        for (int32_t k = lo + lane; k < hi; k += 64) {
            const int r = row_of(R, k);
            if (k > R.p[r] && idx[k] <= idx[k - 1]) bad = 1;
            const bool first = union_first(idx, R, r, k, bs);
            c += first;
            diag |= first && idx[k] / bs == I;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        c += __shfl_xor(c, o, 64);
        bad |= __shfl_xor(bad, o, 64);
        diag |= __shfl_xor(diag, o, 64);
    }
    if (lane == 0) {
        if (bad) *unsorted = 1;
        cnt[I] = c + ((R.rows < bs && !diag) ? 1 : 0);  // padded tail diagonal block
    }
}

__global__ __launch_bounds__(256) void k_bsr_fill(int32_t n, int bs, const int32_t *__restrict__ ptr,
                                                   const int32_t *__restrict__ idx, const double *__restrict__ vals,
                                                   const int32_t *__restrict__ rowptr, int32_t *__restrict__ cols,
                                                   double *__restrict__ bvals)
{
    __shared__ int32_t ucol[4][kUnionCap];
    __shared__ int32_t stage[4][kStageCap];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int32_t nb = (n + bs - 1) / bs;
    const int32_t I = int32_t(blockIdx.x) * 4 + wv;
    if (I >= nb) return;
    const int bb = bs * bs;
    const RowRanges R = block_rows(n, bs, ptr, I);
    const int32_t lo = R.p[0], hi = R.e[R.rows - 1];
    const int32_t base = rowptr[I];
    const int32_t m = rowptr[I + 1] - base;  // distinct block columns (with a padded tail diagonal)
    int32_t *u = ucol[wv];
    if (m <= kUnionCap && hi - lo <= kStageCap) {
        // the row's entries in LDS, positions relative to lo
        int32_t *ix = stage[wv];
        for (int32_t k = lane; k < hi - lo; k += 64) ix[k] = idx[lo + k];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        RowRanges L = R;
        for (int r = 0; r < R.rows; ++r) {
            L.p[r] -= lo;
            L.e[r] -= lo;
        }
        int32_t at = 0;
        for (int32_t k0 = 0; k0 < hi - lo; k0 += 64) {
            const int32_t k = k0 + lane;
            const bool first = k < hi - lo && union_first(ix, L, row_of(L, k), k, bs);
            const unsigned long long mask = __ballot(first);
            if (first) u[at + __popcll(mask & ((1ull << lane) - 1ull))] = ix[k] / bs;
            at += __popcll(mask);
        }
        if (at < m && lane == 0) u[at] = I;  // padded tail diagonal block
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int32_t q = lane; q < m; q += 64) {
            const int32_t c = u[q];
            int32_t rank = 0;
            for (int32_t t = 0; t < m; ++t) rank += u[t] < c;
            cols[base + rank] = c;
            if (c == I && R.rows < bs)
                for (int r = R.rows; r < bs; ++r) bvals[int64_t(base + rank) * bb + r * bs + r] = 1.0;
        }
        for (int32_t k = lane; k < hi - lo; k += 64) {
            const int r = row_of(L, k);
            const int32_t c = ix[k] / bs;
            int32_t rank = 0;
            for (int32_t t = 0; t < m; ++t) rank += u[t] < c;
            bvals[int64_t(base + rank) * bb + r * bs + (ix[k] - c * bs)] = vals[lo + k];
        }
        return;
    }
    if (m <= kUnionCap) {
        // collect the union's block columns in LDS (any order)
        int32_t at = 0;
        for (int32_t k0 = lo; k0 < hi; k0 += 64) {
            const int32_t k = k0 + lane;
            const bool first = k < hi && union_first(idx, R, row_of(R, k), k, bs);
            const unsigned long long mask = __ballot(first);
            if (first) u[at + __popcll(mask & ((1ull << lane) - 1ull))] = idx[k] / bs;
            at += __popcll(mask);
        }
        if (at < m && lane == 0) u[at] = I;  // padded tail diagonal block
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // slot of block column c = number of union columns below c
        for (int32_t q = lane; q < m; q += 64) {
            const int32_t c = u[q];
            int32_t rank = 0;
            for (int32_t t = 0; t < m; ++t) rank += u[t] < c;
            cols[base + rank] = c;
            if (c == I && R.rows < bs)
                for (int r = R.rows; r < bs; ++r) bvals[int64_t(base + rank) * bb + r * bs + r] = 1.0;
        }
        for (int32_t k = lo + lane; k < hi; k += 64) {
            const int r = row_of(R, k);
            const int32_t c = idx[k] / bs;
            int32_t rank = 0;
            for (int32_t t = 0; t < m; ++t) rank += u[t] < c;
            bvals[int64_t(base + rank) * bb + r * bs + (idx[k] - c * bs)] = vals[k];
        }
        return;
    }
    if (lane != 0) return;
    // wide block row: sequential merge of the scalar rows
    int32_t p[3] = {R.p[0], R.p[1], R.p[2]};
    int32_t w = base;
    int32_t diag_slot = -1;
    const bool tail_pending = R.rows < bs;
    while (true) {
        int32_t cur = 0x7fffffff;
        for (int r = 0; r < R.rows; ++r)
            if (p[r] < R.e[r]) cur = min(cur, idx[p[r]] / bs);
        if (tail_pending && I < cur && diag_slot < 0) {  // insert missing diagonal in order
            cols[w] = I;
            diag_slot = w++;
            continue;
        }
        if (cur == 0x7fffffff) break;
        if (cur == I) diag_slot = w;
        cols[w] = cur;
        for (int r = 0; r < R.rows; ++r)
            while (p[r] < R.e[r] && idx[p[r]] / bs == cur) {
                bvals[int64_t(w) * bb + r * bs + (idx[p[r]] - cur * bs)] = vals[p[r]];
                ++p[r];
            }
        ++w;
    }
    if (tail_pending && diag_slot >= 0)
        for (int r = R.rows; r < bs; ++r) bvals[int64_t(diag_slot) * bb + r * bs + r] = 1.0;
}

// One wave per block row: the permuted columns of the row are ranked in LDS,
// then the row's blocks move element by element, so that the wave reads the
// source row and writes the destination row contiguously.  Rows longer than
// kPermCap blocks rank and move one block per lane.
constexpr int kPermCap = 256;

__global__ __launch_bounds__(256) void k_bsr_permute(int32_t nb, int bs, const int32_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ cols,
                                                      const double *__restrict__ vals,
                                                      const int32_t *__restrict__ perm,
                                                      const int32_t *__restrict__ new_rowptr,
                                                      int32_t *__restrict__ new_cols, double *__restrict__ new_vals)
{
    __shared__ int32_t pc[4][kPermCap], rk[4][kPermCap];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int32_t I = int32_t(blockIdx.x) * 4 + wv;
    if (I >= nb) return;
    const int bb = bs * bs;
    const int32_t b = rowptr[I], e = rowptr[I + 1], m = e - b;
    const int32_t dst = new_rowptr[perm[I]];
    if (m <= kPermCap) {
        for (int32_t k = lane; k < m; k += 64) pc[wv][k] = perm[cols[b + k]];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int32_t k = lane; k < m; k += 64) {
            const int32_t ck = pc[wv][k];
            int32_t rank = 0;
            for (int32_t q = 0; q < m; ++q) rank += pc[wv][q] < ck;
            rk[wv][k] = rank;
            new_cols[dst + rank] = ck;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const double *src = vals + int64_t(b) * bb;
        for (int32_t q = lane; q < m * bb; q += 64) {
            const int32_t k = q / bb;
            new_vals[int64_t(dst + rk[wv][k]) * bb + (q - k * bb)] = src[q];
        }
        return;
    }
    for (int32_t k = b + lane; k < e; k += 64) {
        const int32_t ck = perm[cols[k]];
        int32_t rank = 0;
        for (int32_t q = b; q < e; ++q) rank += perm[cols[q]] < ck;
        new_cols[dst + rank] = ck;
        const double *src = vals + int64_t(k) * bb;
        double *out = new_vals + int64_t(dst + rank) * bb;
        for (int t = 0; t < bb; ++t) out[t] = src[t];
    }
}

__global__ void k_sell_width(int32_t nb, int H, const int32_t *__restrict__ rowptr, int32_t *__restrict__ width)
{
    const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t ns = (nb + H - 1) / H;
    if (s >= ns) return;
    int32_t w = 0;
    for (int32_t r = s * H; r < min(nb, (s + 1) * H); ++r) w = max(w, rowptr[r + 1] - rowptr[r]);
    width[s] = w;
}

// bfloat16 bits of an FP32 value, rounded to nearest even (values here are finite)
__device__ __forceinline__ uint16_t to_bf16_bits(float f)
{
    const uint32_t u = __float_as_uint(f);
    return uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float from_bf16_bits(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }

// One wave per slice, lane l = h * bs + rr (column rr of the blocks of the slice's
// block row h): every slot of the slice is written, padding included (column -1,
// zero values), and a slot's values are written by consecutive lanes.  Position
// (c, l) of a slot holds entry (c, rr) of the block, so a lane holds one block
// column and the SpMV gathers one x entry per lane and slot.
__global__ __launch_bounds__(256) void k_sell_fill(int32_t nb, int bs, int H, int32_t nslices,
                                                    const int32_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ cols, const double *__restrict__ vals,
                                                    const int64_t *__restrict__ slice_ptr, int32_t *__restrict__ scols,
                                                    double *__restrict__ svals, float *__restrict__ svals32,
                                                    uint16_t *__restrict__ svals16)
{
    const int lane = threadIdx.x & 63;
    const int32_t sl = int32_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (sl >= nslices || lane >= H * bs) return;
    const int bb = bs * bs;
    const int h = lane / bs, rr = lane - h * bs;
    const int32_t row = sl * H + h;
    const int32_t b = row < nb ? rowptr[row] : 0, len = row < nb ? rowptr[row + 1] - b : 0;
    const int64_t base = slice_ptr[sl];
    const int32_t W = int32_t((slice_ptr[sl + 1] - base) / H);
    for (int32_t j = 0; j < W; ++j) {
        const bool have = j < len;
        if (rr == 0) scols[base + int64_t(j) * H + h] = have ? cols[b + j] : -1;
        const int64_t vbase = (base + int64_t(j) * H) * bb;
        for (int c = 0; c < bs; ++c) {
            const double v = have ? vals[int64_t(b + j) * bb + c * bs + rr] : 0.0;
            const int64_t at = vbase + int64_t(c) * (H * bs) + lane;
            svals[at] = v;
            if (svals32) svals32[at] = static_cast<float>(v);
            if (svals16) svals16[at] = to_bf16_bits(static_cast<float>(v));
        }
    }
}

// One wave per slice.  MODE 0: y = A x; MODE 1: y = sub - A x.
// DOT: per-lane partial of x_i * y_i (p . A p) reduced deterministically.
// K right-hand sides (column c of a vector at v + c * ld) share every column
// index and matrix value load; each column accumulates in the order of K = 1.
//
// Lane (h, r) owns column r of every block of block row h: per slot it gathers the
// single entry x[col * BS + r] and accumulates A(q, r) x_r into BS partial rows q.
// The BS lanes of a block row add their partial rows once, after the last slot (a
// fixed rotation of lane shuffles), so lane (h, q) ends with row q of y.  One x
// load per lane and slot (instead of BS), and fewer registers for the gathered x.
// VT: 0 FP64 values with x; 1 FP32 values, 2 bfloat16 values, both with x32
template <int BS, int MODE, bool DOT, int VT, int K>
__global__ __launch_bounds__(kSpmvThreads) __attribute__((amdgpu_waves_per_eu(6))) void k_spmv(int32_t nb, int32_t nslices, const int64_t *__restrict__ slice_ptr,
                                                        const int32_t *__restrict__ cols, const double *__restrict__ vals,
                                                        const float *__restrict__ vals32,
                                                        const uint16_t *__restrict__ vals16, const double *__restrict__ x,
                                                        const float *__restrict__ x32,
                                                        double *__restrict__ y, const double *__restrict__ sub,
                                                        const int *__restrict__ stop, int64_t ld,
                                                        double *__restrict__ partials)
{
    constexpr int H = 64 / BS;
    constexpr int BB = BS * BS;
    // 5 slots per chunk for one right-hand side: three round trips for the usual
    // width of 15 (14 neighbours and the diagonal block of a tetrahedral mesh node)
    constexpr int C = K >= 4 ? 2 : (K == 2 ? 4 : 5);
    // FP32 and bfloat16 values go with the FP32 copy of x (x32): the
    // preconditioner-internal residual, never the dot-product variant
    constexpr bool F32 = VT != 0;
    static_assert(!(F32 && DOT), "the reduced-precision SpMV has no dot product");
    if (dev::all_stopped(stop, K)) return;
    const int lane = threadIdx.x & 63;
    // wave-uniform: the slice pointer is a scalar load and the chunk loop has scalar control
    const int32_t slice = __builtin_amdgcn_readfirstlane(int(blockIdx.x * (kSpmvThreads / 64) + (threadIdx.x >> 6)));
    const int h = lane / BS, r = lane - h * BS;
    const int32_t row = slice * H + h;
    const bool active = (h < H) && (slice < nslices) && (row < nb);
    // lanes past the last block row (lane 63 for BS = 3) shadow row H-1: their loads
    // and shuffles stay in the slice, their results are discarded
    const int hh = h < H ? h : H - 1;
    double acc[K][BS];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int q = 0; q < BS; ++q) acc[k][q] = 0.0;
    // DOT: x_i of the lane's own row is the x entry it gathered for the diagonal block
    // (column == row), so p . A p needs no second read of p (up to two columns: four
    // would need registers beyond six waves per SIMD)
    constexpr bool DIAGX = DOT && K <= 2;
    double xd[K];
    bool found = false;
#pragma unroll
    for (int k = 0; k < K; ++k) xd[k] = 0.0;
    if (slice < nslices) {
        const int64_t base = slice_ptr[slice];
        const int32_t W = int32_t((slice_ptr[slice + 1] - base) / H);
        // Every load is unconditional: padding slots hold column -1 and zero
        // values, so a C-slot chunk issues all its column, value and x loads back
        // to back.  The last chunk of a slice is a full chunk too: its slots past
        // the slice width W re-read slot W-1 (valid addresses) and are masked to
        // zero, so a slice costs ceil(W / C) memory round trips (a slot-by-slot
        // tail would wait on the column index and then on x for every slot).
        // Values use ordinary loads (measured faster than non-temporal ones).
        const int ll = hh * BS + r;  // lane within the slot's value layout
        // the column indices of a chunk are loaded during the previous chunk, so a chunk
        // waits on one memory latency (values and x together), not on cols then x
        int32_t cn[C];
        if (W > 0) {
#pragma unroll
            for (int u = 0; u < C; ++u)
                cn[u] = __builtin_nontemporal_load(&cols[base + int64_t(min(u, W - 1)) * H + hh]);
        }
        for (int32_t j0 = 0; j0 < W; j0 += C) {
            double a[C][BS], xv[C][K];
#pragma unroll
            for (int u = 0; u < C; ++u) {
                const int64_t vb = (base + int64_t(min(j0 + u, W - 1)) * H) * BB + ll;
#pragma unroll
                for (int q = 0; q < BS; ++q) {
                    if constexpr (VT == 2)
                        a[u][q] = double(from_bf16_bits(vals16[vb + q * (H * BS)]));
                    else if constexpr (VT == 1)
                        a[u][q] = double(vals32[vb + q * (H * BS)]);
                    else
                        a[u][q] = vals[vb + q * (H * BS)];
                }
            }
            bool diag[C];
#pragma unroll
            for (int u = 0; u < C; ++u) {
                const int64_t xc = int64_t(max(cn[u], 0)) * BS + r;  // padding: column -1, zero values
                diag[u] = DIAGX && cn[u] == row && j0 + u < W;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if constexpr (F32)
                        xv[u][k] = double(x32[k * ld + xc]);
                    else
                        xv[u][k] = x[k * ld + xc];
                }
            }
            if (j0 + C < W) {
#pragma unroll
                for (int u = 0; u < C; ++u)
                    cn[u] = __builtin_nontemporal_load(&cols[base + int64_t(min(j0 + C + u, W - 1)) * H + hh]);
            }
#pragma unroll
            for (int u = 0; u < C; ++u) {
                if (j0 + u < W) {  // uniform across the wave
#pragma unroll
                    for (int q = 0; q < BS; ++q)
#pragma unroll
                        for (int k = 0; k < K; ++k) acc[k][q] += a[u][q] * xv[u][k];
                }
                if constexpr (DIAGX) {
                    if (diag[u]) {
                        found = true;
#pragma unroll
                        for (int k = 0; k < K; ++k) xd[k] = xv[u][k];
                    }
                }
            }
        }
    }
    // row r of the block row: sum over the BS lanes (h, r') of their partial row r, in the
    // order r' = r, r + 1, ... (mod BS); in round d lane (h, r') hands out its partial row
    // (r' - d) mod BS to lane (h, (r' - d) mod BS)
    double rowsum[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if constexpr (BS == 1) {
            rowsum[k] = acc[k][0];
        } else {
            double t = acc[k][0];
#pragma unroll
            for (int q = 1; q < BS; ++q) t = (r == q) ? acc[k][q] : t;
#pragma unroll
            for (int d = 1; d < BS; ++d) {
                const int give = (r - d + BS) % BS;
                double v = acc[k][0];
#pragma unroll
                for (int q = 1; q < BS; ++q) v = (give == q) ? acc[k][q] : v;
                t += __shfl(v, hh * BS + (r + d) % BS, 64);
            }
            rowsum[k] = t;
        }
    }
    const int64_t i = int64_t(row) * BS + r;
    double out[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        out[k] = 0.0;
        if (active) {
            out[k] = (MODE == 1) ? sub[k * ld + i] - rowsum[k] : rowsum[k];
            y[k * ld + i] = out[k];
        }
    }
    if constexpr (DOT) {
        if (active && !found) {  // a row without a stored diagonal block
#pragma unroll
            for (int k = 0; k < K; ++k) xd[k] = x[k * ld + i];
        }
        __shared__ double sh[kSpmvThreads / 64];
#pragma unroll
        for (int k = 0; k < K; ++k)
            dev::store_partial(dev::block_sum<kSpmvThreads>(active ? xd[k] * out[k] : 0.0, sh),
                               partials + int64_t(k) * gridDim.x);
    }
}

template <int MODE, bool DOT, int VT, int K>
int launch_spmv_k(const SellView &A, const double *x, const float *x32, double *y, const double *sub,
                  const int *stop, int64_t ld, double *partials, hipStream_t s)
{
    const int grid = div_up(A.nslices, kSpmvThreads / 64);
    if (grid == 0) return 0;
#define LSS_SPMV_CASE(BS)                                                                                      \
    k_spmv<BS, MODE, DOT, VT, K><<<grid, kSpmvThreads, 0, s>>>(A.nb, A.nslices, A.slice_ptr, A.cols, A.vals,   \
                                                               A.vals32, A.vals16, x, x32, y, sub, stop, ld,   \
                                                               partials)
    switch (A.bs) {
        case 1: LSS_SPMV_CASE(1); break;
        case 2: LSS_SPMV_CASE(2); break;
        case 3: LSS_SPMV_CASE(3); break;
        default: fail(LSS_ERR_UNSUPPORTED, "block_size must be 1, 2 or 3");
    }
#undef LSS_SPMV_CASE
    LSS_LAUNCH_CHECK();
    return grid;
}

template <int MODE, bool DOT, int VT>
int launch_spmv(const SellView &A, const double *x, const float *x32, double *y, const double *sub, const int *stop,
                int k, int64_t ld, double *partials, hipStream_t s)
{
    switch (k) {
        case 1: return launch_spmv_k<MODE, DOT, VT, 1>(A, x, x32, y, sub, stop, ld, partials, s);
        case 2: return launch_spmv_k<MODE, DOT, VT, 2>(A, x, x32, y, sub, stop, ld, partials, s);
        case 4: return launch_spmv_k<MODE, DOT, VT, 4>(A, x, x32, y, sub, stop, ld, partials, s);
        default: fail(LSS_ERR_ARG, "a fused SpMV takes 1, 2 or 4 right-hand sides");
    }
}

}  // namespace

void bsr_count(int32_t n, int bs, const int32_t *ptr, const int32_t *idx, int32_t *cnt, int32_t *unsorted_flag,
               hipStream_t s)
{
    const int32_t nb = (n + bs - 1) / bs;
    if (nb == 0) return;
    k_bsr_count<<<div_up(nb, 4), 256, 0, s>>>(n, bs, ptr, idx, cnt, unsorted_flag);
    LSS_LAUNCH_CHECK();
}

void bsr_fill(int32_t n, int bs, const int32_t *ptr, const int32_t *idx, const double *vals, const int32_t *rowptr,
              int32_t *cols, double *bvals, hipStream_t s)
{
    const int32_t nb = (n + bs - 1) / bs;
    if (nb == 0) return;
    k_bsr_fill<<<div_up(nb, 4), 256, 0, s>>>(n, bs, ptr, idx, vals, rowptr, cols, bvals);
    LSS_LAUNCH_CHECK();
}

void bsr_permute(int32_t nb, int bs, const int32_t *rowptr, const int32_t *cols, const double *vals,
                 const int32_t *perm, const int32_t *new_rowptr, int32_t *new_cols, double *new_vals, hipStream_t s)
{
    if (nb == 0) return;
    k_bsr_permute<<<div_up(nb, 4), 256, 0, s>>>(nb, bs, rowptr, cols, vals, perm, new_rowptr, new_cols, new_vals);
    LSS_LAUNCH_CHECK();
}

namespace {
__global__ void k_permute_lengths(int32_t nb, const int32_t *__restrict__ perm, const int32_t *__restrict__ len,
                                  int32_t *__restrict__ plen)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < nb) plen[perm[i]] = len[i];
}
}  // namespace

void permute_lengths(int32_t nb, const int32_t *perm, const int32_t *len, int32_t *plen, hipStream_t s)
{
    if (nb == 0) return;
    k_permute_lengths<<<div_up(nb, 256), 256, 0, s>>>(nb, perm, len, plen);
    LSS_LAUNCH_CHECK();
}

void sell_width(int32_t nb, int bs, const int32_t *rowptr, int32_t *slice_width, hipStream_t s)
{
    const int H = 64 / bs;
    const int32_t ns = div_up(nb, H);
    if (ns == 0) return;
    k_sell_width<<<div_up(ns, 256), 256, 0, s>>>(nb, H, rowptr, slice_width);
    LSS_LAUNCH_CHECK();
}

void sell_fill(int32_t nb, int bs, const int32_t *rowptr, const int32_t *cols, const double *vals,
               const int64_t *slice_ptr, int32_t *scols, double *svals, float *svals32, uint16_t *svals16,
               hipStream_t s)
{
    if (nb == 0) return;
    const int H = 64 / bs;
    const int32_t ns = div_up(nb, H);
    k_sell_fill<<<div_up(ns, 4), 256, 0, s>>>(nb, bs, H, ns, rowptr, cols, vals, slice_ptr, scols, svals, svals32,
                                              svals16);
    LSS_LAUNCH_CHECK();
}

void spmv(const SellView &A, const double *x, double *y, const double *sub, hipStream_t s)
{
    if (sub) launch_spmv<1, false, 0>(A, x, nullptr, y, sub, nullptr, 1, 0, nullptr, s);
    else launch_spmv<0, false, 0>(A, x, nullptr, y, nullptr, nullptr, 1, 0, nullptr, s);
}

int spmv_partials_per_column(const SellView &A) { return div_up(A.nslices, kSpmvThreads / 64); }

int64_t sell_value_count(int64_t slots, int bs) { return slots * bs * bs; }

void spmv_pq(const SellView &A, const double *p, double *q, dev::PcgState *st, double *partials, int iter,
             hipStream_t s, int k, int64_t ld, int finish_op, hipEvent_t after)
{
    const int g = launch_spmv<0, true, 0>(A, p, nullptr, q, nullptr, &st->stop, k, ld, partials, s);
    if (after) LSS_HIP(hipEventRecord(after, s));
    if (g > 0) finish(g, partials, finish_op, st, nullptr, iter, s, k);
}

void spmv_residual(const SellView &A, const double *x, const double *b, double *r, const int *stop, hipStream_t s,
                   bool fp32, int k, int64_t ld, const float *x32)
{
    if (fp32) {
        if (!x32 || (!A.vals32 && !A.vals16))
            fail(LSS_ERR_ARG, "the reduced-precision residual needs FP32 or bfloat16 values and x32");
        if (A.vals16) launch_spmv<1, false, 2>(A, x, x32, r, b, stop, k, ld, nullptr, s);
        else launch_spmv<1, false, 1>(A, x, x32, r, b, stop, k, ld, nullptr, s);
    } else {
        launch_spmv<1, false, 0>(A, x, nullptr, r, b, stop, k, ld, nullptr, s);
    }
}

namespace {
__global__ __launch_bounds__(256) void k_to_f32(int64_t n, const double *__restrict__ x, float *__restrict__ y)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        y[i] = float(x[i]);
}
}  // namespace

void to_f32(int64_t n, const double *x, float *y, hipStream_t s)
{
    if (n <= 0) return;
    k_to_f32<<<int(std::min<int64_t>(div_up(n, 256), 8192)), 256, 0, s>>>(n, x, y);
    LSS_LAUNCH_CHECK();
}

}  // namespace lss::k
