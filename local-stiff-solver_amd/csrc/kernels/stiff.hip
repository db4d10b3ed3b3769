// Stiff-block direct solves on gfx950, FP64 throughout.
//
// Reference: factorize_submatrix (GPUHybrid.cu:993-1411) extracts S_c A S_c^T
// per stiff component and factorizes it with cuDSS (sparse LDL^T) or cuBLAS
// (dense LU + explicit inverse); dss_precond_iter (544-675) applies
// z2 = z1 + S^T (S A S^T)^{-1} S (r - A z1) with gather / solve / scatter.
// Here: dense FP64 Cholesky per component (one workgroup each, blocked,
// diagonal blocks in LDS), then the inverse factor G = L^{-1} (column-parallel
// forward substitution), so that each preconditioner application is two
// dependency-free triangular matrix-vector products x = G^T (G t) instead of
// 2*m/32 dependent block steps of a triangular solve.  The correction computes
// (r - A z1) only on the stiff rows (the other rows of the reference's full
// SpMV are never read) and adds x into z in place.
#include "mas_kernels.hpp"
#include "sparse_kernels.hpp"
#include "../common/runtime.hpp"

namespace lss::k {
namespace {

constexpr int NB = 32;
constexpr int kT = 256;
// Components of at most kSmallMax DOFs are factored and inverted by one
// workgroup each; larger ones take the blocked path (dense_diag_block +
// library GEMMs, dense.cpp) and are skipped here.
constexpr int kSmallMax = 512;

__global__ __launch_bounds__(kT) void k_extract(const int64_t *__restrict__ comp_off, const int64_t *__restrict__ mat_off,
                                                const int32_t *__restrict__ dofs, const int32_t *__restrict__ g2l,
                                                const int32_t *__restrict__ comp_of, const int32_t *__restrict__ ptr,
                                                const int32_t *__restrict__ idx, const double *__restrict__ vals,
                                                double *__restrict__ mats)
{
    // GPUHybrid.cu:1351-1383 semantics: row r of the block = operator row dofs[r]
    const int32_t c = blockIdx.x;
    const int64_t off = comp_off[c];
    const int32_t m = int32_t(comp_off[c + 1] - off);
    double *M = mats + mat_off[c];
    for (int32_t r = threadIdx.x; r < m; r += kT) {
        const int32_t d = dofs[off + r];
        for (int32_t e = ptr[d]; e < ptr[d + 1]; ++e) {
            const int32_t j = idx[e];
            if (comp_of[j] == c) M[int64_t(r) * m + g2l[j]] = vals[e];
        }
    }
}

__global__ __launch_bounds__(kT) void k_cholesky(const int64_t *__restrict__ comp_off, const int64_t *__restrict__ mat_off,
                                                 double *__restrict__ mats, int *__restrict__ status)
{
    __shared__ double Lkk[NB][NB + 1];
    __shared__ int bad;
    const int32_t c = blockIdx.x;
    const int32_t m = int32_t(comp_off[c + 1] - comp_off[c]);
    if (m > kSmallMax) return;
    double *A = mats + mat_off[c];
    if (threadIdx.x == 0) bad = 0;
    for (int32_t k0 = 0; k0 < m; k0 += NB) {
        const int nb = min(NB, m - k0);
        // 1. diagonal block in LDS
        for (int t = threadIdx.x; t < NB * NB; t += kT) {
            const int i = t / NB, j = t % NB;
            Lkk[i][j] = (i < nb && j <= i) ? A[int64_t(k0 + i) * m + k0 + j] : 0.0;
        }
        __syncthreads();
        for (int j = 0; j < nb; ++j) {
            if (threadIdx.x == 0) {
                const double d = Lkk[j][j];
                if (!(d > 0.0) || !isfinite(d)) bad = 1;
                Lkk[j][j] = sqrt(d);
            }
            __syncthreads();
            const double dj = Lkk[j][j];
            for (int i = j + 1 + threadIdx.x; i < nb; i += kT) Lkk[i][j] /= dj;
            __syncthreads();
            for (int t = threadIdx.x; t < NB * NB; t += kT) {
                const int i = t / NB, l = t % NB;
                if (i < nb && l > j && l <= i) Lkk[i][l] -= Lkk[i][j] * Lkk[l][j];
            }
            __syncthreads();
        }
        for (int t = threadIdx.x; t < NB * NB; t += kT) {
            const int i = t / NB, j = t % NB;
            if (i < nb && j < nb) A[int64_t(k0 + i) * m + k0 + j] = (j <= i) ? Lkk[i][j] : 0.0;
        }
        // 2. panel rows below: L[i, blk] = A[i, blk] Lkk^{-T}
        for (int32_t i = k0 + nb + threadIdx.x; i < m; i += kT) {
            double v[NB];
#pragma unroll
            for (int l = 0; l < NB; ++l) {
                if (l < nb) {
                    double s = A[int64_t(i) * m + k0 + l];
                    for (int t = 0; t < l; ++t) s -= v[t] * Lkk[l][t];
                    v[l] = s / Lkk[l][l];
                }
            }
#pragma unroll
            for (int l = 0; l < NB; ++l)
                if (l < nb) A[int64_t(i) * m + k0 + l] = v[l];
        }
        __syncthreads();
        // 3. trailing lower update
        const int32_t T = m - k0 - nb;
        const int64_t pairs = int64_t(T) * (T + 1) / 2;
        for (int64_t q = threadIdx.x; q < pairs; q += kT) {
            // row-major lower enumeration: q -> (i, j), j <= i
            int32_t i = int32_t((sqrt(8.0 * double(q) + 1.0) - 1.0) * 0.5);
            while (int64_t(i) * (i + 1) / 2 > q) --i;
            while (int64_t(i + 1) * (i + 2) / 2 <= q) ++i;
            const int32_t j = int32_t(q - int64_t(i) * (i + 1) / 2);
            const double *Li = A + int64_t(k0 + nb + i) * m + k0;
            const double *Lj = A + int64_t(k0 + nb + j) * m + k0;
            double s = 0.0;
            for (int t = 0; t < nb; ++t) s += Li[t] * Lj[t];
            A[int64_t(k0 + nb + i) * m + k0 + nb + j] -= s;
        }
        __syncthreads();
    }
    // zero the strict upper triangle (never read, kept clean)
    for (int64_t q = threadIdx.x; q < int64_t(m) * m; q += kT) {
        const int32_t i = int32_t(q / m), j = int32_t(q % m);
        if (j > i) A[q] = 0.0;
    }
    if (threadIdx.x == 0 && bad) atomicOr(status, 2);
}

// (r - A z)[p] for the stiff rows p, in FP64 from the block operator: one wave
// per stiff row, lanes over the row's (block, component) entries, a wave
// reduction.  OUT_U: write u[p] (recomputing the stiff rows after a
// reduced-precision SpMV); else work[k] for the stiff correction.  Several
// right-hand sides: grid row y is column y (vectors at stride ld, stiff work
// vectors at stride ns).
template <bool OUT_U>
__global__ __launch_bounds__(256) void k_stiff_residual(int64_t ns, int bs, const int32_t *__restrict__ pdofs,
                                                        const int32_t *__restrict__ rowptr,
                                                        const int32_t *__restrict__ cols,
                                                        const double *__restrict__ bvals, const double *__restrict__ r,
                                                        const double *__restrict__ z, int64_t ld,
                                                        double *__restrict__ out, const int *__restrict__ stop)
{
    if (dev::all_stopped(stop, gridDim.y)) return;
    const int64_t k = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= ns) return;
    r += blockIdx.y * ld;
    const int32_t p = pdofs[k];
    const int32_t I = p / bs;
    const int a = p - I * bs;
    const int bb = bs * bs;
    const int32_t e0 = rowptr[I];
    const int32_t m = z ? (rowptr[I + 1] - e0) * bs : 0;  // z == nullptr: gather r alone
    if (z) z += blockIdx.y * ld;
    double acc = 0.0;
    for (int32_t q = lane; q < m; q += 64) {
        const int32_t e = e0 + q / bs;
        const int c = q - (q / bs) * bs;
        acc += bvals[int64_t(e) * bb + a * bs + c] * z[int64_t(cols[e]) * bs + c];
    }
    acc = dev::wave_sum(acc);
    if (lane == 0) {
        if (OUT_U)
            out[blockIdx.y * ld + p] = r[p] - acc;
        else
            out[blockIdx.y * ns + k] = r[p] - acc;
    }
}

// G = L^{-1} of the components of at most kSmallMax DOFs, stored as M with
// M(i, j) = G(i, j) for j <= i and M(j, i) = G(i, j) above the diagonal (G and
// G^T share one square array, so both triangular products read rows).  One wave
// per panel of 32 columns, the panel G(j0:m, j0:j0+32) kept in LDS.  Rows go in
// order, G(i, j) = -(sum_{k=j}^{i-1} L(i,k) G(k,j)) / L(i,i); lane (h, c) forms
// the dot product of column j0 + c over the entries k of parity h, the two
// halves are added with one cross-half shuffle.
__global__ __launch_bounds__(64) void k_tri_inverse_panel(const int64_t *__restrict__ comp_off,
                                                          const int64_t *__restrict__ mat_off,
                                                          const double *__restrict__ Lall, double *__restrict__ Gall)
{
    extern __shared__ double P[];  // P[(i - j0) * 32 + c] = G(i, j0 + c)
    const int32_t c = blockIdx.x;
    const int32_t m = int32_t(comp_off[c + 1] - comp_off[c]);
    const int32_t j0 = blockIdx.y * 32;
    if (m > kSmallMax || j0 >= m) return;
    const int lane = threadIdx.x, h = lane >> 5, cl = lane & 31;
    const int32_t col = j0 + cl;
    const double *L = Lall + mat_off[c];
    double *G = Gall + mat_off[c];
    for (int32_t i = j0; i < m; ++i) {
        const double *Li = L + int64_t(i) * m;
        double sacc = 0.0;
        if (col < i)
            for (int32_t k = col + h; k < i; k += 2) sacc += Li[k] * P[(k - j0) * 32 + cl];
        sacc += __shfl_xor(sacc, 32, 64);
        double g = 0.0;
        if (col == i) g = 1.0 / Li[i];
        else if (col < i) g = -sacc / Li[i];
        if (h == 0 && col < m) P[(i - j0) * 32 + cl] = g;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (h == 0 && col < m && col <= i) {
            G[int64_t(i) * m + col] = g;  // lower triangle: G
            G[int64_t(col) * m + i] = g;  // upper triangle: G^T (the diagonal twice, same value)
        }
    }
}

// One wave per stiff row (all components at once).  Phase G: u_i = M(i, 0:i) t;
// phase GT: x_j = M(j, j) u_j + M(j, j+1:m) u_{j+1:m}.  Lanes read a row with
// coalesced loads, a wave reduction finishes the dot product.  A unit of the dense
// branch (dense[c] = 1) holds its explicit inverse: phase G is the full row
// u_i = M(i, 0:m) t and phase GT adds u unchanged.
template <bool TRANSPOSED>
__global__ __launch_bounds__(256) void k_stiff_rows(int64_t ns, const int32_t *__restrict__ row_comp,
                                                    const int64_t *__restrict__ comp_off,
                                                    const int64_t *__restrict__ mat_off, const int32_t *__restrict__ pdofs,
                                                    const double *__restrict__ M, const double *__restrict__ in,
                                                    double *__restrict__ out, double *__restrict__ z, int64_t ld,
                                                    const int *__restrict__ stop, float *__restrict__ z32,
                                                    const uint8_t *__restrict__ dense)
{
    if (dev::all_stopped(stop, gridDim.y)) return;
    const int64_t k = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= ns) return;
    in += blockIdx.y * ns;
    if (out) out += blockIdx.y * ns;
    z += blockIdx.y * ld;
    if (z32) z32 += blockIdx.y * ld;
    const int32_t c = row_comp[k];
    const int64_t off = comp_off[c];
    const int32_t m = int32_t(comp_off[c + 1] - off);
    const int32_t i = int32_t(k - off);
    const double *row = M + mat_off[c] + int64_t(i) * m;
    const double *v = in + off;
    const bool full = dense && dense[c];
    double acc = 0.0;
    if (TRANSPOSED && full) {
        acc = v[i];
    } else {
        const int32_t lo = TRANSPOSED ? i : 0, hi = (TRANSPOSED || full) ? m : i + 1;
        for (int32_t j = lo + lane; j < hi; j += 64) acc += row[j] * v[j];
        acc = dev::wave_sum(acc);
    }
    if (lane == 0) {
        if (TRANSPOSED) {
            const int32_t d = pdofs[k];
            const double v = z[d] + acc;  // z2 = z1 + S^T x (GPUHybrid.cu:621-671)
            z[d] = v;
            if (z32) z32[d] = float(v);
        } else
            out[k] = acc;
    }
}

// Explicit inverse M = G^T G = (S A S^T)^{-1} of the listed units (the dense branch,
// GPUHybrid.cu:1385-1405 inverts by LU instead): G is the lower triangle of the stored
// factor, M(i, j) = sum over k >= max(i, j) of G(k, i) G(k, j).  One wave per row i,
// lanes over j <= i; both triangles of M are written.
__global__ __launch_bounds__(256) void k_explicit_inverse(const int32_t *__restrict__ units,
                                                          const int64_t *__restrict__ comp_off,
                                                          const int64_t *__restrict__ mat_off,
                                                          const double *__restrict__ G, double *__restrict__ M)
{
    const int32_t c = units[blockIdx.x];
    const int32_t m = int32_t(comp_off[c + 1] - comp_off[c]);
    const int32_t i = int32_t(blockIdx.y) * 4 + int32_t(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= m) return;
    const double *Gc = G + mat_off[c];
    double *Mc = M + mat_off[c];
    for (int32_t j = lane; j <= i; j += 64) {
        double acc = 0.0;
        for (int32_t k = i; k < m; ++k) acc += Gc[int64_t(k) * m + i] * Gc[int64_t(k) * m + j];
        Mc[int64_t(i) * m + j] = acc;
        Mc[int64_t(j) * m + i] = acc;
    }
}

// Diagonal block of the blocked factorization of one large component: the
// kb x kb block A(k0:k0+kb, k0:k0+kb) (row-major, leading dimension m, lower
// triangle read) is factored A = L L^T in LDS; L goes back in place (strict
// upper triangle of the block zeroed) and L^{-1} into the lower triangle of the
// same block of G.  The block is itself factored in sub-blocks of 32: wave 0
// factors and inverts the 32 x 32 diagonal sub-block in registers (lane i holds
// row i of L and column i of L^{-1}; the other lanes' entries arrive by readlane,
// so no barrier inside the sub-block), then all waves form the panel below it
// (L_IJ = A_IJ G_JJ^T) and the trailing update (A_II -= L_IJ L_IJ^T, 4 x 4 tiles);
// the off-diagonal blocks of the inverse follow by distance from the diagonal,
// G_IJ = -G_II sum_{K=J}^{I-1} L_IK G_KJ.  About 5 barriers per sub-block instead
// of two per column.  LDS: the block with rows padded to DP = kb + 1 doubles
// (lower triangle L, strict upper triangle G^T), G's diagonal, and three 32 x 32
// work blocks.
constexpr int kDenseNB = 128;
constexpr int kDenseSB = 32;

__device__ __forceinline__ double readlane_d(double v, int l)
{
    union {
        double d;
        int i[2];
    } u;
    u.d = v;
    u.i[0] = __builtin_amdgcn_readlane(u.i[0], l);
    u.i[1] = __builtin_amdgcn_readlane(u.i[1], l);
    return u.d;
}

__host__ __device__ constexpr size_t dense_diag_lds(int kb)
{
    return (size_t(kb) * (kb + 1) + kb + 3 * kDenseSB * kDenseSB) * sizeof(double);
}

__device__ __forceinline__ void dense_diag_run(int64_t lda, int k0, int kb, double *__restrict__ A, int64_t ldg,
                                               double *__restrict__ G, int *__restrict__ status)
{
    constexpr int SB = kDenseSB;
    extern __shared__ double lds[];
    const int DP = kb + 1;
    double *M = lds;                    // kb x DP
    double *dg = M + int64_t(kb) * DP;  // diagonal of G
    double *T = dg + kb;                // 3 work blocks of SB x SB
    __shared__ int bad;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0) bad = 0;
    for (int t = tid; t < kb * kb; t += 256) {
        const int i = t / kb, j = t - i * kb;
        if (j <= i) M[i * DP + j] = A[(k0 + i) * lda + k0 + j];
    }
    __syncthreads();
    const int nsub = (kb + SB - 1) / SB;
    for (int J = 0; J < nsub; ++J) {
        const int j0 = J * SB, jb = min(SB, kb - j0);
        if (wid == 0) {
            double row[SB], g[SB];
#pragma unroll
            for (int c = 0; c < SB; ++c) row[c] = (lane < jb && c <= lane) ? M[(j0 + lane) * DP + j0 + c] : 0.0;
            bool ok = true;
#pragma unroll
            for (int k = 0; k < SB; ++k) {
                if (k < jb) {
                    const double dkk = readlane_d(row[k], k);
                    ok = ok && dkk > 0.0 && isfinite(dkk);
                    const double sq = sqrt(dkk), rsq = 1.0 / sq;
                    if (lane == k) row[k] = sq;
                    else if (lane > k) row[k] *= rsq;
                    const double lk = row[k];
#pragma unroll
                    for (int j = k + 1; j < SB; ++j) {
                        const double ljk = readlane_d(row[k], j);
                        if (lane >= j) row[j] -= lk * ljk;
                    }
                }
            }
            // lane c: column c of G_JJ = L_JJ^{-1} by forward substitution
            double rdiag = 0.0;  // lane i: 1 / L_ii
#pragma unroll
            for (int c = 0; c < SB; ++c)
                if (c == lane) rdiag = 1.0 / row[c];
#pragma unroll
            for (int i = 0; i < SB; ++i) {
                if (i < jb) {
                    double acc = lane == i ? 1.0 : 0.0;
#pragma unroll
                    for (int k = 0; k < i; ++k) acc -= readlane_d(row[k], i) * g[k];
                    g[i] = acc * readlane_d(rdiag, i);
                } else {
                    g[i] = 0.0;
                }
            }
            if (lane < jb) {
#pragma unroll
                for (int c = 0; c < SB; ++c) {
                    if (c <= lane) M[(j0 + lane) * DP + j0 + c] = row[c];
                    else if (c < jb) M[(j0 + lane) * DP + j0 + c] = g[c];  // G(c, lane) at (lane, c)
                    if (c == lane) dg[j0 + c] = g[c];
                }
            }
            if (lane == 0 && !ok) bad = 1;
        }
        __syncthreads();
        const int r0 = j0 + jb;
        if (r0 >= kb) continue;
        // panel, one thread per row: L(r, j0 + c) = sum_{k <= c} A(r, j0 + k) G_JJ(c, k)
        for (int r = r0 + tid; r < kb; r += 256) {
            double a[SB];
#pragma unroll
            for (int k = 0; k < SB; ++k) a[k] = k < jb ? M[r * DP + j0 + k] : 0.0;
#pragma unroll
            for (int c = 0; c < SB; ++c) {
                if (c < jb) {
                    double acc = a[c] * dg[j0 + c];
#pragma unroll
                    for (int k = 0; k < c; ++k) acc += a[k] * M[(j0 + k) * DP + j0 + c];
                    M[r * DP + j0 + c] = acc;
                }
            }
        }
        __syncthreads();
        // trailing update in 4 x 4 tiles of the lower triangle of rows/columns r0 .. kb - 1
        const int nt = (kb - r0 + 3) / 4, ntiles = nt * (nt + 1) / 2;
        for (int t = tid; t < ntiles; t += 256) {
            int ti = int((sqrtf(8.0f * float(t) + 1.0f) - 1.0f) * 0.5f);
            while (ti * (ti + 1) / 2 > t) --ti;
            while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
            const int tj = t - ti * (ti + 1) / 2;
            const int rb = r0 + 4 * ti, cb = r0 + 4 * tj;
            double acc[4][4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
            for (int k = 0; k < jb; ++k) {
                double lr[4], lc[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    lr[u] = rb + u < kb ? M[(rb + u) * DP + j0 + k] : 0.0;
                    lc[u] = cb + u < kb ? M[(cb + u) * DP + j0 + k] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[u][v] += lr[u] * lc[v];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = rb + u, c = cb + v;
                    if (r < kb && c <= r) M[r * DP + c] -= acc[u][v];
                }
        }
        __syncthreads();
    }
    // off-diagonal blocks of G, by distance d = I - J from the diagonal, in 4 x 4 tiles:
    // T = sum_{K=J}^{I-1} L_IK G_KJ, then G_IJ = -G_II T
    for (int d = 1; d < nsub; ++d) {
        const int ntile = (nsub - d) * 64;  // 8 x 8 tiles per block pair
        for (int t = tid; t < ntile; t += 256) {
            const int pp = t >> 6, ti = (t >> 3) & 7, tj = t & 7;
            const int ia0 = (pp + d) * SB + 4 * ti, jc0 = pp * SB + 4 * tj;
            const int kend = min((pp + d) * SB, kb);
            double acc[4][4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
#pragma unroll 4
            for (int kk = pp * SB; kk < kend; ++kk) {
                double l[4], gk[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) l[u] = ia0 + u < kb ? M[(ia0 + u) * DP + kk] : 0.0;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int jc = jc0 + v;  // G(kk, jc): lower triangular, G^T stored above the diagonal
                    gk[v] = (jc < kb && kk >= jc) ? (kk == jc ? dg[kk] : M[jc * DP + kk]) : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[u][v] += l[u] * gk[v];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) T[pp * SB * SB + (4 * ti + u) * SB + 4 * tj + v] = acc[u][v];
        }
        __syncthreads();
        for (int t = tid; t < ntile; t += 256) {
            const int pp = t >> 6, ti = (t >> 3) & 7, tj = t & 7;
            const int i0 = (pp + d) * SB, jc0 = pp * SB + 4 * tj;
            double acc[4][4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
            for (int q = 0; q < 4 * ti + 4; ++q) {
                double gq[4], tq[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int a = 4 * ti + u, ia = i0 + a;  // G(ia, i0 + q), q <= a
                    gq[u] = (ia < kb && q <= a) ? (q == a ? dg[ia] : M[(i0 + q) * DP + ia]) : 0.0;
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) tq[v] = T[pp * SB * SB + q * SB + 4 * tj + v];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[u][v] += gq[u] * tq[v];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int ia = i0 + 4 * ti + u, jc = jc0 + v;
                    if (ia < kb && jc < kb) M[jc * DP + ia] = -acc[u][v];  // G(ia, jc) at (jc, ia)
                }
        }
        __syncthreads();
    }
    for (int t = tid; t < kb * kb; t += 256) {
        const int i = t / kb, j = t - i * kb;
        A[(k0 + i) * lda + k0 + j] = j <= i ? M[i * DP + j] : 0.0;
        if (j < i) G[(k0 + i) * ldg + k0 + j] = M[j * DP + i];
        else if (j == i) G[(k0 + i) * ldg + k0 + j] = dg[i];
    }
    if (tid == 0 && bad) atomicOr(status, 2);
}

__global__ __launch_bounds__(256) void k_dense_diag(int64_t m, int k0, int kb, double *__restrict__ A,
                                                    double *__restrict__ G, int *__restrict__ status)
{
    dense_diag_run(m, k0, kb, A, m, G, status);
}

// one workgroup per task: the diagonal block of one front of a batch
__global__ __launch_bounds__(256) void k_dense_diag_batch(const DiagTask *__restrict__ tasks, double *__restrict__ W,
                                                          double *__restrict__ G, int *__restrict__ status)
{
    const DiagTask t = tasks[blockIdx.x];
    dense_diag_run(t.lda, t.k0, t.kb, W + t.a, t.ldg, G + t.g, status);
}

// Upper triangle of G := transpose of its lower triangle (32 x 32 tiles via LDS).
__global__ __launch_bounds__(256) void k_dense_mirror(int64_t m, double *__restrict__ G)
{
    __shared__ double t[32][33];
    const int bx = blockIdx.x, by = blockIdx.y;  // tile column, tile row of the source
    if (bx > by) return;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int64_t i = int64_t(by) * 32 + r, j = int64_t(bx) * 32 + tx;
        t[r][tx] = (i < m && j < m) ? G[i * m + j] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int64_t i = int64_t(bx) * 32 + r, j = int64_t(by) * 32 + tx;  // destination (i, j) = source (j, i)
        if (i < m && j < m && j > i) G[i * m + j] = t[tx][r];
    }
}

}  // namespace

void stiff_extract(int32_t ncomp, const int64_t *comp_off, const int64_t *mat_off, const int32_t *dofs,
                   const int32_t *g2l, const int32_t *comp_of, const int32_t *ptr, const int32_t *idx,
                   const double *vals, double *mats, hipStream_t s)
{
    if (ncomp == 0) return;
    k_extract<<<ncomp, kT, 0, s>>>(comp_off, mat_off, dofs, g2l, comp_of, ptr, idx, vals, mats);
    LSS_LAUNCH_CHECK();
}

void stiff_cholesky(int32_t ncomp, const int64_t *comp_off, const int64_t *mat_off, double *mats,
                    int *status, hipStream_t s)
{
    if (ncomp == 0) return;
    k_cholesky<<<ncomp, kT, 0, s>>>(comp_off, mat_off, mats, status);
    LSS_LAUNCH_CHECK();
}

void stiff_inverse_factor(int32_t ncomp, int32_t max_small_m, const int64_t *comp_off, const int64_t *mat_off,
                          const double *L, double *G, hipStream_t s)
{
    if (ncomp == 0 || max_small_m == 0) return;
    if (max_small_m > kSmallMax) fail(LSS_ERR_ARG, "stiff_inverse_factor: component above the one-workgroup size");
    const size_t lds = size_t(max_small_m) * 32 * sizeof(double);
    static bool attr = false;
    if (!attr) {
        LSS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_tri_inverse_panel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(size_t(kSmallMax) * 32 * sizeof(double))));
        attr = true;
    }
    k_tri_inverse_panel<<<dim3(ncomp, div_up(max_small_m, 32)), 64, lds, s>>>(comp_off, mat_off, L, G);
    LSS_LAUNCH_CHECK();
}

int stiff_small_max() { return kSmallMax; }
int dense_block_size() { return kDenseNB; }

void dense_diag_block(int64_t m, int k0, int kb, double *A, double *G, int *status, hipStream_t s)
{
    if (kb <= 0 || kb > kDenseNB) fail(LSS_ERR_ARG, "dense_diag_block: block size out of range");
    const size_t lds = dense_diag_lds(kb);
    static bool attr = false;
    if (!attr) {
        LSS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_dense_diag),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(dense_diag_lds(kDenseNB))));
        attr = true;
    }
    k_dense_diag<<<1, 256, lds, s>>>(m, k0, kb, A, G, status);
    LSS_LAUNCH_CHECK();
}

void dense_diag_batch(int32_t ntask, const DiagTask *tasks, double *W, double *G, int *status, hipStream_t s)
{
    if (ntask == 0) return;
    static bool attr = false;
    if (!attr) {
        LSS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_dense_diag_batch),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(dense_diag_lds(kDenseNB))));
        attr = true;
    }
    k_dense_diag_batch<<<ntask, 256, dense_diag_lds(kDenseNB), s>>>(tasks, W, G, status);
    LSS_LAUNCH_CHECK();
}

void dense_mirror(int64_t m, double *G, hipStream_t s)
{
    if (m == 0) return;
    const int t = int(div_up(m, 32));
    k_dense_mirror<<<dim3(t, t), 256, 0, s>>>(m, G);
    LSS_LAUNCH_CHECK();
}

void stiff_correct(int32_t ncomp, int64_t ns, int bs, const int64_t *comp_off, const int64_t *mat_off,
                   const int32_t *row_comp, const int32_t *pdofs, const double *G, const int32_t *rowptr,
                   const int32_t *cols,
                   const double *bvals, const double *r, double *z, double *work, double *work2, const int *stop,
                   hipStream_t s, int k, int64_t ld, bool additive, float *z32, const uint8_t *dense)
{
    if (ncomp == 0 || ns == 0) return;
    k_stiff_residual<false><<<dim3(div_up(ns, 4), k), 256, 0, s>>>(ns, bs, pdofs, rowptr, cols, bvals, r,
                                                                  additive ? nullptr : z, ld, work, stop);
    LSS_LAUNCH_CHECK();
    k_stiff_rows<false><<<dim3(div_up(ns, 4), k), 256, 0, s>>>(ns, row_comp, comp_off, mat_off, pdofs, G, work, work2,
                                                              z, ld, stop, nullptr, dense);
    LSS_LAUNCH_CHECK();
    k_stiff_rows<true><<<dim3(div_up(ns, 4), k), 256, 0, s>>>(ns, row_comp, comp_off, mat_off, pdofs, G, work2,
                                                             nullptr, z, ld, stop, z32, dense);
    LSS_LAUNCH_CHECK();
}

void explicit_inverse(int32_t nunits, const int32_t *units, int32_t max_m, const int64_t *comp_off,
                      const int64_t *mat_off, const double *G, double *M, hipStream_t s)
{
    if (nunits == 0 || max_m == 0) return;
    k_explicit_inverse<<<dim3(nunits, div_up(max_m, 4)), 256, 0, s>>>(units, comp_off, mat_off, G, M);
    LSS_LAUNCH_CHECK();
}

void stiff_residual_rows(int64_t ns, int bs, const int32_t *pdofs, const int32_t *rowptr, const int32_t *cols,
                         const double *bvals, const double *r, const double *z, double *out, const int *stop,
                         hipStream_t s, int k, int64_t ld)
{
    if (ns == 0) return;
    k_stiff_residual<false><<<dim3(div_up(ns, 4), k), 256, 0, s>>>(ns, bs, pdofs, rowptr, cols, bvals, r, z, ld, out,
                                                                  stop);
    LSS_LAUNCH_CHECK();
}

void stiff_rows_exact(int64_t ns, int bs, const int32_t *pdofs, const int32_t *rowptr, const int32_t *cols,
                      const double *bvals, const double *r, const double *z, double *u, const int *stop, hipStream_t s,
                      int k, int64_t ld)
{
    if (ns == 0) return;
    k_stiff_residual<true><<<dim3(div_up(ns, 4), k), 256, 0, s>>>(ns, bs, pdofs, rowptr, cols, bvals, r, z, ld, u,
                                                                 stop);
    LSS_LAUNCH_CHECK();
}

}  // namespace lss::k
