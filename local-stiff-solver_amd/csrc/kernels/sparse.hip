// Multifrontal sparse Cholesky of large stiff units: assembly and solve kernels.
//
// The reference factorizes every stiff component with cuDSS (sparse LDL^T,
// GPUHybrid.cu:1231-1249; MAS/BadDofGMMPreconditioner.cu:493-549) and applies it
// by triangular solves (GPUHybrid.cu:575).  Units up to max_dense_component DOFs
// take a dense Cholesky here; larger ones a multifrontal one: a nested-dissection
// ordering and the front structures on the host, then per front (children first)
// assembly of the original entries, extend-add of the children's update matrices,
// and a partial blocked Cholesky: the fronts of one tree level together through the
// batched kernels below (task lists built on the host), large fronts one at a time
// with the dense kernels and rocBLAS (host/sparse_stiff.cpp).
// A front keeps G11 = L11^{-1} (with its transpose in the strict upper triangle)
// and L21, so both triangular solves are products.  The solve walks the tree
// level by level: forward, bottom-up, each front gathers its right-hand side and
// its children's update vectors (in child order), y = G11 r, w = r_U - L21 y;
// backward, top-down, x = G11^T (y - L21^T x_U).  No atomics anywhere: every sum
// has a fixed order, so solves are reproducible bit for bit.
#include "sparse_kernels.hpp"

#include "device_utils.hpp"
#include "../common/runtime.hpp"

namespace lss::k {
namespace {

constexpr int kT = 256;

// Rows of the level steps go to groups of LPR lanes (4 .. 64, a power of two chosen per
// level from the mean row length), summed by a fixed butterfly within the group.
// sum_j a[j] v(j) over j = b + lane, b + lane + LPR, ... < e, in that order; each lane
// issues its loads in batches of U (predicated) before the products, so a row of up to
// U * LPR entries costs one memory round trip
template <int LPR, typename V>
__device__ __forceinline__ double group_dot(const double *__restrict__ a, int b, int e, int lane, V v)
{
    constexpr int U = 16;
    double acc = 0.0;
    for (int j0 = b + lane; j0 < e; j0 += LPR * U) {
        double av[U], bv[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int j = j0 + q * LPR;
            av[q] = j < e ? a[j] : 0.0;
            bv[q] = j < e ? v(j) : 0.0;
        }
#pragma unroll
        for (int q = 0; q < U; ++q) acc += av[q] * bv[q];
    }
    return acc;
}

template <int LPR>
__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ __launch_bounds__(kT) void k_gather_pattern(int32_t m, const int32_t *__restrict__ dofs,
                                                       const int32_t *__restrict__ ptr,
                                                       const int64_t *__restrict__ coff,
                                                       const int32_t *__restrict__ idx, int32_t *__restrict__ out)
{
    const int32_t v = int32_t(blockIdx.x) * (kT / 64) + int32_t(threadIdx.x >> 6);
    if (v >= m) return;
    const int lane = threadIdx.x & 63;
    const int32_t d = dofs[v];
    const int32_t b = ptr[d], e = ptr[d + 1];
    int32_t *o = out + coff[v];
    for (int32_t t = b + lane; t < e; t += 64) o[t - b] = idx[t];
}

__global__ __launch_bounds__(kT) void k_assemble(int64_t ntrip, const int32_t *__restrict__ li,
                                                 const int32_t *__restrict__ lj, const int64_t *__restrict__ src,
                                                 const double *__restrict__ vals, double *__restrict__ F, int64_t ld)
{
    const int64_t t = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (t >= ntrip) return;
    F[int64_t(li[t]) * ld + lj[t]] = vals[src[t]];
}

// one wave per row a of the child's update matrix
__global__ __launch_bounds__(kT) void k_extend_add(int32_t u, const int32_t *__restrict__ rel,
                                                   const double *__restrict__ U, double *__restrict__ F, int64_t ld)
{
    const int32_t a = int32_t(blockIdx.x) * (kT / 64) + int32_t(threadIdx.x >> 6);
    if (a >= u) return;
    const int lane = threadIdx.x & 63;
    const int64_t ra = int64_t(rel[a]) * ld;
    const double *row = U + int64_t(a) * u;
    for (int32_t b = lane; b <= a; b += 64) F[ra + rel[b]] += row[b];
}

__global__ __launch_bounds__(kT) void k_permute_in(int32_t m, const int32_t *__restrict__ order,
                                                   const double *__restrict__ t, int64_t ldt, double *__restrict__ tp,
                                                   const int *__restrict__ stop)
{
    if (dev::all_stopped(stop, gridDim.y)) return;
    const int32_t q = int32_t(blockIdx.x) * kT + int32_t(threadIdx.x);
    if (q >= m) return;
    const int c = blockIdx.y;
    tp[int64_t(c) * m + q] = t[int64_t(c) * ldt + order[q]];
}

__global__ __launch_bounds__(kT) void k_fwd_rhs(int64_t ntask, const int2 *__restrict__ tasks, FrontView fv,
                                                const double *__restrict__ tp, int32_t m,
                                                const double *__restrict__ w, int64_t wtot,
                                                double *__restrict__ rhs, int64_t rtot,
                                                const int *__restrict__ stop)
{
    if (dev::all_stopped(stop, gridDim.y)) return;
    const int64_t t = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (t >= ntask) return;
    const int c = blockIdx.y;
    const int2 task = tasks[t];
    const int f = task.x, a = task.y;
    const int64_t r = fv.roff[f] + a;
    double v = a < fv.np[f] ? tp[int64_t(c) * m + fv.p0[f] + a] : 0.0;
    const double *wc = w + int64_t(c) * wtot;
    for (int64_t q = fv.cptr[r]; q < fv.cptr[r + 1]; ++q) v += wc[fv.cidx[q]];
    rhs[int64_t(c) * rtot + r] = v;
}

// wave per pivot row i: y_i = sum_{j <= i} G11(i, j) rhs_j
template <int LPR>
__global__ __launch_bounds__(kT) void k_fwd_y(int64_t ntask, const int2 *__restrict__ tasks, FrontView fv,
                                              const double *__restrict__ G, const double *__restrict__ rhs,
                                              int64_t rtot, double *__restrict__ ypos, int32_t m,
                                              const int *__restrict__ stop)
{
    if (dev::all_stopped(stop, gridDim.y)) return;
    const int64_t t = (int64_t(blockIdx.x) * kT + threadIdx.x) / LPR;
    if (t >= ntask) return;  // a whole group
    const int lane = threadIdx.x & (LPR - 1), c = blockIdx.y;
    const int2 task = tasks[t];
    const int f = task.x, i = task.y, p = fv.np[f];
    const double *row = G + fv.goff[f] + int64_t(i) * p;
    const double *v = rhs + int64_t(c) * rtot + fv.roff[f];
    double acc = group_dot<LPR>(row, 0, i + 1, lane, [&](int j) { return v[j]; });
    acc = group_sum<LPR>(acc);
    if (lane == 0) ypos[int64_t(c) * m + fv.p0[f] + i] = acc;
}

// wave per update row a: w_a = rhs_{p + a} - sum_j L21(a, j) y_j
template <int LPR>
__global__ __launch_bounds__(kT) void k_fwd_w(int64_t ntask, const int2 *__restrict__ tasks, FrontView fv,
                                              const double *__restrict__ L, const double *__restrict__ rhs,
                                              int64_t rtot, const double *__restrict__ ypos, int32_t m,
                                              double *__restrict__ w, int64_t wtot, const int *__restrict__ stop)
{
    if (dev::all_stopped(stop, gridDim.y)) return;
    const int64_t t = (int64_t(blockIdx.x) * kT + threadIdx.x) / LPR;
    if (t >= ntask) return;  // a whole group
    const int lane = threadIdx.x & (LPR - 1), c = blockIdx.y;
    const int2 task = tasks[t];
    const int f = task.x, a = task.y, p = fv.np[f];
    const double *row = L + fv.loff[f] + int64_t(a) * p;
    const double *y = ypos + int64_t(c) * m + fv.p0[f];
    double acc = group_dot<LPR>(row, 0, p, lane, [&](int j) { return y[j]; });
    acc = group_sum<LPR>(acc);
    if (lane == 0)
        w[int64_t(c) * wtot + fv.woff[f] + a] = rhs[int64_t(c) * rtot + fv.roff[f] + p + a] - acc;
}

// xu[woff[f] + a] = x[U_f(a)]: the update positions' solution entries of a level's fronts,
// gathered once (thread per update row) so that the products below stream contiguous data
__global__ __launch_bounds__(kT) void k_bwd_gather(int64_t ntask, const int2 *__restrict__ tasks, FrontView fv,
                                                   const double *__restrict__ xpos, int32_t m,
                                                   double *__restrict__ xu, int64_t wtot, const int *__restrict__ stop)
{
    if (dev::all_stopped(stop, gridDim.y)) return;
    const int64_t t = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (t >= ntask) return;
    const int c = blockIdx.y;
    const int2 task = tasks[t];
    const int64_t q = fv.woff[task.x] + task.y;
    xu[int64_t(c) * wtot + q] = xpos[int64_t(c) * m + fv.uidx[q]];
}

// row j of L21^T (p x u, row-major) per group: s_j = y_j - sum_a L21(a, j) x[U_f(a)]
template <int LPR>
__global__ __launch_bounds__(kT) void k_bwd_s(int64_t ntask, const int2 *__restrict__ tasks, FrontView fv,
                                              const double *__restrict__ Lt, const double *__restrict__ xu,
                                              int64_t wtot, const double *__restrict__ ypos, int32_t m,
                                              double *__restrict__ sbuf, int64_t rtot, const int *__restrict__ stop)
{
    if (dev::all_stopped(stop, gridDim.y)) return;
    const int64_t t = (int64_t(blockIdx.x) * kT + threadIdx.x) / LPR;
    if (t >= ntask) return;  // a whole group
    const int lane = threadIdx.x & (LPR - 1), c = blockIdx.y;
    const int2 task = tasks[t];
    const int f = task.x, j = task.y, u = fv.nu[f];
    const double *row = Lt + fv.loff[f] + int64_t(j) * u;
    const double *x = xu + int64_t(c) * wtot + fv.woff[f];
    double acc = group_dot<LPR>(row, 0, u, lane, [&](int a) { return x[a]; });
    acc = group_sum<LPR>(acc);
    if (lane == 0) sbuf[int64_t(c) * rtot + fv.roff[f] + j] = ypos[int64_t(c) * m + fv.p0[f] + j] - acc;
}

// dst (c x r) = src (r x c)^T, row-major, 32 x 32 tiles through LDS
__global__ __launch_bounds__(kT) void k_transpose(int64_t r, int64_t c, const double *__restrict__ src,
                                                  double *__restrict__ dst)
{
    __shared__ double tile[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t r0 = int64_t(blockIdx.y) * 32, c0 = int64_t(blockIdx.x) * 32;
    for (int k = ty; k < 32; k += kT / 32)
        if (r0 + k < r && c0 + tx < c) tile[k][tx] = src[(r0 + k) * c + c0 + tx];
    __syncthreads();
    for (int k = ty; k < 32; k += kT / 32)
        if (c0 + k < c && r0 + tx < r) dst[(c0 + k) * r + r0 + tx] = tile[tx][k];
}

// wave per pivot row i: x_i = sum_{j >= i} G11(j, i) s_j (G11^T stored in the strict upper triangle)
template <int LPR>
__global__ __launch_bounds__(kT) void k_bwd_x(int64_t ntask, const int2 *__restrict__ tasks, FrontView fv,
                                              const double *__restrict__ G, const double *__restrict__ sbuf,
                                              int64_t rtot, double *__restrict__ xpos, int32_t m,
                                              const int *__restrict__ stop)
{
    if (dev::all_stopped(stop, gridDim.y)) return;
    const int64_t t = (int64_t(blockIdx.x) * kT + threadIdx.x) / LPR;
    if (t >= ntask) return;  // a whole group
    const int lane = threadIdx.x & (LPR - 1), c = blockIdx.y;
    const int2 task = tasks[t];
    const int f = task.x, i = task.y, p = fv.np[f];
    const double *row = G + fv.goff[f] + int64_t(i) * p;
    const double *v = sbuf + int64_t(c) * rtot + fv.roff[f];
    double acc = group_dot<LPR>(row, i, p, lane, [&](int j) { return v[j]; });
    acc = group_sum<LPR>(acc);
    if (lane == 0) xpos[int64_t(c) * m + fv.p0[f] + i] = acc;
}

__global__ __launch_bounds__(kT) void k_scatter_add(int32_t m, const int32_t *__restrict__ order,
                                                    const int32_t *__restrict__ pdofs,
                                                    const double *__restrict__ xpos, double *__restrict__ z,
                                                    float *__restrict__ z32, int64_t ld, const int *__restrict__ stop)
{
    if (dev::all_stopped(stop, gridDim.y)) return;
    const int32_t q = int32_t(blockIdx.x) * kT + int32_t(threadIdx.x);
    if (q >= m) return;
    const int c = blockIdx.y;
    const int64_t d = int64_t(c) * ld + pdofs[order[q]];
    const double v = z[d] + xpos[int64_t(c) * m + q];
    z[d] = v;
    if (z32) z32[d] = float(v);
}

// ---- batched factorization of a tree level's fronts ----

// one workgroup per front of the list: its original entries (distinct positions of its lower triangle)
__global__ __launch_bounds__(kT) void k_fb_assemble(const int32_t *__restrict__ fronts, const int64_t *__restrict__ toff,
                                                    const int32_t *__restrict__ li, const int32_t *__restrict__ lj,
                                                    const int64_t *__restrict__ src, const int64_t *__restrict__ wb,
                                                    const int32_t *__restrict__ np, const int32_t *__restrict__ nu,
                                                    const double *__restrict__ vals, double *__restrict__ W)
{
    const int32_t f = fronts[blockIdx.x];
    const int64_t m = int64_t(np[f]) + nu[f];
    double *F = W + wb[f];
    for (int64_t t = toff[f] + threadIdx.x; t < toff[f + 1]; t += kT) F[int64_t(li[t]) * m + lj[t]] = vals[src[t]];
}

// one wave per (child, update row); the children of one launch have distinct parents
__global__ __launch_bounds__(kT) void k_fb_extend(int64_t ntask, const ExtendTask *__restrict__ tasks,
                                                  const int32_t *__restrict__ parent, const int64_t *__restrict__ wb,
                                                  const int32_t *__restrict__ np, const int32_t *__restrict__ nu,
                                                  const int64_t *__restrict__ woff, const int64_t *__restrict__ uoff,
                                                  const int32_t *__restrict__ rel, const double *__restrict__ U,
                                                  double *__restrict__ W)
{
    const int64_t t = int64_t(blockIdx.x) * (kT / 64) + (threadIdx.x >> 6);
    if (t >= ntask) return;
    const int lane = threadIdx.x & 63;
    const ExtendTask task = tasks[t];
    const int32_t c = task.c, a = task.a, f = parent[c];
    const int64_t m = int64_t(np[f]) + nu[f];
    const int32_t *rc = rel + woff[c];
    const double *row = U + uoff[c] + int64_t(a) * nu[c];
    double *Fr = W + wb[f] + int64_t(rc[a]) * m;
    for (int32_t b = lane; b <= a; b += 64) Fr[rc[b]] += row[b];
}

// panel rows r0 .. r0 + nr - 1 below diagonal block k0: L_rk = A_rk G_kk^T (G_kk lower triangular,
// zero above its diagonal).  The workgroup holds its rows in LDS, so the product is in place.
__global__ __launch_bounds__(kT) void k_fb_panel(const PanelTask *__restrict__ tasks, double *__restrict__ W,
                                                 const double *__restrict__ G)
{
    constexpr int R = 32, NB = 128, KC = 16;
    __shared__ double As[R][NB + 1];
    __shared__ double Gs[KC][NB + 4];
    const PanelTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x;
    double *a = W + t.a + int64_t(t.r0) * t.lda + t.k0;
    const double *g = G + t.g + int64_t(t.k0) * t.ldg + t.k0;
    for (int q = tid; q < R * NB; q += kT) {
        const int i = q >> 7, c = q & (NB - 1);
        As[i][c] = (i < t.nr && c < t.kb) ? a[int64_t(i) * t.lda + c] : 0.0;
    }
    const int ty = tid >> 5, tx = tid & 31;  // rows 4 ty + u, columns tx + 32 v
    double acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
    for (int k0 = 0; k0 < t.kb; k0 += KC) {
        __syncthreads();
        for (int q = tid; q < KC * NB; q += kT) {
            const int c = q >> 4, kk = q & (KC - 1);
            Gs[kk][c] = (c < t.kb && k0 + kk < t.kb) ? g[int64_t(c) * t.ldg + k0 + kk] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KC; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) av[u] = As[4 * ty + u][k0 + kk];
#pragma unroll
            for (int v = 0; v < 4; ++v) bv[v] = Gs[kk][tx + 32 * v];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] += av[u] * bv[v];
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 4 * ty + u, c = tx + 32 * v;
            if (i < t.nr && c < t.kb) a[int64_t(i) * t.lda + c] = acc[u][v];
        }
}

// one 64 x 64 tile of C = beta C + alpha A op(B), K in chunks of 16 through LDS, 4 x 4 per thread
template <bool NT>
__global__ __launch_bounds__(kT) void k_fb_gemm(const TileTask *__restrict__ tasks, double alpha, double beta,
                                                const double *__restrict__ A, const double *__restrict__ B,
                                                double *__restrict__ C)
{
    constexpr int TM = 64, KC = 16, LD = TM + 4;
    __shared__ double As[KC][LD], Bs[KC][LD];
    const TileTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
    const double *a = A + t.a, *b = B + t.b;
    double acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
    for (int k0 = 0; k0 < t.K; k0 += KC) {
        for (int q = tid; q < TM * KC; q += kT) {
            const int i = q >> 4, kk = q & (KC - 1);
            As[kk][i] = (i < t.M && k0 + kk < t.K) ? a[int64_t(i) * t.lda + k0 + kk] : 0.0;
        }
        if (NT) {
            for (int q = tid; q < TM * KC; q += kT) {
                const int j = q >> 4, kk = q & (KC - 1);
                Bs[kk][j] = (j < t.N && k0 + kk < t.K) ? b[int64_t(j) * t.ldb + k0 + kk] : 0.0;
            }
        } else {
            for (int q = tid; q < TM * KC; q += kT) {
                const int kk = q >> 6, j = q & (TM - 1);
                Bs[kk][j] = (j < t.N && k0 + kk < t.K) ? b[int64_t(k0 + kk) * t.ldb + j] : 0.0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KC; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) av[u] = As[kk][4 * ty + u];
#pragma unroll
            for (int v = 0; v < 4; ++v) bv[v] = Bs[kk][4 * tx + v];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] += av[u] * bv[v];
        }
        __syncthreads();
    }
    double *c = C + t.c;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 4 * ty + u, j = 4 * tx + v;
            if (i < t.M && j < t.N && (!t.lower || j <= i)) {
                double *o = c + int64_t(i) * t.ldc + j;
                *o = beta == 0.0 ? alpha * acc[u][v] : beta * *o + alpha * acc[u][v];
            }
        }
}

__global__ __launch_bounds__(kT) void k_fb_copy(const CopyTask *__restrict__ tasks, const double *__restrict__ src,
                                                double *__restrict__ dst)
{
    __shared__ double tile[32][33];
    const CopyTask t = tasks[blockIdx.x];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const double *s = src + t.src;
    double *d = dst + t.dst;
    if (t.mode == 0) {
        for (int i = ty; i < t.rows; i += kT / 32)
            if (tx < t.cols) d[int64_t(i) * t.ldd + tx] = s[int64_t(i) * t.lds + tx];
        return;
    }
    for (int i = ty; i < 32; i += kT / 32) tile[i][tx] = (i < t.rows && tx < t.cols) ? s[int64_t(i) * t.lds + tx] : 0.0;
    __syncthreads();
    // destination row j (a source column), column i (a source row)
    for (int j = ty; j < t.cols; j += kT / 32)
        if (tx < t.rows && (t.mode == 1 || j < tx)) d[int64_t(j) * t.ldd + tx] = tile[tx][j];
}

int waves_grid(int64_t ntask) { return div_up(ntask, kT / 64); }

}  // namespace

void sparse_gather_pattern(int32_t m, const int32_t *dofs, const int32_t *ptr, const int64_t *coff,
                           const int32_t *idx, int32_t *out, hipStream_t s)
{
    if (m == 0) return;
    k_gather_pattern<<<div_up(m, kT / 64), kT, 0, s>>>(m, dofs, ptr, coff, idx, out);
    LSS_LAUNCH_CHECK();
}

void sparse_assemble(int64_t ntrip, const int32_t *li, const int32_t *lj, const int64_t *src, const double *vals,
                     double *F, int64_t ld, hipStream_t s)
{
    if (ntrip == 0) return;
    k_assemble<<<div_up(ntrip, kT), kT, 0, s>>>(ntrip, li, lj, src, vals, F, ld);
    LSS_LAUNCH_CHECK();
}

void sparse_extend_add(int32_t u, const int32_t *rel, const double *U, double *F, int64_t ld, hipStream_t s)
{
    if (u == 0) return;
    k_extend_add<<<div_up(u, kT / 64), kT, 0, s>>>(u, rel, U, F, ld);
    LSS_LAUNCH_CHECK();
}

void sparse_permute_in(int32_t m, const int32_t *order, const double *t, int64_t ldt, double *tp, int k,
                       const int *stop, hipStream_t s)
{
    if (m == 0) return;
    k_permute_in<<<dim3(div_up(m, kT), k), kT, 0, s>>>(m, order, t, ldt, tp, stop);
    LSS_LAUNCH_CHECK();
}

void sparse_fwd_rhs(int64_t ntask, const int2 *tasks, const FrontView &fv, const double *tp, int32_t m,
                    const double *w, int64_t wtot, double *rhs, int64_t rtot, int k, const int *stop, hipStream_t s)
{
    if (ntask == 0) return;
    k_fwd_rhs<<<dim3(div_up(ntask, kT), k), kT, 0, s>>>(ntask, tasks, fv, tp, m, w, wtot, rhs, rtot, stop);
    LSS_LAUNCH_CHECK();
}

void sparse_fwd_y(int64_t ntask, const int2 *tasks, const FrontView &fv, const double *G, const double *rhs,
                  int64_t rtot, double *ypos, int32_t m, int k, const int *stop, int lpr, hipStream_t s)
{
    if (ntask == 0) return;
    const dim3 grid(div_up(ntask * lpr, kT), k);
    switch (lpr) {
        case 4: k_fwd_y<4><<<grid, kT, 0, s>>>(ntask, tasks, fv, G, rhs, rtot, ypos, m, stop); break;
        case 8: k_fwd_y<8><<<grid, kT, 0, s>>>(ntask, tasks, fv, G, rhs, rtot, ypos, m, stop); break;
        case 16: k_fwd_y<16><<<grid, kT, 0, s>>>(ntask, tasks, fv, G, rhs, rtot, ypos, m, stop); break;
        case 32: k_fwd_y<32><<<grid, kT, 0, s>>>(ntask, tasks, fv, G, rhs, rtot, ypos, m, stop); break;
        case 64: k_fwd_y<64><<<grid, kT, 0, s>>>(ntask, tasks, fv, G, rhs, rtot, ypos, m, stop); break;
        default: fail(LSS_ERR_ARG, "sparse solve: lanes per row must be 4 .. 64");
    }
    LSS_LAUNCH_CHECK();
}

void sparse_fwd_w(int64_t ntask, const int2 *tasks, const FrontView &fv, const double *L, const double *rhs,
                  int64_t rtot, const double *ypos, int32_t m, double *w, int64_t wtot, int k, const int *stop,
                  int lpr, hipStream_t s)
{
    if (ntask == 0) return;
    const dim3 grid(div_up(ntask * lpr, kT), k);
    switch (lpr) {
        case 4: k_fwd_w<4><<<grid, kT, 0, s>>>(ntask, tasks, fv, L, rhs, rtot, ypos, m, w, wtot, stop); break;
        case 8: k_fwd_w<8><<<grid, kT, 0, s>>>(ntask, tasks, fv, L, rhs, rtot, ypos, m, w, wtot, stop); break;
        case 16: k_fwd_w<16><<<grid, kT, 0, s>>>(ntask, tasks, fv, L, rhs, rtot, ypos, m, w, wtot, stop); break;
        case 32: k_fwd_w<32><<<grid, kT, 0, s>>>(ntask, tasks, fv, L, rhs, rtot, ypos, m, w, wtot, stop); break;
        case 64: k_fwd_w<64><<<grid, kT, 0, s>>>(ntask, tasks, fv, L, rhs, rtot, ypos, m, w, wtot, stop); break;
        default: fail(LSS_ERR_ARG, "sparse solve: lanes per row must be 4 .. 64");
    }
    LSS_LAUNCH_CHECK();
}

void sparse_bwd_s(int64_t ntask, const int2 *tasks, const FrontView &fv, const double *Lt, const double *xu,
                  int64_t wtot, const double *ypos, int32_t m, double *sbuf, int64_t rtot, int k, const int *stop,
                  int lpr, hipStream_t s)
{
    if (ntask == 0) return;
    const dim3 grid(div_up(ntask * lpr, kT), k);
    switch (lpr) {
        case 4: k_bwd_s<4><<<grid, kT, 0, s>>>(ntask, tasks, fv, Lt, xu, wtot, ypos, m, sbuf, rtot, stop); break;
        case 8: k_bwd_s<8><<<grid, kT, 0, s>>>(ntask, tasks, fv, Lt, xu, wtot, ypos, m, sbuf, rtot, stop); break;
        case 16: k_bwd_s<16><<<grid, kT, 0, s>>>(ntask, tasks, fv, Lt, xu, wtot, ypos, m, sbuf, rtot, stop); break;
        case 32: k_bwd_s<32><<<grid, kT, 0, s>>>(ntask, tasks, fv, Lt, xu, wtot, ypos, m, sbuf, rtot, stop); break;
        case 64: k_bwd_s<64><<<grid, kT, 0, s>>>(ntask, tasks, fv, Lt, xu, wtot, ypos, m, sbuf, rtot, stop); break;
        default: fail(LSS_ERR_ARG, "sparse solve: lanes per row must be 4 .. 64");
    }
    LSS_LAUNCH_CHECK();
}

void sparse_bwd_gather(int64_t ntask, const int2 *tasks, const FrontView &fv, const double *xpos, int32_t m,
                       double *xu, int64_t wtot, int k, const int *stop, hipStream_t s)
{
    if (ntask == 0) return;
    k_bwd_gather<<<dim3(div_up(ntask, kT), k), kT, 0, s>>>(ntask, tasks, fv, xpos, m, xu, wtot, stop);
    LSS_LAUNCH_CHECK();
}

void sparse_transpose(int64_t r, int64_t c, const double *src, double *dst, hipStream_t s)
{
    if (r == 0 || c == 0) return;
    k_transpose<<<dim3(div_up(c, 32), div_up(r, 32)), kT, 0, s>>>(r, c, src, dst);
    LSS_LAUNCH_CHECK();
}

void sparse_bwd_x(int64_t ntask, const int2 *tasks, const FrontView &fv, const double *G, const double *sbuf,
                  int64_t rtot, double *xpos, int32_t m, int k, const int *stop, int lpr, hipStream_t s)
{
    if (ntask == 0) return;
    const dim3 grid(div_up(ntask * lpr, kT), k);
    switch (lpr) {
        case 4: k_bwd_x<4><<<grid, kT, 0, s>>>(ntask, tasks, fv, G, sbuf, rtot, xpos, m, stop); break;
        case 8: k_bwd_x<8><<<grid, kT, 0, s>>>(ntask, tasks, fv, G, sbuf, rtot, xpos, m, stop); break;
        case 16: k_bwd_x<16><<<grid, kT, 0, s>>>(ntask, tasks, fv, G, sbuf, rtot, xpos, m, stop); break;
        case 32: k_bwd_x<32><<<grid, kT, 0, s>>>(ntask, tasks, fv, G, sbuf, rtot, xpos, m, stop); break;
        case 64: k_bwd_x<64><<<grid, kT, 0, s>>>(ntask, tasks, fv, G, sbuf, rtot, xpos, m, stop); break;
        default: fail(LSS_ERR_ARG, "sparse solve: lanes per row must be 4 .. 64");
    }
    LSS_LAUNCH_CHECK();
}

void sparse_scatter_add(int32_t m, const int32_t *order, const int32_t *pdofs, const double *xpos, double *z,
                        float *z32, int64_t ld, int k, const int *stop, hipStream_t s)
{
    if (m == 0) return;
    k_scatter_add<<<dim3(div_up(m, kT), k), kT, 0, s>>>(m, order, pdofs, xpos, z, z32, ld, stop);
    LSS_LAUNCH_CHECK();
}

void front_panel_batch(int32_t ntask, const PanelTask *tasks, double *W, const double *G, hipStream_t s)
{
    if (ntask == 0) return;
    k_fb_panel<<<ntask, kT, 0, s>>>(tasks, W, G);
    LSS_LAUNCH_CHECK();
}

void front_gemm_batch(int32_t ntask, const TileTask *tasks, bool nt, double alpha, double beta, const double *A,
                      const double *B, double *C, hipStream_t s)
{
    if (ntask == 0) return;
    if (nt) k_fb_gemm<true><<<ntask, kT, 0, s>>>(tasks, alpha, beta, A, B, C);
    else k_fb_gemm<false><<<ntask, kT, 0, s>>>(tasks, alpha, beta, A, B, C);
    LSS_LAUNCH_CHECK();
}

void front_copy_batch(int32_t ntask, const CopyTask *tasks, const double *src, double *dst, hipStream_t s)
{
    if (ntask == 0) return;
    k_fb_copy<<<ntask, kT, 0, s>>>(tasks, src, dst);
    LSS_LAUNCH_CHECK();
}

void front_extend_batch(int64_t ntask, const ExtendTask *tasks, const int32_t *parent, const int64_t *wb,
                        const int32_t *np, const int32_t *nu, const int64_t *woff, const int64_t *uoff,
                        const int32_t *rel, const double *U, double *W, hipStream_t s)
{
    if (ntask == 0) return;
    k_fb_extend<<<waves_grid(ntask), kT, 0, s>>>(ntask, tasks, parent, wb, np, nu, woff, uoff, rel, U, W);
    LSS_LAUNCH_CHECK();
}

void front_assemble_batch(int32_t nfronts, const int32_t *fronts, const int64_t *toff, const int32_t *li,
                          const int32_t *lj, const int64_t *src, const int64_t *wb, const int32_t *np,
                          const int32_t *nu, const double *vals, double *W, hipStream_t s)
{
    if (nfronts == 0) return;
    k_fb_assemble<<<nfronts, kT, 0, s>>>(fronts, toff, li, lj, src, wb, np, nu, vals, W);
    LSS_LAUNCH_CHECK();
}

}  // namespace lss::k
