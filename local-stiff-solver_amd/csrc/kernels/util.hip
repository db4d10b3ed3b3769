// Small setup kernels: row gathers for the host graph steps, stiff index maps,
// the power-iteration start vector and the Jacobi base preconditioner.
#include "util_kernels.hpp"
#include "../common/runtime.hpp"

namespace lss::k {
namespace {

__global__ void k_gather_rows(int64_t m, const int32_t *__restrict__ rows, const int32_t *__restrict__ ptr,
                              const int32_t *__restrict__ idx, const double *__restrict__ vals,
                              const int64_t *__restrict__ out_ptr, int32_t *__restrict__ out_idx,
                              double *__restrict__ out_vals)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int32_t r = rows[k];
    int64_t w = out_ptr[k];
    for (int32_t e = ptr[r]; e < ptr[r + 1]; ++e, ++w) {
        out_idx[w] = idx[e];
        if (out_vals) out_vals[w] = vals[e];
    }
}

__global__ void k_stiff_maps(int32_t ncomp, const int64_t *__restrict__ comp_off, const int32_t *__restrict__ dofs,
                             int32_t *__restrict__ g2l, int32_t *__restrict__ comp_of)
{
    const int32_t c = blockIdx.x;
    const int64_t off = comp_off[c];
    const int32_t m = int32_t(comp_off[c + 1] - off);
    for (int32_t r = threadIdx.x; r < m; r += blockDim.x) {
        g2l[dofs[off + r]] = r;
        comp_of[dofs[off + r]] = c;
    }
}

__global__ void k_fill_i32(int64_t n, int32_t *__restrict__ a, int32_t v)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_hash_vector(int64_t n, double *__restrict__ out)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = double(uint32_t(uint32_t(i) * 2654435761u)) * 0x1p-32 - 0.5;
}

__global__ void k_diag_inverse(int32_t nb, int bs, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ cols,
                               const double *__restrict__ vals, double *__restrict__ dinv)
{
    const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= int64_t(nb) * bs) return;
    const int32_t I = int32_t(g / bs);
    const int a = int(g - int64_t(I) * bs);
    double d = 0.0;
    for (int32_t e = rowptr[I]; e < rowptr[I + 1]; ++e)
        if (cols[e] == I) d = vals[int64_t(e) * bs * bs + a * bs + a];
    dinv[g] = d != 0.0 ? 1.0 / d : 1.0;
}

constexpr int kT = 256;

// grid row y = column y (vectors at stride ld, partials at stride gridDim.x)
template <bool DOT>
__global__ __launch_bounds__(kT) void k_jacobi(int64_t n, const double *__restrict__ dinv, const double *__restrict__ r,
                                               double scale, const double *__restrict__ add, double *__restrict__ z,
                                               const double *__restrict__ dotv, int64_t ld, double *partials,
                                               const int *__restrict__ stop)
{
    if (dev::all_stopped(stop, gridDim.y)) return;
    const int64_t o = blockIdx.y * ld;
    r += o, z += o;
    if (add) add += o;
    if (dotv) dotv += o;
    partials += int64_t(blockIdx.y) * gridDim.x;
    double v = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT) {
        double y = scale * (dinv[i] * r[i]);
        if (add) y += add[i];
        z[i] = y;
        if constexpr (DOT) v += dotv[i] * y;
    }
    if constexpr (DOT) {
        __shared__ double sh[kT / 64];
        dev::store_partial(dev::block_sum<kT>(v, sh), partials);
    }
}

}  // namespace

void gather_rows(int64_t m, const int32_t *rows, const int32_t *ptr, const int32_t *idx, const double *vals,
                 const int64_t *out_ptr, int32_t *out_idx, double *out_vals, hipStream_t s)
{
    if (m == 0) return;
    k_gather_rows<<<div_up(m, 256), 256, 0, s>>>(m, rows, ptr, idx, vals, out_ptr, out_idx, out_vals);
    LSS_LAUNCH_CHECK();
}

void stiff_maps(int32_t ncomp, int32_t n, const int64_t *comp_off, const int32_t *dofs, int32_t *g2l, int32_t *comp_of,
                hipStream_t s)
{
    k_fill_i32<<<div_up(n, 256), 256, 0, s>>>(n, g2l, -1);
    LSS_LAUNCH_CHECK();
    k_fill_i32<<<div_up(n, 256), 256, 0, s>>>(n, comp_of, -1);
    LSS_LAUNCH_CHECK();
    if (ncomp == 0) return;
    k_stiff_maps<<<ncomp, 256, 0, s>>>(ncomp, comp_off, dofs, g2l, comp_of);
    LSS_LAUNCH_CHECK();
}

void hash_vector(int64_t n, double *out, hipStream_t s)
{
    if (n == 0) return;
    k_hash_vector<<<div_up(n, 256), 256, 0, s>>>(n, out);
    LSS_LAUNCH_CHECK();
}

void diag_inverse(int32_t nb, int bs, const int32_t *rowptr, const int32_t *cols, const double *vals, double *dinv,
                  hipStream_t s)
{
    const int64_t m = int64_t(nb) * bs;
    if (m == 0) return;
    k_diag_inverse<<<div_up(m, 256), 256, 0, s>>>(nb, bs, rowptr, cols, vals, dinv);
    LSS_LAUNCH_CHECK();
}

int jacobi_apply(int64_t n, const double *dinv, const double *r, double scale, const double *add, double *z,
                 const double *dotv, dev::PcgState *st, double *partials, const int *stop, hipStream_t s, int k,
                 int64_t ld)
{
    const int grid = int(std::min<int64_t>(std::max<int64_t>(div_up(n, kT * 4), 1), 4096));
    if (st)
        k_jacobi<true><<<dim3(grid, k), kT, 0, s>>>(n, dinv, r, scale, add, z, dotv, ld, partials, stop);
    else
        k_jacobi<false><<<dim3(grid, k), kT, 0, s>>>(n, dinv, r, scale, add, z, dotv, ld, partials, stop);
    LSS_LAUNCH_CHECK();
    return grid;
}

}  // namespace lss::k
