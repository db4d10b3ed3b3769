// BLAS-1 and PCG kernels with device-resident scalars.
//
// The reference keeps alpha/beta/gamma on the host and synchronizes on every
// dot product (GPUHybrid.cu:1455-1511); here every reduction ends in a
// one-workgroup finisher kernel that adds the per-workgroup partials in index
// order (deterministic) and applies the scalar update of GPUHybrid.cu:1461-1507
// on the device, so the PCG loop issues no host sync.
//
// Several right-hand sides: column c of a vector starts at v + c * ld, the
// partials of column c at partials + c * count, its state at st + c.  Vector
// kernels run one grid row (blockIdx.y) per column, finishers one workgroup
// per column, so every column sees exactly the arithmetic of a single solve.
#include "kernels.hpp"
#include "../common/runtime.hpp"

namespace lss::k {
namespace {

constexpr int kT = 256;

inline int red_grid(int64_t n) { return int(std::min<int64_t>(std::max<int64_t>(div_up(n, kT * 4), 1), 4096)); }

__global__ __launch_bounds__(kT) void k_dot(int64_t n, int64_t ld, const double *__restrict__ a,
                                            const double *__restrict__ b, double *partials)
{
    __shared__ double sh[kT / 64];
    a += blockIdx.y * ld;
    b += blockIdx.y * ld;
    double v = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT) v += a[i] * b[i];
    dev::store_partial(dev::block_sum<kT>(v, sh), partials + int64_t(blockIdx.y) * gridDim.x);
}

constexpr int kFinT = 1024;

// Sum `count` partials in index order and apply the scalar update of `op`.
__global__ __launch_bounds__(kFinT) void k_finish(int count, const double *__restrict__ partials, int op,
                                                  dev::PcgState *st, double *out, int iter)
{
    partials += int64_t(blockIdx.x) * count;
    if (st) st += blockIdx.x;
    if (out) out += blockIdx.x;
    if (op != dev::kFinishStore && st->stop) return;
    __shared__ double sh[kFinT / 64];
    double v = 0.0;
    for (int i = threadIdx.x; i < count; i += kFinT) v += partials[i];
    const double total = dev::block_sum<kFinT>(v, sh);
    if (threadIdx.x != 0) return;
    switch (op) {
        case dev::kFinishStore: *out = total; break;
        case dev::kFinishPq: {
            // GPUHybrid.cu:1458-1480
            st->iters = iter + 1;
            st->pq = total;
            if (total == 0.0) {
                st->stop = dev::kZeroSdotp;
            } else {
                const double alpha = st->gamma / total;
                st->alpha = alpha;
                if (alpha <= 0.0)
                    st->stop = dev::kNegativeAlpha;
                else if (alpha < 2.2250738585072014e-308)
                    st->stop = dev::kSubnormalAlpha;
            }
            break;
        }
        case dev::kFinishRr:
            // GPUHybrid.cu:1484-1496
            st->rr = total;
            if (st->rel_eps > 0 && (total / st->bb) < st->rel_eps)
                st->stop = dev::kConvergedRel;
            else if (st->abs_eps > 0 && total < st->abs_eps)
                st->stop = dev::kConvergedAbs;
            break;
        case dev::kFinishRzFirst:
            // GPUHybrid.cu:1450-1451
            st->gamma = total;
            st->gamma_old = total;
            break;
        case dev::kFinishRz:
            // GPUHybrid.cu:1501-1504
            st->gamma = total;
            st->beta = total / st->gamma_old;
            st->gamma_old = total;
            break;
        case dev::kFinishPqMas:
            st->iters = iter + 1;
            st->pq = total;
            st->alpha = fabs(total) < 1e-20 ? 0.0 : st->gamma / total;
            break;
        case dev::kFinishRzMas:
            st->gamma_old = st->gamma;
            st->gamma = total;
            st->beta = fabs(st->gamma_old) < 1e-20 ? 0.0 : total / st->gamma_old;
            break;
    }
}

__global__ void k_axpy_alpha(int64_t n, double sign, const double *__restrict__ x, double *__restrict__ y,
                             const dev::PcgState *st)
{
    const double a = sign * st->alpha;
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT)
        y[i] = a * x[i] + 1.0 * y[i];
}

__global__ __launch_bounds__(kT) void k_update_xr(int64_t n, int64_t ld, double *__restrict__ x,
                                                  double *__restrict__ r, const double *__restrict__ p,
                                                  const double *__restrict__ q, dev::PcgState *st, double *partials)
{
    const int64_t o = blockIdx.y * ld;
    x += o, r += o, p += o, q += o;
    st += blockIdx.y;
    partials += int64_t(blockIdx.y) * gridDim.x;
    if (st->stop) return;
    __shared__ double sh[kT / 64];
    const double alpha = st->alpha;
    double v = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT) {
        x[i] = x[i] + alpha * p[i];
        const double ri = r[i] + (-alpha) * q[i];
        r[i] = ri;
        v += ri * ri;
    }
    dev::store_partial(dev::block_sum<kT>(v, sh), partials);
}

__global__ __launch_bounds__(kT) void k_update_p(int64_t n, int64_t ld, double *__restrict__ p,
                                                 const double *__restrict__ z, const dev::PcgState *st)
{
    p += blockIdx.y * ld;
    z += blockIdx.y * ld;
    st += blockIdx.y;
    if (st->stop) return;
    const double beta = st->beta;
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT)
        p[i] = beta * p[i] + z[i];
}

__global__ void k_copy(int64_t n, const double *__restrict__ a, double *__restrict__ b)
{
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT) b[i] = a[i];
}
__global__ void k_fill(int64_t n, double *__restrict__ a, double v)
{
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT) a[i] = v;
}
__global__ void k_scale(int64_t n, double *__restrict__ a, double v)
{
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT) a[i] *= v;
}

__global__ void k_axpy(int64_t n, double a, const double *__restrict__ x, double *__restrict__ y)
{
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT) y[i] += a * x[i];
}
__global__ void k_xpby(int64_t n, const double *__restrict__ x, double b, double *__restrict__ y)
{
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT)
        y[i] = x[i] + b * y[i];
}

__global__ void k_permute_in(int32_t nb, int bs, int64_t n, const int32_t *__restrict__ perm,
                             const double *__restrict__ src, double *__restrict__ dst)
{
    const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (i >= int64_t(nb) * bs) return;
    const int32_t old = int32_t(i / bs);
    const int c = int(i - int64_t(old) * bs);
    dst[int64_t(perm[old]) * bs + c] = i < n ? src[i] : 0.0;
}

__global__ void k_permute_out(int32_t nb, int bs, int64_t n, const int32_t *__restrict__ perm,
                              const double *__restrict__ src, double *__restrict__ dst)
{
    const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (i >= n) return;
    const int32_t old = int32_t(i / bs);
    const int c = int(i - int64_t(old) * bs);
    dst[i] = src[int64_t(perm[old]) * bs + c];
}

inline int ew_grid(int64_t n) { return int(std::min<int64_t>(std::max<int64_t>(div_up(n, kT * 4), 1), 65536)); }

}  // namespace

void finish(int count, const double *partials, int op, dev::PcgState *st, double *out, int iter, hipStream_t s,
            int k)
{
    if (count <= 0 || k <= 0) return;
    k_finish<<<k, kFinT, 0, s>>>(count, partials, op, st, out, iter);
    LSS_LAUNCH_CHECK();
}

int partials_per_column(int64_t n) { return red_grid(n); }

int dot_partials(int64_t n, const double *a, const double *b, double *partials, hipStream_t s)
{
    const int g = red_grid(n);
    k_dot<<<dim3(g, 1), kT, 0, s>>>(n, 0, a, b, partials);
    LSS_LAUNCH_CHECK();
    return g;
}

void dot(int64_t n, const double *a, const double *b, double *partials, double *out, hipStream_t s, int k,
         int64_t ld)
{
    const int g = red_grid(n);
    k_dot<<<dim3(g, k), kT, 0, s>>>(n, ld, a, b, partials);
    LSS_LAUNCH_CHECK();
    finish(g, partials, dev::kFinishStore, nullptr, out, 0, s, k);
}

void pcg_update_xr(int64_t n, double *x, double *r, const double *p, const double *q, dev::PcgState *st,
                   double *partials, hipStream_t s, int k, int64_t ld)
{
    const int g = red_grid(n);
    k_update_xr<<<dim3(g, k), kT, 0, s>>>(n, ld, x, r, p, q, st, partials);
    LSS_LAUNCH_CHECK();
    finish(g, partials, dev::kFinishRr, st, nullptr, 0, s, k);
}

void pcg_update_p(int64_t n, double *p, const double *z, const dev::PcgState *st, hipStream_t s, int k, int64_t ld)
{
    k_update_p<<<dim3(ew_grid(n), k), kT, 0, s>>>(n, ld, p, z, st);
    LSS_LAUNCH_CHECK();
}

void axpy_alpha(int64_t n, double sign, const double *x, double *y, const dev::PcgState *st, hipStream_t s)
{
    if (n <= 0) return;
    k_axpy_alpha<<<ew_grid(n), kT, 0, s>>>(n, sign, x, y, st);
    LSS_LAUNCH_CHECK();
}

void copy(int64_t n, const double *a, double *b, hipStream_t s)
{
    if (n <= 0) return;
    k_copy<<<ew_grid(n), kT, 0, s>>>(n, a, b);
    LSS_LAUNCH_CHECK();
}
void fill(int64_t n, double *a, double v, hipStream_t s)
{
    if (n <= 0) return;
    k_fill<<<ew_grid(n), kT, 0, s>>>(n, a, v);
    LSS_LAUNCH_CHECK();
}
void scale(int64_t n, double *a, double v, hipStream_t s)
{
    if (n <= 0) return;
    k_scale<<<ew_grid(n), kT, 0, s>>>(n, a, v);
    LSS_LAUNCH_CHECK();
}

void axpy(int64_t n, double a, const double *x, double *y, hipStream_t s)
{
    if (n <= 0) return;
    k_axpy<<<ew_grid(n), kT, 0, s>>>(n, a, x, y);
    LSS_LAUNCH_CHECK();
}
void xpby(int64_t n, const double *x, double b, double *y, hipStream_t s)
{
    if (n <= 0) return;
    k_xpby<<<ew_grid(n), kT, 0, s>>>(n, x, b, y);
    LSS_LAUNCH_CHECK();
}

void permute_in(int32_t nb, int bs, int64_t n, const int32_t *perm, const double *src, double *dst, hipStream_t s)
{
    const int64_t m = int64_t(nb) * bs;
    if (m == 0) return;
    k_permute_in<<<div_up(m, kT), kT, 0, s>>>(nb, bs, n, perm, src, dst);
    LSS_LAUNCH_CHECK();
}

void permute_out(int32_t nb, int bs, int64_t n, const int32_t *perm, const double *src, double *dst, hipStream_t s)
{
    if (n == 0) return;
    k_permute_out<<<div_up(n, kT), kT, 0, s>>>(nb, bs, n, perm, src, dst);
    LSS_LAUNCH_CHECK();
}

}  // namespace lss::k
