// Adjoint of the solve A x = b for L(x): with A^T lam = dL/dx,
//   dL/db = lam,   dL/dA_ij = -lam_i x_j  on the stored pattern,
// and the chain rule to per-element coefficients through the sparse Jacobian
// d vals / d coef (one segment of (slot, value) pairs per element).
// Both kernels are HBM-bound gathers; every output is written by exactly one
// wave in a fixed order, so results are bit-reproducible.
#include "kernels.hpp"
#include "../common/runtime.hpp"

namespace lss::k {
namespace {

constexpr int kAdjThreads = 256;
constexpr int kWavesPerBlock = kAdjThreads / dev::kWave;

// One wave per CSC column j; CSC column j is operator row j (the arrays are
// read as rows, GPUHybrid.cu:401-433), so entry e of column j is A(j, rowidx[e]).
__global__ __launch_bounds__(kAdjThreads) void k_pattern_grad(int32_t n, const int32_t *__restrict__ colptr,
                                                             const int32_t *__restrict__ rowidx,
                                                             const double *__restrict__ lam,
                                                             const double *__restrict__ x, double *__restrict__ grad)
{
    const int lane = threadIdx.x & (dev::kWave - 1);
    const int64_t j = int64_t(blockIdx.x) * kWavesPerBlock + (threadIdx.x / dev::kWave);
    if (j >= n) return;
    const double l = -lam[j];
    const int32_t lo = colptr[j], hi = colptr[j + 1];
    for (int32_t e = lo + lane; e < hi; e += dev::kWave) grad[e] = l * x[rowidx[e]];
}

// out[s] = sum_k coef[k] * g[slot[k]] over segment s; one wave per segment,
// lane-strided partial sums combined by the fixed butterfly of wave_sum.
__global__ __launch_bounds__(kAdjThreads) void k_segment_dot(int64_t nseg, const int64_t *__restrict__ ptr,
                                                            const int32_t *__restrict__ slot,
                                                            const double *__restrict__ coef,
                                                            const double *__restrict__ g, double *__restrict__ out)
{
    const int lane = threadIdx.x & (dev::kWave - 1);
    const int64_t s = int64_t(blockIdx.x) * kWavesPerBlock + (threadIdx.x / dev::kWave);
    if (s >= nseg) return;
    double acc = 0.0;
    for (int64_t k = ptr[s] + lane; k < ptr[s + 1]; k += dev::kWave) acc += coef[k] * g[slot[k]];
    acc = dev::wave_sum(acc);
    if (lane == 0) out[s] = acc;
}

}  // namespace

void pattern_grad(int32_t n, const int32_t *colptr, const int32_t *rowidx, const double *lam, const double *x,
                  double *grad, hipStream_t s)
{
    if (n <= 0) return;
    k_pattern_grad<<<div_up(int64_t(n), kWavesPerBlock), kAdjThreads, 0, s>>>(n, colptr, rowidx, lam, x, grad);
    LSS_LAUNCH_CHECK();
}

void segment_dot(int64_t nseg, const int64_t *ptr, const int32_t *slot, const double *coef, const double *g,
                 double *out, hipStream_t s)
{
    if (nseg <= 0) return;
    k_segment_dot<<<div_up(nseg, int64_t(kWavesPerBlock)), kAdjThreads, 0, s>>>(nseg, ptr, slot, coef, g, out);
    LSS_LAUNCH_CHECK();
}

}  // namespace lss::k
