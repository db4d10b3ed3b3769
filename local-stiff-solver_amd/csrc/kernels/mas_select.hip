// Indicators and sorts for the stiff-DOF selection of the MAS solver
// (MAS/BadDOFPreconditioner.cu:61-143, MAS/BadDofKMeanPreconditioner.cu:178-226,
// MAS/BadDofGMMPreconditioner.cu:260-310).  The reference copies the
// partition-permuted BSR to the host and loops over it; here one thread per
// permuted scalar row walks its block row in the same order (blocks by
// ascending permuted column, then the block's columns), so the sums are the
// reference's sequential sums.  Sorts are stable radix sorts (equal values keep
// ascending positions), which fixes the order std::sort leaves open.
#include <hipcub/hipcub.hpp>

#include "kernels.hpp"
#include "../common/runtime.hpp"

namespace lss::k {
namespace {

// MODE 0: sum |a_ij| over the row (GMM strategy); 1: |a_ii| (elbow); 2: max |a_ij|, j != i (k-means).
// Padding scalars (block pblock, components >= bs - npad) are skipped as rows and columns;
// a padding row writes 0 and the caller drops it.
template <int MODE>
__global__ void k_indicator(int32_t nb, int bs, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ cols,
                            const double *__restrict__ vals, int32_t pblock, int npad, double *__restrict__ out)
{
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= int64_t(nb) * bs) return;
    const int32_t I = int32_t(p / bs);
    const int a = int(p - int64_t(I) * bs);
    const int bb = bs * bs;
    if (I == pblock && a >= bs - npad) {
        out[p] = 0.0;
        return;
    }
    double v = 0.0;
    for (int32_t e = rowptr[I]; e < rowptr[I + 1]; ++e) {
        const int32_t J = cols[e];
        const double *blk = vals + int64_t(e) * bb + a * bs;
        if (MODE == 1) {
            if (J == I) {
                v = fabs(blk[a]);
                break;
            }
            continue;
        }
        for (int c = 0; c < bs; ++c) {
            if (J == pblock && c >= bs - npad) continue;
            if (MODE == 0) v += fabs(blk[c]);
            if (MODE == 2 && !(J == I && c == a)) v = fmax(v, fabs(blk[c]));
        }
    }
    out[p] = v;
}

__global__ void k_iota(int32_t n, int32_t *__restrict__ a)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

}  // namespace

void mas_indicator(int mode, int32_t nb, int bs, const int32_t *rowptr, const int32_t *cols, const double *vals,
                   int32_t pblock, int npad, double *out, hipStream_t s)
{
    const int64_t m = int64_t(nb) * bs;
    if (m == 0) return;
    const int grid = int(div_up(m, 256));
    switch (mode) {
        case 0: k_indicator<0><<<grid, 256, 0, s>>>(nb, bs, rowptr, cols, vals, pblock, npad, out); break;
        case 1: k_indicator<1><<<grid, 256, 0, s>>>(nb, bs, rowptr, cols, vals, pblock, npad, out); break;
        case 2: k_indicator<2><<<grid, 256, 0, s>>>(nb, bs, rowptr, cols, vals, pblock, npad, out); break;
        default: fail(LSS_ERR_ARG, "mas_indicator: unknown mode");
    }
    LSS_LAUNCH_CHECK();
}

void sort_positions(int32_t n, const double *keys, bool descending, double *keys_out, int32_t *pos_out,
                    hipStream_t s)
{
    if (n == 0) return;
    DBuf<int32_t> pos(n);
    k_iota<<<div_up(n, 256), 256, 0, s>>>(n, pos.get());
    LSS_LAUNCH_CHECK();
    size_t tmp_bytes = 0;
    if (descending)
        LSS_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp_bytes, keys, keys_out, pos.get(), pos_out,
                                                             n, 0, 64, s));
    else
        LSS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys_out, pos.get(), pos_out, n, 0, 64,
                                                   s));
    DBuf<uint8_t> tmp(std::max<size_t>(tmp_bytes, 1));
    if (descending)
        LSS_HIP(hipcub::DeviceRadixSort::SortPairsDescending(tmp.get(), tmp_bytes, keys, keys_out, pos.get(), pos_out,
                                                             n, 0, 64, s));
    else
        LSS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.get(), tmp_bytes, keys, keys_out, pos.get(), pos_out, n, 0, 64,
                                                   s));
    LSS_HIP(hipStreamSynchronize(s));  // tmp and pos are released at scope end
}

}  // namespace lss::k
