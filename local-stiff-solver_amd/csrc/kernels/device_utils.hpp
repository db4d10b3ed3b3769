// Device helpers shared by the gfx950 kernels: wave64 reductions, the
// deterministic "last workgroup finishes" grid reduction, and the device-side
// PCG state that replaces the reference's per-iteration host round trips
// (GPUHybrid.cu:1455-1511 reads every scalar back to the host).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lss::dev {

constexpr int kWave = 64;

// Stop codes of the PCG loop (GPUHybrid.cu:1463-1496).
enum StopCode : int {
    kRunning = 0,
    kConvergedRel = 1,
    kConvergedAbs = 2,
    kZeroSdotp = 3,
    kNegativeAlpha = 4,
    kSubnormalAlpha = 5,
};

// Scalars of one PCG instance; lives in device memory, mirrored to pinned host
// memory when the host polls.
struct PcgState {
    double bb;         // b . b
    double rel_eps;    // rtol^2
    double abs_eps;    // atol^2
    double gamma;      // r . z
    double gamma_old;
    double pq;         // p . A p
    double alpha;
    double beta;
    double rr;         // r . r (recurrence residual)
    int stop;          // StopCode
    int iters;         // num_iterations (k + 1 of the last started iteration)
    unsigned ticket[4];
    double scratch[4];
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ float wave_sumf(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Sum over the workgroup in a fixed order; the result is valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < NT / kWave; ++w) r += sh[w];
    }
    __syncthreads();
    return r;
}

__device__ __forceinline__ double load_acquire(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Deterministic grid-wide sum: every workgroup publishes its partial, the last
// one to arrive (by ticket) adds the partials in index order.  Returns true in
// the last workgroup, where *total (thread 0) holds the sum.  The ticket is
// reset by the last workgroup so the kernel can be launched again.
template <int NT>
__device__ __forceinline__ bool grid_sum_last(double partial, double *partials, unsigned *ticket, double *sh,
                                              double &total)
{
    __shared__ bool is_last;
    if (threadIdx.x == 0) {
        __hip_atomic_store(&partials[blockIdx.x], partial, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        is_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return false;
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
    double v = 0.0;
    for (unsigned i = threadIdx.x; i < gridDim.x; i += NT) v += load_acquire(&partials[i]);
    total = block_sum<NT>(v, sh);
    if (threadIdx.x == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

}  // namespace lss::dev
