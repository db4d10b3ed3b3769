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
    kZeroRhs = 6,  // b = 0: x = 0 without iterating
};

// the PCG stopped on a breakdown of the recurrence (not a convergence)
inline __host__ __device__ bool is_breakdown(int code) { return code >= kZeroSdotp && code <= kSubnormalAlpha; }

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

// A solve over k right-hand sides keeps k PcgStates back to back; the stop word of
// column c sits kStateInts ints after that of column c - 1.
constexpr int kStateInts = int(sizeof(PcgState) / sizeof(int));
static_assert(sizeof(PcgState) % sizeof(int) == 0, "PcgState must be a whole number of ints");

// Work shared by all columns is skipped once every column has stopped.
__device__ __forceinline__ bool all_stopped(const int *stop, int k)
{
    if (!stop) return false;
    for (int c = 0; c < k; ++c)
        if (!stop[c * kStateInts]) return false;
    return true;
}

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

// Grid reductions are split in two launches: every workgroup of the producer
// kernel stores its partial (plain store), and a one-workgroup finisher kernel
// adds the partials in index order and applies the scalar update.  The kernel
// boundary makes the partials visible; a per-workgroup agent-scope release
// (needed by a single-kernel "last workgroup" scheme) writes back the XCD's L2
// on gfx950 and costs more than the extra launch.
enum FinishOp : int {
    kFinishStore = 0,  // *out = total
    kFinishPq = 1,     // p.Ap: alpha, breakdown checks, iteration count
    kFinishRr = 2,     // r.r: convergence test
    kFinishRzFirst = 3,  // r.z before the loop: gamma = gamma_old
    kFinishRz = 4,     // r.z in the loop: beta
    // MAS solver PCG (MASSolver.cu:618-744): no breakdown stops, quotients of tiny
    // denominators are 0 (scalar_division, MASSolver.cu:39-50)
    kFinishPqMas = 5,  // alpha = |p.Ap| < 1e-20 ? 0 : rz / p.Ap
    kFinishRzMas = 6,  // rz_old = rz; rz = r.z; beta = |rz_old| < 1e-20 ? 0 : rz / rz_old
};

__device__ __forceinline__ void store_partial(double block_total, double *partials)
{
    if (threadIdx.x == 0) partials[blockIdx.x] = block_total;
}

}  // namespace lss::dev
