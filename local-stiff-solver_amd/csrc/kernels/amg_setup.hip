// Setup of the AMG base preconditioner on the device: l1 norms, classical
// strength, PMIS coarsening, extended+i interpolation truncated to pmax entries,
// transposes and the Galerkin product R (A P).  The reference builds it with
// hypre's BoomerAMG (GPUHybrid.cu:285-339, 373); the methods are the published
// ones (strength of connection and max_row_sum as hypre's classical measure;
// PMIS: De Sterck, Yang, Heys 2006; extended+i: De Sterck, Falgout, Nolting,
// Yang 2008; l1 smoothing: Baker, Falgout, Kolev, Yang 2011).
//
// Every row-wise quantity is computed by one thread in the row's storage order,
// and every entry of a sparse product is the sum, from 0.0, of its products in
// Gustavson order (the order of X's row, then of Y's rows): the expand step
// writes products in that order, a stable radix sort groups them by (row, column)
// without reordering equal keys, and one thread sums each group in sequence.
// With no contraction of products into sums, the hierarchy is bit-identical to a
// sequential restatement of the same steps.
#include <hipcub/hipcub.hpp>

#include "kernels.hpp"
#include "../common/runtime.hpp"

#pragma clang fp contract(off)

namespace lss::k {
namespace {

constexpr int kT = 256;
// largest interpolatory set C-hat_i of one row (3D elasticity: <= 28)
constexpr int kChatMax = 128;

int grid_of(int64_t n) { return int(std::max<int64_t>(1, std::min<int64_t>(div_up(n, kT), 1 << 20))); }

int bits_for(int64_t v)
{
    int b = 1;
    while ((int64_t(1) << b) <= v) ++b;
    return b;
}

__device__ __forceinline__ double point_random(int64_t i)
{
    uint64_t z = uint64_t(i) + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return double(z >> 11) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ double diag_of(const int64_t *ptr, const int32_t *col, const double *val, int32_t i)
{
    for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e)
        if (col[e] == i) return val[e];
    return 0.0;
}

__global__ __launch_bounds__(kT) void k_widen(int64_t n, const int32_t *__restrict__ a, int64_t *__restrict__ b)
{
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT) b[i] = a[i];
}

__global__ __launch_bounds__(kT) void k_l1_inverse(int32_t n, const int64_t *__restrict__ ptr,
                                                   const double *__restrict__ val, double *__restrict__ dinv)
{
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT) {
        double s = 0.0;
        for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) s += fabs(val[e]);
        dinv[i] = s > 0.0 ? 1.0 / s : 0.0;
    }
}

// classical strength: FILL = false counts the strong entries of each row, true writes them
template <bool FILL>
__global__ __launch_bounds__(kT) void k_strength(int32_t n, const int64_t *__restrict__ ptr,
                                                 const int32_t *__restrict__ col, const double *__restrict__ val,
                                                 const int32_t *__restrict__ func, int nf, double theta,
                                                 double max_row_sum, int64_t *__restrict__ cnt,
                                                 const int64_t *__restrict__ sptr, int32_t *__restrict__ scol)
{
    for (int64_t ii = int64_t(blockIdx.x) * kT + threadIdx.x; ii < n; ii += int64_t(gridDim.x) * kT) {
        const int32_t i = int32_t(ii);
        const double diag = diag_of(ptr, col, val, i);
        double scale = 0.0, row_sum = diag;
        const int32_t fi = nf > 1 ? func[i] : 0;
        for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) {
            const int32_t j = col[e];
            if (j == i || (nf > 1 && func[j] != fi)) continue;
            const double a = val[e];
            scale = diag < 0.0 ? fmax(scale, a) : fmin(scale, a);
            row_sum += a;
        }
        int64_t c = 0;
        int64_t w = FILL ? sptr[i] : 0;
        if (!(max_row_sum < 1.0 && fabs(row_sum) > fabs(diag) * max_row_sum)) {
            const double cut = theta * scale;
            for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) {
                const int32_t j = col[e];
                if (j == i || (nf > 1 && func[j] != fi)) continue;
                const double a = val[e];
                if (diag < 0.0 ? a > cut : a < cut) {
                    if constexpr (FILL) scol[w++] = j;
                    ++c;
                }
            }
        }
        if constexpr (!FILL) cnt[i] = c;
    }
}

// row counts of an expanded product X Y
__global__ __launch_bounds__(kT) void k_gemm_count(int32_t n, const int64_t *__restrict__ xptr,
                                                   const int32_t *__restrict__ xcol, const int64_t *__restrict__ yptr,
                                                   int64_t *__restrict__ cnt)
{
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT) {
        int64_t c = 0;
        for (int64_t e = xptr[i]; e < xptr[i + 1]; ++e) {
            const int32_t k = xcol[e];
            c += yptr[k + 1] - yptr[k];
        }
        cnt[i] = c;
    }
}

// products of row i in Gustavson order, keyed (i, j)
__global__ __launch_bounds__(kT) void k_gemm_expand(int32_t n, const int64_t *__restrict__ xptr,
                                                    const int32_t *__restrict__ xcol, const double *__restrict__ xval,
                                                    const int64_t *__restrict__ yptr, const int32_t *__restrict__ ycol,
                                                    const double *__restrict__ yval, const int64_t *__restrict__ off,
                                                    int cb, uint64_t *__restrict__ key, double *__restrict__ prod)
{
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT) {
        int64_t o = off[i];
        const uint64_t hi = uint64_t(i) << cb;
        for (int64_t e = xptr[i]; e < xptr[i + 1]; ++e) {
            const int32_t k = xcol[e];
            const double a = xval[e];
            for (int64_t f = yptr[k]; f < yptr[k + 1]; ++f, ++o) {
                key[o] = hi | uint64_t(uint32_t(ycol[f]));
                prod[o] = a * yval[f];
            }
        }
    }
}

// head[t] = 1 where a run of equal keys starts
__global__ __launch_bounds__(kT) void k_heads(int64_t t, const uint64_t *__restrict__ key, int32_t *__restrict__ head)
{
    for (int64_t q = int64_t(blockIdx.x) * kT + threadIdx.x; q < t; q += int64_t(gridDim.x) * kT)
        head[q] = (q == 0 || key[q] != key[q - 1]) ? 1 : 0;
}

// one thread per run: the run's values summed in order from 0.0 (or the single
// value when SUM is false), the run's row and column
template <bool SUM>
__global__ __launch_bounds__(kT) void k_compress(int64_t t, const uint64_t *__restrict__ key,
                                                 const double *__restrict__ v, const int32_t *__restrict__ uidx,
                                                 int cb, int32_t *__restrict__ orow, int32_t *__restrict__ ocol,
                                                 double *__restrict__ oval)
{
    const uint64_t mask = (uint64_t(1) << cb) - 1;
    for (int64_t q = int64_t(blockIdx.x) * kT + threadIdx.x; q < t; q += int64_t(gridDim.x) * kT) {
        const uint64_t k0 = key[q];
        if (q > 0 && key[q - 1] == k0) continue;
        const int64_t u = int64_t(uidx[q]) - 1;
        orow[u] = int32_t(k0 >> cb);
        ocol[u] = int32_t(k0 & mask);
        if (oval) {
            if constexpr (SUM) {
                double s = 0.0;
                for (int64_t r = q; r < t && key[r] == k0; ++r) s += v[r];
                oval[u] = s;
            } else {
                oval[u] = v[q];
            }
        }
    }
}

// row pointers of a row-sorted entry list: ptr[r] = first entry of row >= r
__global__ __launch_bounds__(kT) void k_rowptr(int64_t u, int32_t n, const int32_t *__restrict__ row,
                                               int64_t *__restrict__ ptr)
{
    for (int64_t q = int64_t(blockIdx.x) * kT + threadIdx.x; q <= u; q += int64_t(gridDim.x) * kT) {
        const int32_t r0 = q == 0 ? -1 : row[q - 1];
        const int32_t r1 = q == u ? n : row[q];
        for (int32_t r = r0 + 1; r <= r1; ++r) ptr[r] = q;
    }
}

// transpose keys: (column, row) of every entry, value index kept
__global__ __launch_bounds__(kT) void k_transpose_keys(int32_t n, const int64_t *__restrict__ ptr,
                                                       const int32_t *__restrict__ col, int rb,
                                                       uint64_t *__restrict__ key)
{
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT)
        for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) key[e] = (uint64_t(uint32_t(col[e])) << rb) | uint64_t(i);
}

__global__ __launch_bounds__(kT) void k_iota64(int64_t n, double *__restrict__ v)
{
    // entry positions carried through a sort as doubles (exact below 2^53)
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT) v[i] = double(i);
}

__global__ __launch_bounds__(kT) void k_gather_vals(int64_t n, const double *__restrict__ pos,
                                                    const double *__restrict__ src, double *__restrict__ dst)
{
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT)
        dst[i] = src[int64_t(pos[i])];
}

// ---------------------------------------------------------------- PMIS
__global__ __launch_bounds__(kT) void k_pmis_init(int32_t n, const int64_t *__restrict__ sptr,
                                                  const int64_t *__restrict__ tptr, double *__restrict__ meas,
                                                  int8_t *__restrict__ cf)
{
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT) {
        const int64_t infl = tptr[i + 1] - tptr[i];
        meas[i] = double(infl) + point_random(i);
        int8_t c = 0;
        if (sptr[i + 1] == sptr[i] && infl == 0) c = -3;
        else if (infl == 0) c = -1;
        cf[i] = c;
    }
}

__device__ __forceinline__ bool beats(const double *meas, int32_t j, int32_t i)
{
    return meas[j] > meas[i] || (meas[j] == meas[i] && j > i);
}

__global__ __launch_bounds__(kT) void k_pmis_select(int32_t n, const int64_t *__restrict__ sptr,
                                                    const int32_t *__restrict__ scol, const int64_t *__restrict__ tptr,
                                                    const int32_t *__restrict__ tcol, const double *__restrict__ meas,
                                                    const int8_t *__restrict__ cf, uint8_t *__restrict__ newc)
{
    for (int64_t ii = int64_t(blockIdx.x) * kT + threadIdx.x; ii < n; ii += int64_t(gridDim.x) * kT) {
        const int32_t i = int32_t(ii);
        if (cf[i] != 0) continue;
        bool top = true;
        for (int64_t e = sptr[i]; e < sptr[i + 1] && top; ++e) {
            const int32_t j = scol[e];
            if (cf[j] == 0 && beats(meas, j, i)) top = false;
        }
        for (int64_t e = tptr[i]; e < tptr[i + 1] && top; ++e) {
            const int32_t j = tcol[e];
            if (cf[j] == 0 && beats(meas, j, i)) top = false;
        }
        newc[i] = top ? 1 : 0;
    }
}

__global__ __launch_bounds__(kT) void k_pmis_setc(int32_t n, uint8_t *__restrict__ newc, int8_t *__restrict__ cf)
{
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT)
        if (newc[i]) {
            cf[i] = 1;
            newc[i] = 0;
        }
}

// undecided points that depend strongly on a C point become F (only C marks are
// read, and only non-C marks are written); the rest are counted
__global__ __launch_bounds__(kT) void k_pmis_setf(int32_t n, const int64_t *__restrict__ sptr,
                                                  const int32_t *__restrict__ scol, int8_t *__restrict__ cf,
                                                  int *__restrict__ left)
{
    int mine = 0;
    for (int64_t ii = int64_t(blockIdx.x) * kT + threadIdx.x; ii < n; ii += int64_t(gridDim.x) * kT) {
        const int32_t i = int32_t(ii);
        if (cf[i] != 0) continue;
        bool f = false;
        for (int64_t e = sptr[i]; e < sptr[i + 1] && !f; ++e) f = cf[scol[e]] == 1;
        if (f) cf[i] = -1;
        else ++mine;
    }
    if (mine) atomicAdd(left, mine);
}

// ---------------------------------------------------------------- interpolation
__global__ __launch_bounds__(kT) void k_is_c(int32_t n, const int8_t *__restrict__ cf, int32_t *__restrict__ flag)
{
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT)
        flag[i] = cf[i] == 1 ? 1 : 0;
}

__global__ __launch_bounds__(kT) void k_coarse_func(int32_t n, const int8_t *__restrict__ cf,
                                                    const int32_t *__restrict__ cidx, const int32_t *__restrict__ func,
                                                    int32_t *__restrict__ cfunc)
{
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT)
        if (cf[i] == 1) cfunc[cidx[i]] = func ? func[i] : 0;
}

// One thread per row: the row's at most pmax interpolation weights into
// pcol/pval[i * pmax ...] and their count; err is raised when C-hat_i exceeds kChatMax.
__global__ __launch_bounds__(kT) void k_interp(int32_t n, const int64_t *__restrict__ ptr,
                                               const int32_t *__restrict__ col, const double *__restrict__ val,
                                               const int64_t *__restrict__ sptr, const int32_t *__restrict__ scol,
                                               const int8_t *__restrict__ cf, const int32_t *__restrict__ cidx,
                                               const int32_t *__restrict__ func, int nf, int pmax,
                                               int32_t *__restrict__ pcnt, int32_t *__restrict__ pcol,
                                               double *__restrict__ pval, int *__restrict__ err)
{
    int32_t chat[kChatMax], srt[kChatMax], sidx[kChatMax];
    double w[kChatMax];
    for (int64_t ii = int64_t(blockIdx.x) * kT + threadIdx.x; ii < n; ii += int64_t(gridDim.x) * kT) {
        const int32_t i = int32_t(ii);
        int32_t *oc = pcol + int64_t(i) * pmax;
        double *ov = pval + int64_t(i) * pmax;
        if (cf[i] == 1) {
            oc[0] = cidx[i];
            ov[0] = 1.0;
            pcnt[i] = 1;
            continue;
        }
        pcnt[i] = 0;
        if (cf[i] == -3 || sptr[i + 1] == sptr[i]) continue;
        // C-hat_i in discovery order: strong neighbours ascending, each strong F
        // neighbour followed by its strong C neighbours
        int m = 0;
        bool over = false;
        auto add = [&](int32_t j) {
            // insert j into the sorted copy unless present
            int lo = 0, hi = m;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (srt[mid] < j) lo = mid + 1;
                else hi = mid;
            }
            if (lo < m && srt[lo] == j) return;
            if (m == kChatMax) {
                over = true;
                return;
            }
            for (int q = m; q > lo; --q) {
                srt[q] = srt[q - 1];
                sidx[q] = sidx[q - 1];
            }
            srt[lo] = j;
            sidx[lo] = m;
            chat[m] = j;
            w[m] = 0.0;
            ++m;
        };
        for (int64_t e = sptr[i]; e < sptr[i + 1]; ++e) {
            const int32_t j = scol[e];
            if (cf[j] == 1) {
                add(j);
            } else {
                for (int64_t f = sptr[j]; f < sptr[j + 1]; ++f) {
                    const int32_t l = scol[f];
                    if (cf[l] == 1) add(l);
                }
            }
        }
        if (over) {
            atomicOr(err, 1);
            continue;
        }
        // Rows are sorted by column, as is srt: membership in C-hat_i is found by
        // walking srt beside each row (the sums keep the rows' order).
        double diag = 0.0;
        const int32_t fi = nf > 1 ? func[i] : 0;
        int64_t se = sptr[i];  // walks S_i (a sorted subsequence of row i) beside the row
        const int64_t send = sptr[i + 1];
        int ri = 0;
        for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) {
            const int32_t j = col[e];
            const double a = val[e];
            while (se < send && scol[se] < j) ++se;
            const bool strong = se < send && scol[se] == j;
            while (ri < m && srt[ri] < j) ++ri;
            if (j == i) {
                diag += a;
            } else if (ri < m && srt[ri] == j) {
                w[sidx[ri]] += a;
            } else if (strong && cf[j] != 1) {
                const double akk = diag_of(ptr, col, val, j);
                double d = 0.0;
                int r = 0;
                for (int64_t f = ptr[j]; f < ptr[j + 1]; ++f) {
                    const int32_t l = col[f];
                    while (r < m && srt[r] < l) ++r;
                    if ((l == i || (r < m && srt[r] == l)) && val[f] * akk < 0.0) d += val[f];
                }
                if (d == 0.0) {
                    diag += a;
                    continue;
                }
                const double fac = a / d;
                r = 0;
                for (int64_t f = ptr[j]; f < ptr[j + 1]; ++f) {
                    const int32_t l = col[f];
                    const double v = val[f];
                    while (r < m && srt[r] < l) ++r;
                    if (v * akk >= 0.0) continue;
                    if (l == i) diag += fac * v;
                    else if (r < m && srt[r] == l) w[sidx[r]] += fac * v;
                }
            } else if (nf == 1 || func[j] == fi) {
                diag += a;
            }
        }
        if (diag == 0.0 || m == 0) continue;
        for (int q = 0; q < m; ++q) w[q] = -w[q] / diag;
        if (m > pmax) {
            // the pmax largest |w| (ties: smaller point first), the row sum kept by one scale
            double all = 0.0;
            for (int q = 0; q < m; ++q) all += w[q];
            int pick[8];
            double kept = 0.0;
            for (int r = 0; r < pmax; ++r) {
                int best = -1;
                for (int q = 0; q < m; ++q) {
                    bool used = false;
                    for (int t = 0; t < r; ++t) used |= pick[t] == q;
                    if (used) continue;
                    if (best < 0) {
                        best = q;
                        continue;
                    }
                    const double x = fabs(w[q]), y = fabs(w[best]);
                    if (x > y || (x == y && chat[q] < chat[best])) best = q;
                }
                pick[r] = best;
                kept += w[best];
            }
            const double sc = kept != 0.0 ? all / kept : 1.0;
            // kept entries by ascending point
            for (int a = 1; a < pmax; ++a)
                for (int b = a; b > 0 && chat[pick[b - 1]] > chat[pick[b]]; --b) {
                    const int t = pick[b];
                    pick[b] = pick[b - 1];
                    pick[b - 1] = t;
                }
            for (int r = 0; r < pmax; ++r) {
                oc[r] = cidx[chat[pick[r]]];
                ov[r] = w[pick[r]] * sc;
            }
            pcnt[i] = pmax;
        } else {
            for (int r = 0; r < m; ++r) {
                oc[r] = cidx[srt[r]];
                ov[r] = w[sidx[r]];
            }
            pcnt[i] = m;
        }
    }
}

__global__ __launch_bounds__(kT) void k_interp_pack(int32_t n, int pmax, const int64_t *__restrict__ ptr,
                                                    const int32_t *__restrict__ pcol, const double *__restrict__ pval,
                                                    int32_t *__restrict__ col, double *__restrict__ val)
{
    for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kT)
        for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) {
            col[e] = pcol[i * pmax + (e - ptr[i])];
            val[e] = pval[i * pmax + (e - ptr[i])];
        }
}

// ---------------------------------------------------------------- helpers
template <class T>
T *grown(DBuf<T> &b, size_t n)
{
    if (b.size() < n) b.resize(n);  // grow only: the workspace keeps its largest size
    return b.get();
}

// out[0..n] = exclusive prefix sums of in[0..n) (out[n] = total); returns the total
template <class T>
T exclusive_scan(const T *in, T *out, int64_t n, AmgWork &w, hipStream_t s)
{
    size_t bytes = 0;
    LSS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n + 1, s));
    LSS_HIP(hipcub::DeviceScan::ExclusiveSum(grown(w.tmp, std::max<size_t>(bytes, 1)), bytes, in, out, n + 1, s));
    T total = 0;
    LSS_HIP(hipMemcpyAsync(&total, out + n, sizeof(T), hipMemcpyDeviceToHost, s));
    LSS_HIP(hipStreamSynchronize(s));
    return total;
}

// sort the t (key, value) pairs of w.key / w.val by the low `bits` bits of the key,
// stably; the result is left in w.key / w.val
void sort_pairs(int64_t t, AmgWork &w, int bits, hipStream_t s)
{
    if (t >= (int64_t(1) << 31)) fail(LSS_ERR_UNSUPPORTED, "AMG setup: more than 2^31 entries in one product");
    uint64_t *k2 = grown(w.key2, size_t(t));
    double *v2 = grown(w.val2, size_t(t));
    size_t bytes = 0;
    LSS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, w.key.get(), k2, w.val.get(), v2, int(t), 0, bits, s));
    LSS_HIP(hipcub::DeviceRadixSort::SortPairs(grown(w.tmp, std::max<size_t>(bytes, 1)), bytes, w.key.get(), k2,
                                               w.val.get(), v2, int(t), 0, bits, s));
    std::swap(w.key, w.key2);
    std::swap(w.val, w.val2);
}

// the t sorted (row, column) keys of w.key -> CSR; SUM: the values of equal keys
// (w.val) summed in order; otherwise no values are written
void keys_to_csr(int64_t t, AmgWork &w, bool sum, int cb, int32_t n, int32_t m, AmgCsrBuf &Z, hipStream_t s)
{
    Z.n = n;
    Z.m = m;
    Z.ptr.resize(size_t(n) + 1);
    if (t == 0) {
        Z.nnz = 0;
        Z.col.resize(1);
        Z.val.resize(1);
        k_rowptr<<<1, kT, 0, s>>>(0, n, nullptr, Z.ptr.get());
        LSS_LAUNCH_CHECK();
        LSS_HIP(hipStreamSynchronize(s));
        return;
    }
    int32_t *head = grown(w.head, size_t(t)), *uidx = grown(w.uidx, size_t(t));
    k_heads<<<grid_of(t), kT, 0, s>>>(t, w.key.get(), head);
    LSS_LAUNCH_CHECK();
    size_t bytes = 0;
    LSS_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, head, uidx, int(t), s));
    LSS_HIP(hipcub::DeviceScan::InclusiveSum(grown(w.tmp, std::max<size_t>(bytes, 1)), bytes, head, uidx, int(t), s));
    int32_t u = 0;
    LSS_HIP(hipMemcpyAsync(&u, uidx + (t - 1), sizeof(int32_t), hipMemcpyDeviceToHost, s));
    LSS_HIP(hipStreamSynchronize(s));
    Z.nnz = u;
    int32_t *row = grown(w.row, size_t(u));
    Z.col.resize(std::max<size_t>(size_t(u), 1));
    Z.val.resize(std::max<size_t>(size_t(u), 1));
    if (sum)
        k_compress<true><<<grid_of(t), kT, 0, s>>>(t, w.key.get(), w.val.get(), uidx, cb, row, Z.col.get(),
                                                    Z.val.get());
    else
        k_compress<false><<<grid_of(t), kT, 0, s>>>(t, w.key.get(), nullptr, uidx, cb, row, Z.col.get(), nullptr);
    LSS_LAUNCH_CHECK();
    k_rowptr<<<grid_of(int64_t(u) + 1), kT, 0, s>>>(u, n, row, Z.ptr.get());
    LSS_LAUNCH_CHECK();
    LSS_HIP(hipStreamSynchronize(s));
}

}  // namespace

void amg_widen_ptr(int64_t count, const int32_t *p32, int64_t *p64, hipStream_t s)
{
    if (count <= 0) return;
    k_widen<<<grid_of(count), kT, 0, s>>>(count, p32, p64);
    LSS_LAUNCH_CHECK();
}

void amg_l1_inverse(const AmgCsr &A, double *dinv, hipStream_t s)
{
    if (A.n <= 0) return;
    k_l1_inverse<<<grid_of(A.n), kT, 0, s>>>(A.n, A.ptr, A.val, dinv);
    LSS_LAUNCH_CHECK();
}

void amg_strength_device(const AmgCsr &A, const int32_t *func, int nf, double theta, double max_row_sum,
                         AmgCsrBuf &S, AmgWork &w, hipStream_t s)
{
    const int32_t n = A.n;
    int64_t *cnt = grown(w.cnt, size_t(n) + 1);
    k_strength<false><<<grid_of(n), kT, 0, s>>>(n, A.ptr, A.col, A.val, func, nf, theta, max_row_sum, cnt, nullptr,
                                                 nullptr);
    LSS_LAUNCH_CHECK();
    S.n = S.m = n;
    S.ptr.resize(size_t(n) + 1);
    S.nnz = exclusive_scan<int64_t>(cnt, S.ptr.get(), n, w, s);
    S.col.resize(std::max<size_t>(size_t(S.nnz), 1));
    S.val.release();
    k_strength<true><<<grid_of(n), kT, 0, s>>>(n, A.ptr, A.col, A.val, func, nf, theta, max_row_sum, nullptr,
                                               S.ptr.get(), S.col.get());
    LSS_LAUNCH_CHECK();
}

void amg_transpose_device(const AmgCsr &P, int32_t m, AmgCsrBuf &R, AmgWork &w, hipStream_t s)
{
    const int64_t t = P.nnz;
    const int rb = bits_for(P.n);
    const int cb = bits_for(m);
    if (t > 0) {
        k_transpose_keys<<<grid_of(P.n), kT, 0, s>>>(P.n, P.ptr, P.col, rb, grown(w.key, size_t(t)));
        LSS_LAUNCH_CHECK();
        k_iota64<<<grid_of(t), kT, 0, s>>>(t, grown(w.val, size_t(t)));
        LSS_LAUNCH_CHECK();
        sort_pairs(t, w, rb + cb, s);
    }
    // keys are (column of P, row of P): the rows of R, their entries by ascending row
    // of P; w.val holds each entry's position in P
    keys_to_csr(t, w, false, rb, m, P.n, R, s);
    if (P.val && t > 0) {
        k_gather_vals<<<grid_of(t), kT, 0, s>>>(t, w.val.get(), P.val, R.val.get());
        LSS_LAUNCH_CHECK();
    }
    LSS_HIP(hipStreamSynchronize(s));
}

void amg_spgemm_device(const AmgCsr &X, const AmgCsr &Y, int32_t m, AmgCsrBuf &Z, AmgWork &w, hipStream_t s)
{
    const int32_t n = X.n;
    int64_t *cnt = grown(w.cnt, size_t(n) + 1), *off = grown(w.off, size_t(n) + 1);
    k_gemm_count<<<grid_of(n), kT, 0, s>>>(n, X.ptr, X.col, Y.ptr, cnt);
    LSS_LAUNCH_CHECK();
    const int64_t t = exclusive_scan<int64_t>(cnt, off, n, w, s);
    const int cb = bits_for(m);
    if (t > 0) {
        k_gemm_expand<<<grid_of(n), kT, 0, s>>>(n, X.ptr, X.col, X.val, Y.ptr, Y.col, Y.val, off, cb,
                                                grown(w.key, size_t(t)), grown(w.val, size_t(t)));
        LSS_LAUNCH_CHECK();
        sort_pairs(t, w, cb + bits_for(n), s);
    }
    keys_to_csr(t, w, true, cb, n, m, Z, s);
}

int32_t amg_pmis_device(const AmgCsr &S, int8_t *cf, AmgWork &w, hipStream_t s)
{
    const int32_t n = S.n;
    AmgCsrBuf T;
    amg_transpose_device(AmgCsr{n, S.ptr, S.col, nullptr, S.nnz}, n, T, w, s);
    DBuf<double> meas(std::max<size_t>(size_t(n), 1));
    DBuf<uint8_t> newc(std::max<size_t>(size_t(n), 1));
    DBuf<int> left(1);
    newc.zero(s);
    k_pmis_init<<<grid_of(n), kT, 0, s>>>(n, S.ptr, T.ptr.get(), meas.get(), cf);
    LSS_LAUNCH_CHECK();
    for (int round = 0;; ++round) {
        if (round > n + 1) fail(LSS_ERR_UNSUPPORTED, "AMG setup: PMIS did not terminate");
        k_pmis_select<<<grid_of(n), kT, 0, s>>>(n, S.ptr, S.col, T.ptr.get(), T.col.get(), meas.get(), cf, newc.get());
        LSS_LAUNCH_CHECK();
        k_pmis_setc<<<grid_of(n), kT, 0, s>>>(n, newc.get(), cf);
        LSS_LAUNCH_CHECK();
        left.zero(s);
        k_pmis_setf<<<grid_of(n), kT, 0, s>>>(n, S.ptr, S.col, cf, left.get());
        LSS_LAUNCH_CHECK();
        int h = 0;
        LSS_HIP(hipMemcpyAsync(&h, left.get(), sizeof(int), hipMemcpyDeviceToHost, s));
        LSS_HIP(hipStreamSynchronize(s));
        if (h == 0) break;
    }
    DBuf<int32_t> flag(static_cast<size_t>(n) + 1), cidx(static_cast<size_t>(n) + 1);
    k_is_c<<<grid_of(n), kT, 0, s>>>(n, cf, flag.get());
    LSS_LAUNCH_CHECK();
    return exclusive_scan<int32_t>(flag.get(), cidx.get(), n, w, s);
}

void amg_interp_device(const AmgCsr &A, const AmgCsr &S, const int8_t *cf, const int32_t *func, int nf, int pmax,
                       AmgCsrBuf &P, int32_t *cfunc, AmgWork &w, hipStream_t s)
{
    const int32_t n = A.n;
    if (pmax < 1 || pmax > 8) fail(LSS_ERR_CONFIG, "AMG setup: pmax must be 1..8");
    DBuf<int32_t> flag(static_cast<size_t>(n) + 1), cidx(static_cast<size_t>(n) + 1);
    k_is_c<<<grid_of(n), kT, 0, s>>>(n, cf, flag.get());
    LSS_LAUNCH_CHECK();
    const int32_t nc = exclusive_scan<int32_t>(flag.get(), cidx.get(), n, w, s);
    flag.release();
    DBuf<int32_t> pcnt(static_cast<size_t>(n) + 1), pcol(static_cast<size_t>(n) * pmax);
    DBuf<double> pval(static_cast<size_t>(n) * pmax);
    DBuf<int> err(1);
    err.zero(s);
    k_interp<<<grid_of(n), kT, 0, s>>>(n, A.ptr, A.col, A.val, S.ptr, S.col, cf, cidx.get(), func, nf, pmax,
                                       pcnt.get(), pcol.get(), pval.get(), err.get());
    LSS_LAUNCH_CHECK();
    int h = 0;
    LSS_HIP(hipMemcpyAsync(&h, err.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    LSS_HIP(hipStreamSynchronize(s));
    if (h) fail(LSS_ERR_UNSUPPORTED, "AMG setup: an interpolatory set C-hat_i exceeds 128 points");
    DBuf<int64_t> cnt64(static_cast<size_t>(n) + 1);
    k_widen<<<grid_of(n), kT, 0, s>>>(n, pcnt.get(), cnt64.get());
    LSS_LAUNCH_CHECK();
    P.n = n;
    P.m = nc;
    P.ptr.resize(size_t(n) + 1);
    P.nnz = exclusive_scan<int64_t>(cnt64.get(), P.ptr.get(), n, w, s);
    P.col.resize(std::max<size_t>(size_t(P.nnz), 1));
    P.val.resize(std::max<size_t>(size_t(P.nnz), 1));
    k_interp_pack<<<grid_of(n), kT, 0, s>>>(n, pmax, P.ptr.get(), pcol.get(), pval.get(), P.col.get(), P.val.get());
    LSS_LAUNCH_CHECK();
    if (cfunc) {
        k_coarse_func<<<grid_of(n), kT, 0, s>>>(n, cf, cidx.get(), func, cfunc);
        LSS_LAUNCH_CHECK();
    }
    LSS_HIP(hipStreamSynchronize(s));
}

}  // namespace lss::k
