// Blocked dense Cholesky and inverse factor of one large stiff component.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lss {

// Holds the rocBLAS handle used by the blocked path (created on first use).
class DenseBlas {
public:
    DenseBlas() = default;
    ~DenseBlas();
    DenseBlas(const DenseBlas &) = delete;
    DenseBlas &operator=(const DenseBlas &) = delete;

    // Workspace (doubles) dense_factor_inverse needs for an m x m component.
    static int64_t work_size(int64_t m);
    // A: m x m row-major symmetric (lower triangle read), overwritten by L (A = L L^T,
    // lower; the strict upper triangle is left unspecified).  G (m x m) receives
    // L^{-1} in the lower triangle and its transpose above, the layout of the
    // one-workgroup path.  status |= 2 when a diagonal block is not positive definite.
    void factor_inverse(int64_t m, double *A, double *G, double *work, int *status, hipStream_t s);
    // Partial factorization of a front (m x m row-major, lower triangle read): the first p
    // columns become [L11; L21], the trailing block A22 - L21 L21^T (lower triangle), and
    // G (leading dimension m) receives L11^{-1} in its leading p x p lower triangle; its
    // first p rows are overwritten (zero above the diagonal).
    void partial_factor_inverse(int64_t m, int64_t p, double *A, double *G, double *work, int *status,
                                hipStream_t s);

private:
    void *handle_ = nullptr;   // rocblas_handle on the caller's stream
    void *handle2_ = nullptr;  // rocblas_handle on s2_: the next diagonal block and panel (look-ahead)
    hipStream_t s2_ = nullptr;
    hipEvent_t ev_col_ = nullptr, ev_pan_ = nullptr;
    void init();
    void inverse(int64_t a, int64_t b, int64_t m, const double *L, double *G, double *work, hipStream_t s);
};

}  // namespace lss
