// Launchers of the small setup kernels (util.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_utils.hpp"

namespace lss::k {

// Copy rows `rows[0..m)` of a CSR into a compact CSR with offsets out_ptr.
void gather_rows(int64_t m, const int32_t *rows, const int32_t *ptr, const int32_t *idx, const double *vals,
                 const int64_t *out_ptr, int32_t *out_idx, double *out_vals, hipStream_t s);
// g2l[dof] = local index inside its component, comp_of[dof] = component (-1 elsewhere).
void stiff_maps(int32_t ncomp, int32_t n, const int64_t *comp_off, const int32_t *dofs, int32_t *g2l, int32_t *comp_of,
                hipStream_t s);
// out[i] = u32(i * 2654435761) / 2^32 - 0.5 (start vector of the lambda_max estimate)
void hash_vector(int64_t n, double *out, hipStream_t s);
void diag_inverse(int32_t nb, int bs, const int32_t *rowptr, const int32_t *cols, const double *vals, double *dinv,
                  hipStream_t s);
// z = scale * D^{-1} r (+ add); with st: per-workgroup partials of dotv . z.
// k right-hand sides at stride ld.  Returns the number of partials per column.
int jacobi_apply(int64_t n, const double *dinv, const double *r, double scale, const double *add, double *z,
                 const double *dotv, dev::PcgState *st, double *partials, const int *stop, hipStream_t s, int k = 1,
                 int64_t ld = 0);

}  // namespace lss::k
