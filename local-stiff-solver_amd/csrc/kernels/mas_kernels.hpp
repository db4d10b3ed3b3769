// Launchers of the MAS preconditioner and stiff-block kernels.
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "device_utils.hpp"

namespace lss::k {

// Storage precision of the MAS bank inverses.
enum class Precision : int { FP64 = 0, FP32 = 1, BF16 = 2, FP16 = 3 };

inline size_t precision_bytes(Precision p) { return p == Precision::FP64 ? 8 : (p == Precision::FP32 ? 4 : 2); }

// Entries stored per bank of a level: the FP32 level-0 factors use a tiled layout
// (whole tiles of the lower triangle, diagonal-tile rows padded to 16 bytes) that
// the level-0 apply stages into LDS with 16-byte stores; every other level and
// precision stores the packed lower triangle, D (D + 1) / 2 entries, D = 32 bs.
int64_t mas_bank_entries(Precision prec, int bs, bool level0);
// Assemble + invert the banks of one level into mas_bank_entries(prec, bs, level0)
// entries per bank.  bank_off: nbanks + 1 node offsets.
void mas_bank_invert(Precision prec, int bs, int32_t nbanks, const int32_t *bank_off, const int32_t *rowptr,
                     const int32_t *cols, const double *vals, void *out, int *status, hipStream_t s, bool level0);
// Galerkin values: coarse slot t = sum of fine blocks src_blk[src_ptr[t] .. src_ptr[t+1]).
void mas_galerkin(int64_t nslots, int bs, const int64_t *src_ptr, const int32_t *src_blk, const double *fine_vals,
                  double *coarse_vals, hipStream_t s);
// Apply kernels take k = 1, 2 or 4 right-hand sides: column c of a vector v is
// v + c * ld (level vectors: ld = nodes * bs of their level); with stop != nullptr
// they do nothing once the stop words of all k columns are set.
// coarse[parent[i]] = sum of fine[i] over the nodes i of each fine bank (banks of at most 32 nodes)
void mas_restrict(int bs, int32_t fine_nbanks, const int32_t *fine_bank_off, const int32_t *fine_parent,
                  const double *fine, double *coarse, const int *stop, hipStream_t s, int k = 1,
                  int64_t ld_fine = 0, int64_t ld_coarse = 0);
void mas_coarse_apply(Precision prec, int bs, int32_t nodes, const void *inv, const double *r, const int32_t *parent,
                      const double *w_parent, double *w, const int *stop, hipStream_t s, int k = 1, int64_t ld = 0,
                      int64_t ld_parent = 0);
// z += scale * w1[parent] over the level-0 scalars (parent: level-0 node -> level-1 node);
// dotv != nullptr: per-workgroup partials of dotv . z (returns their count per column).
// z32 != nullptr: also an FP32 copy of the updated z (same stride ld).
int mas_prolong_add(int64_t np, int bs, const int32_t *parent, const double *w1, double scale, double *z,
                    const double *dotv, double *partials, const int *stop, hipStream_t s, int k = 1,
                    int64_t ld = 0, int64_t ld1 = 0, float *z32 = nullptr);
// st != nullptr: per-bank partials of dotv . z go to partials[c * nbanks + bank] (reduced by k::finish).
// mfma_flush > 0 (FP32 factors only): the bank solve runs on the matrix cores with x and t
// as hi/lo FP32 pairs, adding its FP32 accumulators into FP64 every mfma_flush (4, 8, 16
// or 32) products; 0: the vector-unit solve with FP64 accumulation.
void mas_level0_apply(Precision prec, int bs, int32_t nbanks, const int32_t *part_off, const void *inv,
                      const double *r, const int32_t *parent, const double *w1, double scale, const double *add,
                      double *z, const double *dotv, dev::PcgState *st, double *partials, const int *stop,
                      hipStream_t s, int k = 1, int64_t ld = 0, int64_t ld1 = 0, int mfma_flush = 0);

// ---------------------------------------------------------------- stiff blocks
// Dense S_c A S_c^T per component from operator rows (original numbering).
void stiff_extract(int32_t ncomp, const int64_t *comp_off, const int64_t *mat_off, const int32_t *dofs,
                   const int32_t *g2l, const int32_t *comp_of, const int32_t *ptr, const int32_t *idx,
                   const double *vals, double *mats, hipStream_t s);
// In-place dense Cholesky (lower, row-major) of every component of at most
// stiff_small_max() DOFs; status |= 2 on failure.
void stiff_cholesky(int32_t ncomp, const int64_t *comp_off, const int64_t *mat_off, double *mats, int *status,
                    hipStream_t s);
// G = L^{-1} per component of at most stiff_small_max() DOFs (larger ones are skipped);
// stored with G^T in the strict upper triangle.  max_small_m: largest such component.
void stiff_inverse_factor(int32_t ncomp, int32_t max_small_m, const int64_t *comp_off, const int64_t *mat_off,
                          const double *L, double *G, hipStream_t s);
// stiff_cholesky and stiff_inverse_factor handle components up to this size.
int stiff_small_max();

// Blocked path of one large component (row-major m x m, leading dimension m):
// factor the diagonal block at k0 (kb <= dense_block_size()) in place and write its
// inverse into the same block of G; status |= 2 when it is not positive definite.
int dense_block_size();
void dense_diag_block(int64_t m, int k0, int kb, double *A, double *G, int *status, hipStream_t s);
// G(i, j) := G(j, i) for j > i.
void dense_mirror(int64_t m, double *G, hipStream_t s);
// z[p] += (S_c A S_c^T)^{-1} (r - A z)[p] = G^T G (r - A z)[p] over the stiff rows p of
// each component, (A z) taken from the block operator in the permuted numbering.
// k right-hand sides: r, z at stride ld; work, work2 hold k * ns values each.
// With additive = true the correction is z[p] += G^T G r[p] (the MAS solver's
// additive stiff preconditioner, MASSolver.cu:583-615).  z32 != nullptr: an FP32 copy
// of z (same stride) whose stiff entries are updated with z.
void stiff_correct(int32_t ncomp, int64_t ns, int bs, const int64_t *comp_off, const int64_t *mat_off,
                   const int32_t *row_comp, const int32_t *pdofs, const double *G, const int32_t *rowptr,
                   const int32_t *cols,
                   const double *bvals, const double *r, double *z, double *work, double *work2, const int *stop,
                   hipStream_t s, int k = 1, int64_t ld = 0, bool additive = false, float *z32 = nullptr,
                   const uint8_t *dense = nullptr);
// M = G^T G (explicit inverse, both triangles) of the units listed in units[0 .. nunits),
// each at most max_m DOFs, in the storage of G (comp_off / mat_off).  dense != nullptr
// in stiff_correct marks the units whose storage holds such an M (one product).
void explicit_inverse(int32_t nunits, const int32_t *units, int32_t max_m, const int64_t *comp_off,
                      const int64_t *mat_off, const double *G, double *M, hipStream_t s);

// out[i] = r[pdofs[i]] - (A z)[pdofs[i]] (z == nullptr: r[pdofs[i]]) for i < ns, k columns:
// r, z at stride ld, out at stride ns.
void stiff_residual_rows(int64_t ns, int bs, const int32_t *pdofs, const int32_t *rowptr, const int32_t *cols,
                         const double *bvals, const double *r, const double *z, double *out, const int *stop,
                         hipStream_t s, int k = 1, int64_t ld = 0);
// u[p] = r[p] - (A z)[p] in FP64 for the stiff rows p: after a preconditioner-internal
// SpMV with FP32 values, the rows that hold the stiff entries are recomputed exactly.
void stiff_rows_exact(int64_t ns, int bs, const int32_t *pdofs, const int32_t *rowptr, const int32_t *cols,
                      const double *bvals, const double *r, const double *z, double *u, const int *stop, hipStream_t s,
                      int k = 1, int64_t ld = 0);

}  // namespace lss::k
