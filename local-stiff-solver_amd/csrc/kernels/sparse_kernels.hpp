// Launchers of the multifrontal sparse Cholesky kernels (stiff units too large for
// one dense factor).  Front f of a unit holds the dense matrix of its row structure
// R_f = [pivots p0 .. p0 + p - 1, update positions U_f], m = p + u rows, row-major.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lss::k {

// Per-front tables of the solve, device copies (nfronts entries each).
struct FrontView {
    const int32_t *p0 = nullptr;     // first pivot position
    const int32_t *np = nullptr;     // pivots p
    const int32_t *nu = nullptr;     // update rows u
    const int64_t *roff = nullptr;   // offset of the front's m-vector in the rhs buffer
    const int64_t *woff = nullptr;   // offset of the front's u-vector in the update buffer (and in uidx)
    const int64_t *goff = nullptr;   // offset of G11 (p x p: G lower, G^T strictly upper)
    const int64_t *loff = nullptr;   // offset of L21 (u x p, row-major)
    const int32_t *uidx = nullptr;   // update positions of every front (at woff)
    const int64_t *cptr = nullptr;   // per front row (at roff): contributions of the children
    const int64_t *cidx = nullptr;   // ... as indices into the update buffer, children in order
};

// Batched factorization of the fronts of one tree level (several fronts per launch).
// Fronts live in a level workspace W (front f: m x m row-major at its own offset), G11
// in the factor array G (p x p at goff); offsets below are element offsets.
struct DiagTask {  // factor + invert the kb x kb diagonal block at (k0, k0) of one front
    int64_t a, g;  // front in W, its G11 in G
    int32_t lda, ldg, k0, kb;
};
struct PanelTask {  // rows r0 .. r0 + nr - 1 (nr <= 32) of the panel below block k0: A_rk := A_rk G_kk^T
    int64_t a, g;
    int32_t lda, ldg, k0, kb, r0, nr;
};
struct TileTask {  // one 64 x 64 tile: C = beta C + alpha A op(B), K inner; lower: write j <= i only
    int64_t a, b, c;
    int32_t lda, ldb, ldc, M, N, K, lower, pad;
};
struct CopyTask {  // one 32 x 32 tile: mode 0 copy, 1 transpose, 2 transpose of the strict lower part
    int64_t src, dst;
    int32_t lds, ldd, rows, cols, mode, pad;
};
struct ExtendTask {  // row a of child c's update matrix into its parent's front
    int32_t c, a;
};

void dense_diag_batch(int32_t ntask, const DiagTask *tasks, double *W, double *G, int *status, hipStream_t s);
void front_panel_batch(int32_t ntask, const PanelTask *tasks, double *W, const double *G, hipStream_t s);
// NT: op(B) = B^T (B read by rows j), else op(B) = B (rows k); beta = 0 reads no C
void front_gemm_batch(int32_t ntask, const TileTask *tasks, bool nt, double alpha, double beta, const double *A,
                      const double *B, double *C, hipStream_t s);
void front_copy_batch(int32_t ntask, const CopyTask *tasks, const double *src, double *dst, hipStream_t s);
// W[wb[f] + rel_c[a] * m_f + rel_c[b]] += U[uoff[c] + a * u_c + b], b <= a, f = parent[c] (distinct parents)
void front_extend_batch(int64_t ntask, const ExtendTask *tasks, const int32_t *parent, const int64_t *wb,
                        const int32_t *np, const int32_t *nu, const int64_t *woff, const int64_t *uoff,
                        const int32_t *rel, const double *U, double *W, hipStream_t s);
// the original entries of the listed fronts: W[wb[f] + li[t] * m_f + lj[t]] = vals[src[t]],
// t in [toff[f], toff[f + 1]) (distinct positions per front)
void front_assemble_batch(int32_t nfronts, const int32_t *fronts, const int64_t *toff, const int32_t *li,
                          const int32_t *lj, const int64_t *src, const int64_t *wb, const int32_t *np,
                          const int32_t *nu, const double *vals, double *W, hipStream_t s);

// out[coff[v] + t] = idx[ptr[dofs[v]] + t]: the stored rows of the unit's columns.
void sparse_gather_pattern(int32_t m, const int32_t *dofs, const int32_t *ptr, const int64_t *coff,
                           const int32_t *idx, int32_t *out, hipStream_t s);
// F[li * ld + lj] = vals[src] for the front's original entries (distinct (li, lj)).
void sparse_assemble(int64_t ntrip, const int32_t *li, const int32_t *lj, const int64_t *src, const double *vals,
                     double *F, int64_t ld, hipStream_t s);
// F[rel[a] * ld + rel[b]] += U[a * u + b] for b <= a < u (a child's update matrix).
void sparse_extend_add(int32_t u, const int32_t *rel, const double *U, double *F, int64_t ld, hipStream_t s);

// Solve (L L^T) x = t of one unit for k columns.  Vectors in elimination (position)
// order: tp, ypos, xpos (stride m); rhs buffer (stride rtot), update buffer (stride wtot).
// Level steps take task lists of (front, row) pairs (rows < m, < p, < u); the products
// give each row a group of lpr lanes (4, 8, 16, 32 or 64).
void sparse_permute_in(int32_t m, const int32_t *order, const double *t, int64_t ldt, double *tp, int k,
                       const int *stop, hipStream_t s);
void sparse_fwd_rhs(int64_t ntask, const int2 *tasks, const FrontView &fv, const double *tp, int32_t m,
                    const double *w, int64_t wtot, double *rhs, int64_t rtot, int k, const int *stop, hipStream_t s);
void sparse_fwd_y(int64_t ntask, const int2 *tasks, const FrontView &fv, const double *G, const double *rhs,
                  int64_t rtot, double *ypos, int32_t m, int k, const int *stop, int lpr, hipStream_t s);
void sparse_fwd_w(int64_t ntask, const int2 *tasks, const FrontView &fv, const double *L, const double *rhs,
                  int64_t rtot, const double *ypos, int32_t m, double *w, int64_t wtot, int k, const int *stop,
                  int lpr, hipStream_t s);
// xu (stride wtot) = x at the update positions of the tasks' fronts (tasks: (front, update row))
void sparse_bwd_gather(int64_t ntask, const int2 *tasks, const FrontView &fv, const double *xpos, int32_t m,
                       double *xu, int64_t wtot, int k, const int *stop, hipStream_t s);
// Lt: L21^T of every front (p x u, row-major, at loff); tasks: (front, pivot row) pairs
void sparse_bwd_s(int64_t ntask, const int2 *tasks, const FrontView &fv, const double *Lt, const double *xu,
                  int64_t wtot, const double *ypos, int32_t m, double *sbuf, int64_t rtot, int k, const int *stop,
                  int lpr, hipStream_t s);
// dst (c x r) = src (r x c)^T, both row-major
void sparse_transpose(int64_t r, int64_t c, const double *src, double *dst, hipStream_t s);
void sparse_bwd_x(int64_t ntask, const int2 *tasks, const FrontView &fv, const double *G, const double *sbuf,
                  int64_t rtot, double *xpos, int32_t m, int k, const int *stop, int lpr, hipStream_t s);
// z[pdofs[order[q]]] += xpos[q] (and z32 = float of the new z when given), k columns at stride ld.
void sparse_scatter_add(int32_t m, const int32_t *order, const int32_t *pdofs, const double *xpos, double *z,
                        float *z32, int64_t ld, int k, const int *stop, hipStream_t s);

}  // namespace lss::k
