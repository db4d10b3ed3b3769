// Multifrontal sparse Cholesky of one stiff unit too large for a dense factor.
//
// Role of cuDSS in the reference (sparse LDL^T analysis, factorization and solve of
// S A S^T: GPUHybrid.cu:1231-1249 and 575; MAS/BadDofGMMPreconditioner.cu:493-549).
// Ordering and symbolic analysis on the host (nested dissection by BFS level-set
// separators), numeric factorization front by front on the device, solves as
// level-batched kernels (kernels/sparse.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../common/runtime.hpp"
#include "../kernels/sparse_kernels.hpp"
#include "dense.hpp"

namespace lss {

// Nested-dissection ordering of a symmetric graph and its assembly tree.  Fronts are
// numbered in postorder (children before parents); front f eliminates the positions
// p0[f] .. p0[f] + np[f] - 1 and passes an update matrix on the positions
// uidx[uptr[f] .. uptr[f + 1]) (ascending, all beyond its pivots) to its parent.
struct FrontTree {
    int32_t m = 0;
    std::vector<int32_t> order;  // position -> vertex
    std::vector<int32_t> pos;    // vertex -> position
    std::vector<int32_t> p0, np, parent, height;
    std::vector<std::vector<int32_t>> children;
    std::vector<int64_t> uptr;
    std::vector<int32_t> uidx;
    int32_t fronts() const { return int32_t(p0.size()); }
    int64_t nu(int32_t f) const { return uptr[f + 1] - uptr[f]; }
};

// adj: CSR over vertices 0 .. m - 1, symmetric, without self loops.  Subgraphs of at
// most `leaf` vertices (or without a separating BFS level) become one front.
FrontTree nested_dissection(int32_t m, const int32_t *adj_ptr, const int32_t *adj_idx, int32_t leaf);

class SparseFactor {
public:
    // Factor S A S^T for the unit `dofs` (ascending global DOF ids).  A: the operator as
    // CSC arrays on the device, column j = operator row j; h_ptr: a host copy of the
    // column pointers.  Entry (i, j), i > j in unit order, is taken from operator row
    // dofs[i] (the lower triangle the dense path reads).  status |= 2 when not SPD.  Fronts of
    // at most batch_rows rows are factored in level batches whose workspaces take at most
    // batch_bytes (a batch holds at least one front); larger ones one at a time.
    void factorize(int32_t n, const std::vector<int32_t> &dofs, const std::vector<int32_t> &h_ptr,
                   const int32_t *d_ptr, const int32_t *d_idx, const double *d_vals, int32_t leaf,
                   int64_t batch_rows, int64_t batch_bytes, int max_cols, DenseBlas &blas, int *d_status,
                   hipStream_t s);
    // z[pdofs[i]] += ((S A S^T)^{-1} t)[i] for k columns; t: k columns of m values (unit
    // order) at stride m, z (and z32 when given) at stride ld.
    void solve_add(const double *t, const int32_t *pdofs, double *z, float *z32, int64_t ld, int k, const int *stop,
                   hipStream_t s);
    void release();

    int32_t dofs() const { return m_; }
    int32_t fronts() const { return nf_; }
    int32_t max_front() const { return max_front_; }
    int32_t levels() const { return lv_rhs_.empty() ? 0 : int32_t(lv_rhs_.size()) - 1; }  // tree heights 0 .. levels() - 1
    int64_t factor_entries() const { return int64_t(G_.size() + L_.size()); }  // G11 + L21 (L21^T kept besides)
    int64_t stack_bytes() const { return stack_bytes_; }  // peak update-matrix storage of the factorization
    int32_t batched_fronts() const { return batched_; }    // fronts factored in level batches
    int32_t front_groups() const { return groups_; }       // batches
    double t_symbolic = 0, t_numeric = 0;

private:
    int32_t m_ = 0, nf_ = 0, max_front_ = 0, batched_ = 0, groups_ = 0;
    int64_t rtot_ = 0, wtot_ = 0, stack_bytes_ = 0;
    DBuf<int32_t> order_, p0_, np_, nu_, uidx_;
    DBuf<int64_t> roff_, woff_, goff_, loff_, cptr_, cidx_;
    DBuf<double> G_, L_, Lt_;  // G11 per front; L21 (u x p) and its transpose (p x u)
    // level task lists: tasks of level h at [lv_x[h], lv_x[h + 1])
    DBuf<int2> t_rhs_, t_y_, t_w_, t_x_;
    std::vector<int64_t> lv_rhs_, lv_y_, lv_w_, lv_x_;
    std::vector<int> lpr_y_, lpr_w_, lpr_s_, lpr_x_;  // lanes per row of each level's products
    DBuf<double> tp_, ypos_, xpos_, rhs_, wbuf_;
    k::FrontView view() const;
};

}  // namespace lss
