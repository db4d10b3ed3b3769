// Matrix Market reading and writing for the linear_solve command line tool.
//
// Semantics follow the loaders the reference harness uses
// (tests/linear_solve.cpp:320-343, 352-368; Eigen 3.4 loadMarket /
// loadMarketVector): a sparse matrix is read from the coordinate format with
// 1-based indices, the symmetry field of the header is not interpreted (a
// "symmetric" file gives only its stored triangle), duplicate entries are
// summed and explicit zeros stay stored entries; the result is compressed
// sparse columns with ascending row indices.  A vector is read from the array
// format ("n 1" sizes line, then one value per line).  Dense integer or real
// arrays (positions, elements) are read column by column.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lss::cli {

struct Csc {
    int32_t rows = 0, cols = 0;
    std::vector<int32_t> colptr, rowidx;
    std::vector<double> vals;
    int64_t nnz() const { return int64_t(rowidx.size()); }
};

// Throws std::runtime_error with the file name and line on malformed input.
Csc read_matrix(const std::string &path);
std::vector<double> read_vector(const std::string &path);
// Dense array (rows x cols, column-major values); returns the values and sets rows / cols.
std::vector<double> read_dense(const std::string &path, int64_t &rows, int64_t &cols);
void write_vector(const std::string &path, const std::vector<double> &v);

// A + A^T without the diagonal twice, as --force_symmetry builds it
// (tests/linear_solve.cpp:329-345): every stored off-diagonal entry (r, c) is
// also added at (c, r), so a matrix that already stores both triangles is doubled.
Csc force_symmetry(const Csc &A);

// y = A x for the CSC matrix (column j scatters into its rows).
std::vector<double> multiply(const Csc &A, const std::vector<double> &x);

}  // namespace lss::cli
