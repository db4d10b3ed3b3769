// Host-side graph steps of the stiff DOF selection (integer work on the rows of
// the selected DOFs, copied back from the device).
#pragma once
#include <cstdint>
#include <vector>

namespace lss {

// Rows of the operator for an ascending list of row ids (CSR restricted to them).
struct RowSet {
    std::vector<int32_t> rows;  // ascending
    std::vector<int64_t> ptr;   // rows.size() + 1
    std::vector<int32_t> idx;
    std::vector<double> vals;   // may be empty (pattern only)
    // position of row id r in `rows`, or -1
    int64_t find(int32_t r) const;
};

struct FilterStats {
    int64_t before = 0, after = 0, too_small = 0, too_large = 0, not_poorly_conditioned = 0;
};

// GPUHybrid::filter_subdomains (GPUHybrid.cu:885-970).
std::vector<int32_t> filter_subdomains(int32_t n, const RowSet &rows, const std::vector<int32_t> &set,
                                       int64_t min_size, int64_t max_size, double cond_thr, FilterStats &st);
// GPUHybrid::expand_subdomains_to_strongly_connected (GPUHybrid.cu:972-991).
std::vector<int32_t> expand_subdomains(int32_t n, const RowSet &rows, const std::vector<int32_t> &set);
// GPUHybrid::decompose_subdomains_to_disjoint_subsets (GPUHybrid.cu:691-735); components
// ascending inside, ordered by smallest DOF (the reference uses unordered_map order).
std::vector<std::vector<int32_t>> decompose_subdomains(int32_t n, const RowSet &rows, const std::vector<int32_t> &set,
                                                       int64_t max_size);

}  // namespace lss
