#include "stiff_select.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <unordered_map>

namespace lss {

int64_t RowSet::find(int32_t r) const
{
    auto it = std::lower_bound(rows.begin(), rows.end(), r);
    return (it != rows.end() && *it == r) ? int64_t(it - rows.begin()) : -1;
}

namespace {

// Union by rank with path compression (disjointSet, src/polysolve/linear/Hybrid.hpp:223-260).
struct UnionFind {
    std::vector<int> parent, rank;
    explicit UnionFind(int n) : parent(n), rank(n, 0) { std::iota(parent.begin(), parent.end(), 0); }
    int find(int v)
    {
        int r = v;
        while (parent[r] != r) r = parent[r];
        while (parent[v] != r) {
            const int nx = parent[v];
            parent[v] = r;
            v = nx;
        }
        return r;
    }
    void unite(int x, int y)
    {
        if (x == y) return;
        x = find(x);
        y = find(y);
        if (rank[x] > rank[y])
            parent[y] = x;
        else {
            parent[x] = y;
            if (rank[x] == rank[y]) rank[y]++;
        }
    }
};

// Components of the subgraph induced by `set` over the stored entries of its rows.
std::vector<std::vector<int32_t>> induced_components(int32_t n, const RowSet &rows, const std::vector<int32_t> &set,
                                                     std::vector<int32_t> &g2l)
{
    g2l.assign(n, -1);
    for (size_t c = 0; c < set.size(); ++c) g2l[set[c]] = int32_t(c);
    UnionFind uf(int(set.size()));
    for (int32_t k : set) {
        const int64_t pos = rows.find(k);
        if (pos < 0) continue;
        for (int64_t e = rows.ptr[pos]; e < rows.ptr[pos + 1]; ++e)
            if (g2l[rows.idx[e]] != -1) uf.unite(g2l[rows.idx[e]], g2l[k]);
    }
    std::unordered_map<int, std::vector<int32_t>> chosen;
    for (int32_t v : set) chosen[uf.find(g2l[v])].push_back(v);
    std::vector<std::vector<int32_t>> out;
    out.reserve(chosen.size());
    for (auto &kv : chosen) out.push_back(std::move(kv.second));
    std::sort(out.begin(), out.end(), [](const auto &a, const auto &b) { return a.front() < b.front(); });
    return out;
}

}  // namespace

std::vector<int32_t> filter_subdomains(int32_t n, const RowSet &rows, const std::vector<int32_t> &set,
                                       int64_t min_size, int64_t max_size, double cond_thr, FilterStats &st)
{
    st = FilterStats{};
    st.before = int64_t(set.size());
    std::vector<int32_t> g2l;
    auto comps = induced_components(n, rows, set, g2l);
    std::vector<int32_t> kept;
    for (auto &c : comps) {
        if (int64_t(c.size()) < min_size) {
            ++st.too_small;
            continue;
        }
        if (int64_t(c.size()) > max_size) {
            ++st.too_large;
            continue;
        }
        // Gershgorin bounds over the entries inside the selected set
        double lmin = std::numeric_limits<double>::max(), lmax = 0.0;
        for (int32_t k : c) {
            const int64_t pos = rows.find(k);
            double d = 0.0, off = 0.0;
            for (int64_t e = rows.ptr[pos]; e < rows.ptr[pos + 1]; ++e) {
                const int32_t j = rows.idx[e];
                if (g2l[j] == -1) continue;
                if (j == k)
                    d = rows.vals[e];
                else
                    off += std::fabs(rows.vals[e]);
            }
            lmin = std::min(lmin, d - off);
            lmax = std::max(lmax, d + off);
        }
        if (lmin * lmax < 0.0 || lmax / lmin > cond_thr) {
            kept.insert(kept.end(), c.begin(), c.end());
            continue;
        }
        ++st.not_poorly_conditioned;
    }
    std::sort(kept.begin(), kept.end());
    st.after = int64_t(kept.size());
    return kept;
}

std::vector<int32_t> expand_subdomains(int32_t n, const RowSet &rows, const std::vector<int32_t> &set)
{
    std::vector<uint8_t> mark(n, 0);
    for (int32_t k : set) {
        const int64_t pos = rows.find(k);
        if (pos < 0) continue;
        for (int64_t e = rows.ptr[pos]; e < rows.ptr[pos + 1]; ++e) mark[rows.idx[e]] = 1;
    }
    std::vector<int32_t> out;
    for (int32_t i = 0; i < n; ++i)
        if (mark[i]) out.push_back(i);
    return out;
}

std::vector<std::vector<int32_t>> decompose_subdomains(int32_t n, const RowSet &rows, const std::vector<int32_t> &set,
                                                       int64_t max_size)
{
    std::vector<int32_t> g2l;
    auto comps = induced_components(n, rows, set, g2l);
    std::vector<std::vector<int32_t>> out;
    for (auto &c : comps)
        if (int64_t(c.size()) <= max_size) out.push_back(std::move(c));
    return out;
}

}  // namespace lss
