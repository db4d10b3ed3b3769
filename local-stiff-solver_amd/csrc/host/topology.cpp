#include "topology.hpp"

#include <algorithm>
#include <numeric>
#include <stdexcept>

namespace lss {

namespace {

// Members of `set` (all labelled `label` in lab[]) ordered by (breadth-first
// distance from `start` inside the induced subgraph, node id).  A search that
// runs out restarts from the smallest unreached member, one level beyond the
// deepest level so far.  dist[] is scratch for the members of `set`.
void bfs_order(const Pattern &g, const std::vector<int32_t> &set, int32_t start, int32_t label,
               std::vector<int32_t> &lab, std::vector<int32_t> &dist, std::vector<int32_t> &order)
{
    const int32_t done = -2 - label;  // visited marker, distinct per subproblem
    std::vector<int32_t> queue;
    queue.reserve(set.size());
    int32_t deepest = 0;
    auto visit_from = [&](int32_t s0, int32_t d0) {
        size_t head = queue.size();
        lab[s0] = done;
        dist[s0] = d0;
        deepest = std::max(deepest, d0);
        queue.push_back(s0);
        while (head < queue.size()) {
            const int32_t u = queue[head++];
            const int32_t du = dist[u] + 1;
            for (int32_t e = g.rowptr[u]; e < g.rowptr[u + 1]; ++e) {
                const int32_t w = g.cols[e];
                if (lab[w] == label) {
                    lab[w] = done;
                    dist[w] = du;
                    deepest = std::max(deepest, du);
                    queue.push_back(w);
                }
            }
        }
    };
    visit_from(start, 0);
    if (queue.size() < set.size())
        for (int32_t v : set)  // set is ascending
            if (lab[v] == label) visit_from(v, deepest + 1);
    // stable counting sort of the ascending set by distance
    std::vector<int64_t> cnt(size_t(deepest) + 2, 0);
    for (int32_t v : set) cnt[dist[v] + 1]++;
    for (int32_t d = 0; d <= deepest; ++d) cnt[d + 1] += cnt[d];
    order.resize(set.size());
    for (int32_t v : set) order[cnt[dist[v]]++] = v;
}

void bisect(const Pattern &g, std::vector<int32_t> set, int max_part, int32_t part_base, int32_t label,
            std::vector<int32_t> &lab, std::vector<int32_t> &dist, std::vector<int32_t> &part_of, int depth)
{
    const int64_t m = int64_t(set.size());
    if (m <= max_part) {
        for (int32_t v : set) part_of[v] = part_base;
        return;
    }
    const int64_t k = (m + max_part - 1) / max_part;  // parts in this subtree
    const int64_t kl = (k + 1) / 2;
    const int64_t nl = kl * max_part;                   // left half holds kl full parts
    // pseudo-peripheral start: the farthest node from the smallest member (largest id among ties)
    for (int32_t v : set) lab[v] = label;
    std::vector<int32_t> order;
    bfs_order(g, set, set.front(), label, lab, dist, order);
    const int32_t start = order.back();  // farthest, then largest id
    for (int32_t v : set) lab[v] = label;
    bfs_order(g, set, start, label, lab, dist, order);
    // halves in ascending order: mark the first nl nodes, then one scan of `set`
    const int32_t lbl_l = 2 * label + 1, lbl_r = 2 * label + 2;
    for (int64_t q = 0; q < nl; ++q) lab[order[q]] = lbl_l;
    std::vector<int32_t> L, R;
    L.reserve(nl);
    R.reserve(m - nl);
    for (int32_t v : set) (lab[v] == lbl_l ? L : R).push_back(v);
    std::vector<int32_t>().swap(order);
    std::vector<int32_t>().swap(set);
    if (depth < 12) {
#pragma omp task default(none) firstprivate(L, max_part, part_base, lbl_l, depth) shared(g, lab, dist, part_of)
        bisect(g, std::move(L), max_part, part_base, lbl_l, lab, dist, part_of, depth + 1);
#pragma omp task default(none) firstprivate(R, max_part, part_base, kl, lbl_r, depth) shared(g, lab, dist, part_of)
        bisect(g, std::move(R), max_part, int32_t(part_base + kl), lbl_r, lab, dist, part_of, depth + 1);
#pragma omp taskwait
    } else {
        bisect(g, std::move(L), max_part, part_base, lbl_l, lab, dist, part_of, depth + 1);
        bisect(g, std::move(R), max_part, int32_t(part_base + kl), lbl_r, lab, dist, part_of, depth + 1);
    }
}

}  // namespace

void partition_graph(const Pattern &g, int max_part, std::vector<int32_t> &part_of, int32_t &nparts)
{
    // Recursive breadth-first bisection.  Nodes without off-diagonal entries
    // (e.g. eliminated Dirichlet rows) are set aside into trailing parts: left
    // in place they become singleton components that never merge and fill the
    // coarse levels.  Each subset is split by a search from a pseudo-peripheral
    // node (farthest from its smallest member, then largest id): the first
    // ceil(k/2)*max_part nodes in (distance, id) order go left.  Part ids follow the
    // recursion from left to right, so every run of consecutive parts is one
    // subtree, compact at every scale, as the bank-local coarse levels need.
    const int32_t n = g.n;
    part_of.assign(n, 0);
    std::vector<int32_t> real, iso;
    for (int32_t v = 0; v < n; ++v) {
        bool coupled = false;
        for (int32_t e = g.rowptr[v]; e < g.rowptr[v + 1] && !coupled; ++e) coupled = g.cols[e] != v;
        (coupled ? real : iso).push_back(v);
    }
    const int32_t real_parts = int32_t((int64_t(real.size()) + max_part - 1) / max_part);
    nparts = real_parts + int32_t((int64_t(iso.size()) + max_part - 1) / max_part);
    for (size_t q = 0; q < iso.size(); ++q) part_of[iso[q]] = real_parts + int32_t(q / max_part);
    if (real.empty()) return;
    // labels: subproblem ids (heap numbering) >= 0, visited markers < -1, isolated -1
    std::vector<int32_t> lab(n, 0), dist(n, 0);
    for (int32_t v : iso) lab[v] = -1;
#pragma omp parallel
#pragma omp single
    bisect(g, std::move(real), max_part, 0, 0, lab, dist, part_of, 0);
}

void partition_permutation(const std::vector<int32_t> &part_of, int32_t nparts, std::vector<int32_t> &perm,
                           std::vector<int32_t> &part_off)
{
    const int32_t n = int32_t(part_of.size());
    part_off.assign(nparts + 1, 0);
    for (int32_t i = 0; i < n; ++i) part_off[part_of[i] + 1]++;
    for (int32_t p = 0; p < nparts; ++p) part_off[p + 1] += part_off[p];
    std::vector<int32_t> next(part_off.begin(), part_off.end() - 1);
    perm.assign(n, 0);
    for (int32_t i = 0; i < n; ++i) perm[i] = next[part_of[i]]++;
}

Pattern permute_pattern(const Pattern &g, const std::vector<int32_t> &perm)
{
    Pattern out;
    out.n = g.n;
    out.rowptr.assign(g.n + 1, 0);
    for (int32_t i = 0; i < g.n; ++i) out.rowptr[perm[i] + 1] = g.rowptr[i + 1] - g.rowptr[i];
    for (int32_t i = 0; i < g.n; ++i) out.rowptr[i + 1] += out.rowptr[i];
    out.cols.resize(g.cols.size());
#pragma omp parallel for schedule(dynamic, 4096)
    for (int32_t i = 0; i < g.n; ++i) {
        int32_t *d = out.cols.data() + out.rowptr[perm[i]];
        int32_t k = 0;
        for (int32_t e = g.rowptr[i]; e < g.rowptr[i + 1]; ++e) d[k++] = perm[g.cols[e]];
        std::sort(d, d + k);
    }
    return out;
}

namespace {

// Connected components of one bank (<= 32 nodes) given neighbour masks that
// include the node itself; rank = order of the component's smallest member.
int bank_components(int m, const uint32_t *nbr, int32_t *rank)
{
    uint32_t conn[32];
    for (int l = 0; l < m; ++l) {
        uint32_t c = nbr[l], vis = 1u << l, todo = c ^ vis;
        while (todo) {
            const int v = __builtin_ctz(todo);
            c |= nbr[v];
            vis |= 1u << v;
            todo = c ^ vis;
        }
        conn[l] = c;
    }
    uint32_t leads = 0;
    for (int l = 0; l < m; ++l)
        if (__builtin_ctz(conn[l]) == l) leads |= 1u << l;
    for (int l = 0; l < m; ++l) {
        const int lead = __builtin_ctz(conn[l]);
        rank[l] = __builtin_popcount(leads & ((1u << lead) - 1u));
    }
    return __builtin_popcount(leads);
}

// Bank-local components of a level: returns parent[] and the next level's node count.
int32_t level_components(const Pattern &op, const std::vector<int32_t> &bank_off, std::vector<int32_t> &parent)
{
    const int32_t nbanks = int32_t(bank_off.size()) - 1;
    std::vector<int32_t> count(nbanks);
    parent.assign(op.n, -1);
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t b = 0; b < nbanks; ++b) {
        const int32_t lo = bank_off[b], hi = bank_off[b + 1];
        const int m = hi - lo;
        if (m > 32) throw std::runtime_error("bank larger than 32 nodes");
        uint32_t nbr[32];
        for (int l = 0; l < m; ++l) {
            uint32_t mask = 1u << l;
            const int32_t v = lo + l;
            for (int32_t e = op.rowptr[v]; e < op.rowptr[v + 1]; ++e) {
                const int32_t w = op.cols[e];
                if (w >= lo && w < hi) mask |= 1u << (w - lo);
            }
            nbr[l] = mask;
        }
        int32_t rank[32];
        count[b] = bank_components(m, nbr, rank);
        for (int l = 0; l < m; ++l) parent[lo + l] = rank[l];  // local rank for now
    }
    std::vector<int32_t> base(nbanks + 1, 0);
    for (int32_t b = 0; b < nbanks; ++b) base[b + 1] = base[b] + count[b];
#pragma omp parallel for schedule(static)
    for (int32_t b = 0; b < nbanks; ++b)
        for (int32_t v = bank_off[b]; v < bank_off[b + 1]; ++v) parent[v] += base[b];
    return base[nbanks];
}

// Galerkin pattern R A R^T and per-slot source lists (fine slots in ascending order).
void galerkin(const Pattern &fine, const std::vector<int32_t> &parent, int32_t coarse_n, MasLevel &out)
{
    // children
    out.child_ptr.assign(coarse_n + 1, 0);
    for (int32_t v = 0; v < fine.n; ++v) out.child_ptr[parent[v] + 1]++;
    for (int32_t a = 0; a < coarse_n; ++a) out.child_ptr[a + 1] += out.child_ptr[a];
    out.child.resize(fine.n);
    {
        std::vector<int64_t> fill(out.child_ptr.begin(), out.child_ptr.end() - 1);
        for (int32_t v = 0; v < fine.n; ++v) out.child[fill[parent[v]]++] = v;
    }
    // per coarse row: (coarse col, fine slot) pairs sorted
    std::vector<std::vector<int32_t>> rcols(coarse_n);
    std::vector<std::vector<int32_t>> rsrc(coarse_n);
    std::vector<std::vector<int32_t>> rcnt(coarse_n);
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t a = 0; a < coarse_n; ++a) {
        std::vector<std::pair<int32_t, int32_t>> pr;
        for (int64_t q = out.child_ptr[a]; q < out.child_ptr[a + 1]; ++q) {
            const int32_t v = out.child[q];
            for (int32_t e = fine.rowptr[v]; e < fine.rowptr[v + 1]; ++e) pr.emplace_back(parent[fine.cols[e]], e);
        }
        std::sort(pr.begin(), pr.end());
        auto &c = rcols[a];
        auto &s = rsrc[a];
        auto &k = rcnt[a];
        for (size_t t = 0; t < pr.size(); ++t) {
            if (c.empty() || c.back() != pr[t].first) {
                c.push_back(pr[t].first);
                k.push_back(0);
            }
            ++k.back();
            s.push_back(pr[t].second);
        }
    }
    out.op.n = coarse_n;
    out.op.rowptr.assign(coarse_n + 1, 0);
    for (int32_t a = 0; a < coarse_n; ++a) out.op.rowptr[a + 1] = out.op.rowptr[a] + int32_t(rcols[a].size());
    const int64_t nslots = out.op.rowptr[coarse_n];
    out.op.cols.resize(nslots);
    out.src_ptr.assign(nslots + 1, 0);
    out.src_blk.resize(fine.cols.size());
    // slot source counts, then prefix
    for (int32_t a = 0; a < coarse_n; ++a)
        for (size_t t = 0; t < rcols[a].size(); ++t) {
            out.op.cols[out.op.rowptr[a] + t] = rcols[a][t];
            out.src_ptr[out.op.rowptr[a] + t + 1] = rcnt[a][t];
        }
    for (int64_t t = 0; t < nslots; ++t) out.src_ptr[t + 1] += out.src_ptr[t];
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t a = 0; a < coarse_n; ++a) {
        const int64_t base = out.src_ptr[out.op.rowptr[a]];
        std::copy(rsrc[a].begin(), rsrc[a].end(), out.src_blk.begin() + base);
    }
}

}  // namespace

void build_mas_levels(const Pattern &fine, const std::vector<int32_t> &part_off, int max_coarse,
                      std::vector<MasLevel> &levels)
{
    levels.clear();
    levels.emplace_back();
    MasLevel &l0 = levels.back();
    l0.nodes = fine.n;
    l0.bank_off = part_off;
    l0.op = fine;
    for (int lv = 0; lv < max_coarse; ++lv) {
        MasLevel &cur = levels.back();
        std::vector<int32_t> parent;
        const int32_t next_n = level_components(cur.op, cur.bank_off, parent);
        // stop when nothing merges (MASPreconditioner.cu:449-454); level 0 always coarsens
        if (lv > 0 && next_n == cur.nodes) break;
        MasLevel nxt;
        nxt.nodes = next_n;
        galerkin(cur.op, parent, next_n, nxt);
        const int32_t nb = (next_n + 31) / 32;
        nxt.bank_off.resize(nb + 1);
        for (int32_t b = 0; b <= nb; ++b) nxt.bank_off[b] = std::min<int32_t>(b * 32, next_n);
        levels.back().parent = std::move(parent);
        levels.push_back(std::move(nxt));
    }
}

}  // namespace lss
