// ORACLE — test infrastructure only (see lss_oracle.h for scope and citations).
// CPU FP64 restatement of the reference GPU hybrid solve path.
#include "lss_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <set>
#include <stdexcept>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using clk = std::chrono::steady_clock;
double since(clk::time_point t0) { return std::chrono::duration<double>(clk::now() - t0).count(); }

// ORC_TIMING=1 prints the set-up phases to stderr (for sizing CPU baselines)
void phase(const char *what, clk::time_point &t)
{
    static const bool on = std::getenv("ORC_TIMING") != nullptr;
    if (on) std::fprintf(stderr, "[oracle] %-12s %.3f s\n", what, since(t));
    t = clk::now();
}

constexpr int kChunks = 1024;  // reduction tree shape shared with the device kernels
constexpr int kLanes = 256;
constexpr int kLanczosSteps = 20;  // CG-Lanczos steps of the lambda_max(BA) estimate

template <class F>
double canonical_sum(int64_t n, F f)
{
    if (n <= 0) return 0.0;
    int64_t C = (n + kChunks - 1) / kChunks;
    C = (C + kLanes - 1) / kLanes * kLanes;
    std::vector<double> part(kChunks, 0.0);
#pragma omp parallel for schedule(static)
    for (int b = 0; b < kChunks; ++b) {
        double acc[kLanes];
        const int64_t lo = int64_t(b) * C, hi = std::min<int64_t>(lo + C, n);
        for (int t = 0; t < kLanes; ++t) {
            double a = 0.0;
            for (int64_t i = lo + t; i < hi; i += kLanes) a += f(i);
            acc[t] = a;
        }
        for (int s = kLanes / 2; s > 0; s >>= 1)
            for (int t = 0; t < s; ++t) acc[t] += acc[t + s];
        part[b] = acc[0];
    }
    for (int s = kChunks / 2; s > 0; s >>= 1)
        for (int t = 0; t < s; ++t) part[t] += part[t + s];
    return part[0];
}

// Union by rank + path compression, src/polysolve/linear/Hybrid.hpp:223-260.
struct DisjointSet {
    std::vector<int> parent, rank;
    explicit DisjointSet(int n) : parent(n), rank(n, 0) { std::iota(parent.begin(), parent.end(), 0); }
    int find_set(int v)
    {
        int r = v;
        while (parent[r] != r) r = parent[r];
        while (parent[v] != r) {  // path compression
            const int nx = parent[v];
            parent[v] = r;
            v = nx;
        }
        return r;
    }
    void union_set(int x, int y)
    {
        if (x == y) return;
        x = find_set(x);
        y = find_set(y);
        if (rank[x] > rank[y])
            parent[y] = x;
        else {
            parent[x] = y;
            if (rank[x] == rank[y]) rank[y]++;
        }
    }
};

struct Csr {
    int32_t n = 0;
    const int32_t *ptr = nullptr;
    const int32_t *idx = nullptr;
    const double *vals = nullptr;
};

void spmv(const Csr &A, const double *x, double *y)
{
#pragma omp parallel for schedule(static, 1024)
    for (int32_t i = 0; i < A.n; ++i) {
        double s = 0.0;
        for (int32_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k) s += A.vals[k] * x[A.idx[k]];
        y[i] = s;
    }
}

double dot(int64_t n, const double *a, const double *b)
{
    // fixed chunking, partials added in chunk order: the same bits for any thread count
    constexpr int kChunks = 256;
    double part[kChunks];
    const int64_t len = (n + kChunks - 1) / kChunks;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < kChunks; ++c) {
        double s = 0.0;
        for (int64_t i = c * len; i < std::min(n, (c + 1) * len); ++i) s += a[i] * b[i];
        part[c] = s;
    }
    double s = 0.0;
    for (int c = 0; c < kChunks; ++c) s += part[c];
    return s;
}

// components of the induced subgraph of `set` (GPUHybrid.cu:696-720), keyed by root
std::vector<std::vector<int32_t>> components(const Csr &A, const std::vector<int32_t> &set,
                                             std::vector<int32_t> &g2l)
{
    g2l.assign(A.n, -1);
    for (size_t c = 0; c < set.size(); ++c) g2l[set[c]] = int32_t(c);
    DisjointSet ds(int(set.size()));
    for (int32_t k : set)
        for (int32_t e = A.ptr[k]; e < A.ptr[k + 1]; ++e)
            if (g2l[A.idx[e]] != -1) ds.union_set(g2l[A.idx[e]], g2l[k]);
    std::unordered_map<int, std::vector<int32_t>> chosen;
    for (int32_t v : set) chosen[ds.find_set(g2l[v])].push_back(v);
    std::vector<std::vector<int32_t>> out;
    out.reserve(chosen.size());
    for (auto &kv : chosen) out.push_back(std::move(kv.second));
    // canonical order: by smallest DOF (the reference uses unordered_map order)
    std::sort(out.begin(), out.end(), [](const auto &a, const auto &b) { return a.front() < b.front(); });
    return out;
}

// Dense FP64 Cholesky (row-major, lower), returns false if not SPD.
bool cholesky(int m, std::vector<double> &a)
{
    for (int j = 0; j < m; ++j) {
        double d = a[size_t(j) * m + j];
        for (int k = 0; k < j; ++k) d -= a[size_t(j) * m + k] * a[size_t(j) * m + k];
        if (!(d > 0.0)) return false;
        d = std::sqrt(d);
        a[size_t(j) * m + j] = d;
        for (int i = j + 1; i < m; ++i) {
            double s = a[size_t(i) * m + j];
            for (int k = 0; k < j; ++k) s -= a[size_t(i) * m + k] * a[size_t(j) * m + k];
            a[size_t(i) * m + j] = s / d;
        }
        for (int k = j + 1; k < m; ++k) a[size_t(j) * m + k] = 0.0;
    }
    return true;
}

void chol_solve(int m, const double *L, double *x)
{
    for (int i = 0; i < m; ++i) {
        double s = x[i];
        for (int k = 0; k < i; ++k) s -= L[size_t(i) * m + k] * x[k];
        x[i] = s / L[size_t(i) * m + i];
    }
    for (int i = m - 1; i >= 0; --i) {
        double s = x[i];
        for (int k = i + 1; k < m; ++k) s -= L[size_t(k) * m + i] * x[k];
        x[i] = s / L[size_t(i) * m + i];
    }
}

}  // namespace

// ===================================================================== reductions
extern "C" double orc_canonical_sum(int64_t n, const double *v)
{
    return canonical_sum(n, [v](int64_t i) { return v[i]; });
}

// ===================================================================== selection
extern "C" void orc_row_l1(int32_t n, const int32_t *ptr, const double *vals, double *out)
{
    // GPUHybrid.cu:753-765: sequential sum of |a| in storage order
#pragma omp parallel for schedule(static, 4096)
    for (int32_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int32_t j = ptr[i]; j < ptr[i + 1]; ++j) s += std::fabs(vals[j]);
        out[i] = s;
    }
}

extern "C" int orc_gmm_select(int32_t n, const double *x, double tol, int max_iter, double jump, uint8_t *flags,
                              double *gamma, orc_gmm_info *info)
{
#pragma STDC FP_CONTRACT OFF
    // GPUHybrid.cu:767-785
    const double sum_norms = canonical_sum(n, [x](int64_t i) { return x[i]; });
    const double global_mean = sum_norms / n;
    const double var_sum = canonical_sum(n, [x, global_mean](int64_t i) {
        const double d = x[i] - global_mean;
        return d * d;
    });
    const double global_var = var_sum / n;
    double mean_0 = *std::min_element(x, x + n);
    double mean_1 = *std::max_element(x, x + n);
    double var_0 = global_var, var_1 = global_var, w0 = 0.5, w1 = 0.5;
    const double var_reg = 1e-6;
    std::vector<double> g0(n), g1(n);
    int it;
    // GPUHybrid.cu:787-857
    for (it = 0; it < max_iter; ++it) {
        const double log_w0 = std::log(w0), log_w1 = std::log(w1);
        const double c0 = -0.5 * std::log(2.0 * M_PI * var_0);
        const double c1 = -0.5 * std::log(2.0 * M_PI * var_1);
#pragma omp parallel for schedule(static, 4096)
        for (int32_t i = 0; i < n; ++i) {
            const double xi = x[i];
            const double log_N0 = c0 - 0.5 * (xi - mean_0) * (xi - mean_0) / var_0;
            const double log_N1 = c1 - 0.5 * (xi - mean_1) * (xi - mean_1) / var_1;
            const double log_g0 = log_w0 + log_N0;
            const double log_g1 = log_w1 + log_N1;
            const double mx = std::max(log_g0, log_g1);
            const double log_total = mx + std::log(std::exp(log_g0 - mx) + std::exp(log_g1 - mx));
            g0[i] = std::exp(log_g0 - log_total);
            g1[i] = std::exp(log_g1 - log_total);
        }
        const double sum_g0 = canonical_sum(n, [&](int64_t i) { return g0[i]; });
        const double sum_g1 = canonical_sum(n, [&](int64_t i) { return g1[i]; });
        w0 = sum_g0 / n;
        w1 = sum_g1 / n;
        const double om0 = mean_0, om1 = mean_1, ov0 = var_0, ov1 = var_1;
        mean_0 = canonical_sum(n, [&](int64_t i) { return g0[i] * x[i]; }) / sum_g0;
        mean_1 = canonical_sum(n, [&](int64_t i) { return g1[i] * x[i]; }) / sum_g1;
        const double m0 = mean_0, m1 = mean_1;
        var_0 = canonical_sum(n, [&](int64_t i) { return g0[i] * (x[i] - m0) * (x[i] - m0); }) / sum_g0 + var_reg;
        var_1 = canonical_sum(n, [&](int64_t i) { return g1[i] * (x[i] - m1) * (x[i] - m1); }) / sum_g1 + var_reg;
        // GPUHybrid.cu:850-856 (abs on doubles read as fabs)
        if (std::fabs(mean_0 - om0) / std::fabs(om0) < tol && std::fabs(mean_1 - om1) / std::fabs(om1) < tol &&
            std::fabs(var_0 - ov0) / std::fabs(ov0) < tol && std::fabs(var_1 - ov1) / std::fabs(ov1) < tol)
            break;
    }
    // GPUHybrid.cu:859-873
    const bool jumped = std::fabs(mean_1) / std::fabs(mean_0) > jump;
    int64_t cnt = 0;
    double margin = std::numeric_limits<double>::infinity();
    for (int32_t i = 0; i < n; ++i) {
        const bool sel = jumped && (g0[i] < g1[i]);
        flags[i] = sel ? 1 : 0;
        cnt += sel;
        if (g0[i] > 0 && g1[i] > 0) margin = std::min(margin, std::fabs(std::log(g1[i]) - std::log(g0[i])));
        else margin = std::min(margin, std::numeric_limits<double>::max());
    }
    if (gamma) {
        std::copy(g0.begin(), g0.end(), gamma);
        std::copy(g1.begin(), g1.end(), gamma + n);
    }
    if (info) {
        info->global_mean = global_mean;
        info->global_var = global_var;
        info->mean0 = mean_0;
        info->mean1 = mean_1;
        info->var0 = var_0;
        info->var1 = var_1;
        info->w0 = w0;
        info->w1 = w1;
        info->iters = it;
        info->jumped = jumped;
        info->selected = cnt;
        info->min_margin = margin;
    }
    return 0;
}

extern "C" int orc_filter(int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals, const int32_t *set,
                          int64_t nset, int64_t min_size, int64_t max_size, double cond_thr, int32_t *out,
                          int64_t *nout, int64_t *stats)
{
    // GPUHybrid.cu:885-970
    Csr A{n, ptr, idx, vals};
    std::vector<int32_t> s(set, set + nset), g2l;
    auto comps = components(A, s, g2l);
    std::vector<int32_t> kept;
    int64_t too_small = 0, too_large = 0, well = 0;
    for (auto &c : comps) {
        if (int64_t(c.size()) < min_size) {
            ++too_small;
            continue;
        }
        if (int64_t(c.size()) > max_size) {
            ++too_large;
            continue;
        }
        double lmin = std::numeric_limits<double>::max(), lmax = 0.0;
        for (int32_t k : c) {
            double d = 0.0, off = 0.0;
            for (int32_t e = ptr[k]; e < ptr[k + 1]; ++e) {
                if (g2l[idx[e]] == -1) continue;
                if (idx[e] == k)
                    d = vals[e];
                else
                    off += std::fabs(vals[e]);
            }
            lmin = std::min(lmin, d - off);
            lmax = std::max(lmax, d + off);
        }
        if (lmin * lmax < 0.0 || lmax / lmin > cond_thr) {
            kept.insert(kept.end(), c.begin(), c.end());
            continue;
        }
        ++well;
    }
    std::sort(kept.begin(), kept.end());
    std::copy(kept.begin(), kept.end(), out);
    *nout = int64_t(kept.size());
    if (stats) {
        stats[0] = too_small;
        stats[1] = too_large;
        stats[2] = well;
    }
    return 0;
}

extern "C" int orc_expand(int32_t n, const int32_t *ptr, const int32_t *idx, const int32_t *set, int64_t nset,
                          int32_t *out, int64_t *nout)
{
    // GPUHybrid.cu:972-991: every stored entry of each selected row joins the set
    std::vector<uint8_t> mark(n, 0);
    for (int64_t c = 0; c < nset; ++c)
        for (int32_t e = ptr[set[c]]; e < ptr[set[c] + 1]; ++e) mark[idx[e]] = 1;
    int64_t k = 0;
    for (int32_t i = 0; i < n; ++i)
        if (mark[i]) out[k++] = i;
    *nout = k;
    return 0;
}

extern "C" int orc_decompose(int32_t n, const int32_t *ptr, const int32_t *idx, const int32_t *set, int64_t nset,
                             int64_t max_size, int32_t *comp_dofs, int64_t *comp_off, int64_t *ncomp)
{
    // GPUHybrid.cu:691-735
    Csr A{n, ptr, idx, nullptr};
    std::vector<int32_t> s(set, set + nset), g2l;
    auto comps = components(A, s, g2l);
    int64_t w = 0, c = 0;
    comp_off[0] = 0;
    for (auto &cc : comps) {
        if (int64_t(cc.size()) > max_size) continue;
        std::copy(cc.begin(), cc.end(), comp_dofs + w);
        w += int64_t(cc.size());
        comp_off[++c] = w;
    }
    *ncomp = c;
    return 0;
}

// ===================================================================== partition
namespace {

// members of `set` ordered by (breadth-first distance from `start` inside the
// induced subgraph, node id); when the search runs out, it restarts from the
// smallest unreached member, one level beyond the deepest level so far
std::vector<int32_t> bfs_in(const int64_t *bptr, const int32_t *bidx, const std::vector<int32_t> &set, int32_t start,
                            std::vector<char> &in, std::vector<int32_t> &dist)
{
    for (int32_t v : set) in[v] = 1, dist[v] = -1;
    std::vector<int32_t> queue;
    queue.reserve(set.size());
    int32_t deepest = 0;
    auto from = [&](int32_t s0, int32_t d0) {
        size_t head = queue.size();
        dist[s0] = d0;
        deepest = std::max(deepest, d0);
        queue.push_back(s0);
        while (head < queue.size()) {
            const int32_t u = queue[head++];
            for (int64_t e = bptr[u]; e < bptr[u + 1]; ++e) {
                const int32_t w = bidx[e];
                if (in[w] && dist[w] < 0) {
                    dist[w] = dist[u] + 1;
                    deepest = std::max(deepest, dist[w]);
                    queue.push_back(w);
                }
            }
        }
    };
    from(start, 0);
    for (int32_t v : set)
        if (dist[v] < 0) from(v, deepest + 1);
    // stable counting sort of the ascending set by distance
    std::vector<int64_t> cnt(size_t(deepest) + 2, 0);
    for (int32_t v : set) cnt[dist[v] + 1]++;
    for (int32_t d = 0; d <= deepest; ++d) cnt[d + 1] += cnt[d];
    std::vector<int32_t> order(set.size());
    for (int32_t v : set) order[cnt[dist[v]]++] = v;
    for (int32_t v : set) in[v] = 0;
    return order;
}

void split(const int64_t *bptr, const int32_t *bidx, std::vector<int32_t> set, int max_part, int32_t base,
           int32_t *part_of, std::vector<char> &in, std::vector<int32_t> &seen)
{
    if (int64_t(set.size()) <= max_part) {
        for (int32_t v : set) part_of[v] = base;
        return;
    }
    const int64_t k = (int64_t(set.size()) + max_part - 1) / max_part, kl = (k + 1) / 2;
    const int32_t start = bfs_in(bptr, bidx, set, set.front(), in, seen).back();
    auto order = bfs_in(bptr, bidx, set, start, in, seen);
    std::vector<int32_t> L(order.begin(), order.begin() + kl * max_part), R(order.begin() + kl * max_part, order.end());
    std::sort(L.begin(), L.end());
    std::sort(R.begin(), R.end());
    split(bptr, bidx, std::move(L), max_part, base, part_of, in, seen);
    split(bptr, bidx, std::move(R), max_part, int32_t(base + kl), part_of, in, seen);
}

}  // namespace

extern "C" int orc_partition(int32_t nb, const int64_t *bptr, const int32_t *bidx, int max_part, int32_t *part_of,
                             int32_t *nparts)
{
    // Recursive breadth-first bisection.  Nodes without off-diagonal entries go
    // to trailing parts (32 per part, ascending).  A subset of k parts is
    // ordered by (distance, id) from the last node in that order from its
    // smallest member (a pseudo-peripheral node: farthest, then largest id);
    // its first ceil(k/2)*max_part nodes go left; parts are numbered left to right.
    std::vector<int32_t> real, iso;
    for (int32_t v = 0; v < nb; ++v) {
        bool coupled = false;
        for (int64_t e = bptr[v]; e < bptr[v + 1]; ++e) coupled |= bidx[e] != v;
        (coupled ? real : iso).push_back(v);
    }
    const int32_t real_parts = int32_t((int64_t(real.size()) + max_part - 1) / max_part);
    *nparts = real_parts + int32_t((int64_t(iso.size()) + max_part - 1) / max_part);
    for (size_t q = 0; q < iso.size(); ++q) part_of[iso[q]] = real_parts + int32_t(q / max_part);
    if (real.empty()) return 0;
    std::vector<char> in(nb, 0);
    std::vector<int32_t> dist(nb, -1);
    split(bptr, bidx, std::move(real), max_part, 0, part_of, in, dist);
    return 0;
}

extern "C" void orc_spmv(int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals, const double *x,
                         double *y)
{
    spmv(Csr{n, ptr, idx, vals}, x, y);
}

// ===================================================================== MAS
struct orc_mas {
    static constexpr int BANK = 32;
    static constexpr int MAX_COARSE = 4;
    int32_t n = 0, nb = 0;
    int bs = 1;
    std::vector<int32_t> perm;      // old node -> new node
    std::vector<int32_t> part_off;  // level-0 banks
    int levels = 1;                 // matrix levels
    std::vector<int32_t> counts;    // matrices per level
    std::vector<int32_t> cco;       // coarse nodes per coarse level
    std::vector<int32_t> map;       // new node * MAX_COARSE + lv
    std::vector<std::vector<double>> inv;  // per level: counts[l] * dim * dim Cholesky factors (row-major lower)
    int dim = 32;
    std::vector<int32_t> inv_perm;  // new node -> old node
    // per coarse level: members (new fine nodes, ascending) of each coarse node
    std::vector<std::vector<int64_t>> mem_ptr;
    std::vector<std::vector<int32_t>> mem;
    mutable std::vector<std::vector<double>> rl, yl;  // apply work vectors
};

extern "C" orc_mas *orc_mas_create(int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals, int bs,
                                   const int32_t *perm, const int32_t *part_off, int32_t nparts)
{
    auto M = new orc_mas();
    const int B = orc_mas::BANK;
    M->n = n;
    M->bs = bs;
    M->nb = (n + bs - 1) / bs;
    M->perm.assign(perm, perm + M->nb);
    M->part_off.assign(part_off, part_off + nparts + 1);
    M->dim = B * bs;
    const int32_t nb = M->nb;
    auto tp = clk::now();
    M->inv_perm.assign(nb, 0);
    for (int32_t o = 0; o < nb; ++o) M->inv_perm[perm[o]] = o;
    // block adjacency in the new numbering (pattern of the operator, MASPreconditioner.cu:128-213)
    std::vector<std::vector<int32_t>> adj(nb);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t ob = 0; ob < nb; ++ob) {
        const int32_t I = perm[ob];
        auto &a = adj[I];
        for (int32_t i = ob * bs; i < std::min<int64_t>(int64_t(ob + 1) * bs, n); ++i)
            for (int32_t e = ptr[i]; e < ptr[i + 1]; ++e) {
                const int32_t J = perm[idx[e] / bs];
                if (J != I) a.push_back(J);
            }
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    phase("mas.adj", tp);
    std::vector<int32_t> bank0(nb), local0(nb);
    for (int32_t p = 0; p < nparts; ++p)
        for (int32_t v = part_off[p]; v < part_off[p + 1]; ++v) {
            bank0[v] = p;
            local0[v] = v - part_off[p];
        }
    M->map.assign(size_t(nb) * orc_mas::MAX_COARSE, -1);

    // bank-local connected components, lead lane = smallest local index (MASPreconditioner.cu:194-212)
    auto bank_cco = [&](int nodes_in_bank, const std::vector<uint32_t> &nbr, std::vector<int> &rank_of, int &count) {
        std::vector<uint32_t> conn(nodes_in_bank);
        for (int l = 0; l < nodes_in_bank; ++l) {
            uint32_t c = nbr[l], vis = 1u << l, tv = c ^ vis;
            while (tv) {
                int v = __builtin_ctz(tv);
                c |= nbr[v];
                vis |= 1u << v;
                tv = c ^ vis;
            }
            conn[l] = c;
        }
        uint32_t leads = 0;
        for (int l = 0; l < nodes_in_bank; ++l)
            if (__builtin_ctz(conn[l]) == l) leads |= 1u << l;
        rank_of.assign(nodes_in_bank, 0);
        for (int l = 0; l < nodes_in_bank; ++l) {
            const int lead = __builtin_ctz(conn[l]);
            rank_of[l] = __builtin_popcount(leads & ((1u << lead) - 1u));
        }
        count = __builtin_popcount(leads);
    };

    // level 0 coarse map (MASPreconditioner.cu:128-213, 321-339)
    {
        int32_t base = 0;
        for (int32_t p = 0; p < nparts; ++p) {
            const int sz = part_off[p + 1] - part_off[p];
            std::vector<uint32_t> nbr(sz, 0);
            for (int l = 0; l < sz; ++l) {
                const int32_t v = part_off[p] + l;
                nbr[l] = 1u << l;
                for (int32_t w : adj[v])
                    if (bank0[w] == p) nbr[l] |= 1u << local0[w];
            }
            std::vector<int> rk;
            int cnt;
            bank_cco(sz, nbr, rk, cnt);
            for (int l = 0; l < sz; ++l) M->map[size_t(part_off[p] + l) * orc_mas::MAX_COARSE] = base + rk[l];
            base += cnt;
        }
        M->cco.push_back(base);
    }
    int level_num = 1;
    // coarse levels (MASPreconditioner.cu:403-456)
    for (int lv = 1; lv < orc_mas::MAX_COARSE; ++lv) {
        const int32_t prev = M->cco[lv - 1];
        std::vector<uint32_t> nbr(prev, 0);
        for (int32_t c = 0; c < prev; ++c) nbr[c] = 1u << (c % B);
        for (int32_t v = 0; v < nb; ++v) {
            const int32_t cv = M->map[size_t(v) * orc_mas::MAX_COARSE + lv - 1];
            for (int32_t w : adj[v]) {
                const int32_t cw = M->map[size_t(w) * orc_mas::MAX_COARSE + lv - 1];
                if (cv / B == cw / B) nbr[cv] |= 1u << (cw % B);
            }
        }
        const int32_t banks = (prev + B - 1) / B;
        std::vector<int32_t> newid(prev);
        int32_t base = 0;
        for (int32_t bk = 0; bk < banks; ++bk) {
            const int sz = std::min(B, prev - bk * B);
            std::vector<uint32_t> nb_(nbr.begin() + bk * B, nbr.begin() + bk * B + sz);
            std::vector<int> rk;
            int cnt;
            bank_cco(sz, nb_, rk, cnt);
            for (int l = 0; l < sz; ++l) newid[bk * B + l] = base + rk[l];
            base += cnt;
        }
        if (base == prev) break;
        for (int32_t v = 0; v < nb; ++v)
            M->map[size_t(v) * orc_mas::MAX_COARSE + lv] = newid[M->map[size_t(v) * orc_mas::MAX_COARSE + lv - 1]];
        M->cco.push_back(base);
        level_num = lv + 1;
    }
    M->cco.resize(level_num);
    M->levels = 1 + level_num;
    M->counts.push_back(nparts);
    for (int l = 0; l < level_num; ++l) M->counts.push_back((M->cco[l] + B - 1) / B);

    // matrices (MASPreconditioner.cu:494-591, 779-796) and inverses (799-936)
    const int D = M->dim;
    phase("mas.cco", tp);
    M->inv.resize(M->levels);
    for (int L = 0; L < M->levels; ++L) M->inv[L].assign(size_t(M->counts[L]) * D * D, 0.0);
    // level 0: every scalar row fills its own bank row
#pragma omp parallel for schedule(static, 4096)
    for (int32_t i = 0; i < n; ++i) {
        const int32_t I = perm[i / bs];
        const int a = i % bs;
        double *m = &M->inv[0][size_t(bank0[I]) * D * D + size_t(local0[I] * bs + a) * D];
        for (int32_t e = ptr[i]; e < ptr[i + 1]; ++e) {
            const int32_t J = perm[idx[e] / bs];
            if (bank0[I] == bank0[J]) m[local0[J] * bs + idx[e] % bs] += vals[e];
        }
    }
    // members of each coarse node, ascending in the new numbering
    M->mem_ptr.assign(level_num, {});
    M->mem.assign(level_num, {});
    for (int lv = 0; lv < level_num; ++lv) {
        auto &mp = M->mem_ptr[lv];
        auto &mm = M->mem[lv];
        mp.assign(size_t(M->cco[lv]) + 1, 0);
        for (int32_t v = 0; v < nb; ++v) mp[M->map[size_t(v) * orc_mas::MAX_COARSE + lv] + 1]++;
        for (int32_t c = 0; c < M->cco[lv]; ++c) mp[c + 1] += mp[c];
        mm.resize(nb);
        std::vector<int64_t> nx(mp.begin(), mp.end() - 1);
        for (int32_t v = 0; v < nb; ++v) mm[nx[M->map[size_t(v) * orc_mas::MAX_COARSE + lv]]++] = v;
    }
    // coarse levels: Galerkin sums accumulated in ascending scalar row order, one
    // coarse bank per task (its fine rows gathered in that order)
    for (int lv = 0; lv < level_num; ++lv) {
        const int32_t nbank = M->counts[lv + 1];
        std::vector<std::vector<int32_t>> rows(nbank);
        for (int32_t ob = 0; ob < nb; ++ob) {
            const int32_t ci = M->map[size_t(perm[ob]) * orc_mas::MAX_COARSE + lv];
            rows[ci / B].push_back(ob);
        }
#pragma omp parallel for schedule(dynamic, 4)
        for (int32_t k = 0; k < nbank; ++k) {
            double *mk = &M->inv[lv + 1][size_t(k) * D * D];
            for (int32_t ob : rows[k]) {
                const int32_t ci = M->map[size_t(perm[ob]) * orc_mas::MAX_COARSE + lv];
                for (int32_t i = ob * bs; i < std::min<int64_t>(int64_t(ob + 1) * bs, n); ++i) {
                    const int a = i % bs;
                    for (int32_t e = ptr[i]; e < ptr[i + 1]; ++e) {
                        const int32_t J = perm[idx[e] / bs];
                        const int32_t cj = M->map[size_t(J) * orc_mas::MAX_COARSE + lv];
                        if (ci / B != cj / B) continue;
                        mk[size_t((ci % B) * bs + a) * D + (cj % B) * bs + idx[e] % bs] += vals[e];
                    }
                }
            }
        }
    }
    phase("mas.assemble", tp);
    for (int L = 0; L < M->levels; ++L) {
        const int32_t cnt = M->counts[L];
        bool ok = true;
#pragma omp parallel for schedule(dynamic, 16) reduction(&& : ok)
        for (int32_t k = 0; k < cnt; ++k) {
            std::vector<double> a(M->inv[L].begin() + size_t(k) * D * D, M->inv[L].begin() + size_t(k + 1) * D * D);
            for (int d = 0; d < D; ++d)
                if (a[size_t(d) * D + d] == 0.0) a[size_t(d) * D + d] = 1.0;
            // A_b^{-1} is applied through its Cholesky factor (two triangular solves)
            if (!cholesky(D, a)) ok = false;
            std::copy(a.begin(), a.end(), M->inv[L].begin() + size_t(k) * D * D);
        }
        if (!ok) {
            delete M;
            return nullptr;
        }
    }
    phase("mas.factor", tp);
    return M;
}

extern "C" void orc_mas_destroy(orc_mas *m) { delete m; }

extern "C" void orc_mas_info(const orc_mas *m, int *levels, int32_t *counts, int32_t *cco)
{
    *levels = m->levels;
    for (int l = 0; l < m->levels; ++l) counts[l] = m->counts[l];
    for (int l = 0; l + 1 < m->levels; ++l) cco[l] = m->cco[l];
}

extern "C" int32_t orc_mas_map(const orc_mas *m, int32_t node, int level)
{
    return m->map[size_t(node) * orc_mas::MAX_COARSE + level];
}

extern "C" int orc_mas_inverse(const orc_mas *m, int level, int32_t bank, double *out)
{
    if (level < 0 || level >= m->levels || bank < 0 || bank >= m->counts[level]) return -1;
    const size_t D = m->dim;
    const double *L = m->inv[level].data() + bank * D * D;
    std::vector<double> col(D);
    for (size_t c = 0; c < D; ++c) {
        std::fill(col.begin(), col.end(), 0.0);
        col[c] = 1.0;
        chol_solve(int(D), L, col.data());
        for (size_t r = 0; r < D; ++r) out[r * D + c] = col[r];
    }
    return 0;
}

extern "C" void orc_mas_apply(const orc_mas *m, const double *r, double *z)
{
    // MASPreconditioner.cu:1099-1164 (gather, per-bank inverse, prolongation)
    const int B = orc_mas::BANK, bs = m->bs, D = m->dim;
    const auto &inv_perm = m->inv_perm;
    auto &rl = m->rl, &yl = m->yl;
    rl.resize(m->levels);
    yl.resize(m->levels);
    for (int L = 0; L < m->levels; ++L) {
        rl[L].assign(size_t(m->counts[L]) * D, 0.0);
        yl[L].resize(size_t(m->counts[L]) * D);
    }
    const int32_t nparts = m->counts[0];
    auto fine = [&](int32_t v, int a) {
        const int64_t s = int64_t(inv_perm[v]) * bs + a;
        return s < m->n ? r[s] : 0.0;
    };
#pragma omp parallel for schedule(static, 64)
    for (int32_t p = 0; p < nparts; ++p)
        for (int32_t v = m->part_off[p]; v < m->part_off[p + 1]; ++v)
            for (int a = 0; a < bs; ++a) rl[0][size_t(p) * D + (v - m->part_off[p]) * bs + a] = fine(v, a);
    // restriction: members summed in ascending order
    for (int lv = 0; lv + 1 < m->levels; ++lv) {
        const auto &mp = m->mem_ptr[lv];
        const auto &mm = m->mem[lv];
#pragma omp parallel for schedule(static, 1024)
        for (int32_t c = 0; c < m->cco[lv]; ++c)
            for (int a = 0; a < bs; ++a) {
                double acc = 0.0;
                for (int64_t q = mp[c]; q < mp[c + 1]; ++q) acc += fine(mm[q], a);
                rl[lv + 1][size_t(c) * bs + a] = acc;
            }
    }
    for (int L = 0; L < m->levels; ++L) {
        const int32_t cnt = m->counts[L];
#pragma omp parallel for schedule(static)
        for (int32_t k = 0; k < cnt; ++k) {
            const double *Lf = &m->inv[L][size_t(k) * D * D];
            double *y = &yl[L][size_t(k) * D];
            std::copy(&rl[L][size_t(k) * D], &rl[L][size_t(k) * D] + D, y);
            chol_solve(D, Lf, y);
        }
    }
#pragma omp parallel for schedule(static, 64)
    for (int32_t p = 0; p < nparts; ++p)
        for (int32_t v = m->part_off[p]; v < m->part_off[p + 1]; ++v) {
            const int32_t old = inv_perm[v];
            for (int a = 0; a < bs; ++a) {
                const int64_t s = int64_t(old) * bs + a;
                if (s >= m->n) continue;
                double val = yl[0][size_t(p) * D + (v - m->part_off[p]) * bs + a];
                for (int lv = 0; lv + 1 < m->levels; ++lv) {
                    const int32_t c = m->map[size_t(v) * orc_mas::MAX_COARSE + lv];
                    val += yl[lv + 1][size_t(c) * bs + a];
                }
                z[s] = val;
            }
        }
    (void)B;
}

// ===================================================================== MAS solver selection
// MAS/BadDOFPreconditioner.cu, BadDofKMeanPreconditioner.cu, BadDofGMMPreconditioner.cu:
// indicators over the partition-permuted block operator and the three selections.
extern "C" void orc_mas_indicator(int mode, int32_t n, const int32_t *ptr, const int32_t *idx, const double *vals,
                                  int bs, const int32_t *perm, double *out)
{
    const int32_t nb = (n + bs - 1) / bs;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t I = 0; I < nb; ++I) {
        for (int a = 0; a < bs; ++a) {
            const int64_t pr = int64_t(perm[I]) * bs + a;
            const int64_t i = int64_t(I) * bs + a;
            if (i >= n) {  // padding row
                out[pr] = 0.0;
                continue;
            }
            // the row's entries in (permuted block column, component) order
            std::vector<std::pair<int64_t, double>> e;
            for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k)
                e.emplace_back(int64_t(perm[idx[k] / bs]) * bs + idx[k] % bs, vals[k]);
            std::sort(e.begin(), e.end(), [](const auto &l, const auto &r) { return l.first < r.first; });
            double v = 0.0;
            for (auto &[pc, val] : e) {
                if (mode == 0) v += std::fabs(val);
                else if (mode == 1) {
                    if (pc == pr) v = std::fabs(val);
                } else if (pc != pr) v = std::max(v, std::fabs(val));
            }
            out[pr] = v;
        }
    }
}

namespace {

double rel_change(double now, double before) { return std::fabs(now - before) / std::max(std::fabs(before), 1e-300); }

// union of the windows [p - e, p + e], ascending (BadDofGMMPreconditioner.cu:236-255)
int64_t windows(const std::vector<int64_t> &pos, int64_t n, int e, int64_t *out)
{
    if (e <= 0) {
        std::vector<int64_t> s(pos);
        std::sort(s.begin(), s.end());
        std::copy(s.begin(), s.end(), out);
        return int64_t(s.size());
    }
    std::vector<uint8_t> mark(size_t(n), 0);
    for (int64_t p : pos)
        for (int64_t j = std::max<int64_t>(0, p - e); j <= std::min<int64_t>(n - 1, p + e); ++j) mark[j] = 1;
    int64_t w = 0;
    for (int64_t j = 0; j < n; ++j)
        if (mark[j]) out[w++] = j;
    return w;
}

}  // namespace

extern "C" int64_t orc_mas_gmm_positions(int64_t n, const double *x, double tol, int max_iter, double jump, int expand,
                                         int64_t *out, double *stats)
{
#pragma STDC FP_CONTRACT OFF
    // select_bad_dofs_from_l1_row_norm_gmm, BadDofGMMPreconditioner.cu:86-258, sums in
    // the canonical tree order of the device
    for (int k = 0; k < 8; ++k) stats[k] = 0.0;
    if (n < 2) return 0;
    const double gmean = canonical_sum(n, [x](int64_t i) { return x[i]; }) / n;
    const double gvar = canonical_sum(n, [x, gmean](int64_t i) {
                            const double d = x[i] - gmean;
                            return d * d;
                        }) / n;
    double mean_0 = *std::min_element(x, x + n), mean_1 = *std::max_element(x, x + n);
    const double var_reg = 1e-6;
    double var_0 = std::max(gvar, var_reg), var_1 = std::max(gvar, var_reg), w0 = 0.5, w1 = 0.5;
    auto put = [&](double ratio, int it) {
        stats[0] = mean_0, stats[1] = mean_1, stats[2] = var_0, stats[3] = var_1;
        stats[4] = gmean, stats[5] = gvar, stats[6] = ratio, stats[7] = it;
    };
    put(0.0, 0);
    if (mean_0 == mean_1 || !std::isfinite(gvar)) return 0;
    std::vector<double> g0(n), g1(n);
    const int iterations = std::max(1, max_iter);
    int it = 0;
    for (; it < iterations; ++it) {
        const double lw0 = std::log(std::max(w0, std::numeric_limits<double>::min()));
        const double lw1 = std::log(std::max(w1, std::numeric_limits<double>::min()));
        const double c0 = -0.5 * std::log(2.0 * M_PI * var_0);
        const double c1 = -0.5 * std::log(2.0 * M_PI * var_1);
#pragma omp parallel for schedule(static, 4096)
        for (int64_t i = 0; i < n; ++i) {
            const double xi = x[i];
            const double lg0 = lw0 + (c0 - 0.5 * (xi - mean_0) * (xi - mean_0) / var_0);
            const double lg1 = lw1 + (c1 - 0.5 * (xi - mean_1) * (xi - mean_1) / var_1);
            const double mx = std::max(lg0, lg1);
            const double lt = mx + std::log(std::exp(lg0 - mx) + std::exp(lg1 - mx));
            g0[i] = std::exp(lg0 - lt);
            g1[i] = std::exp(lg1 - lt);
        }
        const double s0 = canonical_sum(n, [&](int64_t i) { return g0[i]; });
        const double s1 = canonical_sum(n, [&](int64_t i) { return g1[i]; });
        if (s0 <= 0.0 || s1 <= 0.0) break;
        w0 = s0 / n;
        w1 = s1 / n;
        const double om0 = mean_0, om1 = mean_1, ov0 = var_0, ov1 = var_1;
        mean_0 = canonical_sum(n, [&](int64_t i) { return g0[i] * x[i]; }) / s0;
        mean_1 = canonical_sum(n, [&](int64_t i) { return g1[i] * x[i]; }) / s1;
        const double m0 = mean_0, m1 = mean_1;
        var_0 = canonical_sum(n, [&](int64_t i) { return g0[i] * (x[i] - m0) * (x[i] - m0); }) / s0 + var_reg;
        var_1 = canonical_sum(n, [&](int64_t i) { return g1[i] * (x[i] - m1) * (x[i] - m1); }) / s1 + var_reg;
        if (rel_change(mean_0, om0) < tol && rel_change(mean_1, om1) < tol && rel_change(var_0, ov0) < tol &&
            rel_change(var_1, ov1) < tol)
            break;
    }
    const int high = std::fabs(mean_0) > std::fabs(mean_1) ? 0 : 1;
    const double hm = high == 0 ? mean_0 : mean_1, lm = high == 0 ? mean_1 : mean_0;
    const double ratio = std::fabs(hm) / std::fabs(lm);
    put(ratio, it);
    if (std::isnan(ratio) || ratio < jump) return 0;
    std::vector<int64_t> pos;
    for (int64_t i = 0; i < n; ++i)
        if (high == 0 ? g0[i] > g1[i] : g1[i] > g0[i]) pos.push_back(i);
    return windows(pos, n, expand, out);
}

// positions sorted by value (stable: equal values keep ascending positions)
static std::vector<int64_t> sorted_positions(int64_t n, const double *x, bool descending)
{
    std::vector<int64_t> o(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) o[i] = i;
    std::stable_sort(o.begin(), o.end(),
                     [&](int64_t a, int64_t b) { return descending ? x[a] > x[b] : x[a] < x[b]; });
    return o;
}

extern "C" int64_t orc_mas_elbow_positions(int64_t n, const double *x, int64_t *out, int64_t *cutoff_out)
{
#pragma STDC FP_CONTRACT OFF
    // select_bad_dofs_from_diagonal, BadDOFPreconditioner.cu:61-100
    if (n == 0) return 0;
    const auto o = sorted_positions(n, x, false);
    int64_t cutoff = 0;
    if (n > 1) {
        std::vector<double> lg(static_cast<size_t>(n));
        for (int64_t i = 0; i < n; ++i) lg[i] = std::log(x[o[i]]);
        const double mn = lg.front(), mx = lg.back();
        double best = -1.0;
        for (int64_t i = 0; i < n; ++i) {
            const double expected = (mx - mn) / (double(n) - 1.0) * double(i) + mn;
            const double dev = std::fabs(expected - lg[i]);
            if (dev > best) {
                best = dev;
                cutoff = i;
            }
        }
    }
    if (cutoff_out) *cutoff_out = cutoff;
    std::vector<int64_t> pos(o.begin() + cutoff, o.end());
    return windows(pos, n, 0, out);
}

extern "C" int64_t orc_mas_kmeans_positions(int64_t n, const double *x, double fraction, double jump, int max_iter,
                                            int expand, int64_t *out, double *stats)
{
#pragma STDC FP_CONTRACT OFF
    // select_bad_dofs_from_diagonal_kmeans, BadDofKMeanPreconditioner.cu:63-172
    for (int k = 0; k < 4; ++k) stats[k] = 0.0;
    if (n < 2) return 0;
    const auto o = sorted_positions(n, x, true);
    const int64_t sub = std::clamp<int64_t>(int64_t(double(n) * fraction), 2, n);
    double c[2] = {x[o[0]], x[o[sub - 1]]};
    int it = 0;
    for (; it < max_iter; ++it) {
        double sum[2] = {0.0, 0.0};
        int64_t cnt[2] = {0, 0};
        for (int64_t i = 0; i < sub; ++i) {
            const double v = x[o[i]];
            const int cid = std::fabs(v - c[0]) < std::fabs(v - c[1]) ? 1 : 0;
            cnt[cid]++;
            sum[cid] += v;
        }
        const double n0 = cnt[0] > 0 ? sum[0] / double(cnt[0]) : c[0];
        const double n1 = cnt[1] > 1 ? sum[1] / double(cnt[1]) : c[1];
        if (n0 == c[0] && n1 == c[1]) break;
        c[0] = n0;
        c[1] = n1;
    }
    if (c[0] < c[1]) std::swap(c[0], c[1]);
    const double ratio = c[0] / c[1];
    stats[0] = c[0], stats[1] = c[1], stats[2] = ratio, stats[3] = it;
    if (std::isnan(ratio) || ratio < jump) return 0;
    const double sep = 0.5 * (c[0] + c[1]);
    std::vector<int64_t> pos;
    for (int64_t i = 0; i < sub; ++i)
        if (x[o[i]] > sep) pos.push_back(o[i]);
    return windows(pos, n, expand, out);
}

// ===================================================================== hybrid
// damping after a breakdown of the damped sweep
constexpr double kSafeDamping = 1.5;

struct orc_hybrid {
    orc_config cfg{};
    Csr A;
    std::vector<int32_t> ptr, idx;
    std::vector<double> vals;
    std::set<int32_t> user_set;
    bool user_set_given = false;
    std::vector<std::vector<int32_t>> comps;
    std::vector<std::vector<double>> chol;  // per component, dense lower (row-major)
    orc_mas *mas = nullptr;
    std::vector<double> jacobi;
    std::vector<int32_t> perm, part_off;
    std::vector<int32_t> given_perm, given_part_off;  // optional partition supplied by the caller
    double omega = 1.0;  // damping of B inside the multiplicative sweep
    double theta = 0.0;  // Lanczos estimate of lambda_max(B A)
    bool damping_retry = false;
    int iters = 0;
    double final_res = 0.0, t_setup = 0.0, t_solve = 0.0;
    int mas_status = 0;  // MAS mode: 1 relative, 2 absolute tolerance, 3 max iterations
    double mas_residual = 0.0;
    ~orc_hybrid() { orc_mas_destroy(mas); }

    void base_apply_raw(const double *r, double *z) const
    {
        const int32_t n = A.n;
        if (cfg.base_precond == 0)
            orc_mas_apply(mas, r, z);
        else if (cfg.base_precond == 1) {
#pragma omp parallel for schedule(static)
            for (int32_t i = 0; i < n; ++i) z[i] = jacobi[i] * r[i];
        } else
            std::copy(r, r + n, z);
    }
    void base_apply(const double *r, double *z) const
    {
        base_apply_raw(r, z);
        if (omega != 1.0) {
            const int32_t n = A.n;
#pragma omp parallel for schedule(static)
            for (int32_t i = 0; i < n; ++i) z[i] *= omega;
        }
    }
    // P t: S^T D^{-1} S t  (GPUHybrid.cu:556-669)
    void stiff_apply(const double *t, double *out) const
    {
#pragma omp parallel for schedule(dynamic, 1)
        for (size_t c = 0; c < comps.size(); ++c) {
            const auto &cc = comps[c];
            std::vector<double> x(cc.size());
            for (size_t r = 0; r < cc.size(); ++r) x[r] = t[cc[r]];
            chol_solve(int(cc.size()), chol[c].data(), x.data());
            for (size_t r = 0; r < cc.size(); ++r) out[cc[r]] = x[r];
        }
    }
    // dense sub-blocks S_c A S_c^T (GPUHybrid.cu:1351-1383 semantics) + Cholesky
    bool factor_comps()
    {
        chol.resize(comps.size());
        bool ok = true;
#pragma omp parallel for schedule(dynamic, 1) reduction(&& : ok)
        for (size_t k = 0; k < comps.size(); ++k) {
            const auto &cc = comps[k];
            const int msz = int(cc.size());
            std::unordered_map<int32_t, int> loc;
            for (int r = 0; r < msz; ++r) loc[cc[r]] = r;
            std::vector<double> D(size_t(msz) * msz, 0.0);
            for (int r = 0; r < msz; ++r)
                for (int32_t e = A.ptr[cc[r]]; e < A.ptr[cc[r] + 1]; ++e) {
                    auto it = loc.find(A.idx[e]);
                    if (it != loc.end()) D[size_t(r) * msz + it->second] = A.vals[e];
                }
            if (!cholesky(msz, D)) ok = false;
            chol[k] = std::move(D);
        }
        return ok;
    }
    // custom_mixed_precond_iter, GPUHybrid.cu:521-542
    void precond(const double *r, double *z, std::vector<double> &buffer, std::vector<double> &z2) const
    {
        const int32_t n = A.n;
        if (cfg.solver_mode == 1) {
            // MASSolverImpl::apply_preconditioner, MASSolver.cu:583-615: z = B r + S^T D^{-1} S r
            base_apply_raw(r, z);
            if (!comps.empty()) {
                std::vector<double> pr(n, 0.0);
                stiff_apply(r, pr.data());
#pragma omp parallel for schedule(static)
                for (int32_t i = 0; i < n; ++i) z[i] += pr[i];
            }
            return;
        }
        if (!cfg.do_hybrid_precond || comps.empty()) {
            base_apply(r, z);
            return;
        }
        if (cfg.composition == 1) {
            // z = P r + y - P A y,  y = B (r - A P r)
            std::vector<double> pr(n, 0.0), t(n), y(n), pay(n, 0.0);
            stiff_apply(r, pr.data());
            spmv(A, pr.data(), t.data());
#pragma omp parallel for schedule(static)
            for (int32_t i = 0; i < n; ++i) t[i] = r[i] - t[i];
            base_apply_raw(t.data(), y.data());
            spmv(A, y.data(), t.data());
            stiff_apply(t.data(), pay.data());
#pragma omp parallel for schedule(static)
            for (int32_t i = 0; i < n; ++i) z[i] = pr[i] + y[i] - pay[i];
            return;
        }
        if (cfg.composition == 2) {
            std::vector<double> pr(n, 0.0);
            stiff_apply(r, pr.data());
            base_apply_raw(r, z);
#pragma omp parallel for schedule(static)
            for (int32_t i = 0; i < n; ++i) z[i] += pr[i];
            return;
        }
        base_apply(r, z);  // z1 = B r
        // dss_precond_iter: next_z = z1 + P (r - A z1)
        spmv(A, z, z2.data());
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < n; ++i) z2[i] = r[i] - z2[i];
        std::vector<double> corr(n, 0.0);
        stiff_apply(z2.data(), corr.data());
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < n; ++i) z2[i] = corr[i] + z[i];
        // z = z2 + B (r - A z2)
        spmv(A, z2.data(), z);
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < n; ++i) buffer[i] = r[i] - z[i];
        base_apply(buffer.data(), z);
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < n; ++i) z[i] += z2[i];
    }
};

namespace {

// Largest eigenvalue of the symmetric tridiagonal (d, e) by Sturm-count bisection.
double tridiag_max_eig(const std::vector<double> &d, const std::vector<double> &e)
{
    const int m = int(d.size());
    double lo = 0.0, hi = 0.0;
    for (int i = 0; i < m; ++i) {
        const double rad = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(e[i]) : 0.0);
        hi = std::max(hi, d[i] + rad);
        lo = std::min(lo, d[i] - rad);
    }
    auto count_below = [&](double x) {  // eigenvalues < x
        int cnt = 0;
        double q = d[0] - x;
        if (q < 0) ++cnt;
        for (int i = 1; i < m; ++i) {
            if (q == 0.0) q = 1e-300;
            q = d[i] - x - e[i - 1] * e[i - 1] / q;
            if (q < 0) ++cnt;
        }
        return cnt;
    };
    for (int it = 0; it < 200 && hi - lo > 1e-13 * std::max(1.0, std::fabs(hi)); ++it) {
        const double mid = 0.5 * (lo + hi);
        if (count_below(mid) >= m) hi = mid;
        else lo = mid;
    }
    return hi;
}

}  // namespace

double lanczos_lambda_max(const orc_hybrid &h, int32_t n)
{
    // CG-Lanczos on A with preconditioner B (undamped), x0 = 0, b = hashed vector
    std::vector<double> r(n), z(n), p(n), q(n);
    for (int32_t i = 0; i < n; ++i) r[i] = double(uint32_t(uint32_t(i) * 2654435761u)) * 0x1p-32 - 0.5;
    h.base_apply_raw(r.data(), z.data());
    p = z;
    double rz = dot(n, r.data(), z.data());
    std::vector<double> alpha, beta;
    for (int j = 0; j < kLanczosSteps; ++j) {
        spmv(h.A, p.data(), q.data());
        const double pq = dot(n, p.data(), q.data());
        if (!(pq > 0.0)) break;
        const double a = rz / pq;
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < n; ++i) r[i] -= a * q[i];
        h.base_apply_raw(r.data(), z.data());
        const double rz2 = dot(n, r.data(), z.data());
        const double b = rz2 / rz;
        alpha.push_back(a);
        beta.push_back(b);
        if (!(rz2 > 0.0)) break;
        rz = rz2;
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < n; ++i) p[i] = z[i] + b * p[i];
    }
    const int m = int(alpha.size());
    std::vector<double> d(m), e(m > 0 ? m - 1 : 0);
    for (int j = 0; j < m; ++j) {
        d[j] = 1.0 / alpha[j] + (j > 0 ? beta[j - 1] / alpha[j - 1] : 0.0);
        if (j + 1 < m) e[j] = std::sqrt(beta[j]) / alpha[j];
    }
    return m > 0 ? tridiag_max_eig(d, e) : 1.0;
}

extern "C" orc_hybrid *orc_hybrid_create(const orc_config *cfg)
{
    auto h = new orc_hybrid();
    h->cfg = *cfg;
    return h;
}
extern "C" void orc_hybrid_destroy(orc_hybrid *h) { delete h; }

extern "C" void orc_hybrid_set_problematic(orc_hybrid *h, const int32_t *dofs, int64_t n)
{
    h->user_set.clear();
    h->user_set.insert(dofs, dofs + n);
    h->user_set_given = true;
}

extern "C" int orc_hybrid_factorize(orc_hybrid *h, int32_t n, const int32_t *ptr, const int32_t *idx,
                                    const double *vals)
{
#ifdef _OPENMP
    if (h->cfg.threads > 0) omp_set_num_threads(h->cfg.threads);
#endif
    auto t0 = clk::now();
    const int64_t nnz = ptr[n];
    h->ptr.assign(ptr, ptr + n + 1);
    h->idx.assign(idx, idx + nnz);
    h->vals.assign(vals, vals + nnz);
    h->A = Csr{n, h->ptr.data(), h->idx.data(), h->vals.data()};
    const Csr &A = h->A;
    const orc_config &c = h->cfg;
    h->comps.clear();
    h->chol.clear();
    auto tp = clk::now();
    if (c.solver_mode == 0 && c.do_hybrid_precond) {
        // GPUHybrid.cu:214-264
        std::vector<int32_t> S;
        if (c.select_bad_dofs_from_l1_norm) {
            std::vector<double> norms(n);
            orc_row_l1(n, A.ptr, A.vals, norms.data());
            std::vector<uint8_t> fl(n);
            orc_gmm_select(n, norms.data(), c.gmm_tol, c.max_gmm_iterations, c.gmm_jump_threshold, fl.data(),
                           nullptr, nullptr);
            for (int32_t i = 0; i < n; ++i)
                if (fl[i]) S.push_back(i);
        } else {
            S.assign(h->user_set.begin(), h->user_set.end());
        }
        std::vector<int32_t> tmp(n);
        int64_t m = 0;
        if (c.decompose_subdomains) {
            int64_t st[3];
            orc_filter(n, A.ptr, A.idx, A.vals, S.data(), int64_t(S.size()), c.min_subdomain_size,
                       c.max_subdomain_size, c.conditioning_threshold, tmp.data(), &m, st);
            S.assign(tmp.begin(), tmp.begin() + m);
        }
        if (c.expand_subdomains) {
            orc_expand(n, A.ptr, A.idx, S.data(), int64_t(S.size()), tmp.data(), &m);
            S.assign(tmp.begin(), tmp.begin() + m);
        }
        if (c.decompose_subdomains) {
            std::vector<int32_t> dofs(S.size());
            std::vector<int64_t> off(S.size() + 1);
            int64_t nc = 0;
            orc_decompose(n, A.ptr, A.idx, S.data(), int64_t(S.size()), c.max_subdomain_size, dofs.data(),
                          off.data(), &nc);
            for (int64_t k = 0; k < nc; ++k) h->comps.emplace_back(dofs.begin() + off[k], dofs.begin() + off[k + 1]);
        } else if (!S.empty()) {
            h->comps.push_back(S);
        }
        phase("select", tp);
        // dense sub-blocks S_c A S_c^T (GPUHybrid.cu:1351-1383 semantics) + Cholesky
        if (!h->factor_comps()) return -5;
        phase("stiff", tp);
    }
    // base preconditioner
    orc_mas_destroy(h->mas);
    h->mas = nullptr;
    const int bs = c.block_size;
    if (c.base_precond == 0) {
        const int32_t nb = (n + bs - 1) / bs;
        std::vector<std::vector<int32_t>> adj(nb);
#pragma omp parallel for schedule(dynamic, 1024)
        for (int32_t I = 0; I < nb; ++I) {
            auto &a = adj[I];
            for (int32_t i = I * bs; i < std::min<int64_t>(int64_t(I + 1) * bs, n); ++i)
                for (int32_t e = A.ptr[i]; e < A.ptr[i + 1]; ++e) a.push_back(A.idx[e] / bs);
            std::sort(a.begin(), a.end());
            a.erase(std::unique(a.begin(), a.end()), a.end());
        }
        std::vector<int64_t> bptr(nb + 1, 0);
        for (int32_t I = 0; I < nb; ++I) bptr[I + 1] = bptr[I] + int64_t(adj[I].size());
        std::vector<int32_t> bidx(bptr[nb]);
#pragma omp parallel for schedule(static, 4096)
        for (int32_t I = 0; I < nb; ++I) std::copy(adj[I].begin(), adj[I].end(), bidx.begin() + bptr[I]);
        std::vector<int32_t> part_of(nb);
        int32_t np = 0;
        if (!h->given_perm.empty() && int32_t(h->given_perm.size()) == nb) {
            np = int32_t(h->given_part_off.size()) - 1;
            for (int32_t I = 0; I < nb; ++I)
                part_of[I] = int32_t(std::upper_bound(h->given_part_off.begin(), h->given_part_off.end(),
                                                      h->given_perm[I]) - h->given_part_off.begin()) - 1;
        } else {
            phase("block_graph", tp);
            orc_partition(nb, bptr.data(), bidx.data(), 32, part_of.data(), &np);
            phase("partition", tp);
        }
        // permutation sorted by part, stable in node id (MASSolver.cu:90-118)
        h->part_off.assign(np + 1, 0);
        for (int32_t I = 0; I < nb; ++I) h->part_off[part_of[I] + 1]++;
        for (int32_t p = 0; p < np; ++p) h->part_off[p + 1] += h->part_off[p];
        std::vector<int32_t> next(h->part_off.begin(), h->part_off.end() - 1);
        h->perm.assign(nb, 0);
        for (int32_t I = 0; I < nb; ++I) h->perm[I] = next[part_of[I]]++;
        h->mas = orc_mas_create(n, A.ptr, A.idx, A.vals, bs, h->perm.data(), h->part_off.data(), np);
        if (!h->mas) return -5;
        phase("mas", tp);
        if (c.solver_mode == 1 && c.use_bad_dof) {
            // MAS solver: the selected positions over the permuted operator (MASSolver.cu:398-432)
            const int64_t npos = int64_t(nb) * bs;
            std::vector<double> ind(npos);
            const int mode = c.bad_dof_strategy == 2 ? 0 : (c.bad_dof_strategy == 1 ? 2 : 1);
            orc_mas_indicator(mode, n, A.ptr, A.idx, A.vals, bs, h->perm.data(), ind.data());
            // positions: non-padding permuted scalars in order
            std::vector<int64_t> pos2p;
            std::vector<double> x;
            const int32_t pad_block = h->perm[nb - 1];
            const int npad = int(npos - n);
            for (int64_t p = 0; p < npos; ++p) {
                if (npad && p / bs == pad_block && p % bs >= bs - npad) continue;
                pos2p.push_back(p);
                x.push_back(ind[p]);
            }
            const int64_t m = int64_t(x.size());
            std::vector<int64_t> sel(m);
            double st[8];
            int64_t cnt = 0;
            if (c.bad_dof_strategy == 2)
                cnt = orc_mas_gmm_positions(m, x.data(), c.mgmm_tol, c.mgmm_max_iter, c.mgmm_jump, c.mgmm_expand,
                                            sel.data(), st);
            else if (c.bad_dof_strategy == 1)
                cnt = orc_mas_kmeans_positions(m, x.data(), c.km_fraction, c.km_jump, c.km_max_iter, c.km_expand,
                                               sel.data(), st);
            else
                cnt = orc_mas_elbow_positions(m, x.data(), sel.data(), nullptr);
            std::vector<int32_t> inv(nb);
            for (int32_t I = 0; I < nb; ++I) inv[h->perm[I]] = I;
            std::vector<int32_t> S;
            for (int64_t k = 0; k < cnt; ++k) {
                const int64_t p = pos2p[sel[k]];
                S.push_back(int32_t(int64_t(inv[p / bs]) * bs + p % bs));
            }
            std::sort(S.begin(), S.end());
            if (!S.empty()) {
                std::vector<int32_t> dofs(S.size());
                std::vector<int64_t> off(S.size() + 1);
                int64_t nc = 0;
                orc_decompose(n, A.ptr, A.idx, S.data(), int64_t(S.size()), std::numeric_limits<int64_t>::max(),
                              dofs.data(), off.data(), &nc);
                for (int64_t k = 0; k < nc; ++k)
                    h->comps.emplace_back(dofs.begin() + off[k], dofs.begin() + off[k + 1]);
            }
            if (!h->factor_comps()) return -5;
            phase("mas_bad_dof", tp);
        }
    } else if (c.base_precond == 1) {
        h->jacobi.assign(n, 1.0);
        for (int32_t i = 0; i < n; ++i)
            for (int32_t e = A.ptr[i]; e < A.ptr[i + 1]; ++e)
                if (A.idx[e] == i && A.vals[e] != 0.0) h->jacobi[i] = 1.0 / A.vals[e];
    }
    // The additive base preconditioner B can have lambda_max(BA) > 2, which makes the
    // multiplicative sweep (I - BA)(I - PA)(I - BA) non-contractive and the composed
    // preconditioner indefinite; damp B by omega = damping / theta, theta the largest
    // Ritz value of B A after kLanczosSteps steps of B-preconditioned CG on A x = v
    // (v a hashed vector), which never exceeds lambda_max(BA) and approaches it fast.
    h->omega = 1.0;
    if (c.solver_mode == 0 && c.do_hybrid_precond && !h->comps.empty() && c.composition == 0) {
        const double theta = lanczos_lambda_max(*h, n);
        h->theta = theta;
        h->omega = (c.damping > 0 ? c.damping : 1.0) / theta;
        phase("lanczos", tp);
    }
    h->t_setup = since(t0);
    return 0;
}

// MASSolverImpl::pcg_solve, MASSolver.cu:618-744 (x0 = 0, MASSolver.cu:500-504)
static int mas_solve(orc_hybrid *h, const double *b, double *xio)
{
    auto t0 = clk::now();
    const Csr &A = h->A;
    const int32_t n = A.n;
    const orc_config &c = h->cfg;
    std::vector<double> x(n, 0.0), r(n), p(n), z(n), Ap(n), buffer(n), z2(n);
    auto div = [](double num, double den) { return std::fabs(den) < 1e-20 ? 0.0 : num / den; };
    spmv(A, x.data(), r.data());
    for (int32_t i = 0; i < n; ++i) r[i] = 1.0 * b[i] + -1.0 * r[i];
    h->precond(r.data(), z.data(), buffer, z2);
    p = z;
    double rz = dot(n, r.data(), z.data());
    if (std::isnan(rz) || !std::isfinite(rz)) return -5;
    const double rr0 = c.precond_norm ? rz : dot(n, r.data(), r.data());
    const double rtol2 = c.rel_tol * c.rel_tol, atol2 = c.abs_tol * c.abs_tol;
    h->mas_status = 0;
    h->mas_residual = 0.0;
    int iters = 0;
    for (int k = 1; k <= c.max_iter; ++k) {
        spmv(A, p.data(), Ap.data());
        const double alpha = div(rz, dot(n, p.data(), Ap.data()));
        for (int32_t i = 0; i < n; ++i) x[i] = alpha * p[i] + 1.0 * x[i];
        if (c.true_residual_period > 0 && k % c.true_residual_period == 0) {
            spmv(A, x.data(), r.data());
            for (int32_t i = 0; i < n; ++i) r[i] = 1.0 * b[i] + -1.0 * r[i];
        } else {
            for (int32_t i = 0; i < n; ++i) r[i] = -alpha * Ap[i] + 1.0 * r[i];
        }
        h->precond(r.data(), z.data(), buffer, z2);
        const double rz_old = rz;
        rz = dot(n, r.data(), z.data());
        iters = k;
        bool converged = false;
        if (k % 10 == 0) {
            const double rr = c.precond_norm ? rz : dot(n, r.data(), r.data());
            h->mas_residual = std::sqrt(rr);
            if (rr <= rtol2 * rr0 || rr <= atol2) {
                h->mas_status = rr <= atol2 ? 2 : 1;
                converged = true;
            }
        }
        if (converged) break;
        const double beta = div(rz, rz_old);
        for (int32_t i = 0; i < n; ++i) p[i] = 1.0 * z[i] + beta * p[i];
    }
    if (iters == c.max_iter) h->mas_status = 3;
    spmv(A, x.data(), buffer.data());
    for (int32_t i = 0; i < n; ++i) buffer[i] -= b[i];
    h->final_res = std::sqrt(dot(n, buffer.data(), buffer.data()));
    h->iters = iters;
    std::copy(x.begin(), x.end(), xio);
    h->t_solve = since(t0);
    return 0;
}

extern "C" int orc_hybrid_mas_status(const orc_hybrid *h, double *residual)
{
    if (residual) *residual = h->mas_residual;
    return h->mas_status;
}

extern "C" void orc_hybrid_set_run(orc_hybrid *h, int threads, int max_iter)
{
    h->cfg.threads = threads;
    h->cfg.max_iter = max_iter;
}

extern "C" int orc_hybrid_solve(orc_hybrid *h, const double *b, double *xio)
{
#ifdef _OPENMP
    if (h->cfg.threads > 0) omp_set_num_threads(h->cfg.threads);
#endif
    if (h->cfg.solver_mode == 1) return mas_solve(h, b, xio);
    // GPUHybrid::pcg_solve, GPUHybrid.cu:1413-1512
    auto t0 = clk::now();
    const Csr &A = h->A;
    const int32_t n = A.n;
    std::vector<double> x(xio, xio + n), r(n), p(n), z(n), z2(n), buffer(n);
    const double bi_prod = dot(n, b, b);
    double rel_eps, abs_eps;
    if (bi_prod > 0.0) {
        rel_eps = h->cfg.rel_tol * h->cfg.rel_tol;
        abs_eps = h->cfg.abs_tol * h->cfg.abs_tol;
    } else {
        std::fill(xio, xio + n, 0.0);
        h->iters = 0;
        h->final_res = 0.0;
        h->t_solve = since(t0);
        return 0;
    }
    // returns true on a breakdown (GPUHybrid.cu:1463-1480)
    auto pcg = [&](int &iters) {
        spmv(A, x.data(), buffer.data());
        for (int32_t i = 0; i < n; ++i) r[i] = b[i] - buffer[i];
        h->precond(r.data(), z.data(), buffer, z2);
        p = z;
        double gamma = dot(n, r.data(), z.data()), old_gamma = gamma;
        iters = 0;
        for (int k = 0; k < h->cfg.max_iter; ++k) {
            iters = k + 1;
            spmv(A, p.data(), buffer.data());
            const double sdotp = dot(n, p.data(), buffer.data());
            if (sdotp == 0.0) return true;
            const double alpha = gamma / sdotp;
            if (alpha <= 0.0) return true;
            if (alpha < std::numeric_limits<double>::min()) return true;
#pragma omp parallel for schedule(static)
            for (int32_t i = 0; i < n; ++i) {
                x[i] += alpha * p[i];
                r[i] += -alpha * buffer[i];
            }
            const double i_prod = dot(n, r.data(), r.data());
            if (rel_eps > 0 && (i_prod / bi_prod) < rel_eps) break;
            if (abs_eps > 0 && i_prod < abs_eps) break;
            h->precond(r.data(), z.data(), buffer, z2);
            gamma = dot(n, r.data(), z.data());
            const double beta = gamma / old_gamma;
            old_gamma = gamma;
#pragma omp parallel for schedule(static)
            for (int32_t i = 0; i < n; ++i) p[i] = beta * p[i] + z[i];
        }
        return false;
    };
    int iters = 0;
    const double omega = h->omega;
    h->damping_retry = false;
    if (pcg(iters) && h->theta > 0 && h->cfg.damping > kSafeDamping) {
        // the damped sweep lost definiteness: again from the same start, safe damping
        h->damping_retry = true;
        h->omega = kSafeDamping / h->theta;
        std::copy(xio, xio + n, x.begin());
        pcg(iters);
        h->omega = omega;
    }
    // true residual, GPUHybrid.cu:383-386
    spmv(A, x.data(), buffer.data());
    for (int32_t i = 0; i < n; ++i) buffer[i] -= b[i];
    h->final_res = std::sqrt(dot(n, buffer.data(), buffer.data()));
    h->iters = iters;
    std::copy(x.begin(), x.end(), xio);
    h->t_solve = since(t0);
    return 0;
}

extern "C" double orc_hybrid_omega(const orc_hybrid *h) { return h->omega; }

// Replace the damping factor found by factorize (tests compare M^{-1} with a device
// whose Lanczos estimate differs in the last digits).
extern "C" void orc_hybrid_set_omega(orc_hybrid *h, double omega) { h->omega = omega; }

extern "C" int orc_hybrid_damping_retry(const orc_hybrid *h) { return h->damping_retry ? 1 : 0; }

extern "C" void orc_hybrid_set_partition(orc_hybrid *h, const int32_t *perm, int32_t nb, const int32_t *part_off,
                                         int32_t nparts)
{
    h->given_perm.assign(perm, perm + nb);
    h->given_part_off.assign(part_off, part_off + nparts + 1);
}

extern "C" void orc_hybrid_info(const orc_hybrid *h, int *iters, double *final_res, double *t_setup, double *t_solve,
                                int64_t *n_stiff, int64_t *n_comp)
{
    *iters = h->iters;
    *final_res = h->final_res;
    *t_setup = h->t_setup;
    *t_solve = h->t_solve;
    int64_t s = 0;
    for (auto &c : h->comps) s += int64_t(c.size());
    *n_stiff = s;
    *n_comp = int64_t(h->comps.size());
}

extern "C" int64_t orc_hybrid_stiff(const orc_hybrid *h, int32_t *dofs, int64_t cap, int64_t *comp_off, int64_t ccap)
{
    int64_t w = 0;
    if (ccap > 0) comp_off[0] = 0;
    for (size_t k = 0; k < h->comps.size(); ++k) {
        for (int32_t d : h->comps[k]) {
            if (w < cap) dofs[w] = d;
            ++w;
        }
        if (int64_t(k + 1) < ccap) comp_off[k + 1] = w;
    }
    return w;
}

extern "C" void orc_hybrid_precond(orc_hybrid *h, const double *r, double *z)
{
    std::vector<double> buffer(h->A.n), z2(h->A.n);
    h->precond(r, z, buffer, z2);
}

extern "C" int32_t orc_hybrid_partition(const orc_hybrid *h, int32_t *perm, int32_t *part_off)
{
    if (perm) std::copy(h->perm.begin(), h->perm.end(), perm);
    if (part_off) std::copy(h->part_off.begin(), h->part_off.end(), part_off);
    return int32_t(h->part_off.size()) - 1;
}
