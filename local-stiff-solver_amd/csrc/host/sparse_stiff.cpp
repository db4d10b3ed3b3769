// Multifrontal sparse Cholesky of one large stiff unit (see sparse_stiff.hpp).
//
// Ordering: nested dissection.  A connected subgraph is split by a level set of a
// breadth-first search from a pseudo-peripheral vertex (two sweeps from its smallest
// vertex): the first level at which the levels so far hold half the vertices, thinned
// to the vertices that touch the next level, is the separator; the levels before it
// (and the thinned-out vertices) and the levels after it are dissected recursively,
// and the separator is numbered after both.  Disconnected subgraphs are dissected per
// component.  Subgraphs of at most `leaf` vertices, or with fewer than three levels,
// are numbered whole.  Each numbered block is one front (a supernode whose columns
// are eliminated together); a front's update positions are its pivots' neighbours
// and its children's update positions beyond its own pivots.
//
// The dissection runs on the quotient graph of indistinguishable vertices (equal closed
// neighbourhoods: the DOFs of one mesh node), each group weighted by its size.
//
// Numeric factorization, level by level of the assembly tree (the fronts of a level are
// independent): zero each m x m front, write the original entries of its pivot columns,
// add the children's update matrices (extend-add, children in order), and factor the
// first p columns, which leaves the Schur complement in the trailing block.  Fronts of up
// to batch_rows rows (4 096 by default) go through batched kernels, one launch per step for all fronts of
// the level: a 128-column step is the diagonal block (factor + inverse, one workgroup per
// front), the panel below it and the trailing update in 64 x 64 tiles; the off-diagonal
// blocks of L11^{-1} follow by block distance.  Larger fronts take the blocked dense path
// (DenseBlas::partial_factor_inverse, rocBLAS products) one at a time.  Kept per front:
// G11 = L11^{-1} (with G11^T in its strict upper triangle), L21 and L21^T.
#include "sparse_stiff.hpp"

#include <algorithm>
#include <atomic>
#include <array>
#include <chrono>
#include <functional>
#include <map>
#include <memory>
#include <numeric>
#include <parallel/algorithm>
#include <string>

#include "../kernels/kernels.hpp"
#include "../kernels/mas_kernels.hpp"

namespace lss {
namespace {

using clk = std::chrono::steady_clock;
double since(clk::time_point t0) { return std::chrono::duration<double>(clk::now() - t0).count(); }

// Largest front (rows) the numeric phase accepts: its dense workspace takes m^2 doubles.
constexpr int32_t kMaxFront = 49152;

int64_t align_up(int64_t n) { return (n + 31) / 32 * 32; }  // 256-byte granules

// First-fit placement of blocks in one array (offsets in doubles); peak() is the extent used.
class FirstFit {
public:
    int64_t take(int64_t n)
    {
        if (n == 0) return 0;
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                const int64_t at = it->first, rest = it->second - n;
                free_.erase(it);
                if (rest) free_[at + n] = rest;
                return at;
            }
        int64_t at = top_;
        if (!free_.empty()) {  // a free block at the top grows into the new one
            auto last = std::prev(free_.end());
            if (last->first + last->second == top_) {
                at = last->first;
                free_.erase(last);
            }
        }
        top_ = at + n;
        peak_ = std::max(peak_, top_);
        return at;
    }
    void release(int64_t at, int64_t n)
    {
        if (n == 0) return;
        auto next = free_.lower_bound(at);
        if (next != free_.end() && at + n == next->first) {
            n += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == at) {
                prev->second += n;
                return;
            }
        }
        free_[at] = n;
    }
    int64_t peak() const { return peak_; }

private:
    std::map<int64_t, int64_t> free_;
    int64_t top_ = 0, peak_ = 0;
};

struct Range {
    int64_t off = 0, n = 0;
};
// Launch ranges of one batch of fronts (into the task lists of the numeric phase).
struct GroupPlan {
    int64_t wsize = 0;                         // workspace doubles
    Range fr;                                  // its fronts (in the front list)
    std::vector<Range> ext;                    // extend-add, one per child slot
    std::vector<std::array<Range, 3>> steps;   // per 128-column step: diagonal block, panel, trailing tiles
    std::vector<std::array<Range, 2>> rounds;  // per block distance of G11: T tiles, G tiles
    Range cl, clt, cu, cm;                     // L21, L21^T, update matrix, G11 mirror
};

// Dissects a graph whose vertices carry weights (the sizes of groups of original
// vertices), breadth-first: all subgraphs of one depth are split by one parallel loop,
// then the tree of splits is numbered in postorder, which gives the recursion's order.
// A split reads and writes the per-vertex arrays of its own vertices only, except `mark`
// of neighbours (read to test membership; every subset has a stamp of its own), so the
// result does not depend on the schedule.
struct Dissector {
    const int32_t *ap, *ai;     // quotient graph
    const int32_t *mptr;        // group sizes: mptr[v + 1] - mptr[v]
    int64_t leaf;
    std::vector<int32_t> mark, lev, vis;
    std::atomic<int32_t> stamp{0};

    // one node of the tree of splits: a front (kLeaf), the components of a disconnected
    // subgraph (kParts), or a separator (verts) with its two sides (kSep)
    enum Kind { kLeaf, kParts, kSep };
    struct Split {
        Kind kind = kLeaf;
        std::vector<int32_t> verts;
        std::vector<std::vector<int32_t>> parts;
    };

    int64_t weight(int32_t v) const { return mptr[v + 1] - mptr[v]; }
    int32_t marked(int32_t w) const { return __atomic_load_n(&mark[w], __ATOMIC_RELAXED); }
    int32_t fresh() { return stamp.fetch_add(1, std::memory_order_relaxed) + 1; }

    // breadth-first search over the current subset from src: q = visit order, lev = levels
    void bfs(int32_t src, int32_t my, std::vector<int32_t> &q)
    {
        const int32_t vs = fresh();
        q.clear();
        q.push_back(src);
        vis[src] = vs;
        lev[src] = 0;
        for (size_t h = 0; h < q.size(); ++h) {
            const int32_t v = q[h];
            for (int32_t e = ap[v]; e < ap[v + 1]; ++e) {
                const int32_t w = ai[e];
                if (marked(w) == my && vis[w] != vs) {
                    vis[w] = vs;
                    lev[w] = lev[v] + 1;
                    q.push_back(w);
                }
            }
        }
    }

    Split split(std::vector<int32_t> V)
    {
        Split out;
        int64_t nv = 0;
        for (int32_t v : V) nv += weight(v);
        if (nv <= leaf) {
            out.verts = std::move(V);
            return out;
        }
        const int32_t my = fresh();
        for (int32_t v : V) __atomic_store_n(&mark[v], my, __ATOMIC_RELAXED);
        std::sort(V.begin(), V.end());
        // connected components, in the order of their smallest vertex
        std::vector<std::vector<int32_t>> comps;
        const int32_t cs = fresh();
        for (int32_t s0 : V) {
            if (vis[s0] == cs) continue;
            comps.emplace_back();
            std::vector<int32_t> &c = comps.back();
            c.push_back(s0);
            vis[s0] = cs;
            for (size_t h = 0; h < c.size(); ++h) {
                const int32_t v = c[h];
                for (int32_t e = ap[v]; e < ap[v + 1]; ++e) {
                    const int32_t w = ai[e];
                    if (marked(w) == my && vis[w] != cs) {
                        vis[w] = cs;
                        c.push_back(w);
                    }
                }
            }
        }
        if (comps.size() > 1) {
            out.kind = kParts;
            out.parts = std::move(comps);
            return out;
        }
        std::vector<int32_t> q;
        bfs(V.front(), my, q);
        bfs(q.back(), my, q);
        const int32_t nlev = lev[q.back()] + 1;
        if (nlev < 3) {
            out.verts = std::move(V);
            return out;
        }
        std::vector<int64_t> cnt(nlev, 0);
        for (int32_t v : q) cnt[lev[v]] += weight(v);
        // the smallest level whose predecessors hold 35-65 % of the weight (the one
        // nearest to half among equals); else the level that reaches half
        const int64_t half = nv / 2;
        int64_t cum = 0, kcum = 0;
        int32_t k = -1, kh = -1;
        for (int32_t l = 1; l < nlev - 1; ++l) {
            cum += cnt[l - 1];  // weight before level l
            if (kh < 0 && cum + cnt[l] >= half) kh = l;
            if (20 * cum < 7 * nv || 20 * cum > 13 * nv) continue;
            if (k < 0 || cnt[l] < cnt[k] ||
                (cnt[l] == cnt[k] && std::llabs(2 * cum - nv) < std::llabs(2 * kcum - nv))) {
                k = l;
                kcum = cum;
            }
        }
        if (k < 0) k = kh < 0 ? 1 : kh;
        std::vector<int32_t> left, right;
        for (int32_t v : q) {
            const int32_t l = lev[v];
            if (l < k) {
                left.push_back(v);
            } else if (l > k) {
                right.push_back(v);
            } else {
                bool cut = false;
                for (int32_t e = ap[v]; e < ap[v + 1] && !cut; ++e) {
                    const int32_t w = ai[e];
                    cut = marked(w) == my && lev[w] == k + 1;
                }
                (cut ? out.verts : left).push_back(v);
            }
        }
        out.kind = kSep;
        out.parts.push_back(std::move(left));
        out.parts.push_back(std::move(right));
        return out;
    }

    // the fronts of the graph V in postorder: verts[f] (quotient vertices), kids[f]
    void dissect(std::vector<int32_t> V, std::vector<std::vector<int32_t>> &verts,
                 std::vector<std::vector<int32_t>> &kids)
    {
        struct Node {
            Kind kind = kLeaf;
            std::vector<int32_t> verts, child;
        };
        std::vector<Node> nodes(1);
        std::vector<std::pair<int32_t, std::vector<int32_t>>> frontier, next;
        frontier.emplace_back(0, std::move(V));
        while (!frontier.empty()) {
            std::vector<Split> res(frontier.size());
#pragma omp parallel for schedule(dynamic, 1)
            for (size_t i = 0; i < frontier.size(); ++i) res[i] = split(std::move(frontier[i].second));
            next.clear();
            for (size_t i = 0; i < frontier.size(); ++i) {
                const int32_t id = frontier[i].first;
                nodes[id].kind = res[i].kind;
                nodes[id].verts = std::move(res[i].verts);
                for (auto &part : res[i].parts) {
                    const int32_t c = int32_t(nodes.size());
                    nodes.emplace_back();
                    nodes[id].child.push_back(c);
                    next.emplace_back(c, std::move(part));
                }
            }
            frontier.swap(next);
        }
        // postorder: the sides (or components) first, then the separator above them
        std::function<std::vector<int32_t>(int32_t)> emit = [&](int32_t id) -> std::vector<int32_t> {
            Node &nd = nodes[id];
            std::vector<int32_t> roots;
            for (int32_t c : nd.child) {
                const std::vector<int32_t> r = emit(c);
                roots.insert(roots.end(), r.begin(), r.end());
            }
            if (nd.kind == kParts) return roots;
            verts.push_back(std::move(nd.verts));
            kids.push_back(std::move(roots));
            return {int32_t(verts.size()) - 1};
        };
        emit(0);
    }
};

uint64_t mix64(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// Groups of indistinguishable vertices (equal closed neighbourhoods N(v) + v), numbered
// by their smallest member: grp[v] = group of v; members listed ascending per group.
int32_t indistinguishable(int32_t m, const int32_t *ap, const int32_t *ai, std::vector<int32_t> &grp,
                          std::vector<int32_t> &mptr, std::vector<int32_t> &mem)
{
    std::vector<uint64_t> h(m);
#pragma omp parallel for schedule(static)
    for (int32_t v = 0; v < m; ++v) {
        uint64_t x = mix64(uint64_t(v));
        for (int32_t e = ap[v]; e < ap[v + 1]; ++e) x += mix64(uint64_t(ai[e]));
        h[v] = x;
    }
    std::vector<int32_t> by(m);
    std::iota(by.begin(), by.end(), 0);
    __gnu_parallel::sort(by.begin(), by.end(), [&](int32_t a, int32_t b) {
        const int32_t da = ap[a + 1] - ap[a], db = ap[b + 1] - ap[b];
        return h[a] != h[b] ? h[a] < h[b] : da != db ? da < db : a < b;
    });
    // closed neighbourhoods of u and v equal: same degree, each in the other's list, and
    // the lists equal once each vertex's own entry takes the other's place
    auto same = [&](int32_t u, int32_t v) {
        if (ap[u + 1] - ap[u] != ap[v + 1] - ap[v]) return false;
        if (!std::binary_search(ai + ap[u], ai + ap[u + 1], v)) return false;
        int32_t i = ap[u], j = ap[v];
        while (i < ap[u + 1] && j < ap[v + 1]) {
            const int32_t a = ai[i] == v ? u : ai[i], b = ai[j] == u ? v : ai[j];
            if (ai[i] == v) { ++i; continue; }  // v in N(u) matches u in N(v): skip both
            if (ai[j] == u) { ++j; continue; }
            if (a != b) return false;
            ++i, ++j;
        }
        return true;
    };
    grp.assign(m, -1);
    std::vector<int32_t> lead;  // smallest member per provisional group
    for (int32_t i = 0; i < m;) {
        int32_t e = i + 1;
        while (e < m && h[by[e]] == h[by[i]] && ap[by[e] + 1] - ap[by[e]] == ap[by[i] + 1] - ap[by[i]]) ++e;
        for (int32_t a = i; a < e; ++a) {
            const int32_t u = by[a];
            if (grp[u] >= 0) continue;
            grp[u] = int32_t(lead.size());
            lead.push_back(u);
            for (int32_t b = a + 1; b < e; ++b)
                if (grp[by[b]] < 0 && same(u, by[b])) grp[by[b]] = grp[u];
        }
        i = e;
    }
    // renumber groups by smallest member (u was the smallest: by[] is ascending within ties)
    const int32_t ng = int32_t(lead.size());
    std::vector<int32_t> rank(ng);
    {
        std::vector<int32_t> ord(ng);
        std::iota(ord.begin(), ord.end(), 0);
        std::sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) { return lead[a] < lead[b]; });
        for (int32_t r = 0; r < ng; ++r) rank[ord[r]] = r;
    }
    mptr.assign(ng + 1, 0);
    for (int32_t v = 0; v < m; ++v) {
        grp[v] = rank[grp[v]];
        ++mptr[grp[v] + 1];
    }
    for (int32_t g = 0; g < ng; ++g) mptr[g + 1] += mptr[g];
    mem.assign(m, 0);
    std::vector<int32_t> cur(mptr.begin(), mptr.end() - 1);
    for (int32_t v = 0; v < m; ++v) mem[cur[grp[v]]++] = v;  // ascending per group
    return ng;
}

}  // namespace

FrontTree nested_dissection(int32_t m, const int32_t *adj_ptr, const int32_t *adj_idx, int32_t leaf)
{
    FrontTree T;
    T.m = m;
    T.order.assign(m, 0);
    T.pos.assign(m, 0);
    // dissect the quotient graph of indistinguishable vertices, weighted by group size
    std::vector<int32_t> grp, mptr, mem, qptr, qidx;
    const int32_t ng = indistinguishable(m, adj_ptr, adj_idx, grp, mptr, mem);
    {
        // neighbour groups of each group's first member (all members have the same ones)
        auto nbs = [&](int32_t g, std::vector<int32_t> &nb) {
            const int32_t r = mem[mptr[g]];
            nb.clear();
            for (int32_t e = adj_ptr[r]; e < adj_ptr[r + 1]; ++e)
                if (grp[adj_idx[e]] != g) nb.push_back(grp[adj_idx[e]]);
            std::sort(nb.begin(), nb.end());
            nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
        };
        qptr.assign(size_t(ng) + 1, 0);
#pragma omp parallel
        {
            std::vector<int32_t> nb;
#pragma omp for schedule(dynamic, 1024)
            for (int32_t g = 0; g < ng; ++g) {
                nbs(g, nb);
                qptr[g + 1] = int32_t(nb.size());
            }
        }
        for (int32_t g = 0; g < ng; ++g) qptr[g + 1] += qptr[g];
        qidx.resize(qptr[ng]);
#pragma omp parallel
        {
            std::vector<int32_t> nb;
#pragma omp for schedule(dynamic, 1024)
            for (int32_t g = 0; g < ng; ++g) {
                nbs(g, nb);
                std::copy(nb.begin(), nb.end(), qidx.begin() + qptr[g]);
            }
        }
    }
    Dissector D{qptr.data(), qidx.data(), mptr.data(), std::max<int64_t>(leaf, 1), {}, {}, {}};
    D.mark.assign(ng, 0);
    D.lev.assign(ng, 0);
    D.vis.assign(ng, 0);
    std::vector<std::vector<int32_t>> fverts, fkids;
    if (ng > 0) {
        std::vector<int32_t> all(ng);
        std::iota(all.begin(), all.end(), 0);
        D.dissect(std::move(all), fverts, fkids);
    }
    // number the fronts' original vertices in postorder, ascending within a front
    {
        int32_t next = 0;
        std::vector<int32_t> verts;
        for (size_t f = 0; f < fverts.size(); ++f) {
            verts.clear();
            for (int32_t v : fverts[f]) verts.insert(verts.end(), mem.begin() + mptr[v], mem.begin() + mptr[v + 1]);
            std::sort(verts.begin(), verts.end());
            T.p0.push_back(next);
            T.np.push_back(int32_t(verts.size()));
            T.parent.push_back(-1);
            T.children.push_back(fkids[f]);
            for (int32_t v : verts) {
                T.pos[v] = next;
                T.order[next] = v;
                ++next;
            }
            for (int32_t c : fkids[f]) T.parent[c] = int32_t(f);
        }
    }
    // update positions: a front's pivots' neighbours and its children's update positions
    // beyond its pivots; fronts of one height are independent
    const int32_t F = T.fronts();
    T.height.assign(F, 0);
    int32_t H = 0;
    for (int32_t f = 0; f < F; ++f) {
        for (int32_t c : T.children[f]) T.height[f] = std::max(T.height[f], T.height[c] + 1);
        H = std::max(H, T.height[f]);
    }
    std::vector<std::vector<int32_t>> bylev(size_t(H) + 1), U(F);
    for (int32_t f = 0; f < F; ++f) bylev[T.height[f]].push_back(f);
    for (int32_t h = 0; h <= H; ++h) {
#pragma omp parallel
        {
            std::vector<int32_t> mk;
#pragma omp for schedule(dynamic, 4)
            for (size_t i = 0; i < bylev[h].size(); ++i) {
                const int32_t f = bylev[h][i], p1 = T.p0[f] + T.np[f];
                if (mk.empty()) mk.assign(m, -1);
                std::vector<int32_t> &u = U[f];
                for (int32_t q = T.p0[f]; q < p1; ++q) {
                    const int32_t v = T.order[q];
                    for (int32_t e = adj_ptr[v]; e < adj_ptr[v + 1]; ++e) {
                        const int32_t qw = T.pos[adj_idx[e]];
                        if (qw >= p1 && mk[qw] != f) {
                            mk[qw] = f;
                            u.push_back(qw);
                        }
                    }
                }
                for (int32_t c : T.children[f])
                    for (int32_t qq : U[c])
                        if (qq >= p1 && mk[qq] != f) {
                            mk[qq] = f;
                            u.push_back(qq);
                        }
                std::sort(u.begin(), u.end());
            }
        }
    }
    T.uptr.assign(size_t(F) + 1, 0);
    for (int32_t f = 0; f < F; ++f) T.uptr[f + 1] = T.uptr[f] + int64_t(U[f].size());
    T.uidx.resize(T.uptr[F]);
    for (int32_t f = 0; f < F; ++f) std::copy(U[f].begin(), U[f].end(), T.uidx.begin() + T.uptr[f]);
    return T;
}

void SparseFactor::release()
{
    m_ = nf_ = max_front_ = batched_ = groups_ = 0;
    rtot_ = wtot_ = 0;
    for (DBuf<int32_t> *b : {&order_, &p0_, &np_, &nu_, &uidx_}) b->release();
    for (DBuf<int64_t> *b : {&roff_, &woff_, &goff_, &loff_, &cptr_, &cidx_}) b->release();
    for (DBuf<double> *b : {&G_, &L_, &Lt_, &tp_, &ypos_, &xpos_, &rhs_, &wbuf_}) b->release();
    for (DBuf<int2> *b : {&t_rhs_, &t_y_, &t_w_, &t_x_}) b->release();
    lv_rhs_.clear(), lv_y_.clear(), lv_w_.clear(), lv_x_.clear();
    lpr_y_.clear(), lpr_w_.clear(), lpr_s_.clear(), lpr_x_.clear();
}

k::FrontView SparseFactor::view() const
{
    k::FrontView v;
    v.p0 = p0_.get(), v.np = np_.get(), v.nu = nu_.get();
    v.roff = roff_.get(), v.woff = woff_.get(), v.goff = goff_.get(), v.loff = loff_.get();
    v.uidx = uidx_.get(), v.cptr = cptr_.get(), v.cidx = cidx_.get();
    return v;
}

void SparseFactor::factorize(int32_t n, const std::vector<int32_t> &dofs, const std::vector<int32_t> &h_ptr,
                             const int32_t *d_ptr, const int32_t *d_idx, const double *d_vals, int32_t leaf,
                             int64_t batch_rows, int64_t batch_bytes, int max_cols, DenseBlas &blas, int *d_status,
                             hipStream_t s)
{
    const auto t0 = clk::now();
    release();
    const int32_t m = int32_t(dofs.size());
    m_ = m;
    if (m == 0) return;
    // 1. stored rows of the unit's columns
    std::vector<int64_t> coff(m + 1, 0);
    for (int32_t v = 0; v < m; ++v) coff[v + 1] = coff[v] + (h_ptr[dofs[v] + 1] - h_ptr[dofs[v]]);
    std::vector<int32_t> rows(coff[m]);
    {
        DBuf<int32_t> d_dofs, d_rows(std::max<int64_t>(coff[m], 1));
        DBuf<int64_t> d_coff;
        d_dofs.upload(dofs.data(), m, s);
        d_coff.upload(coff.data(), m + 1, s);
        k::sparse_gather_pattern(m, d_dofs.get(), d_ptr, d_coff.get(), d_idx, d_rows.get(), s);
        d_rows.download(rows.data(), coff[m], s);
        LSS_HIP(hipStreamSynchronize(s));
    }
    // 2. entries inside the unit per column, by local row; of equal rows the last stored
    //    one counts (as the dense extraction, which overwrites)
    std::vector<int32_t> g2l(n, -1);
    for (int32_t v = 0; v < m; ++v) g2l[dofs[v]] = v;
    std::vector<int64_t> eptr(m + 1, 0);
    std::vector<int32_t> erow;
    std::vector<int64_t> esrc;
    {
        // per column (in parallel) into the column's own input range, then compacted
        // uninitialized: the parallel loop touches its own pages first
        std::unique_ptr<int32_t[]> crow(new int32_t[std::max<int64_t>(coff[m], 1)]);
        std::unique_ptr<int64_t[]> csrc(new int64_t[std::max<int64_t>(coff[m], 1)]);
#pragma omp parallel
        {
            std::vector<std::pair<int32_t, int64_t>> tmp;
#pragma omp for schedule(dynamic, 256)
            for (int32_t v = 0; v < m; ++v) {
                tmp.clear();
                const int64_t base = h_ptr[dofs[v]];
                for (int64_t t = coff[v]; t < coff[v + 1]; ++t) {
                    const int32_t w = g2l[rows[t]];
                    if (w >= 0) tmp.emplace_back(w, base + (t - coff[v]));
                }
                std::stable_sort(tmp.begin(), tmp.end(),
                                 [](const auto &a, const auto &b) { return a.first < b.first; });
                int64_t k = coff[v];
                for (size_t i = 0; i < tmp.size(); ++i) {
                    if (i + 1 < tmp.size() && tmp[i + 1].first == tmp[i].first) continue;
                    crow[k] = tmp[i].first;
                    csrc[k++] = tmp[i].second;
                }
                eptr[v + 1] = k - coff[v];
            }
        }
        for (int32_t v = 0; v < m; ++v) eptr[v + 1] += eptr[v];
        erow.resize(std::max<int64_t>(eptr[m], 1));
        esrc.resize(std::max<int64_t>(eptr[m], 1));
#pragma omp parallel for schedule(dynamic, 1024)
        for (int32_t v = 0; v < m; ++v) {
            std::copy(crow.get() + coff[v], crow.get() + coff[v] + (eptr[v + 1] - eptr[v]), erow.begin() + eptr[v]);
            std::copy(csrc.get() + coff[v], csrc.get() + coff[v] + (eptr[v + 1] - eptr[v]), esrc.begin() + eptr[v]);
        }
    }
    std::vector<int32_t>().swap(rows);
    // 3. symmetric adjacency, no self loops: each column's rows merged with the
    //    transposed pattern (both ascending)
    std::vector<int32_t> aptr(m + 1, 0), aidx;
    {
        std::vector<int64_t> tptr(m + 1, 0);
        for (int64_t i = 0; i < eptr[m]; ++i) ++tptr[erow[i] + 1];
        for (int32_t v = 0; v < m; ++v) tptr[v + 1] += tptr[v];
        std::vector<int32_t> tidx(std::max<int64_t>(eptr[m], 1));
        {
            std::vector<int64_t> cur(tptr.begin(), tptr.end() - 1);
            for (int32_t v = 0; v < m; ++v)
                for (int64_t i = eptr[v]; i < eptr[v + 1]; ++i) tidx[cur[erow[i]]++] = v;
        }
        // per vertex, the merged list (counted, then written at its offset; both in parallel)
        auto merge = [&](int32_t v, int32_t *out) {
            int64_t i = eptr[v], j = tptr[v], k = 0;
            const int64_t ie = eptr[v + 1], je = tptr[v + 1];
            while (i < ie || j < je) {
                int32_t w;
                if (j >= je || (i < ie && erow[i] < tidx[j])) {
                    w = erow[i++];
                } else if (i >= ie || tidx[j] < erow[i]) {
                    w = tidx[j++];
                } else {
                    w = erow[i++];
                    ++j;
                }
                if (w != v) {
                    if (out) out[k] = w;
                    ++k;
                }
            }
            return int32_t(k);
        };
#pragma omp parallel for schedule(dynamic, 1024)
        for (int32_t v = 0; v < m; ++v) aptr[v + 1] = merge(v, nullptr);
        for (int32_t v = 0; v < m; ++v) aptr[v + 1] += aptr[v];
        aidx.resize(std::max<int32_t>(aptr[m], 1));
#pragma omp parallel for schedule(dynamic, 1024)
        for (int32_t v = 0; v < m; ++v) merge(v, aidx.data() + aptr[v]);
    }
    // 4. ordering and front structures
    const FrontTree T = nested_dissection(m, aptr.data(), aidx.data(), leaf);
    const int32_t F = T.fronts();
    nf_ = F;
    std::vector<int64_t> roff(F + 1, 0), woff(F + 1, 0), goff(F + 1, 0), loff(F + 1, 0), toff(F + 1, 0);
    for (int32_t f = 0; f < F; ++f) {
        const int64_t p = T.np[f], u = T.nu(f), mf = p + u;
        if (mf > kMaxFront)
            fail(LSS_ERR_UNSUPPORTED, "sparse stiff factorization: a front of " + std::to_string(mf) +
                                          " rows exceeds " + std::to_string(kMaxFront));
        max_front_ = std::max<int32_t>(max_front_, int32_t(mf));
        roff[f + 1] = roff[f] + mf;
        woff[f + 1] = woff[f] + u;
        goff[f + 1] = goff[f] + p * p;
        loff[f + 1] = loff[f] + u * p;
    }
    rtot_ = roff[F];
    wtot_ = woff[F];
    // front-local rows: pivots first, then the update positions (lmap holds the latter
    // for one front at a time)
    std::vector<int32_t> fof(m), lmap(m, -1);
    for (int32_t f = 0; f < F; ++f)
        for (int32_t q = T.p0[f]; q < T.p0[f] + T.np[f]; ++q) fof[q] = f;
    auto map_front = [&](int32_t f, int32_t on) {
        for (int64_t t = T.uptr[f]; t < T.uptr[f + 1]; ++t) lmap[T.uidx[t]] = on ? T.np[f] + int32_t(t - T.uptr[f]) : -1;
    };
    auto loc = [&](int32_t f, int32_t q) -> int32_t {
        const int32_t r = q >= T.p0[f] && q < T.p0[f] + T.np[f] ? q - T.p0[f] : lmap[q];
        if (r < 0) fail(LSS_ERR_HIP, "sparse stiff factorization: position outside its front (internal)");
        return r;
    };
    // 5. original entries: the pair {a, b} is read from the column of the later unit index
    //    (its row the earlier one; the diagonal from its own column) and belongs to the front
    //    that eliminates the earlier of its two positions: row li (the later position), column
    //    lj (the earlier) of that front's lower triangle
    for (int32_t c = 0; c < m; ++c)
        for (int64_t i = eptr[c]; i < eptr[c + 1] && erow[i] <= c; ++i)
            ++toff[fof[std::min(T.pos[erow[i]], T.pos[c])] + 1];
    for (int32_t f = 0; f < F; ++f) toff[f + 1] += toff[f];
    std::vector<int32_t> tli(std::max<int64_t>(toff[F], 1)), tlj(std::max<int64_t>(toff[F], 1));
    std::vector<int64_t> tsrc(std::max<int64_t>(toff[F], 1));
    {
        std::vector<int64_t> cur(toff.begin(), toff.end() - 1);
        for (int32_t c = 0; c < m; ++c)
            for (int64_t i = eptr[c]; i < eptr[c + 1] && erow[i] <= c; ++i) {
                const int32_t qa = T.pos[erow[i]], qb = T.pos[c], q1 = std::min(qa, qb), f = fof[q1];
                const int64_t k = cur[f]++;
                tli[k] = std::max(qa, qb);  // a position until the loop below
                tlj[k] = q1 - T.p0[f];
                tsrc[k] = esrc[i];
            }
    }
    for (int32_t f = 0; f < F; ++f) {
        map_front(f, 1);
        for (int64_t k = toff[f]; k < toff[f + 1]; ++k) tli[k] = loc(f, tli[k]);
        map_front(f, 0);
    }
    // 6. child -> parent row maps (at the child's woff) and, per parent row, the
    //    children's update entries that land there, children in order
    std::vector<int32_t> rel(std::max<int64_t>(wtot_, 1));
    std::vector<int64_t> cptr(rtot_ + 1, 0);
    for (int32_t f = 0; f < F; ++f) {
        map_front(f, 1);
        for (int32_t c : T.children[f])
            for (int64_t a = 0; a < T.nu(c); ++a) {
                const int32_t r = loc(f, T.uidx[T.uptr[c] + a]);
                rel[woff[c] + a] = r;
                ++cptr[roff[f] + r + 1];
            }
        map_front(f, 0);
    }
    for (int64_t r = 0; r < rtot_; ++r) cptr[r + 1] += cptr[r];
    std::vector<int64_t> cidx(std::max<int64_t>(cptr[rtot_], 1));
    {
        std::vector<int64_t> cur(cptr.begin(), cptr.end() - 1);
        for (int32_t f = 0; f < F; ++f)
            for (int32_t c : T.children[f])
                for (int64_t a = 0; a < T.nu(c); ++a) cidx[cur[roff[f] + rel[woff[c] + a]]++] = woff[c] + a;
    }
    // 7. plan of the numeric phase, level by level (tree heights; the fronts of one level
    //    are independent).  Fronts of at most batch_rows rows are factored together, in
    //    groups whose workspaces fit batch_bytes, one launch per step for the whole
    //    group; larger fronts one at a time with the blocked dense path.  Update matrices
    //    live in a pool from their front's factorization to their parent's assembly; the
    //    offsets come from a first-fit plan in the order the stream reads and writes them.
    int32_t H = 0;
    for (int32_t f = 0; f < F; ++f) H = std::max(H, T.height[f]);
    std::vector<std::vector<int32_t>> bylev(H + 1);
    for (int32_t f = 0; f < F; ++f) bylev[T.height[f]].push_back(f);
    const int32_t NB = k::dense_block_size();
    auto mof = [&](int32_t f) { return int64_t(T.np[f]) + T.nu(f); };
    auto usize = [&](int32_t f) { return align_up(T.nu(f) * T.nu(f)); };
    std::vector<int64_t> wb(F, 0), upoff(F, 0);
    FirstFit pool;
    std::vector<int32_t> gfronts;  // the fronts of every batch, batch after batch
    std::vector<k::ExtendTask> ext;
    std::vector<k::DiagTask> dtask;
    std::vector<k::PanelTask> ptask;
    std::vector<k::TileTask> ttask;
    std::vector<k::CopyTask> ctask;
    std::vector<GroupPlan> groups;
    std::vector<int32_t> stages;  // >= 0: group index; < 0: -1 - front (one-front path)
    int64_t max_w = 1, max_slots = 1, max_m2 = 1, max_pm = 1, max_work = 1;
    auto plan_group = [&](const std::vector<int32_t> &fr, GroupPlan &g) {
        g.fr.off = int64_t(gfronts.size());
        gfronts.insert(gfronts.end(), fr.begin(), fr.end());
        g.fr.n = int64_t(fr.size());
        size_t kmax = 0;
        int32_t nbmax = 0;
        for (int32_t f : fr) {
            kmax = std::max(kmax, T.children[f].size());
            nbmax = std::max(nbmax, (T.np[f] + NB - 1) / NB);
        }
        for (size_t j = 0; j < kmax; ++j) {  // children in order: one launch per child slot
            Range r{int64_t(ext.size()), 0};
            for (int32_t f : fr)
                if (T.children[f].size() > j) {
                    const int32_t c = T.children[f][j];
                    for (int32_t a = 0; a < int32_t(T.nu(c)); ++a) ext.push_back({c, a});
                }
            r.n = int64_t(ext.size()) - r.off;
            if (r.n) g.ext.push_back(r);
        }
        for (int32_t kk = 0; kk < nbmax; ++kk) {
            const int32_t k0 = kk * NB;
            std::array<Range, 3> st;
            st[0].off = int64_t(dtask.size()), st[1].off = int64_t(ptask.size()), st[2].off = int64_t(ttask.size());
            for (int32_t f : fr) {
                const int32_t p = T.np[f], mf = int32_t(mof(f));
                if (p <= k0) continue;
                const int32_t kb = std::min(NB, p - k0), s0 = k0 + kb;
                dtask.push_back({wb[f], goff[f], mf, p, k0, kb});
                for (int32_t r = s0; r < mf; r += 32) ptask.push_back({wb[f], goff[f], mf, p, k0, kb, r, std::min(32, mf - r)});
                for (int32_t i0 = s0; i0 < mf; i0 += 64)
                    for (int32_t j0 = s0; j0 <= i0; j0 += 64)
                        ttask.push_back({wb[f] + int64_t(i0) * mf + k0, wb[f] + int64_t(j0) * mf + k0,
                                         wb[f] + int64_t(i0) * mf + j0, mf, mf, mf, std::min(64, mf - i0),
                                         std::min(64, mf - j0), kb, i0 == j0 ? 1 : 0, 0});
            }
            st[0].n = int64_t(dtask.size()) - st[0].off, st[1].n = int64_t(ptask.size()) - st[1].off;
            st[2].n = int64_t(ttask.size()) - st[2].off;
            g.steps.push_back(st);
        }
        // off-diagonal blocks of G11 by distance d: T = sum_{K=J}^{I-1} L_IK G_KJ, G_IJ = -G_II T
        for (int32_t d = 1; d < nbmax; ++d) {
            std::array<Range, 2> rd;
            int64_t slot = 0;
            rd[0].off = int64_t(ttask.size());
            std::vector<k::TileTask> gt;
            for (int32_t f : fr) {
                const int32_t p = T.np[f], mf = int32_t(mof(f)), nb = (p + NB - 1) / NB;
                for (int32_t I = d; I < nb; ++I, ++slot) {
                    const int32_t J = I - d, kbI = std::min(NB, p - I * NB);
                    const int64_t tb = slot * NB * NB;
                    for (int32_t ti = 0; ti < kbI; ti += 64)
                        for (int32_t tj = 0; tj < NB; tj += 64) {
                            ttask.push_back({wb[f] + int64_t(I * NB + ti) * mf + J * NB,
                                             goff[f] + int64_t(J * NB) * p + J * NB + tj, tb + ti * NB + tj, mf, p,
                                             NB, std::min(64, kbI - ti), 64, d * NB, 0, 0});
                            gt.push_back({goff[f] + int64_t(I * NB + ti) * p + I * NB, tb + tj,
                                          goff[f] + int64_t(I * NB + ti) * p + J * NB + tj, p, NB, p,
                                          std::min(64, kbI - ti), 64, kbI, 0, 0});
                        }
                }
            }
            rd[0].n = int64_t(ttask.size()) - rd[0].off;
            rd[1].off = int64_t(ttask.size());
            ttask.insert(ttask.end(), gt.begin(), gt.end());
            rd[1].n = int64_t(gt.size());
            max_slots = std::max(max_slots, slot);
            g.rounds.push_back(rd);
        }
        // outputs: L21, L21^T, the update matrix (into the pool), G11^T above G11
        Range *cr[4] = {&g.cl, &g.clt, &g.cu, &g.cm};
        for (int mode = 0; mode < 4; ++mode) {
            cr[mode]->off = int64_t(ctask.size());
            for (int32_t f : fr) {
                const int32_t p = T.np[f], u = int32_t(T.nu(f)), mf = int32_t(mof(f));
                if (mode < 2)
                    for (int32_t a0 = 0; a0 < u; a0 += 32)
                        for (int32_t j0 = 0; j0 < p; j0 += 32) {
                            const int64_t src = wb[f] + int64_t(p + a0) * mf + j0;
                            if (mode == 0)
                                ctask.push_back({src, loff[f] + int64_t(a0) * p + j0, mf, p, std::min(32, u - a0),
                                                 std::min(32, p - j0), 0, 0});
                            else
                                ctask.push_back({src, loff[f] + int64_t(j0) * u + a0, mf, u, std::min(32, u - a0),
                                                 std::min(32, p - j0), 1, 0});
                        }
                else if (mode == 2)
                    for (int32_t a0 = 0; a0 < u; a0 += 32)
                        for (int32_t b0 = 0; b0 <= a0; b0 += 32)
                            ctask.push_back({wb[f] + int64_t(p + a0) * mf + p + b0, upoff[f] + int64_t(a0) * u + b0,
                                             mf, u, std::min(32, u - a0), std::min(32, u - b0), 0, 0});
                else
                    for (int32_t i0 = 0; i0 < p; i0 += 32)
                        for (int32_t j0 = 0; j0 <= i0; j0 += 32)
                            ctask.push_back({goff[f] + int64_t(i0) * p + j0, goff[f] + int64_t(j0) * p + i0, p, p,
                                             std::min(32, p - i0), std::min(32, p - j0), i0 == j0 ? 2 : 1, 0});
            }
            cr[mode]->n = int64_t(ctask.size()) - cr[mode]->off;
        }
    };
    for (int32_t h = 0; h <= H; ++h) {
        std::vector<int32_t> small, big;
        for (int32_t f : bylev[h]) (mof(f) <= batch_rows ? small : big).push_back(f);
        for (size_t i = 0; i < small.size();) {
            std::vector<int32_t> fr;
            int64_t w = 0;
            while (i < small.size()) {
                const int64_t sz = align_up(mof(small[i]) * mof(small[i]));
                if (!fr.empty() && (w + sz) * int64_t(sizeof(double)) > batch_bytes) break;
                wb[small[i]] = w;
                w += sz;
                fr.push_back(small[i++]);
            }
            for (int32_t f : fr)
                for (int32_t c : T.children[f]) pool.release(upoff[c], usize(c));
            for (int32_t f : fr) upoff[f] = pool.take(usize(f));
            GroupPlan g;
            g.wsize = w;
            max_w = std::max(max_w, w);
            plan_group(fr, g);
            stages.push_back(int32_t(groups.size()));
            groups.push_back(std::move(g));
            batched_ += int32_t(fr.size());
        }
        for (int32_t f : big) {
            for (int32_t c : T.children[f]) pool.release(upoff[c], usize(c));
            upoff[f] = pool.take(usize(f));
            stages.push_back(-1 - f);
            const int64_t mf = mof(f);
            max_m2 = std::max(max_m2, mf * mf);
            max_pm = std::max(max_pm, int64_t(T.np[f]) * mf);
            max_work = std::max(max_work, DenseBlas::work_size(mf));
        }
    }
    // every task inside its buffers (an internal error otherwise, before any launch)
    {
        const int64_t nW = max_w, nG = goff[F], nL = loff[F], nT = max_slots * NB * NB, nU = pool.peak();
        bool ok = true;
        auto in = [&](int64_t first, int64_t last, int64_t n) { ok = ok && first >= 0 && last < n; };
        auto tiles = [&](const Range &r, bool nt, int64_t na, int64_t nb, int64_t nc) {
            for (int64_t i = r.off; i < r.off + r.n; ++i) {
                const k::TileTask &t = ttask[i];
                ok = ok && t.M > 0 && t.N > 0 && t.K > 0 && t.M <= 64 && t.N <= 64;
                in(t.a, t.a + int64_t(t.M - 1) * t.lda + t.K - 1, na);
                in(t.b, nt ? t.b + int64_t(t.N - 1) * t.ldb + t.K - 1 : t.b + int64_t(t.K - 1) * t.ldb + t.N - 1, nb);
                in(t.c, t.c + int64_t(t.M - 1) * t.ldc + t.N - 1, nc);
            }
        };
        auto copies = [&](const Range &r, int64_t ns, int64_t nd) {
            for (int64_t i = r.off; i < r.off + r.n; ++i) {
                const k::CopyTask &t = ctask[i];
                ok = ok && t.rows > 0 && t.cols > 0 && t.rows <= 32 && t.cols <= 32;
                in(t.src, t.src + int64_t(t.rows - 1) * t.lds + t.cols - 1, ns);
                in(t.dst, t.mode == 0 ? t.dst + int64_t(t.rows - 1) * t.ldd + t.cols - 1
                                      : t.dst + int64_t(t.cols - 1) * t.ldd + t.rows - 1, nd);
            }
        };
        for (const GroupPlan &g : groups) {
            ok = ok && g.wsize <= nW;
            for (int64_t i = g.fr.off; i < g.fr.off + g.fr.n; ++i) {
                const int32_t f = gfronts[i];
                in(wb[f], wb[f] + mof(f) * mof(f) - 1, g.wsize);
            }
            for (const Range &r : g.ext)
                for (int64_t i = r.off; i < r.off + r.n; ++i) {
                    const int32_t c = ext[i].c, a = ext[i].a, f = c >= 0 && c < F ? T.parent[c] : -1;
                    ok = ok && f >= 0 && a >= 0 && a < T.nu(c);
                    if (!ok) break;
                    in(upoff[c], upoff[c] + int64_t(a) * T.nu(c) + a, nU);
                    in(wb[f], wb[f] + mof(f) * mof(f) - 1, g.wsize);
                }
            for (const auto &q : g.steps) {
                for (int64_t i = q[0].off; i < q[0].off + q[0].n; ++i) {
                    const k::DiagTask &t = dtask[i];
                    const int64_t e = t.k0 + t.kb - 1;
                    ok = ok && t.kb > 0 && t.kb <= NB;
                    in(t.a, t.a + e * t.lda + e, g.wsize);
                    in(t.g, t.g + e * t.ldg + e, nG);
                }
                for (int64_t i = q[1].off; i < q[1].off + q[1].n; ++i) {
                    const k::PanelTask &t = ptask[i];
                    const int64_t e = t.k0 + t.kb - 1;
                    ok = ok && t.kb > 0 && t.kb <= NB && t.nr > 0 && t.nr <= 32;
                    in(t.a, t.a + int64_t(t.r0 + t.nr - 1) * t.lda + e, g.wsize);
                    in(t.g, t.g + e * t.ldg + e, nG);
                }
                tiles(q[2], true, g.wsize, g.wsize, g.wsize);
            }
            for (const auto &q : g.rounds) {
                tiles(q[0], false, g.wsize, nG, nT);
                tiles(q[1], false, nG, nT, nG);
            }
            copies(g.cl, g.wsize, nL);
            copies(g.clt, g.wsize, nL);
            copies(g.cu, g.wsize, nU);
            copies(g.cm, nG, nG);
        }
        if (!ok) fail(LSS_ERR_HIP, "sparse stiff factorization: batched plan out of bounds (internal)");
    }
    groups_ = int32_t(groups.size());
    stack_bytes_ = std::max<int64_t>(pool.peak(), 1) * int64_t(sizeof(double));
    t_symbolic = since(t0);
    // 8. numeric factorization
    const auto t1 = clk::now();
    np_.upload(T.np.data(), F, s);
    {
        std::vector<int32_t> nu(F);
        for (int32_t f = 0; f < F; ++f) nu[f] = int32_t(T.nu(f));
        nu_.upload(nu.data(), F, s);
    }
    woff_.upload(woff.data(), F + 1, s);
    G_.resize(std::max<int64_t>(goff[F], 1));
    L_.resize(std::max<int64_t>(loff[F], 1));
    Lt_.resize(std::max<int64_t>(loff[F], 1));
    k::fill(goff[F], G_.get(), 0.0, s);  // the batched steps read G11's upper blocks as zeros
    {
        DBuf<double> upd(std::max<int64_t>(pool.peak(), 1)), W(groups.empty() ? 1 : max_w),
            Tw(groups.empty() ? 1 : max_slots * NB * NB);
        const bool any_big = groups.size() < stages.size();
        DBuf<double> Fw(any_big ? max_m2 : 1), Gt(any_big ? max_pm : 1), work(any_big ? max_work : 1);
        DBuf<int32_t> d_li, d_lj, d_rel, d_parent;
        DBuf<int64_t> d_src, d_wb, d_upoff, d_toff;
        DBuf<int32_t> d_gfr;
        DBuf<k::ExtendTask> d_ext;
        DBuf<k::DiagTask> d_dt;
        DBuf<k::PanelTask> d_pt;
        DBuf<k::TileTask> d_tt;
        DBuf<k::CopyTask> d_ct;
        d_li.upload(tli.data(), tli.size(), s);
        d_lj.upload(tlj.data(), tlj.size(), s);
        d_src.upload(tsrc.data(), tsrc.size(), s);
        d_rel.upload(rel.data(), rel.size(), s);
        d_parent.upload(T.parent.data(), F, s);
        d_wb.upload(wb.data(), F, s);
        d_upoff.upload(upoff.data(), F, s);
        d_toff.upload(toff.data(), toff.size(), s);
        d_gfr.upload(gfronts.data(), gfronts.size(), s);
        d_ext.upload(ext.data(), ext.size(), s);
        d_dt.upload(dtask.data(), dtask.size(), s);
        d_pt.upload(ptask.data(), ptask.size(), s);
        d_tt.upload(ttask.data(), ttask.size(), s);
        d_ct.upload(ctask.data(), ctask.size(), s);
        for (const int32_t st : stages) {
            if (st >= 0) {
                const GroupPlan &g = groups[st];
                double *w = W.get();
                k::fill(g.wsize, w, 0.0, s);
                k::front_assemble_batch(int32_t(g.fr.n), d_gfr.get() + g.fr.off, d_toff.get(), d_li.get(), d_lj.get(),
                                        d_src.get(), d_wb.get(), np_.get(), nu_.get(), d_vals, w, s);
                for (const Range &r : g.ext)
                    k::front_extend_batch(r.n, d_ext.get() + r.off, d_parent.get(), d_wb.get(), np_.get(), nu_.get(),
                                          woff_.get(), d_upoff.get(), d_rel.get(), upd.get(), w, s);
                for (const auto &q : g.steps) {
                    k::dense_diag_batch(int32_t(q[0].n), d_dt.get() + q[0].off, w, G_.get(), d_status, s);
                    k::front_panel_batch(int32_t(q[1].n), d_pt.get() + q[1].off, w, G_.get(), s);
                    k::front_gemm_batch(int32_t(q[2].n), d_tt.get() + q[2].off, true, -1.0, 1.0, w, w, w, s);
                }
                for (const auto &q : g.rounds) {
                    k::front_gemm_batch(int32_t(q[0].n), d_tt.get() + q[0].off, false, 1.0, 0.0, w, G_.get(),
                                        Tw.get(), s);
                    k::front_gemm_batch(int32_t(q[1].n), d_tt.get() + q[1].off, false, -1.0, 0.0, G_.get(),
                                        Tw.get(), G_.get(), s);
                }
                k::front_copy_batch(int32_t(g.cl.n), d_ct.get() + g.cl.off, w, L_.get(), s);
                k::front_copy_batch(int32_t(g.clt.n), d_ct.get() + g.clt.off, w, Lt_.get(), s);
                k::front_copy_batch(int32_t(g.cu.n), d_ct.get() + g.cu.off, w, upd.get(), s);
                k::front_copy_batch(int32_t(g.cm.n), d_ct.get() + g.cm.off, G_.get(), G_.get(), s);
                continue;
            }
            const int32_t f = -1 - st;
            const int64_t p = T.np[f], u = T.nu(f), mf = p + u;
            k::fill(mf * mf, Fw.get(), 0.0, s);
            k::sparse_assemble(toff[f + 1] - toff[f], d_li.get() + toff[f], d_lj.get() + toff[f],
                               d_src.get() + toff[f], d_vals, Fw.get(), mf, s);
            for (int32_t c : T.children[f])
                k::sparse_extend_add(int32_t(T.nu(c)), d_rel.get() + woff[c], upd.get() + upoff[c], Fw.get(), mf, s);
            blas.partial_factor_inverse(mf, p, Fw.get(), Gt.get(), work.get(), d_status, s);
            LSS_HIP(hipMemcpy2DAsync(G_.get() + goff[f], size_t(p) * sizeof(double), Gt.get(),
                                     size_t(mf) * sizeof(double), size_t(p) * sizeof(double), size_t(p),
                                     hipMemcpyDeviceToDevice, s));
            k::dense_mirror(p, G_.get() + goff[f], s);
            if (u > 0) {
                LSS_HIP(hipMemcpy2DAsync(L_.get() + loff[f], size_t(p) * sizeof(double), Fw.get() + p * mf,
                                         size_t(mf) * sizeof(double), size_t(p) * sizeof(double), size_t(u),
                                         hipMemcpyDeviceToDevice, s));
                LSS_HIP(hipMemcpy2DAsync(upd.get() + upoff[f], size_t(u) * sizeof(double), Fw.get() + p * mf + p,
                                         size_t(mf) * sizeof(double), size_t(u) * sizeof(double), size_t(u),
                                         hipMemcpyDeviceToDevice, s));
                k::sparse_transpose(u, p, L_.get() + loff[f], Lt_.get() + loff[f], s);
            }
        }
        LSS_HIP(hipStreamSynchronize(s));  // the workspaces and task lists are released at scope end
    }
    t_numeric = since(t1);
    // 9. solve tables: task lists per tree level
    std::vector<int2> trhs, ty, tw, tx;
    for (auto *lv : {&lv_rhs_, &lv_y_, &lv_w_, &lv_x_}) lv->assign(H + 2, 0);
    // lanes per row of a level's products: about sixteen entries per lane (4 .. 64 lanes), so
    // short rows share a wave and its chain of table loads
    auto lanes = [](int64_t len, int64_t rows) {
        const int64_t per = rows ? (len / rows + 15) / 16 : 1;
        int l = 4;
        while (l < 64 && l < per) l *= 2;
        return l;
    };
    for (auto *lp : {&lpr_y_, &lpr_w_, &lpr_s_, &lpr_x_}) lp->assign(H + 1, 64);
    for (int32_t h = 0; h <= H; ++h) {
        lv_rhs_[h] = int64_t(trhs.size()), lv_y_[h] = int64_t(ty.size()), lv_w_[h] = int64_t(tw.size());
        lv_x_[h] = int64_t(tx.size());
        int64_t ly = 0, lw = 0, ls = 0;  // summed row lengths: G11 halves, L21 rows, L21^T rows
        for (int32_t f : bylev[h]) {
            const int32_t p = T.np[f], u = int32_t(T.nu(f));
            for (int32_t a = 0; a < p + u; ++a) trhs.push_back(make_int2(f, a));
            for (int32_t i = 0; i < p; ++i) ty.push_back(make_int2(f, i)), tx.push_back(make_int2(f, i));
            for (int32_t a = 0; a < u; ++a) tw.push_back(make_int2(f, a));
            ly += int64_t(p) * (p + 1) / 2;
            lw += int64_t(u) * p;
            ls += int64_t(p) * u;
        }
        const int64_t ny = int64_t(ty.size()) - lv_y_[h], nw = int64_t(tw.size()) - lv_w_[h];
        lpr_y_[h] = lpr_x_[h] = lanes(ly, ny);
        lpr_w_[h] = lanes(lw, nw);
        lpr_s_[h] = lanes(ls, ny);
    }
    lv_rhs_[H + 1] = int64_t(trhs.size()), lv_y_[H + 1] = int64_t(ty.size()), lv_w_[H + 1] = int64_t(tw.size());
    lv_x_[H + 1] = int64_t(tx.size());
    t_rhs_.upload(trhs.data(), trhs.size(), s);
    t_y_.upload(ty.data(), ty.size(), s);
    t_w_.upload(tw.data(), tw.size(), s);
    t_x_.upload(tx.data(), tx.size(), s);
    order_.upload(T.order.data(), m, s);
    p0_.upload(T.p0.data(), F, s);
    roff_.upload(roff.data(), F + 1, s);
    goff_.upload(goff.data(), F + 1, s);
    loff_.upload(loff.data(), F + 1, s);
    {
        std::vector<int32_t> ui(T.uidx);
        if (ui.empty()) ui.push_back(0);  // a single front: no update rows, one unread entry
        uidx_.upload(ui.data(), ui.size(), s);
    }
    cptr_.upload(cptr.data(), cptr.size(), s);
    cidx_.upload(cidx.data(), cidx.size(), s);
    const int64_t kc = std::max(max_cols, 1);
    tp_.resize(kc * m);
    ypos_.resize(kc * m);
    xpos_.resize(kc * m);
    rhs_.resize(std::max<int64_t>(kc * rtot_, 1));
    wbuf_.resize(std::max<int64_t>(kc * wtot_, 1));
    LSS_HIP(hipStreamSynchronize(s));  // host tables are released at return
}

void SparseFactor::solve_add(const double *t, const int32_t *pdofs, double *z, float *z32, int64_t ld, int k,
                             const int *stop, hipStream_t s)
{
    if (m_ == 0) return;
    const k::FrontView fv = view();
    const int NL = levels();
    k::sparse_permute_in(m_, order_.get(), t, m_, tp_.get(), k, stop, s);
    for (int h = 0; h < NL; ++h) {
        k::sparse_fwd_rhs(lv_rhs_[h + 1] - lv_rhs_[h], t_rhs_.get() + lv_rhs_[h], fv, tp_.get(), m_, wbuf_.get(),
                          wtot_, rhs_.get(), rtot_, k, stop, s);
        k::sparse_fwd_y(lv_y_[h + 1] - lv_y_[h], t_y_.get() + lv_y_[h], fv, G_.get(), rhs_.get(), rtot_,
                        ypos_.get(), m_, k, stop, lpr_y_[h], s);
        k::sparse_fwd_w(lv_w_[h + 1] - lv_w_[h], t_w_.get() + lv_w_[h], fv, L_.get(), rhs_.get(), rtot_,
                        ypos_.get(), m_, wbuf_.get(), wtot_, k, stop, lpr_w_[h], s);
    }
    for (int h = NL - 1; h >= 0; --h) {
        // x at the level's update positions into the update buffer (free after the forward sweep)
        k::sparse_bwd_gather(lv_w_[h + 1] - lv_w_[h], t_w_.get() + lv_w_[h], fv, xpos_.get(), m_, wbuf_.get(), wtot_,
                             k, stop, s);
        k::sparse_bwd_s(lv_y_[h + 1] - lv_y_[h], t_y_.get() + lv_y_[h], fv, Lt_.get(), wbuf_.get(), wtot_,
                        ypos_.get(), m_, rhs_.get(), rtot_, k, stop, lpr_s_[h], s);
        k::sparse_bwd_x(lv_x_[h + 1] - lv_x_[h], t_x_.get() + lv_x_[h], fv, G_.get(), rhs_.get(), rtot_,
                        xpos_.get(), m_, k, stop, lpr_x_[h], s);
    }
    k::sparse_scatter_add(m_, order_.get(), pdofs, xpos_.get(), z, z32, ld, k, stop, s);
}

}  // namespace lss
