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
// Numeric factorization, children before parents, one front at a time on the stream:
// zero the m x m front, write the original entries of its pivot columns, add the
// children's update matrices (extend-add, children in order), and factor the first p
// columns with the blocked dense path (DenseBlas::partial_factor_inverse), which also
// leaves the Schur complement in the trailing block.  Kept per front: G11 = L11^{-1}
// (with G11^T in its strict upper triangle) and L21.
#include "sparse_stiff.hpp"

#include <algorithm>
#include <chrono>
#include <numeric>
#include <string>

#include "../kernels/kernels.hpp"
#include "../kernels/mas_kernels.hpp"

namespace lss {
namespace {

using clk = std::chrono::steady_clock;
double since(clk::time_point t0) { return std::chrono::duration<double>(clk::now() - t0).count(); }

// Largest front (rows) the numeric phase accepts: its dense workspace takes m^2 doubles.
constexpr int32_t kMaxFront = 49152;

struct Dissector {
    const int32_t *ap, *ai;
    int32_t leaf;
    FrontTree &T;
    std::vector<int32_t> mark, lev, vis;
    int32_t stamp = 0, vstamp = 0, next = 0;

    int32_t emit(std::vector<int32_t> &verts, const std::vector<int32_t> &kids)
    {
        std::sort(verts.begin(), verts.end());
        const int32_t f = T.fronts();
        T.p0.push_back(next);
        T.np.push_back(int32_t(verts.size()));
        T.parent.push_back(-1);
        T.children.push_back(kids);
        for (int32_t v : verts) {
            T.pos[v] = next;
            T.order[next] = v;
            ++next;
        }
        for (int32_t c : kids) T.parent[c] = f;
        return f;
    }

    // breadth-first search over the current subset from src: q = visit order, lev = levels
    void bfs(int32_t src, int32_t my, std::vector<int32_t> &q)
    {
        ++vstamp;
        q.clear();
        q.push_back(src);
        vis[src] = vstamp;
        lev[src] = 0;
        for (size_t h = 0; h < q.size(); ++h) {
            const int32_t v = q[h];
            for (int32_t e = ap[v]; e < ap[v + 1]; ++e) {
                const int32_t w = ai[e];
                if (mark[w] == my && vis[w] != vstamp) {
                    vis[w] = vstamp;
                    lev[w] = lev[v] + 1;
                    q.push_back(w);
                }
            }
        }
    }

    // returns the root fronts of the subgraph induced by V
    std::vector<int32_t> dissect(std::vector<int32_t> V)
    {
        if (int64_t(V.size()) <= leaf) return {emit(V, {})};
        const int32_t my = ++stamp;
        for (int32_t v : V) mark[v] = my;
        std::sort(V.begin(), V.end());
        // connected components, in the order of their smallest vertex
        std::vector<std::vector<int32_t>> comps;
        const int32_t cs = ++vstamp;
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
                    if (mark[w] == my && vis[w] != cs) {
                        vis[w] = cs;
                        c.push_back(w);
                    }
                }
            }
        }
        if (comps.size() > 1) {
            std::vector<int32_t> roots;
            for (auto &c : comps) {
                const std::vector<int32_t> r = dissect(std::move(c));
                roots.insert(roots.end(), r.begin(), r.end());
            }
            return roots;
        }
        std::vector<int32_t> q;
        bfs(V.front(), my, q);
        bfs(q.back(), my, q);
        const int32_t nlev = lev[q.back()] + 1;
        if (nlev < 3) return {emit(V, {})};
        std::vector<int64_t> cnt(nlev, 0);
        for (int32_t v : q) ++cnt[lev[v]];
        // the smallest level whose predecessors hold 35-65 % of the vertices (the one
        // nearest to half among equals); else the level that reaches half
        const int64_t nv = int64_t(V.size()), half = nv / 2;
        int64_t cum = 0, kcum = 0;
        int32_t k = -1, kh = -1;
        for (int32_t l = 1; l < nlev - 1; ++l) {
            cum += cnt[l - 1];  // vertices before level l
            if (kh < 0 && cum + cnt[l] >= half) kh = l;
            if (20 * cum < 7 * nv || 20 * cum > 13 * nv) continue;
            if (k < 0 || cnt[l] < cnt[k] ||
                (cnt[l] == cnt[k] && std::llabs(2 * cum - nv) < std::llabs(2 * kcum - nv))) {
                k = l;
                kcum = cum;
            }
        }
        if (k < 0) k = kh < 0 ? 1 : kh;
        std::vector<int32_t> left, right, sep;
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
                    cut = mark[w] == my && lev[w] == k + 1;
                }
                (cut ? sep : left).push_back(v);
            }
        }
        std::vector<int32_t> kids = dissect(std::move(left));
        const std::vector<int32_t> r = dissect(std::move(right));
        kids.insert(kids.end(), r.begin(), r.end());
        return {emit(sep, kids)};
    }
};

}  // namespace

FrontTree nested_dissection(int32_t m, const int32_t *adj_ptr, const int32_t *adj_idx, int32_t leaf)
{
    FrontTree T;
    T.m = m;
    T.order.assign(m, 0);
    T.pos.assign(m, 0);
    Dissector D{adj_ptr, adj_idx, std::max(leaf, 1), T, {}, {}, {}};
    D.mark.assign(m, 0);
    D.lev.assign(m, 0);
    D.vis.assign(m, 0);
    if (m > 0) {
        std::vector<int32_t> all(m);
        std::iota(all.begin(), all.end(), 0);
        D.dissect(std::move(all));
    }
    const int32_t F = T.fronts();
    T.uptr.assign(1, 0);
    T.height.assign(F, 0);
    std::vector<int32_t> mk(m, -1), U;
    for (int32_t f = 0; f < F; ++f) {
        const int32_t p1 = T.p0[f] + T.np[f];
        U.clear();
        for (int32_t q = T.p0[f]; q < p1; ++q) {
            const int32_t v = T.order[q];
            for (int32_t e = adj_ptr[v]; e < adj_ptr[v + 1]; ++e) {
                const int32_t qw = T.pos[adj_idx[e]];
                if (qw >= p1 && mk[qw] != f) {
                    mk[qw] = f;
                    U.push_back(qw);
                }
            }
        }
        for (int32_t c : T.children[f]) {
            for (int64_t t = T.uptr[c]; t < T.uptr[c + 1]; ++t) {
                const int32_t qq = T.uidx[t];
                if (qq >= p1 && mk[qq] != f) {
                    mk[qq] = f;
                    U.push_back(qq);
                }
            }
            T.height[f] = std::max(T.height[f], T.height[c] + 1);
        }
        std::sort(U.begin(), U.end());
        T.uidx.insert(T.uidx.end(), U.begin(), U.end());
        T.uptr.push_back(int64_t(T.uidx.size()));
    }
    return T;
}

void SparseFactor::release()
{
    m_ = nf_ = max_front_ = 0;
    rtot_ = wtot_ = 0;
    for (DBuf<int32_t> *b : {&order_, &p0_, &np_, &nu_, &uidx_}) b->release();
    for (DBuf<int64_t> *b : {&roff_, &woff_, &goff_, &loff_, &cptr_, &cidx_}) b->release();
    for (DBuf<double> *b : {&G_, &L_, &Lt_, &tp_, &ypos_, &xpos_, &rhs_, &wbuf_}) b->release();
    for (DBuf<int2> *b : {&t_rhs_, &t_y_, &t_w_, &t_x_}) b->release();
    lv_rhs_.clear(), lv_y_.clear(), lv_w_.clear(), lv_x_.clear();
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
                             int max_cols, DenseBlas &blas, int *d_status, hipStream_t s)
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
    erow.reserve(coff[m]);
    esrc.reserve(coff[m]);
    {
        std::vector<std::pair<int32_t, int64_t>> tmp;
        for (int32_t v = 0; v < m; ++v) {
            tmp.clear();
            const int64_t base = h_ptr[dofs[v]];
            for (int64_t t = coff[v]; t < coff[v + 1]; ++t) {
                const int32_t w = g2l[rows[t]];
                if (w >= 0) tmp.emplace_back(w, base + (t - coff[v]));
            }
            std::stable_sort(tmp.begin(), tmp.end(),
                             [](const auto &a, const auto &b) { return a.first < b.first; });
            for (size_t i = 0; i < tmp.size(); ++i) {
                if (i + 1 < tmp.size() && tmp[i + 1].first == tmp[i].first) continue;
                erow.push_back(tmp[i].first);
                esrc.push_back(tmp[i].second);
            }
            eptr[v + 1] = int64_t(erow.size());
        }
    }
    std::vector<int32_t>().swap(rows);
    // 3. symmetric adjacency, no self loops
    std::vector<int32_t> aptr(m + 1, 0), aidx;
    {
        std::vector<int64_t> deg(m + 1, 0);
        for (int32_t v = 0; v < m; ++v)
            for (int64_t i = eptr[v]; i < eptr[v + 1]; ++i)
                if (erow[i] != v) ++deg[v + 1], ++deg[erow[i] + 1];
        for (int32_t v = 0; v < m; ++v) deg[v + 1] += deg[v];
        std::vector<int32_t> raw(deg[m]);
        std::vector<int64_t> cur(deg.begin(), deg.end() - 1);
        for (int32_t v = 0; v < m; ++v)
            for (int64_t i = eptr[v]; i < eptr[v + 1]; ++i)
                if (erow[i] != v) raw[cur[v]++] = erow[i], raw[cur[erow[i]]++] = v;
        aidx.reserve(deg[m]);
        for (int32_t v = 0; v < m; ++v) {
            std::sort(raw.begin() + deg[v], raw.begin() + deg[v + 1]);
            const auto e = std::unique(raw.begin() + deg[v], raw.begin() + deg[v + 1]);
            aidx.insert(aidx.end(), raw.begin() + deg[v], e);
            aptr[v + 1] = int32_t(aidx.size());
        }
    }
    // 4. ordering and front structures
    const FrontTree T = nested_dissection(m, aptr.data(), aidx.data(), leaf);
    const int32_t F = T.fronts();
    nf_ = F;
    std::vector<int64_t> roff(F + 1, 0), woff(F + 1, 0), goff(F + 1, 0), loff(F + 1, 0), uoff(F + 1, 0), toff(F + 1, 0);
    int64_t max_m2 = 1, max_pm = 1, max_work = 1;
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
        max_m2 = std::max(max_m2, mf * mf);
        max_pm = std::max(max_pm, p * mf);
        max_work = std::max(max_work, DenseBlas::work_size(mf));
    }
    rtot_ = roff[F];
    wtot_ = woff[F];
    // update matrices on a stack: in postorder a front's children's updates lie on top,
    // contiguous and in child order; after its extend-add the front's own update takes
    // their place (uoff[f]: its offset; the stack peaks at stack_peak doubles)
    int64_t stack_peak = 1;
    {
        int64_t top = 0;
        for (int32_t f = 0; f < F; ++f) {
            int64_t base = top;
            for (int32_t c : T.children[f]) base -= T.nu(c) * T.nu(c);
            for (size_t i = 0, at = size_t(base); i < T.children[f].size(); ++i) {
                const int32_t c = T.children[f][i];
                if (uoff[c] != int64_t(at))
                    fail(LSS_ERR_HIP, "sparse stiff factorization: update stack out of order (internal)");
                at += size_t(T.nu(c) * T.nu(c));
            }
            uoff[f] = base;
            top = base + T.nu(f) * T.nu(f);
            stack_peak = std::max(stack_peak, top);
        }
    }
    stack_bytes_ = stack_peak * int64_t(sizeof(double));
    // front-local row of position q in front f
    auto loc = [&](int32_t f, int32_t q) -> int32_t {
        const int32_t p1 = T.p0[f] + T.np[f];
        if (q >= T.p0[f] && q < p1) return q - T.p0[f];
        const auto b0 = T.uidx.begin() + T.uptr[f], b1 = T.uidx.begin() + T.uptr[f + 1];
        const auto it = std::lower_bound(b0, b1, q);
        if (q < T.p0[f] || it == b1 || *it != q)
            fail(LSS_ERR_HIP, "sparse stiff factorization: position outside its front (internal)");
        return T.np[f] + int32_t(it - b0);
    };
    // CSC entry of the pair (a, b): operator row max(a, b), column min(a, b)
    auto entry = [&](int32_t a, int32_t b) -> int64_t {
        const int32_t hi = std::max(a, b), lo = std::min(a, b);
        const auto b0 = erow.begin() + eptr[hi], b1 = erow.begin() + eptr[hi + 1];
        const auto it = std::lower_bound(b0, b1, lo);
        return (it != b1 && *it == lo) ? esrc[it - erow.begin()] : -1;
    };
    // 5. original entries of every front's pivot columns (lower triangle of the front)
    std::vector<int32_t> tli, tlj;
    std::vector<int64_t> tsrc;
    for (int32_t f = 0; f < F; ++f) {
        toff[f] = int64_t(tli.size());
        for (int32_t q = T.p0[f]; q < T.p0[f] + T.np[f]; ++q) {
            const int32_t v = T.order[q];
            const int64_t ed = entry(v, v);
            if (ed >= 0) tli.push_back(q - T.p0[f]), tlj.push_back(q - T.p0[f]), tsrc.push_back(ed);
            for (int32_t e = aptr[v]; e < aptr[v + 1]; ++e) {
                const int32_t w = aidx[e], qw = T.pos[w];
                if (qw < q) continue;
                const int64_t ev = entry(v, w);
                if (ev < 0) continue;
                tli.push_back(loc(f, qw));
                tlj.push_back(q - T.p0[f]);
                tsrc.push_back(ev);
            }
        }
    }
    toff[F] = int64_t(tli.size());
    // 6. child -> parent row maps (at the child's woff) and, per parent row, the
    //    children's update entries that land there, children in order
    std::vector<int32_t> rel(std::max<int64_t>(wtot_, 1));
    std::vector<int64_t> cptr(rtot_ + 1, 0);
    for (int32_t f = 0; f < F; ++f)
        for (int32_t c : T.children[f])
            for (int64_t a = 0; a < T.nu(c); ++a) {
                const int32_t r = loc(f, T.uidx[T.uptr[c] + a]);
                rel[woff[c] + a] = r;
                ++cptr[roff[f] + r + 1];
            }
    for (int64_t r = 0; r < rtot_; ++r) cptr[r + 1] += cptr[r];
    std::vector<int64_t> cidx(std::max<int64_t>(cptr[rtot_], 1));
    {
        std::vector<int64_t> cur(cptr.begin(), cptr.end() - 1);
        for (int32_t f = 0; f < F; ++f)
            for (int32_t c : T.children[f])
                for (int64_t a = 0; a < T.nu(c); ++a) cidx[cur[roff[f] + rel[woff[c] + a]]++] = woff[c] + a;
    }
    t_symbolic = since(t0);
    // 7. numeric factorization, children first
    const auto t1 = clk::now();
    G_.resize(std::max<int64_t>(goff[F], 1));
    L_.resize(std::max<int64_t>(loff[F], 1));
    Lt_.resize(std::max<int64_t>(loff[F], 1));
    {
        DBuf<double> upd(stack_peak), Fw(max_m2), Gt(max_pm), work(max_work);
        DBuf<int32_t> d_li, d_lj, d_rel;
        DBuf<int64_t> d_src;
        d_li.upload(tli.data(), tli.size(), s);
        d_lj.upload(tlj.data(), tlj.size(), s);
        d_src.upload(tsrc.data(), tsrc.size(), s);
        d_rel.upload(rel.data(), rel.size(), s);
        for (int32_t f = 0; f < F; ++f) {
            const int64_t p = T.np[f], u = T.nu(f), mf = p + u;
            k::fill(mf * mf, Fw.get(), 0.0, s);
            k::sparse_assemble(toff[f + 1] - toff[f], d_li.get() + toff[f], d_lj.get() + toff[f],
                               d_src.get() + toff[f], d_vals, Fw.get(), mf, s);
            for (int32_t c : T.children[f])
                k::sparse_extend_add(int32_t(T.nu(c)), d_rel.get() + woff[c], upd.get() + uoff[c], Fw.get(), mf, s);
            blas.partial_factor_inverse(mf, p, Fw.get(), Gt.get(), work.get(), d_status, s);
            LSS_HIP(hipMemcpy2DAsync(G_.get() + goff[f], size_t(p) * sizeof(double), Gt.get(),
                                     size_t(mf) * sizeof(double), size_t(p) * sizeof(double), size_t(p),
                                     hipMemcpyDeviceToDevice, s));
            k::dense_mirror(p, G_.get() + goff[f], s);
            if (u > 0) {
                LSS_HIP(hipMemcpy2DAsync(L_.get() + loff[f], size_t(p) * sizeof(double), Fw.get() + p * mf,
                                         size_t(mf) * sizeof(double), size_t(p) * sizeof(double), size_t(u),
                                         hipMemcpyDeviceToDevice, s));
                LSS_HIP(hipMemcpy2DAsync(upd.get() + uoff[f], size_t(u) * sizeof(double), Fw.get() + p * mf + p,
                                         size_t(mf) * sizeof(double), size_t(u) * sizeof(double), size_t(u),
                                         hipMemcpyDeviceToDevice, s));
                k::sparse_transpose(u, p, L_.get() + loff[f], Lt_.get() + loff[f], s);
            }
        }
        LSS_HIP(hipStreamSynchronize(s));  // the workspaces are released at scope end
    }
    t_numeric = since(t1);
    // 8. solve tables: fronts by height, task lists per level
    int32_t H = 0;
    for (int32_t f = 0; f < F; ++f) H = std::max(H, T.height[f]);
    std::vector<std::vector<int32_t>> bylev(H + 1);
    for (int32_t f = 0; f < F; ++f) bylev[T.height[f]].push_back(f);
    std::vector<int2> trhs, ty, tw, tx;
    for (auto *lv : {&lv_rhs_, &lv_y_, &lv_w_, &lv_x_}) lv->assign(H + 2, 0);
    for (int32_t h = 0; h <= H; ++h) {
        lv_rhs_[h] = int64_t(trhs.size()), lv_y_[h] = int64_t(ty.size()), lv_w_[h] = int64_t(tw.size());
        lv_x_[h] = int64_t(tx.size());
        for (int32_t f : bylev[h]) {
            const int32_t p = T.np[f], u = int32_t(T.nu(f));
            for (int32_t a = 0; a < p + u; ++a) trhs.push_back(make_int2(f, a));
            for (int32_t i = 0; i < p; ++i) ty.push_back(make_int2(f, i)), tx.push_back(make_int2(f, i));
            for (int32_t a = 0; a < u; ++a) tw.push_back(make_int2(f, a));
        }
    }
    lv_rhs_[H + 1] = int64_t(trhs.size()), lv_y_[H + 1] = int64_t(ty.size()), lv_w_[H + 1] = int64_t(tw.size());
    lv_x_[H + 1] = int64_t(tx.size());
    t_rhs_.upload(trhs.data(), trhs.size(), s);
    t_y_.upload(ty.data(), ty.size(), s);
    t_w_.upload(tw.data(), tw.size(), s);
    t_x_.upload(tx.data(), tx.size(), s);
    order_.upload(T.order.data(), m, s);
    p0_.upload(T.p0.data(), F, s);
    np_.upload(T.np.data(), F, s);
    std::vector<int32_t> nu(F);
    for (int32_t f = 0; f < F; ++f) nu[f] = int32_t(T.nu(f));
    nu_.upload(nu.data(), F, s);
    roff_.upload(roff.data(), F + 1, s);
    woff_.upload(woff.data(), F + 1, s);
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
                        ypos_.get(), m_, k, stop, s);
        k::sparse_fwd_w(lv_w_[h + 1] - lv_w_[h], t_w_.get() + lv_w_[h], fv, L_.get(), rhs_.get(), rtot_,
                        ypos_.get(), m_, wbuf_.get(), wtot_, k, stop, s);
    }
    for (int h = NL - 1; h >= 0; --h) {
        k::sparse_bwd_s(lv_y_[h + 1] - lv_y_[h], t_y_.get() + lv_y_[h], fv, Lt_.get(), xpos_.get(), ypos_.get(), m_,
                        rhs_.get(), rtot_, k, stop, s);
        k::sparse_bwd_x(lv_x_[h + 1] - lv_x_[h], t_x_.get() + lv_x_[h], fv, G_.get(), rhs_.get(), rtot_,
                        xpos_.get(), m_, k, stop, s);
    }
    k::sparse_scatter_add(m_, order_.get(), pdofs, xpos_.get(), z, z32, ld, k, stop, s);
}

}  // namespace lss
