#include "amg_ref.hpp"

#include <chrono>

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <numeric>

// Every product and sum is rounded on its own (no fused multiply-add): the setup is
// then reproducible by any IEEE implementation of the same operation order.
#pragma clang fp contract(off)

namespace amgref {

double amg_point_random(int64_t i)
{
    // splitmix64 of the index; 53 random bits in [0, 1)
    uint64_t z = uint64_t(i) + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return double(z >> 11) * (1.0 / 9007199254740992.0);
}

namespace {

// rows [0, n) in contiguous chunks of a thread each: fills ptr from per-row counts
void counts_to_ptr(const std::vector<int64_t> &cnt, std::vector<int64_t> &ptr)
{
    ptr.assign(cnt.size() + 1, 0);
    for (size_t i = 0; i < cnt.size(); ++i) ptr[i + 1] = ptr[i] + cnt[i];
}

double diag_of(const CsrHost &A, int32_t i)
{
    for (int64_t e = A.ptr[i]; e < A.ptr[i + 1]; ++e)
        if (A.col[e] == i) return A.val[e];
    return 0.0;
}

}  // namespace

void amg_strength(const CsrHost &A, const std::vector<int32_t> &func, int num_functions, double theta,
                  double max_row_sum, CsrHost &S)
{
    const int32_t n = A.n;
    std::vector<int64_t> cnt(n, 0);
    std::vector<uint8_t> strong(A.col.size(), 0);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int32_t i = 0; i < n; ++i) {
        const double diag = diag_of(A, i);
        double scale = 0.0, row_sum = diag;
        for (int64_t e = A.ptr[i]; e < A.ptr[i + 1]; ++e) {
            const int32_t j = A.col[e];
            if (j == i || (num_functions > 1 && func[j] != func[i])) continue;
            const double a = A.val[e];
            scale = diag < 0.0 ? std::max(scale, a) : std::min(scale, a);
            row_sum += a;
        }
        if (max_row_sum < 1.0 && std::fabs(row_sum) > std::fabs(diag) * max_row_sum) continue;  // all weak
        const double cut = theta * scale;
        int64_t c = 0;
        for (int64_t e = A.ptr[i]; e < A.ptr[i + 1]; ++e) {
            const int32_t j = A.col[e];
            if (j == i || (num_functions > 1 && func[j] != func[i])) continue;
            const double a = A.val[e];
            if (diag < 0.0 ? a > cut : a < cut) {
                strong[e] = 1;
                ++c;
            }
        }
        cnt[i] = c;
    }
    S.n = S.m = n;
    counts_to_ptr(cnt, S.ptr);
    S.col.resize(S.ptr[n]);
    S.val.clear();
#pragma omp parallel for schedule(dynamic, 4096)
    for (int32_t i = 0; i < n; ++i) {
        int64_t w = S.ptr[i];
        for (int64_t e = A.ptr[i]; e < A.ptr[i + 1]; ++e)
            if (strong[e]) S.col[w++] = A.col[e];
    }
}

void amg_transpose(const CsrHost &P, CsrHost &R)
{
    R.n = P.m;
    R.m = P.n;
    std::vector<int64_t> cnt(R.n, 0);
    for (int64_t e = 0; e < int64_t(P.col.size()); ++e) ++cnt[P.col[e]];
    counts_to_ptr(cnt, R.ptr);
    R.col.resize(P.col.size());
    const bool vals = !P.val.empty();
    R.val.resize(vals ? P.val.size() : 0);
    std::vector<int64_t> at(R.ptr.begin(), R.ptr.end() - 1);
    for (int32_t i = 0; i < P.n; ++i)  // rows in order: columns of R ascend
        for (int64_t e = P.ptr[i]; e < P.ptr[i + 1]; ++e) {
            const int64_t w = at[P.col[e]]++;
            R.col[w] = i;
            if (vals) R.val[w] = P.val[e];
        }
}

void amg_pmis(const CsrHost &S, std::vector<int8_t> &cf)
{
    const int32_t n = S.n;
    CsrHost ST;
    amg_transpose(S, ST);
    std::vector<double> measure(n);
    cf.assign(n, 0);
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        const int64_t infl = ST.ptr[i + 1] - ST.ptr[i];  // points that depend strongly on i
        measure[i] = double(infl) + amg_point_random(i);
        if (S.ptr[i + 1] == S.ptr[i] && infl == 0) cf[i] = -3;  // no strong connection at all
        else if (infl == 0) cf[i] = -1;                          // influences nobody: F
    }
    auto beats = [&](int32_t j, int32_t i) {
        return measure[j] > measure[i] || (measure[j] == measure[i] && j > i);
    };
    std::vector<int32_t> undecided;
    for (int32_t i = 0; i < n; ++i)
        if (cf[i] == 0) undecided.push_back(i);
    std::vector<uint8_t> newc(n, 0);
    while (!undecided.empty()) {
        // independent set: an undecided point beats every undecided strong neighbour
        const int64_t u = int64_t(undecided.size());
#pragma omp parallel for schedule(dynamic, 2048)
        for (int64_t q = 0; q < u; ++q) {
            const int32_t i = undecided[q];
            bool top = true;
            for (int64_t e = S.ptr[i]; e < S.ptr[i + 1] && top; ++e) {
                const int32_t j = S.col[e];
                if (cf[j] == 0 && beats(j, i)) top = false;
            }
            for (int64_t e = ST.ptr[i]; e < ST.ptr[i + 1] && top; ++e) {
                const int32_t j = ST.col[e];
                if (cf[j] == 0 && beats(j, i)) top = false;
            }
            newc[i] = top;
        }
        for (int64_t q = 0; q < u; ++q)
            if (newc[undecided[q]]) cf[undecided[q]] = 1;
        // undecided points that depend strongly on a C point become F
#pragma omp parallel for schedule(dynamic, 2048)
        for (int64_t q = 0; q < u; ++q) {
            const int32_t i = undecided[q];
            if (cf[i] != 0) continue;
            for (int64_t e = S.ptr[i]; e < S.ptr[i + 1]; ++e)
                if (cf[S.col[e]] == 1) {
                    newc[i] = 2;  // made F below (this pass reads C marks only)
                    break;
                }
        }
        std::vector<int32_t> next;
        for (int64_t q = 0; q < u; ++q) {
            const int32_t i = undecided[q];
            if (newc[i] == 2) cf[i] = -1;
            if (cf[i] == 0) next.push_back(i);
            newc[i] = 0;
        }
        undecided.swap(next);
    }
}

void amg_interp_extpi(const CsrHost &A, const CsrHost &S, const std::vector<int8_t> &cf,
                      const std::vector<int32_t> &func, int num_functions, int pmax, CsrHost &P)
{
    const int32_t n = A.n;
    std::vector<int32_t> cidx(n, -1);
    int32_t nc = 0;
    for (int32_t i = 0; i < n; ++i)
        if (cf[i] == 1) cidx[i] = nc++;
    std::vector<std::vector<int32_t>> rcol(n);
    std::vector<std::vector<double>> rval(n);
    const int nt = omp_get_max_threads();
    // per thread: position of a point in the current row's interpolatory set, or -1;
    // strong-F marks for the current row
    std::vector<std::vector<int32_t>> posv(nt, std::vector<int32_t>(n, -1));
    std::vector<std::vector<uint8_t>> sfv(nt, std::vector<uint8_t>(n, 0));
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t i = 0; i < n; ++i) {
        if (cf[i] == 1) {
            rcol[i] = {cidx[i]};
            rval[i] = {1.0};
            continue;
        }
        if (cf[i] == -3 || S.ptr[i + 1] == S.ptr[i]) continue;  // no strong neighbours: no interpolation
        std::vector<int32_t> &pos = posv[omp_get_thread_num()];
        std::vector<uint8_t> &isf = sfv[omp_get_thread_num()];
        std::vector<int32_t> chat;  // fine indices of C-hat_i
        for (int64_t e = S.ptr[i]; e < S.ptr[i + 1]; ++e) {
            const int32_t j = S.col[e];
            if (cf[j] == 1) {
                if (pos[j] < 0) {
                    pos[j] = int32_t(chat.size());
                    chat.push_back(j);
                }
            } else {
                isf[j] = 1;
                for (int64_t f = S.ptr[j]; f < S.ptr[j + 1]; ++f) {
                    const int32_t l = S.col[f];
                    if (cf[l] == 1 && pos[l] < 0) {
                        pos[l] = int32_t(chat.size());
                        chat.push_back(l);
                    }
                }
            }
        }
        std::vector<double> w(chat.size(), 0.0);
        double diag = 0.0;
        for (int64_t e = A.ptr[i]; e < A.ptr[i + 1]; ++e) {
            const int32_t j = A.col[e];
            const double a = A.val[e];
            if (j == i) {
                diag += a;
            } else if (pos[j] >= 0) {
                w[pos[j]] += a;  // direct connection to C-hat (strong or weak)
            } else if (isf[j]) {
                // strong F neighbour k = j: distribute a_ik over C-hat_i and i by the
                // entries of row k of sign opposite to a_kk
                const double akk = diag_of(A, j);
                double d = 0.0;
                for (int64_t f = A.ptr[j]; f < A.ptr[j + 1]; ++f) {
                    const int32_t l = A.col[f];
                    if ((l == i || pos[l] >= 0) && A.val[f] * akk < 0.0) d += A.val[f];
                }
                if (d == 0.0) {
                    diag += a;
                    continue;
                }
                const double fac = a / d;
                for (int64_t f = A.ptr[j]; f < A.ptr[j + 1]; ++f) {
                    const int32_t l = A.col[f];
                    const double v = A.val[f];
                    if (v * akk >= 0.0) continue;
                    if (l == i) diag += fac * v;
                    else if (pos[l] >= 0) w[pos[l]] += fac * v;
                }
            } else if (num_functions == 1 || func[j] == func[i]) {
                diag += a;  // weak connection outside C-hat: lumped
            }
        }
        for (int32_t j : chat) pos[j] = -1;
        for (int64_t e = S.ptr[i]; e < S.ptr[i + 1]; ++e) isf[S.col[e]] = 0;
        if (diag == 0.0 || chat.empty()) continue;
        std::vector<int32_t> order(chat.size());
        std::iota(order.begin(), order.end(), 0);
        for (auto &v : w) v = -v / diag;
        if (int(chat.size()) > pmax) {
            // the pmax largest weights, the row sum kept by one scale
            std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
                const double x = std::fabs(w[a]), y = std::fabs(w[b]);
                return x > y || (x == y && chat[a] < chat[b]);
            });
            double all = 0.0, kept = 0.0;
            for (double v : w) all += v;
            for (int q = 0; q < pmax; ++q) kept += w[order[q]];
            order.resize(pmax);
            const double sc = kept != 0.0 ? all / kept : 1.0;
            for (int32_t q : order) w[q] *= sc;
        }
        std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return chat[a] < chat[b]; });
        for (int32_t q : order) {
            rcol[i].push_back(cidx[chat[q]]);
            rval[i].push_back(w[q]);
        }
    }
    P.n = n;
    P.m = nc;
    P.ptr.assign(n + 1, 0);
    for (int32_t i = 0; i < n; ++i) P.ptr[i + 1] = P.ptr[i] + int64_t(rcol[i].size());
    P.col.resize(P.ptr[n]);
    P.val.resize(P.ptr[n]);
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        std::copy(rcol[i].begin(), rcol[i].end(), P.col.begin() + P.ptr[i]);
        std::copy(rval[i].begin(), rval[i].end(), P.val.begin() + P.ptr[i]);
    }
}

void amg_spgemm(const CsrHost &X, const CsrHost &Y, CsrHost &Z)
{
    // Gustavson, one row at a time: products summed in the order of X's row and
    // Y's rows, columns of the result sorted.  Rows are taken in chunks; a chunk's
    // rows go to buffers of its own (no allocation per row), then every chunk is
    // copied to its place in Z.
    const int32_t n = X.n, m = Y.m;
    constexpr int32_t kChunk = 256;
    const int32_t nchunks = (n + kChunk - 1) / kChunk;
    const int nt = omp_get_max_threads();
    std::vector<std::vector<int32_t>> ccol(nchunks);
    std::vector<std::vector<double>> cval(nchunks);
    Z.n = n;
    Z.m = m;
    Z.ptr.assign(size_t(n) + 1, 0);
    std::vector<std::vector<int32_t>> posv(nt, std::vector<int32_t>(size_t(m), -1));
#pragma omp parallel
    {
        std::vector<int32_t> &pos = posv[omp_get_thread_num()];
        std::vector<int32_t> cols;
        std::vector<double> acc;
#pragma omp for schedule(dynamic, 1)
        for (int32_t ch = 0; ch < nchunks; ++ch) {
            const int32_t i0 = ch * kChunk, i1 = std::min(n, i0 + kChunk);
            std::vector<int32_t> &oc = ccol[ch];
            std::vector<double> &ov = cval[ch];
            for (int32_t i = i0; i < i1; ++i) {
                cols.clear();
                acc.clear();
                for (int64_t e = X.ptr[i]; e < X.ptr[i + 1]; ++e) {
                    const int32_t k = X.col[e];
                    const double a = X.val[e];
                    for (int64_t f = Y.ptr[k]; f < Y.ptr[k + 1]; ++f) {
                        const int32_t j = Y.col[f];
                        int32_t q = pos[j];
                        if (q < 0) {
                            q = pos[j] = int32_t(cols.size());
                            cols.push_back(j);
                            acc.push_back(0.0);
                        }
                        acc[q] += a * Y.val[f];
                    }
                }
                const size_t base = oc.size();
                oc.insert(oc.end(), cols.begin(), cols.end());
                std::sort(oc.begin() + base, oc.end());
                for (size_t q = base; q < oc.size(); ++q) {
                    const int32_t j = oc[q];
                    ov.push_back(acc[pos[j]]);
                    pos[j] = -1;
                }
                Z.ptr[size_t(i) + 1] = int64_t(cols.size());
            }
        }
    }
    for (int32_t i = 0; i < n; ++i) Z.ptr[size_t(i) + 1] += Z.ptr[i];
    Z.col.resize(size_t(Z.ptr[n]));
    Z.val.resize(size_t(Z.ptr[n]));
#pragma omp parallel for schedule(dynamic, 16)
    for (int32_t ch = 0; ch < nchunks; ++ch) {
        const int64_t o = Z.ptr[size_t(ch) * kChunk];
        std::copy(ccol[ch].begin(), ccol[ch].end(), Z.col.begin() + o);
        std::copy(cval[ch].begin(), cval[ch].end(), Z.val.begin() + o);
        std::vector<int32_t>().swap(ccol[ch]);
        std::vector<double>().swap(cval[ch]);
    }
}

void amg_setup_host(CsrHost A0, const AmgOptions &opt, std::vector<AmgLevelHost> &levels, AmgSetupTimes *times)
{
    AmgSetupTimes tm;
    using clk = std::chrono::steady_clock;
    auto t = clk::now();
    auto lap = [&](double &acc) {
        const auto u = clk::now();
        acc += std::chrono::duration<double>(u - t).count();
        t = u;
    };
    levels.clear();
    levels.emplace_back();
    const int32_t n0 = A0.n;
    levels[0].A = std::move(A0);
    levels[0].func.resize(n0);
    for (int32_t i = 0; i < n0; ++i) levels[0].func[i] = i % std::max(opt.num_functions, 1);
    for (int lv = 0;; ++lv) {
        AmgLevelHost &L = levels[lv];
        const int32_t n = L.A.n;
        L.dinv.resize(n);
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < n; ++i) {
            double s = 0.0;
            for (int64_t e = L.A.ptr[i]; e < L.A.ptr[i + 1]; ++e) s += std::fabs(L.A.val[e]);
            L.dinv[i] = s > 0.0 ? 1.0 / s : 0.0;
        }
        lap(tm.norms);
        if (lv + 1 >= opt.max_levels || n <= opt.min_coarse) break;
        CsrHost S;
        amg_strength(L.A, L.func, opt.num_functions, opt.theta, opt.max_row_sum, S);
        lap(tm.strength);
        amg_pmis(S, L.cf);
        lap(tm.pmis);
        int32_t nc = 0;
        for (int8_t c : L.cf) nc += c == 1;
        if (nc == 0 || nc == n) {
            L.cf.clear();
            break;
        }
        amg_interp_extpi(L.A, S, L.cf, L.func, opt.num_functions, opt.pmax, L.P);
        lap(tm.interp);
        amg_transpose(L.P, L.R);
        lap(tm.transpose);
        AmgLevelHost next;
        CsrHost AP;
        amg_spgemm(L.A, L.P, AP);
        amg_spgemm(L.R, AP, next.A);
        lap(tm.rap);
        next.func.resize(nc);
        for (int32_t i = 0, c = 0; i < n; ++i)
            if (L.cf[i] == 1) next.func[c++] = L.func[i];
        levels.push_back(std::move(next));
    }
    if (times) *times = tm;
}

}  // namespace amgref
