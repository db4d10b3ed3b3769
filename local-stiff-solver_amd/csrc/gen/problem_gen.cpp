// Synthetic locally-stiff SPD systems (see problem_gen.h for the model).
#include "problem_gen.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

// Kuhn split of the unit cell: tet t visits corners 000 -> e_s1 -> e_s1+e_s2 -> 111.
constexpr int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

inline std::array<int, 4> tet_corners(int t)
{
    const int s1 = kPerm[t][0], s2 = kPerm[t][1];
    return {0, 1 << s1, (1 << s1) | (1 << s2), 7};
}

struct Sliver {
    int64_t elem;
    double X[4][3];
};

struct Contact {
    int32_t i, j;
    double K[9];  // kappa * n n^T, exactly symmetric
};

// Barycentric gradients and volume of a tet.
bool tet_gradients(const double X[4][3], double g[4][3], double &vol)
{
    double M[3][3];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) M[r][c] = X[c + 1][r] - X[0][r];
    const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                       M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                       M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    if (det == 0.0) return false;
    const double id = 1.0 / det;
    double inv[3][3];
    inv[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) * id;
    inv[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) * id;
    inv[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) * id;
    inv[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) * id;
    inv[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) * id;
    inv[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * id;
    inv[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) * id;
    inv[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) * id;
    inv[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * id;
    // grad lambda_{k+1} = row k of M^{-1}
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) g[k + 1][d] = inv[k][d];
    for (int d = 0; d < 3; ++d) g[0][d] = -(g[1][d] + g[2][d] + g[3][d]);
    vol = std::fabs(det) / 6.0;
    return true;
}

// Element matrix per unit coefficient. bs = 1: 4x4 Laplacian; bs = 3: 12x12 elasticity.
void element_matrix(const double X[4][3], int bs, double nu, double *K)
{
    double g[4][3], vol;
    if (!tet_gradients(X, g, vol)) throw std::runtime_error("degenerate tetrahedron");
    const int m = 4 * bs;
    if (bs == 1) {
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
                K[a * m + b] = vol * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
        return;
    }
    const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = 1.0 / (2.0 * (1.0 + nu));
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            const double gg = g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double v = lam * g[a][i] * g[b][j] + mu * g[a][j] * g[b][i];
                    if (i == j) v += mu * gg;
                    K[(a * 3 + i) * m + (b * 3 + j)] = vol * v;
                }
        }
}

}  // namespace

struct lssgen_problem {
    lssgen_spec spec{};
    int bs = 1;
    int N = 0;
    double h = 0;
    int64_t n_nodes = 0, n_elem = 0;
    std::vector<double> coef;               // per element
    std::vector<Sliver> slivers;            // sorted by elem
    std::vector<Contact> contacts;
    std::vector<int64_t> contact_ptr;       // per node CSR into contact_idx
    std::vector<int32_t> contact_idx;
    std::vector<double> Kref;               // 6 reference element matrices (unit cell, coef 1)
    // block pattern
    std::vector<int64_t> rowptr;            // n_nodes + 1
    std::vector<int32_t> cols;              // sorted per row
    std::vector<double> blocks;             // bs*bs per slot, row-major
    std::vector<uint8_t> dirichlet;         // per node
    int64_t nnz = 0;

    int corner_node(int64_t cube, int corner) const
    {
        const int64_t Nc = N - 1;
        const int64_t ci = cube % Nc, cj = (cube / Nc) % Nc, ck = cube / (Nc * Nc);
        return static_cast<int>((ci + (corner & 1)) + N * ((cj + ((corner >> 1) & 1)) + N * (ck + ((corner >> 2) & 1))));
    }
    const Sliver *find_sliver(int64_t e) const
    {
        auto it = std::lower_bound(slivers.begin(), slivers.end(), e,
                                   [](const Sliver &s, int64_t v) { return s.elem < v; });
        return (it != slivers.end() && it->elem == e) ? &*it : nullptr;
    }
    void element_vertices(int64_t e, double X[4][3]) const
    {
        const int64_t cube = e / 6;
        const int t = static_cast<int>(e % 6);
        const auto cs = tet_corners(t);
        for (int a = 0; a < 4; ++a) {
            const int node = corner_node(cube, cs[a]);
            const int i = node % N, j = (node / N) % N, k = node / (N * N);
            X[a][0] = i * h;
            X[a][1] = j * h;
            X[a][2] = k * h;
        }
    }
    // element matrix per unit coefficient (m x m, m = 4 bs)
    void unit_element(int64_t e, double *K) const
    {
        const int m = 4 * bs;
        if (const Sliver *s = find_sliver(e)) {
            element_matrix(s->X, bs, spec.nu, K);
            return;
        }
        const double *R = &Kref[(e % 6) * m * m];
        for (int q = 0; q < m * m; ++q) K[q] = R[q] * h;  // P1 stiffness scales with h in 3D
    }
    int64_t slot(int64_t p, int32_t q) const
    {
        const int32_t *b = cols.data() + rowptr[p], *en = cols.data() + rowptr[p + 1];
        const int32_t *it = std::lower_bound(b, en, q);
        if (it == en || *it != q) throw std::runtime_error("pattern slot missing");
        return it - cols.data();
    }
    void assemble();
};

void lssgen_problem::assemble()
{
    const int bb = bs * bs;
    const int m = 4 * bs;
    std::fill(blocks.begin(), blocks.end(), 0.0);
    const int64_t Nc = N - 1;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t p = 0; p < n_nodes; ++p) {
        const int pi = static_cast<int>(p % N), pj = static_cast<int>((p / N) % N), pk = static_cast<int>(p / (int64_t(N) * N));
        std::vector<double> K(m * m);
        for (int c = 0; c < 8; ++c) {
            const int oi = pi - (c & 1), oj = pj - ((c >> 1) & 1), ok = pk - ((c >> 2) & 1);
            if (oi < 0 || oj < 0 || ok < 0 || oi >= Nc || oj >= Nc || ok >= Nc) continue;
            const int64_t cube = oi + Nc * (oj + Nc * int64_t(ok));
            for (int t = 0; t < 6; ++t) {
                const auto cs = tet_corners(t);
                int a = -1;
                for (int v = 0; v < 4; ++v)
                    if (cs[v] == c) a = v;
                if (a < 0) continue;
                const int64_t e = cube * 6 + t;
                unit_element(e, K.data());
                const double ce = coef[e];
                for (int b = 0; b < 4; ++b) {
                    const int32_t q = corner_node(cube, cs[b]);
                    double *blk = &blocks[slot(p, q) * bb];
                    for (int i = 0; i < bs; ++i)
                        for (int j = 0; j < bs; ++j) blk[i * bs + j] += ce * K[(a * bs + i) * m + (b * bs + j)];
                }
            }
        }
        for (int64_t c = contact_ptr[p]; c < contact_ptr[p + 1]; ++c) {
            const Contact &ct = contacts[contact_idx[c]];
            const int32_t other = (ct.i == p) ? ct.j : ct.i;
            double *dg = &blocks[slot(p, static_cast<int32_t>(p)) * bb];
            double *od = &blocks[slot(p, other) * bb];
            for (int q = 0; q < 9; ++q) {
                dg[q] += ct.K[q];
                od[q] -= ct.K[q];
            }
        }
    }
    // Mirror so that block(q, p) == block(p, q)^T bit for bit.
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t q = 0; q < n_nodes; ++q) {
        for (int64_t s = rowptr[q]; s < rowptr[q + 1]; ++s) {
            const int32_t p = cols[s];
            double *dst = &blocks[s * bb];
            if (p < q) {
                const double *src = &blocks[slot(p, static_cast<int32_t>(q)) * bb];
                for (int i = 0; i < bs; ++i)
                    for (int j = 0; j < bs; ++j) dst[i * bs + j] = src[j * bs + i];
            } else if (p == q) {
                for (int i = 0; i < bs; ++i)
                    for (int j = 0; j < i; ++j) dst[i * bs + j] = dst[j * bs + i];
            }
        }
    }
}

extern "C" {

lssgen_problem *lssgen_create(const lssgen_spec *spec, char *err, int errcap)
{
    try {
        auto P = new lssgen_problem();
        P->spec = *spec;
        P->N = spec->N;
        if (P->N < 2) throw std::runtime_error("N must be >= 2");
        P->bs = (spec->kind == LSSGEN_ELASTICITY) ? 3 : 1;
        const int N = P->N, bs = P->bs;
        P->h = 1.0 / (N - 1);
        P->n_nodes = int64_t(N) * N * N;
        if (P->n_nodes * bs >= (int64_t(1) << 31)) throw std::runtime_error("problem too large for int32 indices");
        const int64_t Nc = N - 1;
        P->n_elem = 6 * Nc * Nc * Nc;
        const double h = P->h;

        // reference element matrices (unit cell, coefficient 1)
        const int m = 4 * bs;
        P->Kref.assign(6 * m * m, 0.0);
        for (int t = 0; t < 6; ++t) {
            const auto cs = tet_corners(t);
            double X[4][3];
            for (int a = 0; a < 4; ++a)
                for (int d = 0; d < 3; ++d) X[a][d] = (cs[a] >> d) & 1;
            element_matrix(X, bs, spec->nu, &P->Kref[t * m * m]);
        }

        std::mt19937 rng(spec->seed);
        std::uniform_real_distribution<double> U01(0.0, 1.0);

        // per-element coefficients
        P->coef.assign(P->n_elem, spec->E);
        if (spec->sphere_radius > 0) {
            const double cx = U01(rng), cy = U01(rng), cz = U01(rng);
            const double r2 = spec->sphere_radius * spec->sphere_radius;
            for (int64_t e = 0; e < P->n_elem; ++e) {
                double X[4][3];
                P->element_vertices(e, X);
                const double gx = 0.25 * (X[0][0] + X[1][0] + X[2][0] + X[3][0]) - cx;
                const double gy = 0.25 * (X[0][1] + X[1][1] + X[2][1] + X[3][1]) - cy;
                const double gz = 0.25 * (X[0][2] + X[1][2] + X[2][2] + X[3][2]) - cz;
                if (gx * gx + gy * gy + gz * gz <= r2) P->coef[e] = spec->sphere_coef;
            }
        }
        if (spec->E_log10_max > spec->E_log10_min) {
            std::uniform_real_distribution<double> UE(spec->E_log10_min, spec->E_log10_max);
            for (int64_t e = 0; e < P->n_elem; ++e) P->coef[e] = std::pow(10.0, UE(rng));
        }

        // slivers: distinct elements via a partial Fisher-Yates shuffle
        if (spec->sliver_frac > 0) {
            const int64_t cnt = std::llround(spec->sliver_frac * double(P->n_elem));
            std::vector<int64_t> idx(P->n_elem);
            for (int64_t e = 0; e < P->n_elem; ++e) idx[e] = e;
            for (int64_t s = 0; s < cnt; ++s) {
                std::uniform_int_distribution<int64_t> pick(s, P->n_elem - 1);
                std::swap(idx[s], idx[pick(rng)]);
            }
            std::uniform_int_distribution<int> vpick(0, 3);
            for (int64_t s = 0; s < cnt; ++s) {
                Sliver sl;
                sl.elem = idx[s];
                P->element_vertices(sl.elem, sl.X);
                const int a = vpick(rng);
                // plane through the other three vertices
                int f[3], c = 0;
                for (int v = 0; v < 4; ++v)
                    if (v != a) f[c++] = v;
                double u[3], w[3], nrm[3];
                for (int d = 0; d < 3; ++d) {
                    u[d] = sl.X[f[1]][d] - sl.X[f[0]][d];
                    w[d] = sl.X[f[2]][d] - sl.X[f[0]][d];
                }
                nrm[0] = u[1] * w[2] - u[2] * w[1];
                nrm[1] = u[2] * w[0] - u[0] * w[2];
                nrm[2] = u[0] * w[1] - u[1] * w[0];
                const double nl = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
                for (int d = 0; d < 3; ++d) nrm[d] /= nl;
                double dist = 0;
                for (int d = 0; d < 3; ++d) dist += (sl.X[a][d] - sl.X[f[0]][d]) * nrm[d];
                const double target = (dist > 0 ? 1.0 : -1.0) * spec->sliver_eps * h;
                for (int d = 0; d < 3; ++d) sl.X[a][d] -= (dist - target) * nrm[d];
                P->slivers.push_back(sl);
            }
            std::sort(P->slivers.begin(), P->slivers.end(),
                      [](const Sliver &x, const Sliver &y) { return x.elem < y.elem; });
        }

        auto node_id = [N](int i, int j, int k) { return i + N * (j + N * k); };
        P->dirichlet.assign(P->n_nodes, 0);
        if (spec->dirichlet_axis >= 0 && spec->dirichlet_axis < 3) {
            for (int64_t p = 0; p < P->n_nodes; ++p) {
                const int c[3] = {int(p % N), int((p / N) % N), int(p / (int64_t(N) * N))};
                if (c[spec->dirichlet_axis] == 0) P->dirichlet[p] = 1;
            }
        }

        // contact pairs on the top face z = 1
        if (spec->contact_frac > 0) {
            if (bs != 3) throw std::runtime_error("contact pairs need block size 3");
            std::vector<int32_t> top;
            for (int j = 0; j < N; ++j)
                for (int i = 0; i < N; ++i) {
                    const int32_t p = node_id(i, j, N - 1);
                    if (!P->dirichlet[p]) top.push_back(p);
                }
            const int64_t cnt = std::min<int64_t>(std::llround(spec->contact_frac * double(N) * N), top.size());
            for (int64_t s = 0; s < cnt; ++s) {
                std::uniform_int_distribution<int64_t> pick(s, int64_t(top.size()) - 1);
                std::swap(top[s], top[pick(rng)]);
            }
            const double R = spec->contact_radius_h;
            const int Ri = static_cast<int>(std::floor(R));
            std::uniform_real_distribution<double> Uk(spec->contact_log10_kmin, spec->contact_log10_kmax);
            std::uniform_real_distribution<double> Uz(-1.0, 1.0);
            std::uniform_real_distribution<double> Uphi(0.0, 2.0 * M_PI);
            for (int64_t s = 0; s < cnt; ++s) {
                const int32_t p = top[s];
                const int pi = p % N, pj = (p / N) % N, pk = p / (N * N);
                std::vector<int32_t> cand;
                for (int dk = -Ri; dk <= Ri; ++dk)
                    for (int dj = -Ri; dj <= Ri; ++dj)
                        for (int di = -Ri; di <= Ri; ++di) {
                            if (!di && !dj && !dk) continue;
                            if (di * di + dj * dj + dk * dk > R * R) continue;
                            const int qi = pi + di, qj = pj + dj, qk = pk + dk;
                            if (qi < 0 || qj < 0 || qk < 0 || qi >= N || qj >= N || qk >= N) continue;
                            const int32_t q = node_id(qi, qj, qk);
                            if (P->dirichlet[q]) continue;
                            cand.push_back(q);
                        }
                if (cand.empty()) continue;
                std::uniform_int_distribution<int64_t> cp(0, int64_t(cand.size()) - 1);
                const int32_t q = cand[cp(rng)];
                const double kappa = std::pow(10.0, Uk(rng));
                const double z = Uz(rng), phi = Uphi(rng);
                const double rr = std::sqrt(std::max(0.0, 1.0 - z * z));
                const double n[3] = {rr * std::cos(phi), rr * std::sin(phi), z};
                Contact ct;
                ct.i = p;
                ct.j = q;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) ct.K[a * 3 + b] = kappa * (a <= b ? n[a] * n[b] : n[b] * n[a]);
                P->contacts.push_back(ct);
            }
        }
        // per-node contact lists
        P->contact_ptr.assign(P->n_nodes + 1, 0);
        for (const auto &c : P->contacts) {
            P->contact_ptr[c.i + 1]++;
            P->contact_ptr[c.j + 1]++;
        }
        for (int64_t p = 0; p < P->n_nodes; ++p) P->contact_ptr[p + 1] += P->contact_ptr[p];
        P->contact_idx.assign(P->contact_ptr[P->n_nodes], 0);
        {
            std::vector<int64_t> fill(P->contact_ptr.begin(), P->contact_ptr.end() - 1);
            for (size_t c = 0; c < P->contacts.size(); ++c) {
                P->contact_idx[fill[P->contacts[c].i]++] = int32_t(c);
                P->contact_idx[fill[P->contacts[c].j]++] = int32_t(c);
            }
        }

        // block pattern: 14-neighbour Kuhn stencil + self + contact partners
        static const int offs[15][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
                                        {0, 0, 1},  {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {1, 0, 1},
                                        {-1, 0, -1}, {0, 1, 1},  {0, -1, -1}, {1, 1, 1},  {-1, -1, -1}};
        P->rowptr.assign(P->n_nodes + 1, 0);
        std::vector<int32_t> cnts(P->n_nodes);
#pragma omp parallel for schedule(static)
        for (int64_t p = 0; p < P->n_nodes; ++p) {
            const int pi = p % N, pj = (p / N) % N, pk = p / (int64_t(N) * N);
            int c = 0;
            for (auto &o : offs) {
                const int qi = pi + o[0], qj = pj + o[1], qk = pk + o[2];
                if (qi < 0 || qj < 0 || qk < 0 || qi >= N || qj >= N || qk >= N) continue;
                ++c;
            }
            c += static_cast<int>(P->contact_ptr[p + 1] - P->contact_ptr[p]);
            cnts[p] = c;
        }
        for (int64_t p = 0; p < P->n_nodes; ++p) P->rowptr[p + 1] = P->rowptr[p] + cnts[p];
        P->cols.assign(P->rowptr[P->n_nodes], 0);
#pragma omp parallel for schedule(static)
        for (int64_t p = 0; p < P->n_nodes; ++p) {
            const int pi = p % N, pj = (p / N) % N, pk = p / (int64_t(N) * N);
            int32_t *c = &P->cols[P->rowptr[p]];
            int k = 0;
            for (auto &o : offs) {
                const int qi = pi + o[0], qj = pj + o[1], qk = pk + o[2];
                if (qi < 0 || qj < 0 || qk < 0 || qi >= N || qj >= N || qk >= N) continue;
                c[k++] = node_id(qi, qj, qk);
            }
            for (int64_t t = P->contact_ptr[p]; t < P->contact_ptr[p + 1]; ++t) {
                const Contact &ct = P->contacts[P->contact_idx[t]];
                c[k++] = (ct.i == p) ? ct.j : ct.i;
            }
            std::sort(c, c + k);
            cnts[p] = static_cast<int32_t>(std::unique(c, c + k) - c);
        }
        // compact duplicates (contact partner inside the stencil, repeated pairs)
        {
            std::vector<int64_t> nr(P->n_nodes + 1, 0);
            for (int64_t p = 0; p < P->n_nodes; ++p) nr[p + 1] = nr[p] + cnts[p];
            std::vector<int32_t> nc(nr[P->n_nodes]);
#pragma omp parallel for schedule(static)
            for (int64_t p = 0; p < P->n_nodes; ++p)
                std::copy(&P->cols[P->rowptr[p]], &P->cols[P->rowptr[p]] + cnts[p], &nc[nr[p]]);
            P->rowptr.swap(nr);
            P->cols.swap(nc);
        }
        P->blocks.assign(P->cols.size() * bs * bs, 0.0);
        P->assemble();

        // stored scalar entries: full blocks between free nodes, diagonal only for Dirichlet nodes
        int64_t nnz = 0;
        for (int64_t p = 0; p < P->n_nodes; ++p) {
            if (P->dirichlet[p]) {
                nnz += bs;
                continue;
            }
            for (int64_t s = P->rowptr[p]; s < P->rowptr[p + 1]; ++s)
                if (!P->dirichlet[P->cols[s]]) nnz += bs * bs;
        }
        if (nnz >= (int64_t(1) << 31)) throw std::runtime_error("nnz exceeds int32 range");
        P->nnz = nnz;
        return P;
    } catch (const std::exception &ex) {
        if (err && errcap > 0) std::snprintf(err, errcap, "%s", ex.what());
        return nullptr;
    }
}

void lssgen_destroy(lssgen_problem *p) { delete p; }
int64_t lssgen_n(const lssgen_problem *p) { return p->n_nodes * p->bs; }
int64_t lssgen_nnz(const lssgen_problem *p) { return p->nnz; }
int64_t lssgen_n_elements(const lssgen_problem *p) { return p->n_elem; }
int lssgen_block_size(const lssgen_problem *p) { return p->bs; }
int64_t lssgen_n_contacts(const lssgen_problem *p) { return int64_t(p->contacts.size()); }
int64_t lssgen_n_slivers(const lssgen_problem *p) { return int64_t(p->slivers.size()); }

int lssgen_sliver_geometry(const lssgen_problem *p, int64_t *elem, double *X, double *X0)
{
    for (size_t q = 0; q < p->slivers.size(); ++q) {
        const Sliver &sl = p->slivers[q];
        elem[q] = sl.elem;
        double orig[4][3];
        p->element_vertices(sl.elem, orig);
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 3; ++d) {
                X[q * 12 + a * 3 + d] = sl.X[a][d];
                X0[q * 12 + a * 3 + d] = orig[a][d];
            }
    }
    return 0;
}

int lssgen_fill_csc(const lssgen_problem *P, int32_t *colptr, int32_t *rowidx, double *vals)
{
    const int bs = P->bs, bb = bs * bs;
    const int64_t n = P->n_nodes * bs;
    // column c of the CSC is scalar row c of the (exactly symmetric) assembled matrix
    std::vector<int64_t> cnt(n, 0);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < P->n_nodes; ++p) {
        int64_t c = 0;
        if (P->dirichlet[p])
            c = 1;
        else
            for (int64_t s = P->rowptr[p]; s < P->rowptr[p + 1]; ++s)
                if (!P->dirichlet[P->cols[s]]) c += bs;
        for (int r = 0; r < bs; ++r) cnt[p * bs + r] = c;
    }
    colptr[0] = 0;
    for (int64_t c = 0; c < n; ++c) colptr[c + 1] = static_cast<int32_t>(colptr[c] + cnt[c]);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < P->n_nodes; ++p) {
        for (int r = 0; r < bs; ++r) {
            const int64_t col = p * bs + r;
            int64_t dst = colptr[col];
            if (P->dirichlet[p]) {
                const int64_t s = P->slot(p, static_cast<int32_t>(p));
                rowidx[dst] = static_cast<int32_t>(col);
                vals[dst] = P->blocks[s * bb + r * bs + r];
                continue;
            }
            for (int64_t s = P->rowptr[p]; s < P->rowptr[p + 1]; ++s) {
                const int32_t q = P->cols[s];
                if (P->dirichlet[q]) continue;
                for (int c = 0; c < bs; ++c) {
                    rowidx[dst] = static_cast<int32_t>(int64_t(q) * bs + c);
                    vals[dst] = P->blocks[s * bb + r * bs + c];
                    ++dst;
                }
            }
        }
    }
    return 0;
}

void lssgen_random_vector(int64_t n, uint32_t seed, double lo, double hi, double *out)
{
    std::mt19937 g(seed);
    std::uniform_real_distribution<double> d(lo, hi);
    for (int64_t i = 0; i < n; ++i) out[i] = d(g);
}

int lssgen_element_coef(const lssgen_problem *p, double *out)
{
    std::copy(p->coef.begin(), p->coef.end(), out);
    return 0;
}

int lssgen_element_nodes(const lssgen_problem *P, int32_t *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < P->n_elem; ++e) {
        const auto cs = tet_corners(static_cast<int>(e % 6));
        for (int a = 0; a < 4; ++a) out[4 * e + a] = P->corner_node(e / 6, cs[a]);
    }
    return 0;
}

int lssgen_grad_coef(const lssgen_problem *P, const double *lam, const double *x, double *out)
{
    const int bs = P->bs, m = 4 * bs;
#pragma omp parallel
    {
        std::vector<double> K(m * m);
#pragma omp for schedule(static)
        for (int64_t e = 0; e < P->n_elem; ++e) {
            P->unit_element(e, K.data());
            const auto cs = tet_corners(static_cast<int>(e % 6));
            int nodes[4];
            for (int a = 0; a < 4; ++a) nodes[a] = P->corner_node(e / 6, cs[a]);
            double s = 0.0;
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) {
                    const bool masked = P->dirichlet[nodes[a]] || P->dirichlet[nodes[b]];
                    if (masked && nodes[a] != nodes[b]) continue;
                    for (int i = 0; i < bs; ++i)
                        for (int j = 0; j < bs; ++j) {
                            if (masked && i != j) continue;
                            s += lam[int64_t(nodes[a]) * bs + i] * K[(a * bs + i) * m + (b * bs + j)] *
                                 x[int64_t(nodes[b]) * bs + j];
                        }
                }
            out[e] = -s;
        }
    }
    return 0;
}

extern "C++" {
namespace {
// Visit the unmasked entries (row, col, value) of element e's unit matrix in a
// fixed order (a, i, b, j); masking as in lssgen_grad_coef.
template <class F>
void for_element_entries(const lssgen_problem *P, int64_t e, const double *K, F &&f)
{
    const int bs = P->bs, m = 4 * bs;
    const auto cs = tet_corners(static_cast<int>(e % 6));
    int nodes[4];
    for (int a = 0; a < 4; ++a) nodes[a] = P->corner_node(e / 6, cs[a]);
    for (int a = 0; a < 4; ++a)
        for (int i = 0; i < bs; ++i)
            for (int b = 0; b < 4; ++b) {
                const bool masked = P->dirichlet[nodes[a]] || P->dirichlet[nodes[b]];
                if (masked && nodes[a] != nodes[b]) continue;
                for (int j = 0; j < bs; ++j) {
                    if (masked && i != j) continue;
                    f(int64_t(nodes[a]) * bs + i, int64_t(nodes[b]) * bs + j, K[(a * bs + i) * m + (b * bs + j)]);
                }
            }
}
}  // namespace
}  // extern "C++"

int64_t lssgen_element_jacobian_nnz(const lssgen_problem *P)
{
    const int m = 4 * P->bs;
    int64_t total = 0;
#pragma omp parallel
    {
        std::vector<double> K(m * m, 0.0);
#pragma omp for schedule(static) reduction(+ : total)
        for (int64_t e = 0; e < P->n_elem; ++e)
            for_element_entries(P, e, K.data(), [&](int64_t, int64_t, double) { ++total; });
    }
    return total;
}

int lssgen_element_jacobian(const lssgen_problem *P, int64_t *seg_ptr, int32_t *slot, double *val)
{
    const int m = 4 * P->bs;
    const int64_t n = P->n_nodes * P->bs;
    std::vector<int32_t> colptr(n + 1), rowidx(P->nnz);
    std::vector<double> tmp(P->nnz);
    lssgen_fill_csc(P, colptr.data(), rowidx.data(), tmp.data());
    seg_ptr[0] = 0;
    {
        std::vector<double> K(m * m, 0.0);
        for (int64_t e = 0; e < P->n_elem; ++e) {
            int64_t c = 0;
            for_element_entries(P, e, K.data(), [&](int64_t, int64_t, double) { ++c; });
            seg_ptr[e + 1] = seg_ptr[e] + c;
        }
    }
    int bad = 0;
#pragma omp parallel
    {
        std::vector<double> K(m * m);
#pragma omp for schedule(static) reduction(| : bad)
        for (int64_t e = 0; e < P->n_elem; ++e) {
            P->unit_element(e, K.data());
            int64_t w = seg_ptr[e];
            // the CSC column of operator row r holds that row's entries
            for_element_entries(P, e, K.data(), [&](int64_t r, int64_t c, double v) {
                const int32_t *b = rowidx.data() + colptr[r], *en = rowidx.data() + colptr[r + 1];
                const int32_t *it = std::lower_bound(b, en, static_cast<int32_t>(c));
                if (it == en || *it != c) bad = 1;
                slot[w] = static_cast<int32_t>(it - rowidx.data());
                val[w] = v;
                ++w;
            });
        }
    }
    return bad ? -1 : 0;
}

int lssgen_fill_vals_with_coef(lssgen_problem *P, const double *coef, double *vals)
{
    std::copy(coef, coef + P->n_elem, P->coef.begin());
    P->assemble();
    std::vector<int32_t> colptr(P->n_nodes * P->bs + 1);
    std::vector<int32_t> rowidx(P->nnz);
    return lssgen_fill_csc(P, colptr.data(), rowidx.data(), vals);
}

}  // extern "C"
