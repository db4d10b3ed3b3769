// Test shim: runs the C++ AMG restatement (oracle/amg_ref.cpp) on a CSR matrix and
// writes every level (C/F splitting, operator, interpolation) to a binary file for
// the tests (built by tests/amg_ref.py).
#include "amg_ref.hpp"

#include <cstdio>
#include <vector>

namespace {
template <class T> void put(FILE *fp, const std::vector<T> &v)
{
    const int64_t n = int64_t(v.size());
    std::fwrite(&n, sizeof n, 1, fp);
    if (n) std::fwrite(v.data(), sizeof(T), v.size(), fp);
}
}  // namespace

extern "C" int amg_shim_setup(int32_t n, const int64_t *ptr, const int32_t *col, const double *val, int32_t bs,
                              double theta, const char *path)
{
    amgref::CsrHost A;
    A.n = A.m = n;
    A.ptr.assign(ptr, ptr + n + 1);
    A.col.assign(col, col + ptr[n]);
    A.val.assign(val, val + ptr[n]);
    amgref::AmgOptions opt;
    opt.num_functions = bs;
    opt.theta = theta;
    std::vector<amgref::AmgLevelHost> levels;
    amgref::amg_setup_host(A, opt, levels);
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return 1;
    const int64_t nl = int64_t(levels.size());
    std::fwrite(&nl, sizeof nl, 1, fp);
    for (const auto &l : levels) {
        put(fp, l.cf);
        put(fp, l.A.ptr);
        put(fp, l.A.col);
        put(fp, l.A.val);
        put(fp, l.P.ptr);
        put(fp, l.P.col);
        put(fp, l.P.val);
        const int64_t m = l.P.m;
        std::fwrite(&m, sizeof m, 1, fp);
    }
    std::fclose(fp);
    return 0;
}
