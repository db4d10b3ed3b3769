// Matrix Market reading and writing (see market.hpp).
#include "market.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace lss::cli {
namespace {

struct Entry {
    int32_t r, c;
    double v;
};

[[noreturn]] void bad(const std::string &path, int64_t line, const std::string &what)
{
    throw std::runtime_error(path + ":" + std::to_string(line) + ": " + what);
}

bool is_comment_or_blank(const std::string &s)
{
    const size_t p = s.find_first_not_of(" \t\r\n");
    return p == std::string::npos || s[p] == '%';
}

// entries -> CSC: sort by (column, row), sum duplicates
Csc build(int32_t rows, int32_t cols, std::vector<Entry> &e)
{
    std::sort(e.begin(), e.end(), [](const Entry &a, const Entry &b) { return a.c != b.c ? a.c < b.c : a.r < b.r; });
    Csc A;
    A.rows = rows;
    A.cols = cols;
    A.colptr.assign(size_t(cols) + 1, 0);
    A.rowidx.reserve(e.size());
    A.vals.reserve(e.size());
    for (size_t k = 0; k < e.size(); ++k) {
        if (k > 0 && e[k].r == e[k - 1].r && e[k].c == e[k - 1].c) {
            A.vals.back() += e[k].v;
            continue;
        }
        A.rowidx.push_back(e[k].r);
        A.vals.push_back(e[k].v);
        A.colptr[size_t(e[k].c) + 1]++;
    }
    if (A.rowidx.size() > size_t(INT32_MAX)) throw std::runtime_error("matrix has more than 2^31 - 1 entries");
    for (int32_t j = 0; j < cols; ++j) A.colptr[size_t(j) + 1] += A.colptr[j];
    return A;
}

}  // namespace

Csc read_matrix(const std::string &path)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::string line;
    int64_t ln = 0;
    bool banner = false;
    while (std::getline(in, line)) {
        ++ln;
        if (line.rfind("%%MatrixMarket", 0) == 0) {
            banner = true;
            std::string s = line;
            std::transform(s.begin(), s.end(), s.begin(), ::tolower);
            if (s.find("coordinate") == std::string::npos) bad(path, ln, "expected the coordinate format");
            if (s.find("complex") != std::string::npos || s.find("pattern") != std::string::npos)
                bad(path, ln, "only real or integer values are supported");
            continue;
        }
        if (!is_comment_or_blank(line)) break;
    }
    (void)banner;
    int64_t M = -1, N = -1, NNZ = -1;
    {
        std::istringstream ss(line);
        if (!(ss >> M >> N >> NNZ) || M <= 0 || N <= 0 || NNZ < 0) bad(path, ln, "bad size line");
        if (M > INT32_MAX || N > INT32_MAX) bad(path, ln, "dimension above 2^31 - 1");
    }
    std::vector<Entry> e;
    e.reserve(size_t(NNZ));
    while (std::getline(in, line)) {
        ++ln;
        if (is_comment_or_blank(line)) continue;
        const char *p = line.c_str();
        char *end = nullptr;
        const long long i = std::strtoll(p, &end, 10);
        if (end == p) bad(path, ln, "bad entry");
        p = end;
        const long long j = std::strtoll(p, &end, 10);
        if (end == p) bad(path, ln, "bad entry");
        p = end;
        const double v = std::strtod(p, &end);
        if (end == p) bad(path, ln, "missing value");
        if (i < 1 || j < 1 || i > M || j > N) bad(path, ln, "index out of range");
        e.push_back({int32_t(i - 1), int32_t(j - 1), v});
    }
    if (int64_t(e.size()) != NNZ)
        throw std::runtime_error(path + ": " + std::to_string(e.size()) + " entries read, the size line says " +
                                 std::to_string(NNZ));
    return build(int32_t(M), int32_t(N), e);
}

std::vector<double> read_vector(const std::string &path)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::string line;
    int64_t ln = 0;
    while (std::getline(in, line)) {
        ++ln;
        if (!is_comment_or_blank(line)) break;
    }
    int64_t n = 0, c = 0;
    {
        std::istringstream ss(line);
        if (!(ss >> n >> c) || n <= 0 || c <= 0) bad(path, ln, "bad size line");
        int64_t extra;
        if (ss >> extra) bad(path, ln, "expected the array format, found a coordinate size line");
    }
    std::vector<double> v;
    v.reserve(size_t(n));
    while (int64_t(v.size()) < n && std::getline(in, line)) {
        ++ln;
        if (is_comment_or_blank(line)) continue;
        const char *p = line.c_str();
        char *end = nullptr;
        const double x = std::strtod(p, &end);
        if (end == p) bad(path, ln, "bad value");
        v.push_back(x);
    }
    if (int64_t(v.size()) != n) throw std::runtime_error(path + ": unable to read all " + std::to_string(n) + " values");
    return v;
}

std::vector<double> read_dense(const std::string &path, int64_t &rows, int64_t &cols)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::string line;
    int64_t ln = 0;
    while (std::getline(in, line)) {
        ++ln;
        if (!is_comment_or_blank(line)) break;
    }
    {
        std::istringstream ss(line);
        if (!(ss >> rows >> cols) || rows < 0 || cols < 0) bad(path, ln, "bad size line");
        int64_t extra;
        if (ss >> extra) bad(path, ln, "expected the array format, found a coordinate size line");
    }
    std::vector<double> v;
    v.reserve(size_t(rows * cols));
    double x;
    while (int64_t(v.size()) < rows * cols && (in >> x)) v.push_back(x);
    if (int64_t(v.size()) != rows * cols) throw std::runtime_error(path + ": unexpected end of data");
    return v;
}

void write_vector(const std::string &path, const std::vector<double> &v)
{
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + path);
    std::fprintf(f, "%%%%MatrixMarket matrix array real general\n%zu 1\n", v.size());
    for (double x : v) std::fprintf(f, "%.17g\n", x);
    std::fclose(f);
}

Csc force_symmetry(const Csc &A)
{
    std::vector<Entry> e;
    e.reserve(size_t(A.nnz()) * 2);
    for (int32_t j = 0; j < A.cols; ++j)
        for (int32_t k = A.colptr[j]; k < A.colptr[size_t(j) + 1]; ++k) {
            const int32_t i = A.rowidx[k];
            e.push_back({i, j, A.vals[k]});
            if (i != j) e.push_back({j, i, A.vals[k]});
        }
    return build(A.rows, A.cols, e);
}

std::vector<double> multiply(const Csc &A, const std::vector<double> &x)
{
    std::vector<double> y(size_t(A.rows), 0.0);
    for (int32_t j = 0; j < A.cols; ++j)
        for (int32_t k = A.colptr[j]; k < A.colptr[size_t(j) + 1]; ++k) y[A.rowidx[k]] += A.vals[k] * x[j];
    return y;
}

}  // namespace lss::cli
