// linear_solve: the reference's Matrix Market harness (tests/linear_solve.cpp)
// over the C ABI of this library.
//
//   linear_solve -A matrix.mtx [-b rhs.mtx | --rand [seed]] [-j spec.json]
//                [-i info.mtx [-t threshold] [--less_than]] [-p positions.mtx]
//                [-e elements.mtx] [-w warmup] [-r repeat] [--force_symmetry]
//                [-o x.mtx] [--info] [--check]
//
// The options of the reference keep their meaning (tests/linear_solve.cpp:184-
// 241): the right-hand side defaults to uniform(-1, 1) values from mt19937 with
// seed 0 (112-132), -i / -t / --less_than turn a per-DOF indicator into the
// caller's stiff set (396-429), every repetition creates a new solver and runs
// factorize + solve from x = 0, and warm-up repetitions print nothing
// (439-472).  -p / -e are read and checked but, as for GPUHybrid, not used.
// Differences: the default configuration is {"solver": "GPUHybrid"} (this
// library implements that solver only); -o writes the last solution, --info
// prints the solver's info JSON after each timed repetition, --check loads and
// summarizes the inputs without touching the GPU.
#include <sys/resource.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <optional>
#include <random>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "lss.h"
#include "market.hpp"

namespace {

using lss::cli::Csc;

struct Args {
    std::string A, b, j, i, p, e, o;
    bool rand = false;
    std::optional<long long> seed;
    double threshold = 0.0;
    bool less_than = false, force_symmetry = false, info = false, check = false;
    int warmup = 1, repeat = 1;
};

[[noreturn]] void usage(const char *msg)
{
    std::fprintf(stderr,
                 "error: %s\nusage: linear_solve -A matrix.mtx [-b rhs.mtx | --rand [seed]] [-j spec.json]\n"
                 "                    [-i info.mtx [-t threshold] [--less_than]] [-p positions.mtx] [-e elements.mtx]\n"
                 "                    [-w warmup] [-r repeat] [--force_symmetry] [-o x.mtx] [--info] [--check]\n",
                 msg);
    std::exit(2);
}

Args parse(int argc, char **argv)
{
    Args a;
    auto need = [&](int &k) -> std::string {
        if (k + 1 >= argc) usage((std::string("missing value for ") + argv[k]).c_str());
        return argv[++k];
    };
    for (int k = 1; k < argc; ++k) {
        const std::string s = argv[k];
        if (s == "-A") a.A = need(k);
        else if (s == "-b") a.b = need(k);
        else if (s == "--rand") {
            a.rand = true;
            if (k + 1 < argc && argv[k + 1][0] != '-') a.seed = std::stoll(argv[++k]);
        } else if (s == "-j") a.j = need(k);
        else if (s == "-i") a.i = need(k);
        else if (s == "-t" || s == "--threshold") a.threshold = std::stod(need(k));
        else if (s == "--less_than") a.less_than = true;
        else if (s == "-p") a.p = need(k);
        else if (s == "-e") a.e = need(k);
        else if (s == "-w") a.warmup = std::stoi(need(k));
        else if (s == "-r") a.repeat = std::stoi(need(k));
        else if (s == "--force_symmetry") a.force_symmetry = true;
        else if (s == "-o") a.o = need(k);
        else if (s == "--info") a.info = true;
        else if (s == "--check") a.check = true;
        else usage(("unknown option " + s).c_str());
    }
    if (a.A.empty()) usage("-A is required");
    if (!a.b.empty() && a.rand) usage("-b and --rand are mutually exclusive");
    if (a.warmup < 0) throw std::runtime_error("warmup count must be non-negative");
    if (a.repeat < 1) throw std::runtime_error("repeat count must be at least 1");
    if (a.seed && (*a.seed < 0 || *a.seed > 0xffffffffLL))
        throw std::runtime_error("random seed must be in [0, 2^32)");
    return a;
}

std::vector<double> random_rhs(int64_t n, std::optional<uint32_t> seed)
{
    std::mt19937 gen;
    if (seed) gen.seed(*seed);
    else gen.seed(std::random_device{}());
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    std::vector<double> b(static_cast<size_t>(n));
    for (auto &v : b) v = dist(gen);
    return b;
}

double norm(const std::vector<double> &v)
{
    double s = 0.0;
    for (double x : v) s += x * x;
    return std::sqrt(s);
}

std::string read_text(const std::string &path)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("failed to open json config: " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

size_t peak_rss()
{
    struct rusage u;
    getrusage(RUSAGE_SELF, &u);
    return size_t(u.ru_maxrss) * 1024;
}

void check_rc(int rc, lss_solver *s, const char *what)
{
    if (rc != LSS_OK) throw std::runtime_error(std::string(what) + ": " + lss_last_error(s));
}

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int run(const Args &a)
{
    const std::string config = a.j.empty() ? std::string("{\"solver\": \"GPUHybrid\"}") : read_text(a.j);
    Csc A = lss::cli::read_matrix(a.A);
    if (A.rows != A.cols) throw std::runtime_error("matrix must be square");
    if (a.force_symmetry) A = lss::cli::force_symmetry(A);
    const int64_t n = A.rows;
    std::vector<double> b;
    if (!a.b.empty()) {
        b = lss::cli::read_vector(a.b);
        if (int64_t(b.size()) != n) throw std::runtime_error("rhs dimension mismatch");
    } else {
        std::optional<uint32_t> seed = uint32_t(0);  // the reference's default seed
        if (a.rand) seed = a.seed ? std::optional<uint32_t>(uint32_t(*a.seed)) : std::nullopt;
        b = random_rhs(n, seed);
    }
    int64_t pr = 0, pc = 0, er = 0, ec = 0;
    if (!a.p.empty()) lss::cli::read_dense(a.p, pr, pc);
    if (!a.e.empty()) lss::cli::read_dense(a.e, er, ec);
    std::vector<int32_t> stiff;
    if (!a.i.empty()) {
        const std::vector<double> info = lss::cli::read_vector(a.i);
        if (int64_t(info.size()) != n) throw std::runtime_error("problematic info dimension mismatch");
        for (int64_t k = 0; k < n; ++k)
            if (a.less_than ? info[k] < a.threshold : info[k] > a.threshold) stiff.push_back(int32_t(k));
    }
    if (a.check) {
        double sum = 0.0;
        for (double v : A.vals) sum += v;
        std::printf("{\"n\": %lld, \"nnz\": %lld, \"value_sum\": %.17g, \"rhs_norm\": %.17g, \"stiff_dofs\": %zu, "
                    "\"positions\": [%lld, %lld], \"elements\": [%lld, %lld]}\n",
                    (long long)n, (long long)A.nnz(), sum, norm(b), stiff.size(), (long long)pr, (long long)pc,
                    (long long)er, (long long)ec);
        return 0;
    }
    // validate the configuration before the repetitions (tests/linear_solve.cpp:292-296)
    {
        lss_solver *s = nullptr;
        const int rc = lss_create(config.c_str(), &s);
        if (rc != LSS_OK) {
            const std::string msg = s ? lss_last_error(s) : lss_last_error(nullptr);
            lss_destroy(s);
            throw std::runtime_error("invalid solver configuration: " + msg);
        }
        lss_destroy(s);
    }
    std::vector<double> x(size_t(n), 0.0);
    for (int it = 0; it < a.warmup + a.repeat; ++it) {
        const bool quiet = it < a.warmup;
        lss_solver *s = nullptr;
        check_rc(lss_create(config.c_str(), &s), s, "create");
        try {
            check_rc(lss_set_problematic_dofs(s, stiff.data(), int64_t(stiff.size())), s, "set_problematic_dofs");
            const char *name = lss_name(s);
            if (!quiet) std::printf("[%s] [matrix_info] [size=%lld] [nnzs=%lld]\n", name, (long long)n, (long long)A.nnz());
            std::fill(x.begin(), x.end(), 0.0);
            auto t0 = std::chrono::steady_clock::now();
            check_rc(lss_factorize_csc(s, int32_t(n), A.nnz(), A.colptr.data(), A.rowidx.data(), A.vals.data()), s,
                     "factorize");
            const double tf = seconds_since(t0);
            t0 = std::chrono::steady_clock::now();
            check_rc(lss_solve(s, b.data(), x.data()), s, "solve");
            const double ts = seconds_since(t0);
            const std::vector<double> Ax = lss::cli::multiply(A, x);
            std::vector<double> r(static_cast<size_t>(n));
            for (int64_t k = 0; k < n; ++k) r[k] = b[k] - Ax[k];
            if (!quiet) {
                size_t need = 0;
                lss_get_info(s, nullptr, 0, &need);
                std::string info(need + 1, '\0');
                lss_get_info(s, info.data(), info.size(), &need);
                info.resize(std::strlen(info.c_str()));
                std::printf("[%s] [factorize] [%.6f]\n[%s] [solve] [%.6f]\n", name, tf, name, ts);
                std::printf("[%s] [residual=%.17g] [relative_residual=%.6g] [peak_memory=%zu]\n", name, norm(r),
                            norm(r) / std::max(norm(b), 1e-300), peak_rss());
                if (a.info) std::printf("%s\n", info.c_str());
                std::fflush(stdout);
            }
        } catch (...) {
            lss_destroy(s);
            throw;
        }
        lss_destroy(s);
    }
    if (!a.o.empty()) lss::cli::write_vector(a.o, x);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    try {
        return run(parse(argc, argv));
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
