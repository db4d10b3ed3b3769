// ORACLE — test infrastructure only (built by the tests, never linked into the
// library).  Sequential C++ restatement of the AMG setup of the GPU AMG base
// preconditioner, bit-identical to oracle/amg.py (same operation order, no
// contraction of products into sums); the device setup is compared with it at
// sizes the numpy restatement would take minutes on.
//
// The reference's GPUHybrid preconditions with hypre's BoomerAMG (v3.1.0,
// cmake/recipes/hypre.cmake; options SRC/GPUHybrid.cu:285-339): PMIS coarsening
// (type 8), strength threshold theta (0.25; `theta` for block size > 1, with the
// "unknown" approach of num_functions = block size), extended+i interpolation
// (type 6) truncated to Pmax = 4 entries per row, Galerkin coarse operators
// R A P with R = P^T, l1-Jacobi smoothing (type 18, one sweep), at most 25
// levels down to 5 points, one V-cycle per application (SRC/GPUHybrid.cu:677-689).
// hypre's sources are not part of the reference; this is the published method:
//  * strength: hypre's classical measure (row scale = most negative same-function
//    off-diagonal, strong if a_ij < theta * scale; all weak when
//    |row sum| > max_row_sum |a_ii|, max_row_sum 0.9),
//  * PMIS: De Sterck, Yang, Heys, SIAM J. Matrix Anal. Appl. 27 (2006) 1019,
//  * extended+i: De Sterck, Falgout, Nolting, Yang, Numer. Linear Algebra Appl.
//    15 (2008) 115 (same-function weak entries lumped into the diagonal),
//  * truncation: the Pmax largest |weights| (ties: the smaller point), scaled by
//    one factor that keeps the row sum,
//  * l1-Jacobi: D_ii = sum_j |a_ij| (Baker, Falgout, Kolev, Yang 2011).
// The random part of PMIS's measure is a hash of the point index (hypre draws it
// from a generator), so the splitting is reproducible.
#pragma once
#include <cstdint>
#include <vector>

namespace amgref {

// Scalar CSR with ascending columns per row.
struct CsrHost {
    int32_t n = 0, m = 0;  // rows, columns
    std::vector<int64_t> ptr;
    std::vector<int32_t> col;
    std::vector<double> val;
};

struct AmgOptions {
    double theta = 0.25;
    double max_row_sum = 0.9;
    int pmax = 4;
    int max_levels = 25;
    int min_coarse = 5;
    int num_functions = 1;
};

struct AmgLevelHost {
    CsrHost A;                 // operator of this level
    CsrHost P, R;              // interpolation to this level from the next (empty on the coarsest), R = P^T
    std::vector<double> dinv;  // 1 / l1 row norms
    std::vector<int8_t> cf;    // 1 C point, -1 F point, -3 F point without strong connections
    std::vector<int32_t> func; // function (DOF component) of each point
};

// Random part of the PMIS measure of point i, in [0, 1).
double amg_point_random(int64_t i);

// Seconds spent in each step of the setup, summed over the levels.
struct AmgSetupTimes {
    double strength = 0, pmis = 0, interp = 0, transpose = 0, rap = 0, norms = 0;
};

// A0 is moved into levels[0].A.
void amg_setup_host(CsrHost A0, const AmgOptions &opt, std::vector<AmgLevelHost> &levels,
                    AmgSetupTimes *times = nullptr);

// Pieces, exposed for the tests of the setup.
void amg_strength(const CsrHost &A, const std::vector<int32_t> &func, int num_functions, double theta,
                  double max_row_sum, CsrHost &S);
void amg_pmis(const CsrHost &S, std::vector<int8_t> &cf);
void amg_interp_extpi(const CsrHost &A, const CsrHost &S, const std::vector<int8_t> &cf,
                      const std::vector<int32_t> &func, int num_functions, int pmax, CsrHost &P);
void amg_transpose(const CsrHost &P, CsrHost &R);
void amg_spgemm(const CsrHost &X, const CsrHost &Y, CsrHost &Z);

}  // namespace amgref
