/*
 * Synthetic locally-stiff SPD systems on structured tetrahedral grids.
 *
 * The workloads are P1 finite elements on an N^3-node grid spanning the unit
 * cube; every grid cell is split into the 6 Kuhn (Freudenthal) tetrahedra that
 * share the cell's main diagonal, so the mesh is conforming and every interior
 * node couples to 14 neighbours.  Features that create local stiffness:
 *   - a high-conductivity sphere (Laplacian),
 *   - sliver elements (elasticity),
 *   - contact-barrier pair blocks kappa * n n^T (elasticity),
 *   - per-element log-uniform Young's modulus (elasticity).
 * Dirichlet DOFs keep their assembled diagonal and lose every other entry of
 * their row and column, so the matrix stays symmetric positive definite and the
 * DOF count is bs * N^3.
 *
 * The output is an exactly symmetric matrix in compressed sparse column form
 * (0-based int32 indices, ascending row indices per column), the layout of
 * polysolve's StiffnessMatrix (src/polysolve/Types.hpp:162-166).  The random
 * right-hand side follows make_random_rhs (tests/linear_solve.cpp:112-132):
 * std::mt19937 seeded with `seed` feeding
 * std::uniform_real_distribution<double>(-1, 1), one draw per DOF.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { LSSGEN_LAPLACIAN = 0, LSSGEN_ELASTICITY = 1 };

typedef struct lssgen_spec {
    int kind;                 /* LSSGEN_LAPLACIAN (block size 1) or LSSGEN_ELASTICITY (block size 3) */
    int N;                    /* nodes per axis (>= 2) */
    double E;                 /* Young's modulus / base conductivity */
    double nu;                /* Poisson ratio (elasticity) */
    int dirichlet_axis;       /* clamp the face {coord[axis] == 0}; -1 = none */
    double sphere_radius;     /* > 0: elements whose centroid is inside get sphere_coef */
    double sphere_coef;
    double sliver_frac;       /* fraction of elements collapsed into slivers */
    double sliver_eps;        /* moved vertex ends at sliver_eps * h from the opposite face */
    double contact_frac;      /* fraction of top-face (z = 1) nodes that get a contact pair */
    double contact_log10_kmin;
    double contact_log10_kmax;
    double contact_radius_h;  /* partner chosen uniformly among nodes within this many h */
    double E_log10_min;       /* E_log10_max > E_log10_min: per-element E = 10^U[min, max] */
    double E_log10_max;
    uint32_t seed;
} lssgen_spec;

typedef struct lssgen_problem lssgen_problem;

/* Assemble the problem; returns NULL and writes a message into err on failure. */
lssgen_problem *lssgen_create(const lssgen_spec *spec, char *err, int errcap);
void lssgen_destroy(lssgen_problem *p);

int64_t lssgen_n(const lssgen_problem *p);          /* scalar DOFs */
int64_t lssgen_nnz(const lssgen_problem *p);        /* stored scalar entries */
int64_t lssgen_n_elements(const lssgen_problem *p);
int lssgen_block_size(const lssgen_problem *p);
int64_t lssgen_n_contacts(const lssgen_problem *p);
int64_t lssgen_n_slivers(const lssgen_problem *p);
/* the sliver elements (ascending), their vertices after and before the collapse:
   X, X0 of n_slivers * 4 * 3 doubles (vertex-major) */
int lssgen_sliver_geometry(const lssgen_problem *p, int64_t *elem, double *X, double *X0);

/* Fill CSC arrays: colptr[n+1], rowidx[nnz], vals[nnz].  Returns 0 on success. */
int lssgen_fill_csc(const lssgen_problem *p, int32_t *colptr, int32_t *rowidx, double *vals);

/* b_i ~ U[-1, 1] from std::mt19937(seed), i = 0 .. n-1. */
void lssgen_random_vector(int64_t n, uint32_t seed, double lo, double hi, double *out);

/* Per-element coefficient (conductivity or Young's modulus). */
int lssgen_element_coef(const lssgen_problem *p, double *out);

/* Element connectivity: out[4*e + a] = node of local vertex a. */
int lssgen_element_nodes(const lssgen_problem *p, int32_t *out);

/*
 * dL/dcoef_e for L(x(coef)) with A(coef) x = b and adjoint lam (A^T lam = dL/dx):
 * dL/dcoef_e = -lam^T (dA/dcoef_e) x, where dA/dcoef_e is the element matrix per
 * unit coefficient with the Dirichlet masking applied.  lam, x have n entries.
 */
int lssgen_grad_coef(const lssgen_problem *p, const double *lam, const double *x, double *out);

/*
 * Sparse Jacobian d vals / d coef in element-major form: element e owns
 * entries [seg_ptr[e], seg_ptr[e+1]) of (slot into the CSC arrays of
 * lssgen_fill_csc, value of its unit-coefficient element matrix), with the
 * Dirichlet masking applied.  vals = J coef + (contact terms).
 * lssgen_element_jacobian_nnz gives the entry count; returns 0 on success.
 */
int64_t lssgen_element_jacobian_nnz(const lssgen_problem *p);
int lssgen_element_jacobian(const lssgen_problem *p, int64_t *seg_ptr, int32_t *slot, double *val);

/*
 * Rebuild the matrix values with per-element coefficients `coef` (same pattern),
 * then refill vals[nnz] in the CSC order of lssgen_fill_csc.
 */
int lssgen_fill_vals_with_coef(lssgen_problem *p, const double *coef, double *vals);

#ifdef __cplusplus
}
#endif
