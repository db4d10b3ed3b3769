"""Error paths of the C ABI, the full preconditioner against the oracle, and the
reported true residual against an independent evaluation."""
import ctypes

import numpy as np
import pytest

from lss_amd import _lib, problems
from lss_amd.solver import Solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LSS_ERR_SIZE, LSS_ERR_NOT_FACTORIZED, LSS_ERR_ARG = 2, 3, 8


def _cfg(bs, **kw):
    g = {"block_size": bs, "relative_tolerance": 1e-8, "gmm_jump_threshold": 10.0, "min_subdomain_size": 4}
    g.update(kw)
    return {"solver": "GPUHybrid", "GPUHybrid": g}


def _ocfg(bs, **kw):
    base = dict(block_size=bs, rel_tol=1e-8, gmm_jump_threshold=10.0, min_subdomain_size=4)
    base.update(kw)
    return O.config(**base)


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def test_solve_before_factorize_status(gpu):
    # through the C ABI itself (the Python wrapper raises before calling it)
    L = _lib.load()
    s = Solver.create(_cfg(3))
    b = np.ones(30)
    x = np.zeros(30)
    assert L.lss_solve(s._h, _p(b), _p(x)) == LSS_ERR_NOT_FACTORIZED
    assert "factorize" in L.lss_last_error(s._h).decode()
    assert L.lss_debug_precond_apply(s._h, _p(b), _p(x)) == LSS_ERR_NOT_FACTORIZED


def test_size_mismatch_status(gpu):
    L = _lib.load()
    A = problems.make("laplace_sphere", N=6).csc()
    s = Solver.create(_cfg(1))
    # colptr[n] disagrees with nnz
    rc = L.lss_factorize_csc(s._h, A.n, A.nnz + 1, _p(A.colptr), _p(A.rowidx), _p(A.vals))
    assert rc == LSS_ERR_SIZE
    s.factorize(A)
    out = np.empty(A.n + 5)
    assert L.lss_debug_row_norms(s._h, _p(out), A.n + 5) == LSS_ERR_SIZE
    perm = np.empty(A.n + 1, np.int32)
    assert L.lss_debug_partition(s._h, _p(perm), A.n + 1, None, 0, None) == LSS_ERR_SIZE
    # the handle stays usable after the errors
    x = s.solve(np.ones(A.n))
    assert s.get_info()["stop_reason"] == "converged_rel"
    assert np.all(np.isfinite(x))


@pytest.mark.parametrize("change", [{"precision": "fp64"}, {"base_preconditioner": "Jacobi"},
                                    {"max_coarse_levels": 1}, {"do_hybrid_precond": False}])
def test_set_parameters_after_factorize_drops_the_factorization(gpu, change):
    # the device buffers were built for the old keys (FP32 bank layout, MAS levels,
    # stiff blocks): a solve must report LSS_ERR_NOT_FACTORIZED, not read them
    L = _lib.load()
    prob = problems.make("elastic_contact", N=6)
    A = prob.csc()
    s = Solver.create(_cfg(3, precision="fp32"))
    s.factorize(A)
    s.set_parameters({"GPUHybrid": change})
    b = prob.rhs(0)
    x = np.zeros(A.n)
    assert L.lss_solve(s._h, _p(b), _p(x)) == LSS_ERR_NOT_FACTORIZED
    # refactorizing with the new keys works
    s.factorize(A)
    s.solve(b)
    assert s.get_info()["stop_reason"] == "converged_rel"


def test_set_parameters_solve_keys_keep_the_factorization(gpu):
    prob = problems.make("elastic_contact", N=6)
    A = prob.csc()
    s = Solver.create(_cfg(3))
    s.factorize(A)
    b = prob.rhs(0)
    s.solve(b)
    it8 = s.get_info()["num_iterations"]
    s.set_parameters({"GPUHybrid": {"relative_tolerance": 1e-4, "max_iter": 500, "check_every": 2}})
    s.solve(b)
    info = s.get_info()
    assert info["stop_reason"] == "converged_rel" and info["num_iterations"] < it8


@pytest.mark.parametrize("name,N,precision", [("elastic_contact", 10, "fp64"), ("laplace_sphere", 12, "fp64"),
                                              ("elastic_logE", 8, "fp64"), ("elastic_contact", 10, "fp32")])
def test_precond_apply_matches_oracle(gpu, name, N, precision):
    # M^{-1} r = the B-P-B sweep (GPUHybrid.cu:521-542) with the device's partition,
    # stiff set and damping in the oracle
    prob = problems.make(name, N=N)
    A = prob.csc()
    bs = prob.block_size
    s = Solver.create(_cfg(bs, precision=precision))
    s.factorize(A)
    info = s.get_info()
    H = O.Hybrid(_ocfg(bs))
    perm, off = s.partition(A.n // bs)
    H.set_partition(perm, off)
    H.factorize(A.colptr, A.rowidx, A.vals)
    H.set_omega(info["omega"])
    assert [c.size for c in s.stiff_components()] == [c.size for c in H.stiff_components()]
    for seed in (1, 2):
        r = problems.random_vector(A.n, seed)
        z = s.precond_apply(r)
        zr = H.precond(r)
        # FP64: bank factors from a different algorithm, contact banks of condition ~1e8;
        # FP32 banks and inner SpMV values: rounding of the factors
        tol = 1e-8 if precision == "fp64" else 2e-4
        assert np.linalg.norm(z - zr) <= tol * np.linalg.norm(zr), (seed, np.linalg.norm(z - zr) / np.linalg.norm(zr))


def test_final_res_norm_is_the_true_residual(gpu):
    # get_info's final_res_norm (GPUHybrid.cu:383-386) against ||b - A x|| evaluated
    # on the host: both are FP64 evaluations, equal to within their rounding
    for name, N in (("elastic_contact", 10), ("laplace_sphere", 14)):
        prob = problems.make(name, N=N)
        A = prob.csc()
        S = A.to_scipy()
        s = Solver.create(_cfg(prob.block_size))
        s.factorize(A)
        b = prob.rhs(3)
        x = s.solve(b)
        rep = s.get_info()["final_res_norm"]
        host = np.linalg.norm(b - S @ x)
        floor = np.finfo(np.float64).eps * np.linalg.norm(abs(S) @ np.abs(x))
        assert abs(rep - host) <= 20 * floor + 1e-12 * host, (name, rep, host, floor)
        # and a solve from a poor initial guess stopped early reports its larger residual
        s.set_parameters({"GPUHybrid": {"max_iter": 3}})
        x3 = s.solve(b)
        rep3 = s.get_info()["final_res_norm"]
        assert rep3 == pytest.approx(np.linalg.norm(b - S @ x3), rel=1e-10)
        assert rep3 > 100 * rep
