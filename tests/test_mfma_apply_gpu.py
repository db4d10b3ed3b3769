"""Level-0 bank solve on the matrix cores ("bank_apply": "mfma") against the oracle.

The MFMA solve reads the same FP32 factors as the vector-unit solve; only its
arithmetic differs (x and t as hi/lo FP32 pairs, FP32 accumulation over at most
`mfma_flush` products, then FP64).  So:
  * against the oracle's FP64 MAS it agrees to the FP32 rounding of the factors,
    like the vector FP32 path (an entry in the wrong place would be an O(1) error);
  * against the vector FP32 path (same factors, FP64 accumulation) it agrees to
    FP32 accumulation rounding, far inside the factor rounding;
  * scaling r by 2 scales z by 2 bit for bit (every product and sum scales
    exactly), and the operator stays symmetric to rounding;
  * solves with it converge to the oracle's solution.
"""
import numpy as np
import pytest

from lss_amd import problems
from lss_amd.solver import Solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _cfg(bs, **kw):
    g = {"block_size": bs, "relative_tolerance": 1e-8, "gmm_jump_threshold": 10.0, "min_subdomain_size": 4}
    g.update(kw)
    return {"solver": "GPUHybrid", "GPUHybrid": g}


@pytest.mark.parametrize("flush", [4, 32])
@pytest.mark.parametrize("name,N,bs", [("laplace_sphere", 12, 1), ("laplace_sphere", 12, 2),
                                       ("laplace_sphere", 12, 3), ("elastic_logE", 8, 3)])
def test_mfma_apply_matches_oracle(gpu, name, N, bs, flush):
    prob = problems.make(name, N=N)
    A = prob.csc()
    vec = Solver.create(_cfg(bs, do_hybrid_precond=False, precision="fp32"))
    vec.factorize(A)
    mf = Solver.create(_cfg(bs, do_hybrid_precond=False, precision="fp32", bank_apply="mfma", mfma_flush=flush))
    mf.factorize(A)
    assert mf.get_info()["bank_apply"] == "mfma" and mf.get_info()["mfma_flush"] == flush
    perm, off = mf.partition(A.n // bs)
    mas = O.Mas(A.colptr, A.rowidx, A.vals, bs, perm, off)
    for seed in (3, 4):
        r = problems.random_vector(A.n, seed)
        zr = mas.apply(r)
        z = mf.base_apply(r)
        zv = vec.base_apply(r)
        err_oracle = np.linalg.norm(z - zr) / np.linalg.norm(zr)
        err_vec = np.linalg.norm(z - zv) / np.linalg.norm(zv)
        print(f"{name} N={N} bs={bs} flush={flush} seed={seed}: |z-z_oracle|/|z_oracle| = {err_oracle:.2e}, "
              f"|z-z_vector|/|z_vector| = {err_vec:.2e}")
        assert err_oracle <= 1e-4, seed
        assert err_vec <= 1e-5, seed
        assert np.array_equal(mf.base_apply(2.0 * r), 2.0 * z)


@pytest.mark.parametrize("flush", [4, 32])
def test_mfma_apply_symmetric(gpu, flush):
    prob = problems.make("elastic_logE", N=8)
    A = prob.csc()
    s = Solver.create(_cfg(3, do_hybrid_precond=False, precision="fp32", bank_apply="mfma", mfma_flush=flush))
    s.factorize(A)
    x = problems.random_vector(A.n, 5)
    y = problems.random_vector(A.n, 6)
    xMy, yMx = x @ s.base_apply(y), y @ s.base_apply(x)
    assert abs(xMy - yMx) <= 1e-5 * max(abs(xMy), abs(yMx))
    assert x @ s.base_apply(x) > 0.0


@pytest.mark.parametrize("name,N", [("laplace_sphere", 12), ("elastic_contact", 10), ("elastic_slivers", 8),
                                    ("elastic_logE", 8)])
def test_mfma_solve_matches_oracle(gpu, name, N):
    prob = problems.make(name, N=N)
    A = prob.csc()
    b = prob.rhs(0)
    s = Solver.create(_cfg(prob.block_size, precision="fp32", bank_apply="mfma"))
    s.factorize(A)
    x = s.solve(b)
    info = s.get_info()
    vec = Solver.create(_cfg(prob.block_size, precision="fp32"))
    vec.factorize(A)
    vec.solve(b)
    vit = vec.get_info()["num_iterations"]
    H = O.Hybrid(O.config(block_size=prob.block_size, rel_tol=1e-8, gmm_jump_threshold=10.0, min_subdomain_size=4))
    H.factorize(A.colptr, A.rowidx, A.vals)
    xr = H.solve(b)
    S = A.to_scipy()
    rel = np.linalg.norm(S @ x - b) / np.linalg.norm(b)
    floor = 50 * np.finfo(np.float64).eps * np.linalg.norm(abs(S) @ np.abs(x)) / np.linalg.norm(b)
    print(f"{name} N={N}: mfma {info['num_iterations']} iterations, vector FP32 {vit}, "
          f"oracle {H.info()['num_iterations']}; rel residual {rel:.2e}")
    assert info["stop_reason"] == "converged_rel"
    assert rel <= max(1.01e-8, floor)
    assert np.linalg.norm(x - xr) <= 1e-5 * np.linalg.norm(xr)
    assert info["num_iterations"] <= vit + max(2, 0.05 * vit)
