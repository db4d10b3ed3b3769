"""Fused multi-right-hand-side solve (lss_solve_multi_device) against single solves.

Every column of a fused solve runs the arithmetic of a single-right-hand-side
solve (same kernels with the column count as a template argument, per-column
partials, finishers and stop words), so the results must be bit-identical to
solving the columns one at a time, including the iteration counts, columns that
stop earlier than others and zero columns.
"""
import numpy as np
import pytest

from lss_amd import problems
from lss_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _cfg(bs, **kw):
    g = {"block_size": bs, "relative_tolerance": 1e-8, "gmm_jump_threshold": 10.0, "min_subdomain_size": 4}
    g.update(kw)
    return {"solver": "GPUHybrid", "GPUHybrid": g}


def _rhs(n, k, seed):
    # columns of different character converge after different iteration counts
    B = np.stack([problems.random_vector(n, seed + c) for c in range(k)])
    if k > 2:
        B[2] = B[2] * 1e-3 + 1.0
    return B


def _fused(s, B, X0):
    import torch

    dB = torch.from_numpy(np.ascontiguousarray(B)).cuda()
    dX = torch.from_numpy(np.ascontiguousarray(X0)).cuda()
    s.solve_multi_device(dB, dX, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    return dX.cpu().numpy(), s.get_info()


def _singles(s, B, X0):
    xs, its, res = [], [], []
    for c in range(B.shape[0]):
        xs.append(s.solve(B[c], X0[c].copy()))
        info = s.get_info()
        its.append(info["num_iterations"])
        res.append(info["final_res_norm"])
    return np.stack(xs), its, res


@pytest.mark.parametrize(
    "name,N,k,kw",
    [
        ("elastic_contact", 10, 2, {}),
        ("elastic_contact", 10, 4, {}),
        ("elastic_contact", 10, 5, {}),
        ("elastic_contact", 10, 3, {"precision": "fp32"}),
        ("elastic_contact", 10, 2, {"precision": "fp32"}),
        ("laplace_sphere", 12, 4, {"precision": "fp32"}),
        # no coarse MAS levels: the level-0 apply reduces r.z itself (its dot variants)
        ("elastic_contact", 10, 2, {"precision": "fp32", "max_coarse_levels": 0}),
        ("laplace_sphere", 12, 4, {"precision": "fp32", "max_coarse_levels": 0}),
        ("elastic_contact", 10, 2, {"max_coarse_levels": 0}),
        ("laplace_sphere", 12, 2, {}),
        ("elastic_logE", 8, 4, {"base_preconditioner": "Jacobi"}),
        ("elastic_contact", 8, 2, {"do_hybrid_precond": False}),
        # the AMG base preconditioner: one V-cycle per column
        ("elastic_contact", 10, 3, {"base_preconditioner": "AMG"}),
        ("laplace_sphere", 12, 2, {"base_preconditioner": "AMG", "do_hybrid_precond": False}),
        # the matrix-core bank solve: each output column of an MFMA step depends only on
        # its own B column, so fused columns equal single solves here too
        ("elastic_contact", 10, 2, {"precision": "fp32", "bank_apply": "mfma"}),
        ("elastic_contact", 10, 4, {"precision": "fp32", "bank_apply": "mfma", "mfma_flush": 32}),
        ("laplace_sphere", 12, 3, {"precision": "fp32", "bank_apply": "mfma", "max_coarse_levels": 0}),
        # stiff units through the multifrontal sparse factor (per-column level solves)
        ("elastic_contact", 10, 3, {"stiff_factorization": "sparse", "sparse_leaf_size": 16}),
    ],
)
def test_fused_equals_single_solves(gpu, name, N, k, kw):
    prob = problems.make(name, N=N)
    A = prob.csc()
    s = Solver.create(_cfg(prob.block_size, **kw))
    s.factorize(A)
    B = _rhs(A.n, k, 11)
    X0 = np.zeros_like(B)
    X0[0] = problems.random_vector(A.n, 99)  # a non-zero initial guess in one column
    X, info = _fused(s, B, X0)
    Xs, its, res = _singles(s, B, X0)
    cols = info["columns"]
    assert len(cols) == k
    for c in range(k):
        assert np.array_equal(X[c], Xs[c]), f"column {c}: max diff {np.max(np.abs(X[c] - Xs[c]))}"
        assert cols[c]["num_iterations"] == its[c]
        assert cols[c]["final_res_norm"] == res[c]
        assert cols[c]["stop_reason"] == "converged_rel"
    assert info["num_iterations"] == max(its)


def test_fused_zero_column(gpu):
    prob = problems.make("elastic_contact", N=8)
    A = prob.csc()
    s = Solver.create(_cfg(3))
    s.factorize(A)
    B = _rhs(A.n, 4, 5)
    B[1] = 0.0
    X0 = np.ones_like(B)
    X, info = _fused(s, B, X0)
    assert np.all(X[1] == 0)
    assert info["columns"][1]["num_iterations"] == 0
    assert info["columns"][1]["stop_reason"] == "zero_rhs"
    Xs, its, _ = _singles(s, B, X0)
    for c in (0, 2, 3):
        assert np.array_equal(X[c], Xs[c])
        assert info["columns"][c]["num_iterations"] == its[c]
    # all-zero block: nothing to iterate
    X, info = _fused(s, np.zeros((2, A.n)), np.ones((2, A.n)))
    assert np.all(X == 0) and info["num_iterations"] == 0


def test_fused_damping_retry_column(gpu):
    # damping 3.5 makes the sweep indefinite; a column that breaks down is solved
    # again alone with damping 1.5, as a single solve would be
    prob = problems.make("elastic_contact", N=10)
    A = prob.csc()
    s = Solver.create(_cfg(3, damping=3.5))
    s.factorize(A)
    B = _rhs(A.n, 2, 21)
    X0 = np.zeros_like(B)
    X, info = _fused(s, B, X0)
    assert info["damping_retry"]
    Xs, its, _ = _singles(s, B, X0)
    for c in range(2):
        assert np.array_equal(X[c], Xs[c])
        assert info["columns"][c]["num_iterations"] == its[c]
        assert info["columns"][c]["stop_reason"] == "converged_rel"


# Each fused column against the oracle's single solve of the same column (FP64 banks in
# the oracle): the same stiff set and damping, so FP64 columns take the oracle's
# iteration count (within the partition-independent slack of the single-solve parity
# tests) and every column, FP32 and BF16 bank storage included, reaches the oracle's
# solution to within the tolerance the FP64 outer PCG guarantees.
@pytest.mark.parametrize("precision,k", [("fp64", 2), ("fp32", 2), ("bf16", 2), ("fp32", 4), ("bf16", 4)])
def test_fused_columns_match_oracle(gpu, precision, k):
    from oracle import oracle as O

    prob = problems.make("elastic_contact", N=10)
    A = prob.csc()
    s = Solver.create(_cfg(3, precision=precision))
    s.factorize(A)
    B = _rhs(A.n, k, 31)
    X, info = _fused(s, B, np.zeros_like(B))
    H = O.Hybrid(O.config(block_size=3, rel_tol=1e-8, gmm_jump_threshold=10.0, min_subdomain_size=4))
    H.factorize(A.colptr, A.rowidx, A.vals)
    assert [c.size for c in s.stiff_components()] == [c.size for c in H.stiff_components()]
    S = A.to_scipy()
    for c in range(k):
        xr = H.solve(B[c])
        oi = H.info()
        col = info["columns"][c]
        assert col["stop_reason"] == "converged_rel"
        assert np.linalg.norm(X[c] - xr) <= 1e-5 * np.linalg.norm(xr), c
        floor = 50 * np.finfo(np.float64).eps * np.linalg.norm(abs(S) @ np.abs(X[c]))
        assert np.linalg.norm(S @ X[c] - B[c]) <= max(1.01e-8 * np.linalg.norm(B[c]), floor)
        if precision == "fp64":
            assert abs(col["num_iterations"] - oi["num_iterations"]) <= max(2, 0.05 * oi["num_iterations"])
        else:
            # reduced-precision bank factors cost iterations, never accuracy
            assert col["num_iterations"] <= 2.5 * oi["num_iterations"] + 3
