"""The standalone MAS solver (SRC/MASSolver.cu) on the GPU against the oracle:
the stiff set of each strategy, the additive preconditioner and the PCG loop."""
import numpy as np
import pytest

from lss_amd import problems
from lss_amd.solver import Solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _dev_cfg(bs, **kw):
    m = {"block_dim": bs}
    m.update(kw)
    return {"solver": "MAS", "MAS": m}


def _orc_cfg(bs, **kw):
    base = dict(block_size=bs, solver="MAS", rel_tol=kw.pop("relative_tolerance", 1e-6), abs_tol=1e-20,
                max_iter=kw.pop("max_iter", 100000))
    base.update(kw)
    return O.config(**base)


CASES = [("elastic_contact", 10), ("laplace_sphere", 12), ("elastic_logE", 8)]
STRATEGIES = [{}, {"use_bad_dof_precond": True, "bad_dof_strategy": "gmm"},
              {"use_bad_dof_precond": True, "bad_dof_strategy": "elbow"},
              {"use_bad_dof_precond": True, "bad_dof_strategy": "kmean", "bad_dof_kmeans_expand_neighbors": 2}]


@pytest.mark.parametrize("kw", STRATEGIES)
@pytest.mark.parametrize("name,N", CASES)
def test_mas_solver_matches_oracle(gpu, name, N, kw):
    prob = problems.make(name, N=N)
    A = prob.csc()
    bs = prob.block_size
    s = Solver.create(_dev_cfg(bs, **kw))
    assert s.name() == "MAS"
    s.factorize(A)
    H = O.Hybrid(_orc_cfg(bs, **kw))
    H.factorize(A.colptr, A.rowidx, A.vals)
    dc, oc = s.stiff_components(), H.stiff_components()
    assert len(dc) == len(oc)
    for a, b in zip(dc, oc):
        assert np.array_equal(a, b)
    b = prob.rhs(2)
    x = s.solve(b, np.ones(A.n))  # the initial guess is ignored (x0 = 0)
    info = s.get_info()
    xr = H.solve(b)
    oi = H.info()
    assert info["solver_status"] == "Reach relative tolerance" and oi["mas_status"] == 1
    assert info["solver_iter"] == info["num_iterations"]
    # Both test every 10 iterations.  On the contact system the count is not a smooth
    # function of the rounding: the oracle alone moves between 80 and 100 (elbow) or 100
    # and 120 (k-means) under 1e-13 relative perturbations of b, so the device count must
    # lie in the oracle's range over such perturbations, give or take one test period.
    its = [oi["num_iterations"]]
    rng = np.random.default_rng(0)
    for _ in range(4):
        H.solve(b * (1.0 + 1e-13 * rng.standard_normal(b.size)))
        its.append(H.info()["num_iterations"])
    assert min(its) - 10 <= info["solver_iter"] <= max(its) + 10, its
    S = A.to_scipy()
    assert np.linalg.norm(S @ x - b) <= 2e-6 * np.linalg.norm(b)
    assert np.linalg.norm(x - xr) <= 1e-4 * np.linalg.norm(xr)
    if kw.get("bad_dof_strategy") == "gmm":
        st = info["bad_dof"]
        perm, _ = H.partition()
        g = O.mas_indicator(0, A.colptr, A.rowidx, A.vals, bs, perm)
        _, ost = O.mas_gmm_positions(g)
        for k in ("mean0", "mean1", "var0", "var1"):
            assert st[k] == pytest.approx(ost[k], rel=1e-12), k


@pytest.mark.parametrize("kw", STRATEGIES[:2])
def test_mas_solver_without_residual_replacement(gpu, kw):
    # "true_residual_period": 0 keeps the recurrence residual (MASSolver.cu:655-666 replaces
    # it by b - A x every 4 iterations); the device loop against the oracle's same loop
    prob = problems.make("elastic_contact", N=10)
    A = prob.csc()
    s = Solver.create(_dev_cfg(3, true_residual_period=0, **kw))
    s.factorize(A)
    H = O.Hybrid(_orc_cfg(3, true_residual_period=0, **kw))
    H.factorize(A.colptr, A.rowidx, A.vals)
    b = prob.rhs(2)
    x = s.solve(b)
    info = s.get_info()
    its = []
    rng = np.random.default_rng(0)
    for q in range(5):
        H.solve(b if q == 0 else b * (1.0 + 1e-13 * rng.standard_normal(b.size)))
        its.append(H.info()["num_iterations"])
    assert info["solver_status"] == "Reach relative tolerance"
    assert min(its) - 10 <= info["solver_iter"] <= max(its) + 10, (info["solver_iter"], its)
    S = A.to_scipy()
    assert np.linalg.norm(S @ x - b) <= 2e-6 * np.linalg.norm(b)


@pytest.mark.parametrize("solver", ["MAS", "GPUHybrid"])
def test_lazy_partitioning_reuses_the_first_partition(gpu, solver):
    # MASSolver.cu:389-405: with lazy_partitioning (the MAS solver's default) a later
    # factorization with the same block count keeps the first partition
    import scipy.sparse as sp

    prob = problems.make("elastic_contact", N=10)
    a = prob.csc()
    nb = a.n // 3
    # b: the same system with its nodes renumbered (same block count, another pattern)
    node = np.random.default_rng(3).permutation(nb)
    dof = (3 * node[:, None] + np.arange(3)).ravel()
    Sb = sp.csc_matrix(a.to_scipy()[dof][:, dof])
    Sb.sort_indices()
    b = type(a)(a.n, Sb.indptr.astype(np.int32), Sb.indices.astype(np.int32), Sb.data.copy())
    if solver == "MAS":
        cfg, cfg_off = _dev_cfg(3), _dev_cfg(3, lazy_partitioning=False)
    else:
        g = {"block_size": 3, "relative_tolerance": 1e-8, "gmm_jump_threshold": 10.0, "min_subdomain_size": 4}
        cfg = {"solver": "GPUHybrid", "GPUHybrid": dict(g, lazy_partitioning=True)}
        cfg_off = {"solver": "GPUHybrid", "GPUHybrid": g}
    s, s_off = Solver.create(cfg), Solver.create(cfg_off)
    s.factorize(a)
    assert not s.get_info()["partition_reused"]
    pa, oa = s.partition(nb)
    s.factorize(b)
    assert s.get_info()["partition_reused"]
    pb, ob = s.partition(nb)
    assert np.array_equal(pa, pb) and np.array_equal(oa, ob)
    s_off.factorize(a)
    s_off.factorize(b)
    assert not s_off.get_info()["partition_reused"]
    pc, _ = s_off.partition(nb)
    assert not np.array_equal(pc, pa)  # b's own partition differs
    # the kept partition still gives a convergent solve of b
    rhs = prob.rhs(0)[dof]
    x = s.solve(rhs)
    S = b.to_scipy()
    tol = 2e-6 if solver == "MAS" else 1.01e-8
    assert np.linalg.norm(S @ x - rhs) <= tol * np.linalg.norm(rhs)


def test_mas_solver_status_and_selection(gpu):
    prob = problems.make("elastic_contact", N=8)
    A = prob.csc()
    b = prob.rhs(0)
    # priority list: the first solver this library implements
    s = Solver.create({"solver": ["Eigen::SimplicialLDLT", "MAS", "GPUHybrid"], "MAS": {"max_iter": 7}})
    s.factorize(A)
    s.solve(b)
    info = s.get_info()
    assert s.name() == "MAS" and info["solver_status"] == "Reach max iterations" and info["solver_iter"] == 7
    # the contact rows are the GMM strategy's high cluster; windows of +-32 positions
    s2 = Solver.create(_dev_cfg(3, use_bad_dof_precond=True, bad_dof_strategy="gmm"))
    s2.factorize(A)
    st = s2.get_info()["bad_dof"]
    assert st["selected"] > 0 and st["expanded"] >= st["selected"]
    assert sum(c.size for c in s2.stiff_components()) == st["expanded"]
    # multi-RHS on the MAS solver: columns one at a time, equal to single solves
    import torch

    B = np.stack([prob.rhs(c) for c in range(2)])
    dB = torch.from_numpy(B).cuda()
    dX = torch.zeros_like(dB)
    s2.solve_multi_device(dB, dX, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    X = dX.cpu().numpy()
    for c in range(2):
        assert np.array_equal(X[c], s2.solve(B[c]))
