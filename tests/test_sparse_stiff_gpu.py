"""Stiff units factored by the multifrontal sparse Cholesky (the cuDSS role,
GPUHybrid.cu:1231-1249 and 575).

  * Every DOF stiff (caller-given set, no filtering): the stiff correction of the
    composed preconditioner is the exact inverse, so PCG stops within two iterations
    at the direct solution; any error in the ordering, the front structures, the
    assembly, the extend-add, the partial factorizations or the level-batched
    triangular solves shows.  Leaf sizes from 8 (deep trees) to whole-graph fronts; batch
    limits that split tree levels into many batches, mix batched and one-at-a-time fronts
    within a level, or batch nothing.
  * On GMM-selected units the sparse factor and the dense factor give the same
    preconditioner to rounding, and the same solve.
  * Units above max_dense_component take the sparse path by themselves ("auto") and
    are rejected with "dense"; a unit that is not SPD is reported.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from lss_amd import problems
from lss_amd.solver import LssError, Solver

pytestmark = pytest.mark.gpu


def _cfg(bs, **kw):
    g = {"block_size": bs, "relative_tolerance": 1e-8, "select_bad_dofs_from_l1_norm": False,
         "decompose_subdomains": False}
    g.update(kw)
    return {"solver": "GPUHybrid", "GPUHybrid": g}


@pytest.mark.parametrize("name,N,leaf", [("elastic_contact", 5, 8), ("elastic_contact", 8, 64),
                                         ("laplace_sphere", 11, 16), ("elastic_slivers", 10, 256),
                                         ("elastic_logE", 16, 128), ("laplace_sphere", 9, 100000)])
def test_all_stiff_sparse_is_a_direct_solve(gpu, name, N, leaf):
    prob = problems.make(name, N=N)
    A = prob.csc()
    n = A.n
    s = Solver.create(_cfg(prob.block_size, stiff_factorization="sparse", sparse_leaf_size=leaf))
    s.set_problematic_dofs(np.arange(n, dtype=np.int32))
    s.factorize(A)
    info = s.get_info()
    sp_info = info["stiff"]["sparse"]
    assert sp_info["units"] == 1 and sp_info["dofs"] == [n]
    if leaf < n:
        assert sp_info["fronts"][0] > 1
    b = prob.rhs(2)
    x = s.solve(b)
    info = s.get_info()
    print(f"{name} N={N} n={n} leaf={leaf}: fronts {sp_info['fronts']} levels {sp_info['levels']} max front "
          f"{sp_info['max_front']} factor entries {sp_info['factor_entries']} iterations {info['num_iterations']}")
    assert info["stop_reason"] == "converged_rel"
    assert info["num_iterations"] <= 2
    S = A.to_scipy().tocsc()
    xd = spla.spsolve(S, b)
    floor = 50 * np.finfo(np.float64).eps * np.linalg.norm(abs(S) @ np.abs(x))
    assert np.linalg.norm(S @ x - b) <= max(1e-8 * np.linalg.norm(b), floor)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)


@pytest.mark.parametrize("rows,nbytes", [(200, 200_000), (100_000, 1), (0, 1 << 32)])
def test_batch_limits_keep_the_direct_solve(gpu, rows, nbytes):
    # small batch limits: tree levels split into several batches (one front per batch with a
    # 1-byte workspace limit), batched and one-at-a-time fronts within a level, or no batches
    prob = problems.make("elastic_logE", N=16)
    A = prob.csc()
    n = A.n
    ref = Solver.create(_cfg(3, stiff_factorization="sparse", sparse_leaf_size=128))
    s = Solver.create(_cfg(3, stiff_factorization="sparse", sparse_leaf_size=128, sparse_batch_rows=rows,
                           sparse_batch_bytes=nbytes))
    for x in (ref, s):
        x.set_problematic_dofs(np.arange(n, dtype=np.int32))
        x.factorize(A)
    sp, spr = s.get_info()["stiff"]["sparse"], ref.get_info()["stiff"]["sparse"]
    assert sp["fronts"] == spr["fronts"]
    nb, ng = sp["batched_fronts"][0], sp["front_groups"][0]
    print(f"rows {rows} bytes {nbytes}: {nb} of {sp['fronts'][0]} fronts batched in {ng} batches "
          f"(defaults: {spr['batched_fronts'][0]} in {spr['front_groups'][0]})")
    if rows == 0:
        assert nb == 0 and ng == 0
    elif nbytes == 1:
        assert nb == sp["fronts"][0] and ng == nb
    else:
        assert 0 < nb < sp["fronts"][0] and ng > spr["front_groups"][0]
    b = prob.rhs(2)
    x = s.solve(b)
    assert s.get_info()["num_iterations"] <= 2
    S = A.to_scipy().tocsc()
    xd = spla.spsolve(S, b)
    floor = 50 * np.finfo(np.float64).eps * np.linalg.norm(abs(S) @ np.abs(x))
    assert np.linalg.norm(S @ x - b) <= max(1e-8 * np.linalg.norm(b), floor)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)


@pytest.mark.parametrize("name,N", [("elastic_contact", 12), ("laplace_sphere", 12)])
def test_sparse_units_equal_dense_units(gpu, name, N):
    prob = problems.make(name, N=N)
    A = prob.csc()
    g = {"block_size": prob.block_size, "relative_tolerance": 1e-8, "gmm_jump_threshold": 10.0,
         "min_subdomain_size": 4}
    dense = Solver.create({"solver": "GPUHybrid", "GPUHybrid": dict(g, stiff_factorization="dense")})
    sparse = Solver.create({"solver": "GPUHybrid", "GPUHybrid": dict(g, stiff_factorization="sparse",
                                                                     sparse_leaf_size=8)})
    dense.factorize(A)
    sparse.factorize(A)
    comps = dense.stiff_components()
    assert len(comps) >= 1
    assert sparse.get_info()["stiff"]["sparse"]["units"] == len(comps)
    # both factor S A S^T in FP64 with different elimination orders: they agree to the
    # rounding of the stiff solve, eps * cond(S A S^T) (contact units reach cond ~1e10)
    S = A.to_scipy().toarray()
    cond = max(np.linalg.cond(S[np.ix_(c, c)]) for c in comps)
    tol = max(1e-12, 100 * np.finfo(np.float64).eps * cond)
    for seed in (1, 2):
        r = problems.random_vector(A.n, seed)
        zd, zs = dense.precond_apply(r), sparse.precond_apply(r)
        rel = np.linalg.norm(zs - zd) / np.linalg.norm(zd)
        print(f"{name} N={N}: max cond(S A S^T) {cond:.1e}, |M_sparse r - M_dense r| / |M_dense r| = {rel:.2e}")
        assert rel <= tol, seed
    b = prob.rhs(0)
    xd, xs = dense.solve(b), sparse.solve(b)
    itd, its = dense.get_info()["num_iterations"], sparse.get_info()["num_iterations"]
    assert abs(itd - its) <= 2
    assert np.linalg.norm(xs - xd) <= 1e-6 * np.linalg.norm(xd)


def test_large_unit_takes_the_sparse_path(gpu):
    prob = problems.make("elastic_slivers", N=9)
    A = prob.csc()
    dofs = np.arange(A.n, dtype=np.int32)
    s = Solver.create(_cfg(3, max_dense_component=500))
    s.set_problematic_dofs(dofs)
    s.factorize(A)
    assert s.get_info()["stiff"]["sparse"]["units"] == 1
    b = prob.rhs(1)
    x = s.solve(b)
    assert s.get_info()["num_iterations"] <= 2
    xd = spla.spsolve(A.to_scipy().tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)
    rej = Solver.create(_cfg(3, max_dense_component=500, stiff_factorization="dense"))
    rej.set_problematic_dofs(dofs)
    with pytest.raises(LssError) as e:
        rej.factorize(A)
    assert e.value.code == 7  # LSS_ERR_UNSUPPORTED


def test_sparse_not_spd_is_reported(gpu):
    prob = problems.make("laplace_sphere", N=9)
    A = prob.csc()
    vals = A.vals.copy()
    j = 600
    e = A.colptr[j] + int(np.nonzero(A.rowidx[A.colptr[j]:A.colptr[j + 1]] == j)[0][0])
    vals[e] = -vals[e]
    B = problems.CSC(A.n, A.colptr, A.rowidx, vals)
    s = Solver.create(_cfg(1, stiff_factorization="sparse", sparse_leaf_size=32))
    s.set_problematic_dofs(np.arange(A.n, dtype=np.int32))
    with pytest.raises(LssError) as ei:
        s.factorize(B)
    assert ei.value.code == 5  # LSS_ERR_NOT_SPD


@pytest.mark.parametrize("kw", [{"use_bad_dof_precond": True, "bad_dof_strategy": "elbow"},
                                {"use_bad_dof_precond": True, "bad_dof_strategy": "gmm"}])
def test_mas_solver_sparse_units(gpu, kw):
    # the MAS solver's additive correction S^T (S A S^T)^{-1} S r through the sparse factor
    prob = problems.make("elastic_contact", N=10)
    A = prob.csc()
    dense = Solver.create({"solver": "MAS", "MAS": dict(block_dim=3, **kw)})
    sparse = Solver.create({"solver": "MAS", "MAS": dict(block_dim=3, stiff_factorization="sparse",
                                                         sparse_leaf_size=16, **kw)})
    dense.factorize(A)
    sparse.factorize(A)
    assert sparse.get_info()["stiff"]["sparse"]["units"] == len(dense.stiff_components()) >= 1
    r = problems.random_vector(A.n, 3)
    zd, zs = dense.precond_apply(r), sparse.precond_apply(r)
    assert np.linalg.norm(zs - zd) <= 1e-10 * np.linalg.norm(zd)
    b = prob.rhs(2)
    xd, xs = dense.solve(b), sparse.solve(b)
    assert abs(dense.get_info()["num_iterations"] - sparse.get_info()["num_iterations"]) <= 10
    assert np.linalg.norm(xs - xd) <= 1e-4 * np.linalg.norm(xd)


def test_all_stiff_direct_at_size(gpu):
    # 48^3 log-uniform-E elasticity (331 776 DOF), every DOF stiff: one unit, factored by
    # the multifrontal path, so the hybrid PCG must stop at the direct solution
    import torch

    from lss_amd.solver import residual_norm

    prob = problems.make("elastic_logE", N=48)
    A = prob.csc()
    n = A.n
    dev = torch.device("cuda:0")
    dA = [torch.from_numpy(a).to(dev) for a in (A.colptr, A.rowidx, A.vals)]
    b = torch.from_numpy(prob.rhs(0)).to(dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    s = Solver.create(_cfg(3))
    s.set_problematic_dofs(np.arange(n, dtype=np.int32))
    s.factorize_device(n, *dA, stream=torch.cuda.current_stream())
    sp_info = s.get_info()["stiff"]["sparse"]
    assert sp_info["units"] == 1 and sp_info["dofs"] == [n] and sp_info["fronts"][0] > 1000
    s.solve_device(b, x, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    info = s.get_info()
    r, bn = residual_norm(*dA, b, x)
    print(f"48^3 all-stiff: fronts {sp_info['fronts']} max front {sp_info['max_front']} factor entries "
          f"{sp_info['factor_entries']} symbolic {sp_info['t_symbolic']} numeric {sp_info['t_numeric']}; "
          f"{info['num_iterations']} iterations, true rel residual {r / bn:.2e}")
    assert info["stop_reason"] == "converged_rel" and info["num_iterations"] <= 2
    assert r / bn <= 1e-8


@pytest.mark.parametrize("name,N,strategy", [("elastic_logE", 48, "elbow"), ("elastic_slivers", 64, "gmm")])
def test_mas_solver_large_unit_at_size(gpu, name, N, strategy):
    # the MAS solver's window-expanded sets at the workloads' stated sizes form one unit of
    # 0.3-0.5 M DOF (above the dense limit): the additive correction goes through the
    # multifrontal factor and the loop reaches the MAS solver's tolerance (1e-6)
    import torch

    from lss_amd.solver import residual_norm

    prob = problems.make(name, N=N)
    A = prob.csc()
    n = A.n
    dev = torch.device("cuda:0")
    dA = [torch.from_numpy(a).to(dev) for a in (A.colptr, A.rowidx, A.vals)]
    b = torch.from_numpy(prob.rhs(0)).to(dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    s = Solver.create({"solver": "MAS", "MAS": {"block_dim": 3, "max_iter": 3000, "use_bad_dof_precond": True,
                                                "bad_dof_strategy": strategy}})
    s.factorize_device(n, *dA, stream=torch.cuda.current_stream())
    info = s.get_info()
    sp_info = info["stiff"]["sparse"]
    assert sp_info["units"] == 1 and sp_info["dofs"][0] > 16384
    s.solve_device(b, x, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    info = s.get_info()
    r, bn = residual_norm(*dA, b, x)
    print(f"{name} {N}^3 MAS+{strategy}: unit {sp_info['dofs']} fronts {sp_info['fronts']} max front "
          f"{sp_info['max_front']}; {info['solver_iter']} iterations, true rel residual {r / bn:.2e}")
    assert info["solver_status"] == "Reach relative tolerance"
    assert r / bn <= 2e-6
