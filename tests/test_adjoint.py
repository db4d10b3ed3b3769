"""Adjoint of the hybrid solve: dL/db, dL/dA on the pattern, dL/dE per element.

CPU tests pin the element Jacobian and the adjoint formulas with scipy direct
solves and central finite differences; GPU tests check the HIP kernels and the
torch.autograd.Function against them.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from lss_amd import problems


def _direct(prob, coef, b):
    """x = A(coef)^{-1} b with scipy (operator = CSC arrays read as rows; symmetric)."""
    A = prob.csc()
    vals = prob.vals_with_coef(coef)
    import scipy.sparse as sp

    K = sp.csr_matrix((vals, A.rowidx, A.colptr), shape=(A.n, A.n)).tocsc()
    return spla.spsolve(K, b), K


def _loss(x, xt):
    return 0.5 * float(np.dot(x - xt, x - xt))


def _segment_sums(seg, slot, val, g):
    contrib = val * g[slot]
    return np.add.reduceat(contrib, seg[:-1]) if contrib.size else np.zeros(seg.size - 1)


@pytest.mark.parametrize("name,N", [("elastic_logE", 5), ("laplace_sphere", 6)])
def test_element_jacobian_reassembles_values(name, N):
    prob = problems.make(name, N=N)
    A = prob.csc()
    seg, slot, val = prob.element_jacobian()
    coef = prob.element_coef()
    per_entry = val * np.repeat(coef, np.diff(seg))
    v = np.zeros(A.nnz)
    np.add.at(v, slot, per_entry)
    assert np.max(np.abs(v - A.vals)) <= 1e-12 * np.max(np.abs(A.vals))


def test_element_chain_matches_generator_gradient():
    prob = problems.make("elastic_logE", N=5)
    A = prob.csc()
    seg, slot, val = prob.element_jacobian()
    lam = problems.random_vector(A.n, 1)
    x = problems.random_vector(A.n, 2)
    cols = np.repeat(np.arange(A.n), np.diff(A.colptr))
    gvals = -lam[cols] * x[A.rowidx]
    ge = _segment_sums(seg, slot, val, gvals)
    ref = prob.grad_coef(lam, x)
    assert np.max(np.abs(ge - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_adjoint_formulas_by_finite_differences():
    """dL/dE from lam = A^{-1}(x - x_t) and the Jacobian, against central differences."""
    prob = problems.make("elastic_logE", N=5)
    coef = prob.element_coef()
    b = prob.rhs(0)
    xt = problems.random_vector(prob.n, 9)
    x, K = _direct(prob, coef, b)
    lam = spla.spsolve(K, x - xt)
    A = prob.csc()
    cols = np.repeat(np.arange(A.n), np.diff(A.colptr))
    gvals = -lam[cols] * x[A.rowidx]
    seg, slot, val = prob.element_jacobian()
    gE = _segment_sums(seg, slot, val, gvals)
    rng = np.random.default_rng(0)
    for _ in range(4):
        d = rng.standard_normal(coef.size) * coef
        h = 1e-6
        lp = _loss(_direct(prob, coef + h * d, b)[0], xt)
        lm = _loss(_direct(prob, coef - h * d, b)[0], xt)
        fd = (lp - lm) / (2 * h)
        assert np.dot(gE, d) == pytest.approx(fd, rel=1e-5)
    # dL/db = lam
    db = rng.standard_normal(b.size)
    h = 1e-6
    fd = (_loss(spla.spsolve(K, b + h * db), xt) - _loss(spla.spsolve(K, b - h * db), xt)) / (2 * h)
    assert np.dot(lam, db) == pytest.approx(fd, rel=1e-6)


# ---------------------------------------------------------------- GPU
def _dev(prob, coef=None):
    import torch

    A = prob.csc()
    vals = A.vals if coef is None else prob.vals_with_coef(coef)
    d = torch.device("cuda:0")
    return (torch.from_numpy(A.colptr).to(d), torch.from_numpy(A.rowidx).to(d),
            torch.from_numpy(vals).to(d), A)


def _cfg(bs, rtol):
    return {"solver": "GPUHybrid", "GPUHybrid": {"block_size": bs, "relative_tolerance": rtol,
                                                  "gmm_jump_threshold": 10.0, "min_subdomain_size": 4}}


@pytest.mark.gpu
def test_pattern_grad_and_chain_kernels(gpu):
    import torch

    from lss_amd import autograd as ag

    prob = problems.make("elastic_logE", N=6)
    cp, ri, vals, A = _dev(prob)
    lam = problems.random_vector(A.n, 1)
    x = problems.random_vector(A.n, 2)
    d = torch.device("cuda:0")
    g = ag.pattern_grad(cp, ri, torch.from_numpy(lam).to(d), torch.from_numpy(x).to(d)).cpu().numpy()
    cols = np.repeat(np.arange(A.n), np.diff(A.colptr))
    assert np.array_equal(g, -lam[cols] * x[A.rowidx])
    seg, slot, val = prob.element_jacobian()
    ge = ag.element_grad(torch.from_numpy(seg).to(d), torch.from_numpy(slot).to(d), torch.from_numpy(val).to(d),
                         torch.from_numpy(g).to(d)).cpu().numpy()
    ref = _segment_sums(seg, slot, val, g)
    assert np.max(np.abs(ge - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autograd_matches_direct_adjoint(gpu, precision):
    import torch

    from lss_amd import autograd as ag
    from lss_amd.solver import Solver

    prob = problems.make("elastic_logE", N=8)
    coef = prob.element_coef()
    cp, ri, vals, A = _dev(prob)
    d = torch.device("cuda:0")
    b = torch.from_numpy(prob.rhs(0)).to(d).requires_grad_(True)
    v = vals.clone().requires_grad_(True)
    xt = problems.random_vector(A.n, 9)
    cfg = _cfg(3, 1e-12)
    cfg["GPUHybrid"]["precision"] = precision
    s = Solver.create(cfg)
    x = ag.hybrid_solve(v, b, cp, ri, solver=s)
    L = 0.5 * ((x - torch.from_numpy(xt).to(d)) ** 2).sum()
    L.backward()
    xd, K = _direct(prob, coef, prob.rhs(0))
    lam = spla.spsolve(K, xd - xt)
    assert np.linalg.norm(x.detach().cpu().numpy() - xd) <= 1e-9 * np.linalg.norm(xd)
    gb = b.grad.cpu().numpy()
    assert np.linalg.norm(gb - lam) <= 1e-8 * np.linalg.norm(lam)
    cols = np.repeat(np.arange(A.n), np.diff(A.colptr))
    gv_ref = -lam[cols] * xd[A.rowidx]
    gv = v.grad.cpu().numpy()
    assert np.linalg.norm(gv - gv_ref) <= 1e-8 * np.linalg.norm(gv_ref)
    # per-element chain and central differences in E
    seg, slot, val = prob.element_jacobian()
    gE = ag.element_grad(torch.from_numpy(seg).to(d), torch.from_numpy(slot).to(d), torch.from_numpy(val).to(d),
                         v.grad).cpu().numpy()
    assert np.max(np.abs(gE - prob.grad_coef(lam, xd))) <= 1e-7 * np.max(np.abs(gE))
    rng = np.random.default_rng(1)
    for _ in range(3):
        dE = rng.standard_normal(coef.size) * coef
        h = 1e-6
        fd = (_loss(_direct(prob, coef + h * dE, prob.rhs(0))[0], xt)
              - _loss(_direct(prob, coef - h * dE, prob.rhs(0))[0], xt)) / (2 * h)
        assert np.dot(gE, dE) == pytest.approx(fd, rel=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("name,N", [("elastic_logE", 12), ("elastic_contact", 10)])
def test_autograd_matches_oracle_solver_gradient(gpu, name, N):
    """The GPU gradients against a gradient built from the oracle's CPU hybrid solver (the
    restated reference GPUHybrid algorithm, FP64): x and lambda = A^{-1}(x - x_t) from two
    oracle solves to the same relative tolerance, then dL/db = lambda, dL/dA = -lambda x^T on
    the pattern and dL/dE per element from them."""
    import torch

    from lss_amd import autograd as ag
    from lss_amd.solver import Solver
    from oracle import oracle as O

    rtol = 1e-12
    prob = problems.make(name, N=N)
    cp, ri, vals, A = _dev(prob)
    d = torch.device("cuda:0")
    b_np = prob.rhs(0)
    b = torch.from_numpy(b_np).to(d).requires_grad_(True)
    v = vals.clone().requires_grad_(True)
    xt = problems.random_vector(A.n, 9)
    s = Solver.create(_cfg(prob.block_size, rtol))
    x = ag.hybrid_solve(v, b, cp, ri, solver=s)
    L = 0.5 * ((x - torch.from_numpy(xt).to(d)) ** 2).sum()
    L.backward()

    H = O.Hybrid(O.config(block_size=prob.block_size, rel_tol=rtol, gmm_jump_threshold=10.0, min_subdomain_size=4))
    H.factorize(A.colptr, A.rowidx, A.vals)
    x_cpu = H.solve(b_np)
    lam_cpu = H.solve(x_cpu - xt)
    assert H.info()["num_iterations"] > 0
    # both solves stop at a relative residual of 1e-12; measured agreement: 1e-14 (log-E),
    # 2e-10 to 5e-10 (contact entries up to 1e8)
    xg = x.detach().cpu().numpy()
    rel = lambda a, r: float(np.linalg.norm(a - r) / np.linalg.norm(r))  # noqa: E731
    print(f"{name} {N}: x {rel(xg, x_cpu):.1e}, dL/db {rel(b.grad.cpu().numpy(), lam_cpu):.1e}")
    assert np.linalg.norm(xg - x_cpu) <= 1e-8 * np.linalg.norm(x_cpu)
    gb = b.grad.cpu().numpy()
    assert np.linalg.norm(gb - lam_cpu) <= 1e-8 * np.linalg.norm(lam_cpu)
    cols = np.repeat(np.arange(A.n), np.diff(A.colptr))
    gv_cpu = -lam_cpu[cols] * x_cpu[A.rowidx]
    gv = v.grad.cpu().numpy()
    assert np.linalg.norm(gv - gv_cpu) <= 1e-8 * np.linalg.norm(gv_cpu)
    if name == "elastic_logE":  # per-element E (the contact system has no element coefficients of its own)
        seg, slot, val = prob.element_jacobian()
        gE = ag.element_grad(torch.from_numpy(seg).to(d), torch.from_numpy(slot).to(d),
                             torch.from_numpy(val).to(d), v.grad).cpu().numpy()
        gE_cpu = prob.grad_coef(lam_cpu, x_cpu)
        assert np.linalg.norm(gE - gE_cpu) <= 1e-8 * np.linalg.norm(gE_cpu)


@pytest.mark.gpu
def test_autograd_gradient_by_gpu_finite_differences(gpu):
    """Central differences of the GPU forward solve itself (step 1e-6 relative in E)."""
    import torch

    from lss_amd import autograd as ag
    from lss_amd.solver import Solver

    prob = problems.make("elastic_logE", N=10)
    coef = prob.element_coef()
    cp, ri, vals, A = _dev(prob)
    d = torch.device("cuda:0")
    b = torch.from_numpy(prob.rhs(0)).to(d)
    xt = torch.from_numpy(problems.random_vector(A.n, 9)).to(d)
    s = Solver.create(_cfg(3, 1e-13))
    v = vals.clone().requires_grad_(True)
    x = ag.hybrid_solve(v, b, cp, ri, solver=s)
    L = 0.5 * ((x - xt) ** 2).sum()
    L.backward()
    seg, slot, val = prob.element_jacobian()
    gE = ag.element_grad(torch.from_numpy(seg).to(d), torch.from_numpy(slot).to(d), torch.from_numpy(val).to(d),
                         v.grad).cpu().numpy()

    def loss_at(c):
        s2 = Solver.create(_cfg(3, 1e-13))
        with torch.no_grad():
            xx = ag.hybrid_solve(torch.from_numpy(prob.vals_with_coef(c)).to(d), b, cp, ri, solver=s2)
        return float(0.5 * ((xx - xt) ** 2).sum())

    rng = np.random.default_rng(2)
    for _ in range(3):
        dE = rng.standard_normal(coef.size) * coef
        h = 1e-6
        fd = (loss_at(coef + h * dE) - loss_at(coef - h * dE)) / (2 * h)
        assert np.dot(gE, dE) == pytest.approx(fd, rel=1e-4)


@pytest.mark.gpu
def test_adjoint_at_full_size_by_gpu_finite_differences(gpu):
    """The differentiable-solve workload at its stated size: 48^3 elasticity with
    log-uniform E in [1, 1e5] (331 776 DOF), L = 0.5 ||x - x_t||^2, dL/dE from the
    adjoint against central differences (step 1e-6 relative in E) of the GPU forward
    solve in 64 random directions, FP64, solves to a recurrence residual of 1e-13."""
    import torch

    from lss_amd import autograd as ag
    from lss_amd.solver import Solver

    prob = problems.make("elastic_logE", N=48)
    coef = prob.element_coef()
    cp, ri, vals, A = _dev(prob)
    d = torch.device("cuda:0")
    b = torch.from_numpy(prob.rhs(0)).to(d)
    xt = torch.from_numpy(problems.random_vector(A.n, 9)).to(d)
    s = Solver.create(_cfg(3, 1e-13))
    v = vals.clone().requires_grad_(True)
    x = ag.hybrid_solve(v, b, cp, ri, solver=s)
    L = 0.5 * ((x - xt) ** 2).sum()
    L.backward()
    seg, slot, val = prob.element_jacobian()
    gE = ag.element_grad(torch.from_numpy(seg).to(d), torch.from_numpy(slot).to(d), torch.from_numpy(val).to(d),
                         v.grad).cpu().numpy()
    s2 = Solver.create(_cfg(3, 1e-13))

    def loss_at(c):
        with torch.no_grad():
            xx = ag.hybrid_solve(torch.from_numpy(prob.vals_with_coef(c)).to(d), b, cp, ri, solver=s2)
        return float(0.5 * ((xx - xt) ** 2).sum())

    rng = np.random.default_rng(2)
    an, fd = [], []
    for _ in range(64):
        dE = rng.standard_normal(coef.size) * coef
        h = 1e-6
        fd.append((loss_at(coef + h * dE) - loss_at(coef - h * dE)) / (2 * h))
        an.append(float(np.dot(gE, dE)))
    an, fd = np.array(an), np.array(fd)
    scale = np.median(np.abs(fd))
    # every direction to 1e-5 of the typical directional derivative (a direction whose
    # derivative happens to be near zero is judged on that scale), most to 1e-5 relative
    assert np.abs(an - fd).max() <= 1e-5 * scale, np.abs(an - fd).max() / scale
    assert np.median(np.abs(an - fd) / np.abs(fd)) <= 1e-5
