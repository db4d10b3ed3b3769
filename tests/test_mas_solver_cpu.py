"""The oracle's restatement of the standalone MAS solver (SRC/MASSolver.cu) and its
stiff-DOF strategies (MAS/BadDOF*Preconditioner.cu), pinned by independent numpy
implementations, Python loops in the reference's order and closed forms."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import independent as I
from lss_amd import problems
from oracle import oracle as O


def _contact(N=8):
    prob = problems.make("elastic_contact", N=N)
    A = prob.csc()
    return prob, A, I.operator_rows(A.colptr, A.rowidx, A.vals, A.n)


def _perm(A, bs):
    H = O.Hybrid(O.config(block_size=bs, do_hybrid_precond=False))
    H.factorize(A.colptr, A.rowidx, A.vals)
    perm, off = H.partition()
    return perm


def test_indicators_follow_the_reference_loop_order():
    _, A, R = _contact(6)
    bs = 3
    perm = _perm(A, bs)
    Rp, _ = I.permuted_rows(R, bs, perm)
    Rp.sort_indices()
    for mode in (0, 1, 2):
        got = O.mas_indicator(mode, A.colptr, A.rowidx, A.vals, bs, perm)
        want = np.zeros(A.n)
        for p in range(A.n):
            cols = Rp.indices[Rp.indptr[p]:Rp.indptr[p + 1]]
            vals = Rp.data[Rp.indptr[p]:Rp.indptr[p + 1]]
            acc = 0.0
            for c, v in zip(cols, vals):  # ascending permuted column = block order, then component
                if mode == 0:
                    acc += abs(v)
                elif mode == 1 and c == p:
                    acc = abs(v)
                elif mode == 2 and c != p:
                    acc = max(acc, abs(v))
            want[p] = acc
        assert np.array_equal(got, want), mode


def _planted(n=5000, seed=4, lo=1e4, hi=1e5):
    rng = np.random.default_rng(seed)
    x = rng.uniform(1.0, 2.0, n)
    hot = rng.choice(n, 40, replace=False)
    x[hot] = rng.uniform(lo, hi, 40)
    return x, np.sort(hot)


@pytest.mark.parametrize("expand", [0, 3, 32])
def test_gmm_strategy_matches_numpy(expand):
    x, hot = _planted()
    got, st = O.mas_gmm_positions(x, expand=expand)
    want, wst = I.gmm_mas_numpy(x, expand=expand)
    assert np.array_equal(got, want)
    for k in ("mean0", "mean1", "var0", "var1", "ratio"):
        assert st[k] == pytest.approx(wst[k], rel=1e-9), k
    assert int(st["iters"]) == wst["iters"]
    if expand == 0:
        assert np.array_equal(got, hot)  # the planted outliers, nothing else
    # no jump: a unimodal indicator selects nothing
    assert O.mas_gmm_positions(np.random.default_rng(1).uniform(1, 2, 999))[0].size == 0
    assert O.mas_gmm_positions(np.ones(10))[0].size == 0  # mean0 == mean1


def test_elbow_strategy_matches_numpy():
    x, hot = _planted(seed=9)
    got, cut = O.mas_elbow_positions(x)
    want, wcut = I.elbow_numpy(x)
    assert cut == wcut and np.array_equal(got, want)
    # ties at the cut are broken by position: equal values keep ascending order
    y = np.array([1.0, 1.0, 1.0, 5.0, 5.0, 1e6, 1e6, 1e6])
    g2, _ = O.mas_elbow_positions(y)
    w2, _ = I.elbow_numpy(y)
    assert np.array_equal(g2, w2)


@pytest.mark.parametrize("expand", [0, 5])
def test_kmeans_strategy_matches_numpy(expand):
    # outliers of one magnitude: the two means separate them from the rest of the top 2 %
    x, hot = _planted(seed=7, lo=1e5, hi=1.1e5)
    got, st = O.mas_kmeans_positions(x, fraction=0.02, expand=expand)
    want, wst = I.kmeans_numpy(x, fraction=0.02, expand=expand)
    assert np.array_equal(got, want)
    assert st["ratio"] == pytest.approx(wst["ratio"], rel=1e-12)
    if expand == 0:
        assert np.array_equal(got, hot)


def test_strategies_on_the_contact_indicators():
    _, A, R = _contact(8)
    perm = _perm(A, 3)
    g = O.mas_indicator(0, A.colptr, A.rowidx, A.vals, 3, perm)
    assert np.array_equal(O.mas_gmm_positions(g)[0], I.gmm_mas_numpy(g)[0])
    d = O.mas_indicator(1, A.colptr, A.rowidx, A.vals, 3, perm)
    assert np.array_equal(O.mas_elbow_positions(d)[0], I.elbow_numpy(d)[0])
    m = O.mas_indicator(2, A.colptr, A.rowidx, A.vals, 3, perm)
    assert np.array_equal(O.mas_kmeans_positions(m)[0], I.kmeans_numpy(m)[0])


def _mas_cfg(**kw):
    # the MAS solver's defaults (MASSolver.cu:148-151); its true residual, recomputed every
    # 4 iterations, cannot go below eps ||A|| ||x|| ~ 1e-7 relative on the contact systems
    base = dict(block_size=3, solver="MAS", rel_tol=1e-6, abs_tol=1e-20, max_iter=100000)
    base.update(kw)
    return O.config(**base)


@pytest.mark.parametrize("strategy", ["gmm", "elbow", "kmean"])
def test_additive_preconditioner_closed_form(strategy):
    # z = B r + S^T (S A S^T)^{-1} S r over the selected components (MASSolver.cu:583-615)
    _, A, R = _contact(6)
    H = O.Hybrid(_mas_cfg(use_bad_dof_precond=True, bad_dof_strategy=strategy))
    H.factorize(A.colptr, A.rowidx, A.vals)
    comps = H.stiff_components()
    assert len(comps) > 0
    perm, off = H.partition()
    M = O.Mas(A.colptr, A.rowidx, A.vals, 3, perm, off)
    r = problems.random_vector(A.n, 5)
    want = M.apply(r)
    Rd = R.toarray()
    for c in comps:
        want[c] += np.linalg.solve(Rd[np.ix_(c, c)], r[c])
    got = H.precond(r)
    assert np.linalg.norm(got - want) <= 1e-9 * np.linalg.norm(want)


@pytest.mark.parametrize("kw", [{}, {"use_bad_dof_precond": True, "bad_dof_strategy": "gmm"},
                                {"use_preconditioned_residual_norm": True}])
def test_mas_pcg_reaches_the_direct_solution(kw):
    prob, A, R = _contact(6)
    H = O.Hybrid(_mas_cfg(**kw))
    H.factorize(A.colptr, A.rowidx, A.vals)
    b = prob.rhs(1)
    x = H.solve(b, np.ones(A.n))  # the MAS solver ignores the initial guess (x0 = 0)
    info = H.info()
    xd = spla.spsolve(R.tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-3 * np.linalg.norm(xd)
    assert np.linalg.norm(R @ x - b) <= 1e-6 * np.linalg.norm(b) * 1.0001
    assert info["mas_status"] == 1  # relative tolerance
    assert info["num_iterations"] % 10 == 0  # tested every 10 iterations
    x2 = H.solve(b)
    assert np.array_equal(x, x2)
    # max_iter reached: status 3
    H2 = O.Hybrid(_mas_cfg(max_iter=7, **kw))
    H2.factorize(A.colptr, A.rowidx, A.vals)
    H2.solve(b)
    assert H2.info()["mas_status"] == 3 and H2.info()["num_iterations"] == 7
