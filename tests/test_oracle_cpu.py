"""Pins the oracle (CPU restatement of the reference GPU hybrid path) on the CPU.

Each check compares the oracle with an independent implementation, a closed
form or an invariant; none re-types the oracle's own formula.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import independent as ind
from lss_amd import problems
from oracle import oracle as O


def _small(name, N):
    prob = problems.make(name, N=N)
    return prob, prob.csc()


# ---------------------------------------------------------------- reductions
@pytest.mark.parametrize("n", [1, 255, 256, 257, 3000, 70001])
def test_canonical_sum_is_the_tree(n):
    v = problems.random_vector(n, 11) * 1e3
    got = O.canonical_sum(v)
    if n <= 3000:
        assert got == ind.tree_sum(v)  # exact emulation of the documented order
    exact, bound = ind.fsum_bound(v)
    assert abs(got - exact) <= bound


# ---------------------------------------------------------------- generator / RHS
def test_rhs_is_mt19937_uniform():
    # make_random_rhs (tests/linear_solve.cpp:112-132) on libstdc++
    for seed in (0, 1, 12345):
        assert np.array_equal(problems.random_vector(1000, seed), ind.mt19937_uniform(1000, seed))


@pytest.mark.parametrize("name", ["laplace_sphere", "elastic_slivers", "elastic_contact", "elastic_logE"])
def test_generator_spd_symmetric(name):
    prob, A = _small(name, 5)
    S = A.to_scipy().toarray()
    assert np.array_equal(S, S.T)
    assert np.linalg.eigvalsh(S)[0] > 0


def test_elasticity_rigid_modes():
    prob = problems.Problem("elasticity", 5, dirichlet_axis=-1)
    S = prob.csc().to_scipy()
    N, h = 5, 0.25
    idx = np.arange(N ** 3)
    X = np.stack([idx % N, (idx // N) % N, idx // N // N], 1) * h
    for d in range(3):
        u = np.zeros((N ** 3, 3))
        u[:, d] = 1.0
        assert np.abs(S @ u.ravel()).max() < 1e-12
    u = np.stack([-X[:, 1], X[:, 0], 0 * X[:, 0]], 1)
    assert np.abs(S @ u.ravel()).max() < 1e-12


# ---------------------------------------------------------------- stiff DOF selection
CASES = [("laplace_sphere", 10), ("elastic_contact", 10), ("elastic_slivers", 8), ("elastic_logE", 8),
         ("laplace_sphere", 32)]


@pytest.mark.parametrize("name,N", CASES[:4])
def test_row_l1_storage_order_exact(name, N):
    prob, A = _small(name, N)
    assert np.array_equal(O.row_l1(A.colptr, A.vals), ind.row_l1_loop(A.colptr, A.vals))


@pytest.mark.parametrize("name,N", CASES)
def test_gmm_matches_independent_em(name, N):
    prob, A = _small(name, N)
    x = O.row_l1(A.colptr, A.vals)
    got = O.gmm_select(x, 1e-3, 20, 10.0)
    ref = ind.gmm_numpy(x, 1e-3, 20, 10.0)
    for k in ("mean0", "mean1", "var0", "var1", "w0", "w1"):
        assert got[k] == pytest.approx(ref[k], rel=1e-9), k
    assert got["iters"] == ref["iters"]
    assert bool(got["jumped"]) == ref["jumped"]
    assert np.array_equal(got["set"], ref["set"])
    assert got["w0"] + got["w1"] == pytest.approx(1.0, abs=1e-12)


def test_gmm_separated_point_masses():
    x = np.concatenate([np.full(900, 2.0), np.full(100, 5e6)])
    got = O.gmm_select(x, 1e-3, 20, 10.0)
    assert got["mean0"] == pytest.approx(2.0, rel=1e-12)
    assert got["mean1"] == pytest.approx(5e6, rel=1e-12)
    assert np.array_equal(got["set"], np.arange(900, 1000))


def test_gmm_no_jump_selects_nothing():
    x = 1.0 + 0.01 * problems.random_vector(5000, 3)
    got = O.gmm_select(x, 1e-3, 20, 10.0)
    assert not got["jumped"] and got["set"].size == 0


@pytest.mark.parametrize("name,N", CASES)
def test_filter_expand_decompose(name, N):
    prob, A = _small(name, N)
    R = ind.operator_rows(A.colptr, A.rowidx, A.vals, A.n)
    S0 = O.gmm_select(O.row_l1(A.colptr, A.vals), 1e-3, 20, 10.0)["set"]
    if S0.size == 0:  # force a non-trivial set
        S0 = np.unique(np.concatenate([np.arange(0, A.n, 7), np.arange(5, min(A.n, 400))])).astype(np.int32)
    S1, _ = O.filter_set(A.colptr, A.rowidx, A.vals, S0, 4, 10_000_000, 100.0)
    assert np.array_equal(S1, ind.filter_numpy(R, S0, 4, 10_000_000, 100.0))
    S2 = O.expand_set(A.colptr, A.rowidx, S1)
    assert np.array_equal(S2, ind.expand_numpy(R, S1))
    comps = O.decompose_set(A.colptr, A.rowidx, S2)
    ref = ind.induced_components(R, S2)
    assert len(comps) == len(ref)
    for a, b in zip(comps, ref):
        assert np.array_equal(a, b)


def test_filter_size_limits():
    prob, A = _small("elastic_contact", 10)
    R = ind.operator_rows(A.colptr, A.rowidx, A.vals, A.n)
    S0 = np.arange(0, 300, dtype=np.int32)
    for lo, hi in [(1, 10), (50, 10_000), (400, 10_000)]:
        S1, st = O.filter_set(A.colptr, A.rowidx, A.vals, S0, lo, hi, 100.0)
        assert np.array_equal(S1, ind.filter_numpy(R, S0, lo, hi, 100.0))


# ---------------------------------------------------------------- partition
@pytest.mark.parametrize("name,N", CASES[:4])
def test_partition_invariants(name, N):
    prob, A = _small(name, N)
    bs = prob.block_size
    nb = A.n // bs
    G = sp.csr_matrix((np.ones(A.nnz), A.rowidx, A.colptr), shape=(A.n, A.n))
    Rb = sp.csr_matrix((np.ones(A.n), (np.arange(A.n) // bs, np.arange(A.n))), shape=(nb, A.n))
    B = (Rb @ G @ Rb.T).tocsr()
    B.sort_indices()
    part, nparts = O.partition(B.indptr.astype(np.int64), B.indices, 32)
    sizes = np.bincount(part, minlength=nparts)
    assert sizes.min() >= 1 and sizes.max() <= 32 and sizes.sum() == nb
    deg = np.diff(B.indptr) - (B.diagonal() != 0)
    iso = int((deg == 0).sum())
    assert nparts == -(-(nb - iso) // 32) + -(-iso // 32)


# ---------------------------------------------------------------- MAS
def _dense_mas(A, bs, perm, off, mas):
    """sum_l R_l^T (R_l A R_l^T)_bank^{-1} R_l as a dense matrix, from the coarse maps."""
    n, nb = A.shape[0], A.shape[0] // bs
    Ad = A.toarray()
    M = np.zeros((n, n))
    inv_perm = np.argsort(perm)
    # level 0: parts
    for p in range(off.size - 1):
        nodes = inv_perm[off[p]:off[p + 1]]
        dofs = (nodes[:, None] * bs + np.arange(bs)).ravel()
        M[np.ix_(dofs, dofs)] += np.linalg.inv(Ad[np.ix_(dofs, dofs)])
    for lv in range(mas.levels - 1):
        cmap = np.array([mas.map(perm[v], lv) for v in range(nb)])
        nc = mas.cco[lv]
        Rl = np.zeros((nc * bs, n))
        for v in range(nb):
            for a in range(bs):
                Rl[cmap[v] * bs + a, v * bs + a] = 1.0
        Ac = Rl @ Ad @ Rl.T
        for k in range(-(-nc // 32)):
            c = np.arange(k * 32 * bs, min((k + 1) * 32, nc) * bs)
            blk = Ac[np.ix_(c, c)]
            Mc = np.zeros_like(Ac)
            Mc[np.ix_(c, c)] = np.linalg.inv(blk)
            M += Rl.T @ Mc @ Rl
    return M


@pytest.mark.parametrize("name,N", [("laplace_sphere", 6), ("elastic_contact", 5), ("elastic_logE", 4)])
def test_mas_matches_dense_construction(name, N):
    prob, A = _small(name, N)
    bs = prob.block_size
    H = O.Hybrid(O.config(block_size=bs, do_hybrid_precond=False))
    H.factorize(A.colptr, A.rowidx, A.vals)
    perm, off = H.partition()
    mas = O.Mas(A.colptr, A.rowidx, A.vals, bs, perm, off)
    S = A.to_scipy()
    M = _dense_mas(S, bs, perm, off, mas)
    r = problems.random_vector(A.n, 5)
    z = mas.apply(r)
    # contact banks have condition numbers near 1e8: dense inverses agree with the
    # oracle's Cholesky solves to about eps * cond
    assert np.linalg.norm(z - M @ r) <= 1e-7 * np.linalg.norm(M @ r)
    # B is symmetric positive definite
    r2 = problems.random_vector(A.n, 6)
    assert r2 @ mas.apply(r) == pytest.approx(r @ mas.apply(r2), rel=1e-9)
    assert r @ z > 0


def _block_graph(A, bs):
    nb = -(-A.n // bs)
    G = sp.csr_matrix((np.ones(A.nnz), A.rowidx, A.colptr), shape=(A.n, A.n))
    Rb = sp.csr_matrix((np.ones(A.n), (np.arange(A.n) // bs, np.arange(A.n))), shape=(nb, A.n))
    B = (Rb @ G @ Rb.T).tocsr()
    B.sort_indices()
    return B


def _random_forest(seed, n=2500):
    """SPD M-matrix on random trees plus isolated nodes: parts whose induced graphs fall
    apart into several components, coarse banks that merge partly, and a hierarchy that
    stops before the level cap."""
    rng = np.random.default_rng(seed)
    rows, cols, base = [], [], 0
    for m in (2, 5, 40, 300, 1200, n - 1547 - 60):
        par = np.array([rng.integers(0, v) for v in range(1, m)], dtype=np.int64)
        rows.append(base + np.arange(1, m))
        cols.append(base + par)
        base += m
    nn = base + 60
    r, c = np.concatenate(rows), np.concatenate(cols)
    w = rng.uniform(0.1, 1.0, r.size)
    W = sp.coo_matrix((np.r_[w, w], (np.r_[r, c], np.r_[c, r])), shape=(nn, nn)).tocsr()
    p = rng.permutation(nn)
    W = W[p][:, p]
    S = (sp.diags(np.asarray(W.sum(axis=1)).ravel() + 1.0) - W).tocsc()
    S.sort_indices()
    return problems.CSC(nn, S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.astype(np.float64))


# The coarse-space map of the oracle's MAS against an independent construction (scipy
# connected components per bank of each level, components ranked by smallest member,
# stop when a level merges nothing, at most four coarse levels:
# MASPreconditioner.cu:128-213, 376-481), for every block node and level.
@pytest.mark.parametrize("case", [("laplace_sphere", 10, 1), ("laplace_sphere", 12, 3), ("elastic_contact", 10, 3),
                                  ("elastic_slivers", 8, 3), ("elastic_logE", 9, 3), ("forest", 0, 1),
                                  ("forest", 1, 1), ("elastic_contact", 16, 3)])
def test_mas_coarse_map_matches_components(case):
    name, N, bs = case
    A = _random_forest(N) if name == "forest" else problems.make(name, N=N).csc()
    H = O.Hybrid(O.config(block_size=bs, do_hybrid_precond=False))
    H.factorize(A.colptr, A.rowidx, A.vals)
    perm, off = H.partition()
    mas = O.Mas(A.colptr, A.rowidx, A.vals, bs, perm, off)
    B = _block_graph(A, bs).tocoo()
    nb = B.shape[0]
    Bp = sp.csr_matrix((B.data, (perm[B.row], perm[B.col])), shape=(nb, nb))
    maps, counts = ind.mas_coarse_maps(Bp, off)
    assert mas.levels == len(maps) + 1
    assert list(mas.cco) == counts
    for lv, m in enumerate(maps):
        got = np.array([mas.map(v, lv) for v in range(nb)])
        assert np.array_equal(got, m), lv
    # parts are the level-0 banks: no coarse node spans two parts, and every level
    # maps each bank of 32 coarse nodes into one coarse node range
    part = np.searchsorted(off, np.arange(nb), side="right") - 1
    for c in range(counts[0]):
        assert np.unique(part[maps[0] == c]).size == 1


def test_mas_coarse_map_catches_a_wrong_rule():
    # the independent construction tells a rule that ignores bank boundaries from the
    # right one (a plausible mistake: components of the whole graph at level 1)
    A = problems.make("laplace_sphere", N=12).csc()
    H = O.Hybrid(O.config(block_size=1, do_hybrid_precond=False))
    H.factorize(A.colptr, A.rowidx, A.vals)
    perm, off = H.partition()
    B = _block_graph(A, 1).tocoo()
    nb = B.shape[0]
    Bp = sp.csr_matrix((B.data, (perm[B.row], perm[B.col])), shape=(nb, nb))
    maps, counts = ind.mas_coarse_maps(Bp, off)
    wrong, _ = ind.mas_coarse_maps(Bp, off, bank=10 ** 9)
    assert len(maps) >= 2
    assert not np.array_equal(wrong[1], maps[1])


# ---------------------------------------------------------------- hybrid preconditioner and PCG
def _dense_precond(H, n):
    return np.column_stack([H.precond(np.eye(n)[:, j]) for j in range(n)])


@pytest.mark.parametrize("name,N", [("laplace_sphere", 7), ("elastic_contact", 5)])
def test_bpb_composition_closed_form(name, N):
    prob, A = _small(name, N)
    bs = prob.block_size
    cfg = O.config(block_size=bs, gmm_jump_threshold=10.0, min_subdomain_size=1)
    H = O.Hybrid(cfg)
    H.factorize(A.colptr, A.rowidx, A.vals)
    comps = H.stiff_components()
    if not comps:
        pytest.skip("no stiff set on this instance")
    n = A.n
    Ad = A.to_scipy().toarray()
    omega = H.info()["omega"]
    # B (undamped MAS) and P = S^T (S A S^T)^{-1} S assembled densely
    Hb = O.Hybrid(O.config(block_size=bs, do_hybrid_precond=False))
    Hb.factorize(A.colptr, A.rowidx, A.vals)
    B = _dense_precond(Hb, n)
    Pm = np.zeros((n, n))
    for c in comps:
        Pm[np.ix_(c, c)] = np.linalg.inv(Ad[np.ix_(c, c)])
    I = np.eye(n)
    Ai = np.linalg.inv(Ad)
    Bw = omega * B
    Minv = Ai - (I - Bw @ Ad) @ (I - Pm @ Ad) @ (I - Bw @ Ad) @ Ai  # GPUHybrid.cu:521-542
    M = _dense_precond(H, n)
    assert np.linalg.norm(M - Minv) <= 1e-7 * np.linalg.norm(Minv)
    assert np.linalg.norm(M - M.T) <= 1e-7 * np.linalg.norm(M)
    assert np.linalg.eigvalsh(0.5 * (M + M.T))[0] > 0


def test_lanczos_damping_bounds_lambda_max():
    prob, A = _small("elastic_contact", 5)
    H = O.Hybrid(O.config(block_size=3, gmm_jump_threshold=10.0, min_subdomain_size=1, damping=1.0))
    H.factorize(A.colptr, A.rowidx, A.vals)
    if not H.stiff_components():
        pytest.skip("no stiff set")
    theta = 1.0 / H.info()["omega"]
    Hb = O.Hybrid(O.config(block_size=3, do_hybrid_precond=False))
    Hb.factorize(A.colptr, A.rowidx, A.vals)
    B = _dense_precond(Hb, A.n)
    lam = np.max(np.linalg.eigvals(B @ A.to_scipy().toarray()).real)
    assert theta <= lam * (1 + 1e-9)
    assert theta >= 0.9 * lam


@pytest.mark.parametrize("name,N", [("laplace_sphere", 12), ("elastic_contact", 8), ("elastic_logE", 7)])
def test_hybrid_solve_matches_direct(name, N):
    prob, A = _small(name, N)
    b = prob.rhs(0)
    H = O.Hybrid(O.config(block_size=prob.block_size, rel_tol=1e-10, gmm_jump_threshold=10.0, min_subdomain_size=4))
    H.factorize(A.colptr, A.rowidx, A.vals)
    x = H.solve(b)
    xd = spla.spsolve(A.to_scipy().tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)
    inf = H.info()
    assert inf["final_res_norm"] == pytest.approx(np.linalg.norm(A.to_scipy() @ x - b), rel=1e-6)


def test_zero_rhs_gives_zero_solution():
    prob, A = _small("elastic_contact", 6)
    H = O.Hybrid(O.config(block_size=3))
    H.factorize(A.colptr, A.rowidx, A.vals)
    x = H.solve(np.zeros(A.n), np.ones(A.n))
    assert np.all(x == 0) and H.info()["num_iterations"] == 0


def test_damping_retry_recovers_from_breakdown():
    # damping 3.5 puts omega * lambda_max(BA) far above 2: the damped sweep is
    # indefinite, the first PCG breaks down, the retry with damping 1.5 converges
    prob, A = _small("elastic_contact", 8)
    b = prob.rhs(0)
    H = O.Hybrid(O.config(block_size=3, rel_tol=1e-10, gmm_jump_threshold=10.0, min_subdomain_size=4, damping=3.5))
    H.factorize(A.colptr, A.rowidx, A.vals)
    x = H.solve(b)
    assert H.info()["damping_retry"]
    xd = spla.spsolve(A.to_scipy().tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)
