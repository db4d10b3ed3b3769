"""Nested-dissection ordering and multifrontal structure of the sparse stiff
factorization (host code of liblss.so, no GPU).

Pinned against an independent symbolic Cholesky: for the graph permuted by the
returned order, the column structures of L are computed by elimination-tree
recursion (struct L(:, j) = upper neighbours of j, plus the structures of its
etree children, minus j).  Then
  * every column j of front f has its structure inside the front: later pivots of f
    or f's update positions (a missing row would drop fill and give a wrong factor);
  * the union of the structures of f's pivot columns beyond its pivots is exactly
    f's update set (no superfluous rows);
  * fronts are contiguous in postorder, parents come after children, update
    positions are ascending and belong to ancestors;
  * on a 3D 7-point grid the ordering fills far less than the natural order.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from lss_amd.solver import nested_dissection


def _grid(n, bs=1, stencil=27):
    """Node graph of an n^3 grid (7- or 27-point neighbourhoods) with bs coupled DOFs per node."""
    idx = np.arange(n ** 3).reshape(n, n, n)
    rows, cols = [], []
    offs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)
            if (a, b, c) != (0, 0, 0) and (stencil == 27 or abs(a) + abs(b) + abs(c) == 1)]
    for a, b, c in offs:
        src = idx[max(0, -a):n - max(0, a), max(0, -b):n - max(0, b), max(0, -c):n - max(0, c)]
        dst = idx[max(0, a):n + min(0, a) or n, max(0, b):n + min(0, b) or n, max(0, c):n + min(0, c) or n]
        rows.append(src.ravel())
        cols.append(dst.ravel())
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    G = sp.coo_matrix((np.ones(r.size), (r, c)), shape=(n ** 3, n ** 3)).tocsr()
    if bs > 1:
        G = sp.kron(G + sp.eye(n ** 3), np.ones((bs, bs))).tocsr()
        G.setdiag(0)
        G.eliminate_zeros()
    return _sym(G)


def _sym(G):
    G = (G + G.T).tocsr()
    G.setdiag(0)
    G.eliminate_zeros()
    G.sort_indices()
    return G


def _random(n, deg, seed, pieces=1):
    rng = np.random.default_rng(seed)
    blocks = []
    for p in range(pieces):
        k = n // pieces
        r = rng.integers(0, k, size=k * deg // 2)
        c = np.clip(r + rng.integers(-20, 21, size=r.size), 0, k - 1)  # banded-ish, like a mesh
        blocks.append(sp.coo_matrix((np.ones(r.size), (r, c)), shape=(k, k)))
    return _sym(sp.block_diag(blocks).tocsr())


def _lstruct(G, order):
    """Column structures of the Cholesky factor of G permuted by order (position -> vertex)."""
    m = G.shape[0]
    pos = np.empty(m, np.int64)
    pos[order] = np.arange(m)
    B = G[order][:, order].tocsr()
    upper = [set(B.indices[B.indptr[j]:B.indptr[j + 1]][B.indices[B.indptr[j]:B.indptr[j + 1]] > j].tolist())
             for j in range(m)]
    struct = [None] * m
    kids = [[] for _ in range(m)]
    for j in range(m):
        s = set(upper[j])
        for c in kids[j]:
            s |= struct[c]
        s.discard(j)
        struct[j] = s
        if s:
            kids[min(s)].append(j)
    return struct


def _check(G, leaf):
    m = G.shape[0]
    r = nested_dissection(G.indptr, G.indices, leaf)
    order, p0, npv, parent = r["order"], r["p0"], r["np"], r["parent"]
    sptr, sidx = r["struct_ptr"], r["struct_idx"]
    assert np.array_equal(np.sort(order), np.arange(m))
    F = p0.size
    assert F >= 1 and p0[0] == 0 and np.all(npv >= 1)
    assert np.array_equal(p0[1:], p0[:-1] + npv[:-1]) and p0[-1] + npv[-1] == m
    assert np.all((parent == -1) | (parent > np.arange(F)))
    front_of = np.repeat(np.arange(F), npv)
    U = [sidx[sptr[f]:sptr[f + 1]] for f in range(F)]
    for f in range(F):
        u = U[f]
        assert np.all(np.diff(u) > 0)
        assert np.all(u >= p0[f] + npv[f])
        # update positions are pivots of ancestors
        anc, a = set(), parent[f]
        while a != -1:
            anc.add(a)
            a = parent[a]
        assert all(front_of[q] in anc for q in u.tolist())
    struct = _lstruct(G, order)
    for f in range(F):
        p1 = p0[f] + npv[f]
        allowed = set(range(p0[f], p1)) | set(U[f].tolist())
        union = set()
        for j in range(p0[f], p1):
            assert struct[j] <= allowed, (f, j)
            union |= {q for q in struct[j] if q >= p1}
        assert union == set(U[f].tolist()), f
    nnz_l = sum(len(s) for s in struct) + m
    return r, nnz_l


@pytest.mark.parametrize("n,bs,stencil,leaf", [(6, 1, 7, 8), (7, 1, 27, 16), (5, 3, 27, 24), (6, 3, 27, 64),
                                               (4, 2, 27, 1)])
def test_grid_structure_matches_symbolic_cholesky(n, bs, stencil, leaf):
    G = _grid(n, bs, stencil)
    r, _ = _check(G, leaf)
    assert r["p0"].size > 1  # the grid was dissected


def test_dissection_reduces_fill_on_grid():
    G = _grid(16, 1, 7)
    _, nnz_nd = _check(G, 64)
    nnz_nat = sum(len(s) for s in _lstruct(G, np.arange(G.shape[0]))) + G.shape[0]
    assert nnz_nd < 0.5 * nnz_nat


@pytest.mark.parametrize("n,deg,seed,pieces,leaf", [(400, 6, 0, 1, 16), (600, 8, 1, 3, 32), (300, 4, 2, 5, 300),
                                                    (257, 10, 3, 2, 7)])
def test_random_graph_structure_matches_symbolic_cholesky(n, deg, seed, pieces, leaf):
    _check(_random(n, deg, seed, pieces), leaf)


def test_edge_cases():
    # no edges: every vertex its own component
    r, _ = _check(sp.csr_matrix((50, 50)), 4)
    assert r["struct_idx"].size == 0
    # one vertex, a path, a star (the centre separates every leaf: no fill at all)
    _check(sp.csr_matrix((1, 1)), 1)
    path = _sym(sp.diags([np.ones(99)], [1], shape=(100, 100)).tocsr())
    _check(path, 3)
    star = _sym(sp.coo_matrix((np.ones(40), (np.zeros(40, int), np.arange(1, 41))), shape=(41, 41)).tocsr())
    r, nnz_star = _check(star, 4)
    assert r["order"][-1] == 0 and nnz_star == 41 + 40
    # a clique has no separating level: one front
    clique = _sym(sp.csr_matrix(np.ones((30, 30))))
    r, _ = _check(clique, 4)
    assert r["p0"].size == 1
    # the same graph gives the same ordering
    G = _grid(5, 1, 27)
    assert np.array_equal(nested_dissection(G.indptr, G.indices, 8)["order"],
                          nested_dissection(G.indptr, G.indices, 8)["order"])


def test_same_ordering_from_new_threads():
    # the stiff factorization runs on a worker thread created per factorization; the
    # dissection's parallel phase must give the main thread's ordering from any thread
    import threading

    G = _grid(20, 3, 27)
    ref = nested_dissection(G.indptr, G.indices, 32)
    assert ref["p0"].size > 64  # deep enough for the parallel phase below the top levels
    out = []
    for _ in range(3):
        t = threading.Thread(target=lambda: out.append(nested_dissection(G.indptr, G.indices, 32)))
        t.start()
        t.join()
    assert len(out) == 3
    for r in out:
        for key in ("order", "p0", "np", "parent", "struct_ptr", "struct_idx"):
            assert np.array_equal(r[key], ref[key]), key
