"""Independent numpy / scipy implementations used to pin the oracle.

Written from the mathematics of each step (and the reference's semantics named
in each docstring), not from the oracle's code, so that a dropped term, a wrong
index or a transposed operand in the oracle shows up as a mismatch.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

# Reduction tree of the device kernels and the oracle: 1024 contiguous chunks,
# 256 lanes per chunk summing strided elements in order, pairwise trees.
CHUNKS, LANES = 1024, 256


def tree_sum(values) -> float:
    """Pure-Python emulation of the canonical reduction tree (exact IEEE order)."""
    v = [float(a) for a in values]
    n = len(v)
    if n == 0:
        return 0.0
    clen = -(-n // CHUNKS)
    clen = -(-clen // LANES) * LANES
    part = []
    for b in range(CHUNKS):
        lo, hi = b * clen, min(b * clen + clen, n)
        acc = []
        for t in range(LANES):
            a = 0.0
            for i in range(lo + t, hi, LANES):
                a += v[i]
            acc.append(a)
        s = LANES // 2
        while s > 0:
            for t in range(s):
                acc[t] += acc[t + s]
            s //= 2
        part.append(acc[0])
    s = CHUNKS // 2
    while s > 0:
        for t in range(s):
            part[t] += part[t + s]
        s //= 2
    return part[0]


def row_l1_loop(colptr, vals) -> np.ndarray:
    """Row l1 norms summed sequentially in storage order (GPUHybrid.cu:753-765)."""
    out = np.empty(len(colptr) - 1)
    for i in range(len(colptr) - 1):
        s = 0.0
        for k in range(colptr[i], colptr[i + 1]):
            s += abs(float(vals[k]))
        out[i] = s
    return out


def gmm_numpy(x, tol=1e-3, max_iter=20, jump=10.0):
    """Two-component 1-D Gaussian mixture EM (GPUHybrid.cu:767-883), vectorized numpy."""
    x = np.asarray(x, np.float64)
    n = x.size
    mean = x.sum() / n
    var = ((x - mean) ** 2).sum() / n
    m = np.array([x.min(), x.max()])
    v = np.array([var, var])
    w = np.array([0.5, 0.5])
    it = 0
    g = None
    while it < max_iter:
        logp = np.log(w)[:, None] - 0.5 * np.log(2 * np.pi * v)[:, None] - 0.5 * (x[None, :] - m[:, None]) ** 2 / v[:, None]
        mx = logp.max(0)
        tot = mx + np.log(np.exp(logp - mx).sum(0))
        g = np.exp(logp - tot)
        sg = g.sum(1)
        w = sg / n
        om, ov = m.copy(), v.copy()
        m = (g * x).sum(1) / sg
        v = (g * (x[None, :] - m[:, None]) ** 2).sum(1) / sg + 1e-6
        if np.all(np.abs(m - om) / np.abs(om) < tol) and np.all(np.abs(v - ov) / np.abs(ov) < tol):
            break
        it += 1  # index of the converging pass, or max_iter when none converged
    jumped = abs(m[1]) / abs(m[0]) > jump
    sel = np.flatnonzero(g[0] < g[1]) if jumped else np.zeros(0, np.int64)
    return {"mean0": m[0], "mean1": m[1], "var0": v[0], "var1": v[1], "w0": w[0], "w1": w[1], "iters": it,
            "jumped": bool(jumped), "set": sel}


def operator_rows(colptr, rowidx, vals, n):
    """The solver's operator: CSC column j read as row j (GPUHybrid.cu:401-433)."""
    return sp.csr_matrix((vals, rowidx, colptr), shape=(n, n))


def induced_components(R: sp.csr_matrix, s: np.ndarray):
    """Connected components of the stored-entry graph induced on the set s."""
    s = np.asarray(s, np.int64)
    sub = R[s][:, s].tocsr()
    pat = sp.csr_matrix((np.ones(sub.nnz), sub.indices, sub.indptr), shape=sub.shape)
    nc, lab = connected_components(pat, directed=False)
    comps = [np.sort(s[lab == c]) for c in range(nc)]
    comps.sort(key=lambda c: c[0])
    return comps


def filter_numpy(R: sp.csr_matrix, s, min_size, max_size, cond):
    """Component filter with the Gershgorin bound (GPUHybrid.cu:885-970)."""
    s = np.asarray(s, np.int64)
    inset = np.zeros(R.shape[0], bool)
    inset[s] = True
    kept = []
    for c in induced_components(R, s):
        if c.size < min_size or c.size > max_size:
            continue
        lmin, lmax = np.inf, 0.0
        for k in c:
            cols = R.indices[R.indptr[k]:R.indptr[k + 1]]
            v = R.data[R.indptr[k]:R.indptr[k + 1]]
            msk = inset[cols]
            d = v[msk & (cols == k)]
            d = d[-1] if d.size else 0.0
            off = np.abs(v[msk & (cols != k)]).sum()
            lmin, lmax = min(lmin, d - off), max(lmax, d + off)
        if lmin * lmax < 0 or lmax / lmin > cond:
            kept.append(c)
    return np.sort(np.concatenate(kept)) if kept else np.zeros(0, np.int64)


def expand_numpy(R: sp.csr_matrix, s):
    """One-ring closure over stored entries (GPUHybrid.cu:972-991)."""
    s = np.asarray(s, np.int64)
    if s.size == 0:
        return s
    return np.unique(np.concatenate([R.indices[R.indptr[k]:R.indptr[k + 1]] for k in s]))


def mt19937_uniform(n: int, seed: int, lo=-1.0, hi=1.0) -> np.ndarray:
    """std::uniform_real_distribution<double>(lo, hi) on std::mt19937(seed), libstdc++.

    The engine is the standard MT19937 seeded by init_genrand (numpy's legacy
    RandomState integer seeding); generate_canonical<double, 53> combines two
    32-bit outputs as (g1 + g2 * 2^32) / 2^64, clamped below 1.
    """
    rs = np.random.RandomState(seed)
    raw = rs.randint(0, 2**32, size=2 * n, dtype=np.uint64).astype(np.float64)
    g1, g2 = raw[0::2], raw[1::2]
    u = (g1 + g2 * 4294967296.0) / 18446744073709551616.0
    u = np.where(u >= 1.0, np.nextafter(1.0, 0.0), u)
    return (hi - lo) * u + lo


def fsum_bound(values) -> tuple[float, float]:
    """Exact sum and a first-order bound for any summation order."""
    v = np.asarray(values, np.float64)
    return math.fsum(v.tolist()), float(np.abs(v).sum()) * v.size * np.finfo(np.float64).eps


# ---------------------------------------------------------------- MAS solver selection
def permuted_rows(R: sp.csr_matrix, bs: int, perm):
    """Operator R renumbered so that node I becomes perm[I] (scalar p = perm[I]*bs + a)."""
    n = R.shape[0]
    perm = np.asarray(perm, np.int64)
    sp_perm = (perm[np.arange(n) // bs] * bs + np.arange(n) % bs)
    P = sp.csr_matrix((np.ones(n), (sp_perm, np.arange(n))), shape=(n, n))
    return (P @ R @ P.T).tocsr(), sp_perm


def gmm_mas_numpy(x, tol=1e-3, max_iter=5, jump=10.0, expand=32):
    """select_bad_dofs_from_l1_row_norm_gmm (BadDofGMMPreconditioner.cu:86-258), vectorized numpy."""
    x = np.asarray(x, np.float64)
    n = x.size
    if n < 2:
        return np.zeros(0, np.int64), {}
    gm = x.sum() / n
    gv = ((x - gm) ** 2).sum() / n
    m = np.array([x.min(), x.max()])
    v = np.array([max(gv, 1e-6)] * 2)
    w = np.array([0.5, 0.5])
    if m[0] == m[1] or not np.isfinite(gv):
        return np.zeros(0, np.int64), {}
    it, g = 0, None
    while it < max(1, max_iter):
        logp = (np.log(np.maximum(w, np.finfo(float).tiny))[:, None] - 0.5 * np.log(2 * np.pi * v)[:, None]
                - 0.5 * (x[None, :] - m[:, None]) ** 2 / v[:, None])
        mx = logp.max(0)
        g = np.exp(logp - (mx + np.log(np.exp(logp - mx).sum(0))))
        sg = g.sum(1)
        if sg[0] <= 0 or sg[1] <= 0:
            break
        w = sg / n
        om, ov = m.copy(), v.copy()
        m = (g * x).sum(1) / sg
        v = (g * (x[None, :] - m[:, None]) ** 2).sum(1) / sg + 1e-6
        rc = lambda a, b: np.abs(a - b) / np.maximum(np.abs(b), 1e-300)  # noqa: E731
        if np.all(rc(m, om) < tol) and np.all(rc(v, ov) < tol):
            break
        it += 1
    hi = 0 if abs(m[0]) > abs(m[1]) else 1
    ratio = abs(m[hi]) / abs(m[1 - hi])
    st = {"mean0": m[0], "mean1": m[1], "var0": v[0], "var1": v[1], "ratio": ratio, "iters": it}
    if np.isnan(ratio) or ratio < jump:
        return np.zeros(0, np.int64), st
    sel = np.flatnonzero(g[hi] > g[1 - hi])
    return window_union(sel, n, expand), st


def window_union(sel, n, e):
    """Positions within e of a selected position (BadDofGMMPreconditioner.cu:236-255)."""
    sel = np.asarray(sel, np.int64)
    if e <= 0 or sel.size == 0:
        return np.sort(sel)
    mark = np.zeros(n + 1, np.int64)
    np.add.at(mark, np.maximum(sel - e, 0), 1)
    np.add.at(mark, np.minimum(sel + e, n - 1) + 1, -1)
    return np.flatnonzero(np.cumsum(mark)[:n] > 0)


def elbow_numpy(x):
    """select_bad_dofs_from_diagonal (BadDOFPreconditioner.cu:61-100); ties by position."""
    x = np.asarray(x, np.float64)
    n = x.size
    o = np.argsort(x, kind="stable")
    if n < 2:
        return np.sort(o), 0
    lg = np.log(x[o])
    expected = (lg[-1] - lg[0]) / (n - 1.0) * np.arange(n) + lg[0]
    cut = int(np.argmax(np.abs(expected - lg)))  # first maximum
    return np.sort(o[cut:]), cut


def kmeans_numpy(x, fraction=0.02, jump=10.0, max_iter=32, expand=0):
    """select_bad_dofs_from_diagonal_kmeans (BadDofKMeanPreconditioner.cu:63-172)."""
    x = np.asarray(x, np.float64)
    n = x.size
    if n < 2:
        return np.zeros(0, np.int64), {}
    o = np.argsort(-x, kind="stable")
    sub = int(np.clip(int(n * fraction), 2, n))
    v = x[o[:sub]]
    c = np.array([v[0], v[-1]])
    it = 0
    while it < max_iter:
        cid = (np.abs(v - c[0]) < np.abs(v - c[1])).astype(int)
        cnt = np.bincount(cid, minlength=2)
        s = np.array([v[cid == 0].sum(), v[cid == 1].sum()])
        nxt = np.array([s[0] / cnt[0] if cnt[0] > 0 else c[0], s[1] / cnt[1] if cnt[1] > 1 else c[1]])
        if np.array_equal(nxt, c):
            break
        c = nxt
        it += 1
    c = np.sort(c)[::-1]
    ratio = c[0] / c[1]
    st = {"c0": c[0], "c1": c[1], "ratio": ratio, "iters": it}
    if np.isnan(ratio) or ratio < jump:
        return np.zeros(0, np.int64), st
    sel = o[:sub][v > 0.5 * (c[0] + c[1])]
    return window_union(sel, n, expand), st


def mas_coarse_maps(Bg: sp.csr_matrix, part_off, max_coarse=4, bank=32):
    """Coarse-space maps of the MAS, from the rule of MASPreconditioner.cu:128-213, 376-481.

    Bg: block graph in partition order (node v of part p lies in
    [part_off[p], part_off[p+1])).  Level 0: the coarse nodes are the connected
    components of each part's induced subgraph.  Level l >= 1: the level l-1
    coarse nodes form banks of `bank` consecutive ids, the graph is the quotient
    of the block graph, and the coarse nodes are the components of each bank's
    induced subgraph.  Inside a bank (and so overall, banks being ascending id
    ranges) components are numbered by their smallest member.  A level that
    merges nothing ends the construction; at most `max_coarse` levels.
    Returns (maps, counts): maps[l][v] = level-l coarse id of block node v.
    Uses scipy's connected_components per level, not the lane masks."""
    nb = Bg.shape[0]
    part_off = np.asarray(part_off, np.int64)
    G = sp.coo_matrix(Bg)
    u, w = G.row.astype(np.int64), G.col.astype(np.int64)
    maps, counts = [], []
    cur = np.arange(nb, dtype=np.int64)  # level l-1 id of every block node (identity before level 0)
    ncur = nb
    bank_of_fine = np.searchsorted(part_off, np.arange(nb), side="right") - 1
    for lv in range(max_coarse):
        if lv == 0:
            bk = bank_of_fine  # parts are the level-0 banks
        else:
            bk = np.arange(ncur) // bank
        a, c = cur[u], cur[w]
        keep = bk[a] == bk[c]
        E = sp.coo_matrix((np.ones(int(keep.sum())), (a[keep], c[keep])), shape=(ncur, ncur))
        ncomp, lab = connected_components(E, directed=False)
        mins = np.full(ncomp, ncur, np.int64)
        np.minimum.at(mins, lab, np.arange(ncur))
        rank = np.empty(ncomp, np.int64)
        rank[np.argsort(mins, kind="stable")] = np.arange(ncomp)
        newid = rank[lab]
        if lv > 0 and ncomp == ncur:
            break
        cur = newid[cur]
        maps.append(cur.copy())
        counts.append(ncomp)
        ncur = ncomp
    return maps, counts
