"""ORACLE — test infrastructure only.

CPU restatement of the AMG base preconditioner of the GPU hybrid path: hypre's
BoomerAMG with the options of src/polysolve/linear/GPUHybrid.cu:285-339 (PMIS
coarsening, strength threshold theta with num_functions = block size, extended+i
interpolation truncated to 4 entries, Galerkin R A P with R = P^T, l1-Jacobi
smoothing, one V-cycle, GPUHybrid.cu:677-689).  hypre v3.1.0 is not part of the
reference tree (cmake/recipes/hypre.cmake fetches it); this follows the published
methods it implements (references in local-stiff-solver_amd/csrc/host/amg.hpp),
with the random part of PMIS's measure a hash of the point index so that the
splitting is reproducible.  Plain Python sets and loops over rows: for systems of
a few thousand unknowns.  Only tests/ import it.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

MASK64 = (1 << 64) - 1


def splitmix64(i: int) -> int:
    """Vigna's splitmix64 output for the state i: mix(i + golden gamma)."""
    z = (i + 0x9E3779B97F4A7C15) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def point_random(i: int) -> float:
    """splitmix64(i) scaled to [0, 1) with 53 bits."""
    return (splitmix64(i) >> 11) / 9007199254740992.0


def _rows(A: sp.csr_matrix):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return [(A.indices[A.indptr[i]:A.indptr[i + 1]], A.data[A.indptr[i]:A.indptr[i + 1]]) for i in range(A.shape[0])]


def strength(A, func, num_functions, theta, max_row_sum=0.9):
    """S[i] = set of j that i depends on strongly."""
    rows = _rows(A)
    S = []
    for i, (cols, vals) in enumerate(rows):
        d = {int(j): float(v) for j, v in zip(cols, vals)}
        diag = d.get(i, 0.0)
        same = [(j, v) for j, v in d.items() if j != i and (num_functions == 1 or func[j] == func[i])]
        if diag < 0:
            scale = max([0.0] + [v for _, v in same])
        else:
            scale = min([0.0] + [v for _, v in same])
        row_sum = diag
        for _, v in same:
            row_sum += v
        if max_row_sum < 1.0 and abs(row_sum) > abs(diag) * max_row_sum:
            S.append(set())
            continue
        cut = theta * scale
        S.append({j for j, v in same if (v > cut if diag < 0 else v < cut)})
    return S


def pmis(S):
    """cf[i] = 1 (C), -1 (F), -3 (no strong connection)."""
    n = len(S)
    ST = [set() for _ in range(n)]
    for i, s in enumerate(S):
        for j in s:
            ST[j].add(i)
    meas = [len(ST[i]) + point_random(i) for i in range(n)]
    cf = [0] * n
    for i in range(n):
        if not S[i] and not ST[i]:
            cf[i] = -3
        elif not ST[i]:
            cf[i] = -1

    def key(i):
        return (meas[i], i)

    while any(c == 0 for c in cf):
        und = [i for i in range(n) if cf[i] == 0]
        new_c = [i for i in und if all(cf[j] != 0 or key(j) < key(i) for j in S[i] | ST[i])]
        for i in new_c:
            cf[i] = 1
        for i in und:
            if cf[i] == 0 and any(cf[j] == 1 for j in S[i]):
                cf[i] = -1
    return cf


def interp_extpi(A, S, cf, func, num_functions, pmax=4):
    """Extended+i interpolation, truncated to pmax entries per row (row sum kept)."""
    rows = _rows(A)
    n = len(rows)
    cidx = {}
    for i in range(n):
        if cf[i] == 1:
            cidx[i] = len(cidx)
    A_d = [dict(zip(map(int, c), map(float, v))) for c, v in rows]
    P_rows = []
    for i in range(n):
        if cf[i] == 1:
            P_rows.append({cidx[i]: 1.0})
            continue
        if cf[i] == -3 or not S[i]:
            P_rows.append({})
            continue
        # C-hat_i in the order points are met: strong neighbours ascending, each strong
        # F neighbour followed by its strong C neighbours
        Fs = set()
        chat = []
        inchat = set()
        for j in sorted(S[i]):
            if cf[j] == 1:
                if j not in inchat:
                    inchat.add(j)
                    chat.append(j)
            else:
                Fs.add(j)
                for l in sorted(S[j]):
                    if cf[l] == 1 and l not in inchat:
                        inchat.add(l)
                        chat.append(l)
        # one pass over row i in column order, each product and sum as it comes
        w = {j: 0.0 for j in chat}
        diag = 0.0
        for j, a in A_d[i].items():
            if j == i:
                diag += a
            elif j in inchat:
                w[j] += a
            elif j in Fs:
                akk = A_d[j].get(j, 0.0)
                d = 0.0
                for l, v in A_d[j].items():
                    if (l == i or l in inchat) and v * akk < 0.0:
                        d += v
                if d == 0.0:
                    diag += a
                    continue
                fac = a / d
                for l, v in A_d[j].items():
                    if v * akk >= 0.0:
                        continue
                    if l == i:
                        diag += fac * v
                    elif l in inchat:
                        w[l] += fac * v
            elif num_functions == 1 or func[j] == func[i]:
                diag += a
        if diag == 0.0 or not chat:
            P_rows.append({})
            continue
        w = {j: -w[j] / diag for j in chat}
        if len(chat) > pmax:
            order = sorted(chat, key=lambda j: (-abs(w[j]), j))
            total = 0.0
            for j in chat:
                total += w[j]
            kept = 0.0
            for j in order[:pmax]:
                kept += w[j]
            scale = total / kept if kept != 0.0 else 1.0
            w = {j: w[j] * scale for j in order[:pmax]}
        P_rows.append({cidx[j]: v for j, v in w.items()})
    nc = len(cidx)
    r, c, v = [], [], []
    for i, row in enumerate(P_rows):
        for j in sorted(row):
            r.append(i)
            c.append(j)
            v.append(row[j])
    return sp.csr_matrix((v, (r, c)), shape=(n, nc))


class Hierarchy:
    """AMG levels: A[l], P[l] (to level l from l + 1), cf[l], dinv[l]."""

    def __init__(self, A, block_size=1, theta=None, max_levels=25, min_coarse=5, pmax=4):
        nf = block_size
        theta = (0.25 if block_size == 1 else 0.5) if theta is None else theta
        A = sp.csr_matrix(A, dtype=np.float64)
        func = [i % nf for i in range(A.shape[0])]
        self.A, self.P, self.cf, self.dinv = [], [], [], []
        while True:
            self.A.append(A)
            self.dinv.append(l1_dinv(A))
            n = A.shape[0]
            if len(self.A) >= max_levels or n <= min_coarse:
                break
            S = strength(A, func, nf, theta)
            cf = pmis(S)
            nc = sum(1 for c in cf if c == 1)
            if nc == 0 or nc == n:
                break
            P = interp_extpi(A, S, cf, func, nf, pmax)
            self.cf.append(cf)
            self.P.append(P)
            func = [func[i] for i in range(n) if cf[i] == 1]
            A = sp.csr_matrix(P.T.tocsr() @ (A @ P))

    def vcycle(self, b):
        return vcycle(self.A, self.P, self.dinv, b)


def vcycle(As, Ps, dinvs, b, level=0):
    """One V-cycle from x = 0: l1-Jacobi before and after the coarse correction, one
    l1-Jacobi sweep on the coarsest level (GPUHybrid.cu:677-689, one BoomerAMG
    iteration)."""
    A, dinv = As[level], dinvs[level]
    x = dinv * b
    if level == len(As) - 1:
        return x
    P = Ps[level]
    xc = vcycle(As, Ps, dinvs, P.T @ (b - A @ x), level + 1)
    x = x + P @ xc
    return x + dinv * (b - A @ x)


def l1_dinv(A):
    """1 / sum_j |a_ij| per row (l1-Jacobi)."""
    return 1.0 / np.asarray(abs(sp.csr_matrix(A)).sum(axis=1)).ravel()
