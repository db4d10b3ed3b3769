"""Writes the golden fixtures for the reference's own example systems.

Inputs are the reference's data files mats/A_contact.mtx (4374 DOF, block 3)
and mats/A_pq.mtx (130569 DOF, block 1), read as tests/linear_solve.cpp:320-346
reads them (Eigen loadMarket: compressed column storage, rows ascending in each
column, duplicates summed; no forced symmetry).  Every golden value is computed
by the independent numpy/scipy implementations in tests/independent.py and
checked against the oracle before it is written; a disagreement aborts.

Outputs (committed):
  A_contact.npz          the CSC arrays of A_contact and its reference RHS
                         (tests/linear_solve.cpp:112-132, seed 0)
  A_contact_golden.npz   row l1 norms, the stiff set after each selection stage,
                         the component offsets and a direct solution
  A_pq.npz, A_pq_golden.npz   the same for A_pq
  golden.json            scalar goldens for both matrices (GMM parameters,
                         stage sizes, component sizes, set digests)

Run from the repository root with the reference mounted:
  python tests/golden/make_golden.py [/root/reference/mats]
"""
import hashlib
import json
import os
import sys

import numpy as np
import scipy.io
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import independent as ind  # noqa: E402
from oracle import oracle as O  # noqa: E402

# spec/GPUHybrid.json:1-19
SPEC = {"gmm_jump_threshold": 10.0, "gmm_tol": 1e-3, "max_gmm_iterations": 20, "min_subdomain_size": 4,
        "max_subdomain_size": 10_000_000, "conditioning_threshold": 100.0}
MATS = {"A_contact": 3, "A_pq": 1}


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.int64).tobytes()).hexdigest()[:16]


def load_csc(path):
    A = scipy.io.mmread(path).tocsc()
    A.sum_duplicates()
    A.sort_indices()
    return A


def stages(A: sp.csc_matrix):
    n = A.shape[0]
    colptr, rowidx, vals = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data
    R = ind.operator_rows(colptr, rowidx, vals, n)
    l1 = ind.row_l1_loop(colptr, vals)
    assert np.array_equal(l1, O.row_l1(colptr, vals))
    g = ind.gmm_numpy(l1, SPEC["gmm_tol"], SPEC["max_gmm_iterations"], SPEC["gmm_jump_threshold"])
    go = O.gmm_select(l1, SPEC["gmm_tol"], SPEC["max_gmm_iterations"], SPEC["gmm_jump_threshold"])
    for k in ("mean0", "mean1", "var0", "var1", "w0", "w1"):
        assert abs(g[k] - go[k]) <= 1e-9 * abs(g[k]), k
    assert g["iters"] == go["iters"] and np.array_equal(g["set"], go["set"])
    s0 = g["set"]
    s1 = ind.filter_numpy(R, s0, SPEC["min_subdomain_size"], SPEC["max_subdomain_size"],
                          SPEC["conditioning_threshold"])
    o1, _ = O.filter_set(colptr, rowidx, vals, s0.astype(np.int32), SPEC["min_subdomain_size"],
                         SPEC["max_subdomain_size"], SPEC["conditioning_threshold"])
    assert np.array_equal(s1, o1)
    s2 = ind.expand_numpy(R, s1)
    assert np.array_equal(s2, O.expand_set(colptr, rowidx, s1.astype(np.int32)))
    comps = ind.induced_components(R, s2)
    oc = O.decompose_set(colptr, rowidx, s2.astype(np.int32))
    assert len(comps) == len(oc) and all(np.array_equal(a, b) for a, b in zip(comps, oc))
    return l1, g, s0, s1, s2, comps


def main(mats_dir):
    out = {}
    for name, bs in MATS.items():
        A = load_csc(os.path.join(mats_dir, name + ".mtx"))
        n = A.shape[0]
        l1, g, s0, s1, s2, comps = stages(A)
        b = ind.mt19937_uniform(n, 0)
        rec = {
            "n": n, "nnz": int(A.nnz), "block_size": bs,
            "row_l1_fsum": float(np.sum(np.sort(l1))), "row_l1_digest": hashlib.sha256(l1.tobytes()).hexdigest()[:16],
            "gmm": {k: float(g[k]) for k in ("mean0", "mean1", "var0", "var1", "w0", "w1")},
            "gmm_iters": int(g["iters"]), "jumped": bool(g["jumped"]),
            "n_selected": int(s0.size), "selected_digest": digest(s0),
            "n_filtered": int(s1.size), "filtered_digest": digest(s1),
            "n_expanded": int(s2.size), "expanded_digest": digest(s2),
            "component_sizes": [int(c.size) for c in comps],
            "rhs_digest": hashlib.sha256(b.tobytes()).hexdigest()[:16],
        }
        # operator = CSC arrays read as rows (GPUHybrid.cu:401-433), i.e. A^T
        x = spla.spsolve(A.T.tocsc(), b)
        rec["direct_rel_residual"] = float(np.linalg.norm(A.T @ x - b) / np.linalg.norm(b))
        out[name] = rec
        np.savez_compressed(os.path.join(HERE, name + ".npz"), n=np.int64(n), colptr=A.indptr.astype(np.int32),
                            rowidx=A.indices.astype(np.int32), vals=A.data, rhs=b)
        off = np.cumsum([0] + [c.size for c in comps]).astype(np.int64)
        np.savez_compressed(os.path.join(HERE, name + "_golden.npz"), row_l1=l1, selected=s0, filtered=s1,
                            expanded=s2, comp_off=off, comp_dofs=np.concatenate(comps) if comps else s2,
                            x_direct=x)
        print(name, json.dumps({k: v for k, v in rec.items() if k != "component_sizes"}), flush=True)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/mats")
