"""The oracle against the golden fixtures made from the reference's own example
systems (mats/A_contact.mtx, mats/A_pq.mtx; see tests/golden/make_golden.py)."""
import hashlib
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import independent as ind
from lss_amd import problems
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MATS = "/root/reference/mats"
SPEC = dict(gmm_jump_threshold=10.0, min_subdomain_size=4)  # spec/GPUHybrid.json:1-19


def _gold():
    with open(os.path.join(GOLD, "golden.json")) as f:
        return json.load(f)


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.int64).tobytes()).hexdigest()[:16]


@pytest.fixture(scope="module")
def contact():
    d = np.load(os.path.join(GOLD, "A_contact.npz"))
    g = np.load(os.path.join(GOLD, "A_contact_golden.npz"))
    return d, g


def test_contact_fixture_is_the_mtx(contact):
    d, _ = contact
    gold = _gold()["A_contact"]
    assert int(d["n"]) == gold["n"] and d["vals"].size == gold["nnz"]
    assert hashlib.sha256(d["rhs"].tobytes()).hexdigest()[:16] == gold["rhs_digest"]
    # tests/linear_solve.cpp:112-132 with seed 0, from the library's own generator too
    assert np.array_equal(d["rhs"], problems.random_vector(int(d["n"]), 0))


def test_contact_selection_stages(contact):
    d, g = contact
    gold = _gold()["A_contact"]
    l1 = O.row_l1(d["colptr"], d["vals"])
    assert np.array_equal(l1, g["row_l1"])
    assert hashlib.sha256(l1.tobytes()).hexdigest()[:16] == gold["row_l1_digest"]
    gm = O.gmm_select(l1, 1e-3, 20, 10.0)
    for k, v in gold["gmm"].items():
        assert gm[k] == pytest.approx(v, rel=1e-9), k
    assert gm["iters"] == gold["gmm_iters"] and bool(gm["jumped"]) == gold["jumped"]
    assert np.array_equal(gm["set"], g["selected"])
    s1, _ = O.filter_set(d["colptr"], d["rowidx"], d["vals"], gm["set"], 4, 10_000_000, 100.0)
    assert np.array_equal(s1, g["filtered"])
    s2 = O.expand_set(d["colptr"], d["rowidx"], s1)
    assert np.array_equal(s2, g["expanded"])
    comps = O.decompose_set(d["colptr"], d["rowidx"], s2)
    assert [c.size for c in comps] == gold["component_sizes"]
    assert np.array_equal(np.concatenate(comps), g["comp_dofs"])


def test_contact_hybrid_selects_golden_set(contact):
    d, g = contact
    H = O.Hybrid(O.config(block_size=3, **SPEC))
    H.factorize(d["colptr"], d["rowidx"], d["vals"])
    comps = H.stiff_components()
    off = np.cumsum([0] + [c.size for c in comps])
    assert np.array_equal(off, g["comp_off"])
    assert np.array_equal(np.concatenate(comps), g["comp_dofs"])


def test_contact_hybrid_solution(contact):
    d, g = contact
    n = int(d["n"])
    H = O.Hybrid(O.config(block_size=3, rel_tol=1e-8, **SPEC))
    H.factorize(d["colptr"], d["rowidx"], d["vals"])
    x = H.solve(d["rhs"])
    # the operator is the CSC arrays read as rows (GPUHybrid.cu:401-433)
    R = ind.operator_rows(d["colptr"], d["rowidx"], d["vals"], n)
    rel = np.linalg.norm(R @ x - d["rhs"]) / np.linalg.norm(d["rhs"])
    # row sums reach 6e11 against O(1) right-hand sides: the FP64 floor is near 1e-6,
    # and the direct solution itself leaves 9e-7
    assert rel < 2e-6
    xd = g["x_direct"]
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)


@pytest.fixture(scope="module")
def pq():
    d = np.load(os.path.join(GOLD, "A_pq.npz"))
    g = np.load(os.path.join(GOLD, "A_pq_golden.npz"))
    return d, g


def test_pq_fixture_is_the_mtx(pq):
    d, _ = pq
    gold = _gold()["A_pq"]
    assert int(d["n"]) == gold["n"] and d["vals"].size == gold["nnz"]
    assert hashlib.sha256(d["rhs"].tobytes()).hexdigest()[:16] == gold["rhs_digest"]
    assert np.array_equal(d["rhs"], problems.random_vector(int(d["n"]), 0))


@pytest.mark.skipif(not os.path.exists(os.path.join(MATS, "A_pq.mtx")), reason="reference data files not mounted")
def test_pq_fixture_arrays_equal_the_mtx(pq):
    import scipy.io

    d, _ = pq
    A = scipy.io.mmread(os.path.join(MATS, "A_pq.mtx")).tocsc()
    A.sum_duplicates()
    A.sort_indices()
    assert np.array_equal(A.indptr, d["colptr"]) and np.array_equal(A.indices, d["rowidx"])
    assert np.array_equal(A.data, d["vals"])


def test_pq_selection_stages(pq):
    d, g = pq
    gold = _gold()["A_pq"]
    cp, ri, v = d["colptr"], d["rowidx"], d["vals"]
    l1 = O.row_l1(cp, v)
    assert np.array_equal(l1, g["row_l1"])
    assert hashlib.sha256(l1.tobytes()).hexdigest()[:16] == gold["row_l1_digest"]
    gm = O.gmm_select(l1, 1e-3, 20, 10.0)
    for k, val in gold["gmm"].items():
        assert gm[k] == pytest.approx(val, rel=1e-9), k
    assert _digest(gm["set"]) == gold["selected_digest"]
    s1, _ = O.filter_set(cp, ri, v, gm["set"], 4, 10_000_000, 100.0)
    assert _digest(s1) == gold["filtered_digest"]
    s2 = O.expand_set(cp, ri, s1)
    assert _digest(s2) == gold["expanded_digest"]
    comps = O.decompose_set(cp, ri, s2)
    assert [c.size for c in comps] == gold["component_sizes"]
    assert np.array_equal(np.concatenate(comps), g["comp_dofs"])

