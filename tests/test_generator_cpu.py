"""The synthetic workloads' geometry: the sliver model of the 64^3 sliver system
("0.5 % of the elements collapsed into slivers by moving one vertex to within
1e-3 h of the opposite face") checked element by element on the generator's output."""
import numpy as np
import pytest

from lss_amd import problems


def _signed_volume(X):
    return np.einsum("...i,...i->...", np.cross(X[..., 1, :] - X[..., 0, :], X[..., 2, :] - X[..., 0, :]),
                     X[..., 3, :] - X[..., 0, :]) / 6.0


@pytest.mark.parametrize("N", [12, 20])
def test_sliver_elements(N):
    prob = problems.make("elastic_slivers", N=N)
    h = 1.0 / (N - 1)
    elem, X, X0 = prob.sliver_geometry()
    assert prob.n_elements == 6 * (N - 1) ** 3
    assert elem.size == round(0.005 * prob.n_elements) > 0
    assert np.all(np.diff(elem) > 0) and elem[0] >= 0 and elem[-1] < prob.n_elements
    moved = np.abs(X - X0).max(axis=2) > 0
    assert np.all(moved.sum(axis=1) == 1)  # one vertex per element
    a = moved.argmax(axis=1)
    for q in range(elem.size):
        f = [v for v in range(4) if v != a[q]]
        nrm = np.cross(X[q, f[1]] - X[q, f[0]], X[q, f[2]] - X[q, f[0]])
        nrm /= np.linalg.norm(nrm)
        height = abs(np.dot(X[q, a[q]] - X[q, f[0]], nrm))
        assert height == pytest.approx(1e-3 * h, rel=1e-9)
    # the collapse keeps each element's orientation (no inverted element)
    v, v0 = _signed_volume(X), _signed_volume(X0)
    assert np.all(np.sign(v) == np.sign(v0))
    assert np.all(np.abs(v) < 1e-2 * np.abs(v0))


def test_sliver_rows_stand_out():
    prob = problems.make("elastic_slivers", N=16)
    A = prob.csc()
    norms = np.add.reduceat(np.abs(A.vals), A.colptr[:-1])
    elem, _, _ = prob.sliver_geometry()
    nodes = prob.element_nodes().reshape(-1, 4)[elem]
    rows = np.unique((3 * nodes[:, :, None] + np.arange(3)).ravel())
    # sliver element matrices scale like 1 / height: their rows lie far above the bulk
    assert np.median(norms[rows]) > 5 * np.median(norms)
