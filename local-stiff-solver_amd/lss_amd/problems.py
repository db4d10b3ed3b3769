"""Synthetic locally-stiff SPD systems (ctypes front end of liblss_gen.so).

The named workloads are the five benchmark configurations of this solver:

* ``laplace_sphere``   P1 Laplacian, 32^3 nodes, conductivity 1e6 inside a random
                       sphere (radius 0.1), Dirichlet on z = 0.
* ``elastic_slivers``  linear elasticity (E = 1, nu = 0.3), 64^3 nodes, 0.5 % of the
                       elements collapsed into slivers (vertex at 1e-3 h from the
                       opposite face), clamped on x = 0.
* ``elastic_contact``  elasticity on 128^3 nodes plus contact-barrier pair blocks
                       kappa n n^T on 2 % of the top-face nodes, kappa log-uniform in
                       [1e4, 1e8], partner within 2h, clamped on x = 0.
* ``elastic_logE``     elasticity on 48^3 nodes with per-element E log-uniform in
                       [1, 1e5], clamped on x = 0.
* ``elastic_contact_xl`` the contact system on 160^3 nodes.

Right-hand sides follow make_random_rhs (tests/linear_solve.cpp:112-132):
std::mt19937(seed) feeding std::uniform_real_distribution<double>(-1, 1).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_LIB_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib")


class _Spec(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("N", ctypes.c_int),
        ("E", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("dirichlet_axis", ctypes.c_int),
        ("sphere_radius", ctypes.c_double),
        ("sphere_coef", ctypes.c_double),
        ("sliver_frac", ctypes.c_double),
        ("sliver_eps", ctypes.c_double),
        ("contact_frac", ctypes.c_double),
        ("contact_log10_kmin", ctypes.c_double),
        ("contact_log10_kmax", ctypes.c_double),
        ("contact_radius_h", ctypes.c_double),
        ("E_log10_min", ctypes.c_double),
        ("E_log10_max", ctypes.c_double),
        ("seed", ctypes.c_uint32),
    ]


_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(_LIB_DIR, "liblss_gen.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    lib.lssgen_create.restype = P
    lib.lssgen_create.argtypes = [ctypes.POINTER(_Spec), ctypes.c_char_p, ctypes.c_int]
    lib.lssgen_destroy.argtypes = [P]
    for name in ("lssgen_n", "lssgen_nnz", "lssgen_n_elements", "lssgen_n_contacts", "lssgen_n_slivers"):
        getattr(lib, name).restype = ctypes.c_int64
        getattr(lib, name).argtypes = [P]
    lib.lssgen_block_size.restype = ctypes.c_int
    lib.lssgen_block_size.argtypes = [P]
    lib.lssgen_fill_csc.argtypes = [P, P, P, P]
    lib.lssgen_random_vector.argtypes = [ctypes.c_int64, ctypes.c_uint32, ctypes.c_double, ctypes.c_double, P]
    lib.lssgen_element_coef.argtypes = [P, P]
    lib.lssgen_element_nodes.argtypes = [P, P]
    lib.lssgen_sliver_geometry.argtypes = [P, P, P, P]
    lib.lssgen_grad_coef.argtypes = [P, P, P, P]
    lib.lssgen_fill_vals_with_coef.argtypes = [P, P, P]
    lib.lssgen_element_jacobian_nnz.restype = ctypes.c_int64
    lib.lssgen_element_jacobian_nnz.argtypes = [P]
    lib.lssgen_element_jacobian.argtypes = [P, P, P, P]
    _lib = lib
    return lib


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def random_vector(n: int, seed: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    """std::mt19937(seed) + std::uniform_real_distribution<double>(lo, hi), n draws."""
    out = np.empty(n, dtype=np.float64)
    _load().lssgen_random_vector(n, seed, lo, hi, _ptr(out))
    return out


@dataclass
class CSC:
    """Compressed sparse column matrix (0-based int32), polysolve's StiffnessMatrix layout."""

    n: int
    colptr: np.ndarray
    rowidx: np.ndarray
    vals: np.ndarray
    block_size: int = 1
    meta: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.colptr[-1])

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csc_matrix((self.vals, self.rowidx, self.colptr), shape=(self.n, self.n))


class Problem:
    """An assembled synthetic problem; keeps the element model for d/dcoef gradients."""

    def __init__(self, kind: str, N: int, seed: int = 0, **kw):
        lib = _load()
        s = _Spec()
        s.kind = 0 if kind == "laplacian" else 1
        s.N = N
        s.E = kw.get("E", 1.0)
        s.nu = kw.get("nu", 0.3)
        s.dirichlet_axis = kw.get("dirichlet_axis", 2 if kind == "laplacian" else 0)
        s.sphere_radius = kw.get("sphere_radius", 0.0)
        s.sphere_coef = kw.get("sphere_coef", 1.0)
        s.sliver_frac = kw.get("sliver_frac", 0.0)
        s.sliver_eps = kw.get("sliver_eps", 1e-3)
        s.contact_frac = kw.get("contact_frac", 0.0)
        s.contact_log10_kmin = kw.get("contact_log10_kmin", 4.0)
        s.contact_log10_kmax = kw.get("contact_log10_kmax", 8.0)
        s.contact_radius_h = kw.get("contact_radius_h", 2.0)
        s.E_log10_min = kw.get("E_log10_min", 0.0)
        s.E_log10_max = kw.get("E_log10_max", 0.0)
        s.seed = seed
        err = ctypes.create_string_buffer(512)
        h = lib.lssgen_create(ctypes.byref(s), err, 512)
        if not h:
            raise RuntimeError("problem generation failed: " + err.value.decode())
        self._h = ctypes.c_void_p(h)
        self.kind = kind
        self.N = N
        self.seed = seed
        self.block_size = lib.lssgen_block_size(self._h)
        self.n = lib.lssgen_n(self._h)
        self.nnz = lib.lssgen_nnz(self._h)
        self.n_elements = lib.lssgen_n_elements(self._h)
        self.n_contacts = lib.lssgen_n_contacts(self._h)
        self.n_slivers = lib.lssgen_n_slivers(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.lssgen_destroy(h)
            self._h = None

    def csc(self, colptr=None, rowidx=None, vals=None) -> CSC:
        """Fill (optionally caller-provided, e.g. pinned) arrays with the matrix."""
        colptr = np.empty(self.n + 1, np.int32) if colptr is None else colptr
        rowidx = np.empty(self.nnz, np.int32) if rowidx is None else rowidx
        vals = np.empty(self.nnz, np.float64) if vals is None else vals
        _load().lssgen_fill_csc(self._h, _ptr(colptr), _ptr(rowidx), _ptr(vals))
        return CSC(self.n, colptr, rowidx, vals, self.block_size,
                   {"kind": self.kind, "N": self.N, "seed": self.seed})

    def rhs(self, seed: int = 0) -> np.ndarray:
        return random_vector(self.n, seed)

    def element_coef(self) -> np.ndarray:
        out = np.empty(self.n_elements, np.float64)
        _load().lssgen_element_coef(self._h, _ptr(out))
        return out

    def sliver_geometry(self):
        """(elements, vertices after, vertices before the collapse) of the sliver elements."""
        k = self.n_slivers
        elem = np.zeros(max(k, 1), np.int64)
        X = np.zeros((max(k, 1), 4, 3))
        X0 = np.zeros((max(k, 1), 4, 3))
        _load().lssgen_sliver_geometry(self._h, _ptr(elem), _ptr(X), _ptr(X0))
        return elem[:k], X[:k], X0[:k]

    def element_nodes(self) -> np.ndarray:
        out = np.empty((self.n_elements, 4), np.int32)
        _load().lssgen_element_nodes(self._h, _ptr(out))
        return out

    def grad_coef(self, lam: np.ndarray, x: np.ndarray) -> np.ndarray:
        """-lam^T (dA/dcoef_e) x per element."""
        lam = np.ascontiguousarray(lam, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        out = np.empty(self.n_elements, np.float64)
        _load().lssgen_grad_coef(self._h, _ptr(lam), _ptr(x), _ptr(out))
        return out

    def element_jacobian(self):
        """d vals / d coef, element-major: (seg_ptr[n_elem+1] int64, slot int32, value float64)."""
        lib = _load()
        m = lib.lssgen_element_jacobian_nnz(self._h)
        seg = np.empty(self.n_elements + 1, np.int64)
        slot = np.empty(m, np.int32)
        val = np.empty(m, np.float64)
        if lib.lssgen_element_jacobian(self._h, _ptr(seg), _ptr(slot), _ptr(val)) != 0:
            raise RuntimeError("element Jacobian: an element entry is missing from the pattern")
        return seg, slot, val

    def vals_with_coef(self, coef: np.ndarray) -> np.ndarray:
        """Matrix values (CSC order) re-assembled with per-element coefficients."""
        coef = np.ascontiguousarray(coef, np.float64)
        vals = np.empty(self.nnz, np.float64)
        _load().lssgen_fill_vals_with_coef(self._h, _ptr(coef), _ptr(vals))
        return vals


CONFIGS = {
    "laplace_sphere": dict(kind="laplacian", N=32, sphere_radius=0.1, sphere_coef=1e6),
    "elastic_slivers": dict(kind="elasticity", N=64, sliver_frac=0.005, sliver_eps=1e-3),
    "elastic_contact": dict(kind="elasticity", N=128, contact_frac=0.02, contact_log10_kmin=4.0,
                            contact_log10_kmax=8.0, contact_radius_h=2.0),
    "elastic_logE": dict(kind="elasticity", N=48, E_log10_min=0.0, E_log10_max=5.0),
    "elastic_contact_xl": dict(kind="elasticity", N=160, contact_frac=0.02, contact_log10_kmin=4.0,
                               contact_log10_kmax=8.0, contact_radius_h=2.0),
}


def make(name: str, seed: int = 0, N: int | None = None) -> Problem:
    """Build a named configuration; N overrides the grid size (smaller test instances)."""
    cfg = dict(CONFIGS[name])
    if N is not None:
        cfg["N"] = N
    kind = cfg.pop("kind")
    n = cfg.pop("N")
    return Problem(kind, n, seed=seed, **cfg)
