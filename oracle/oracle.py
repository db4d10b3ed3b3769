"""ORACLE — test infrastructure only.

ctypes front end of liblss_oracle.so, the CPU FP64 restatement of the reference
GPU hybrid solve path (see lss_oracle.h for what each function restates and how
it is pinned).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg import this module; the product library never does.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


class GmmInfo(ctypes.Structure):
    _fields_ = [
        ("global_mean", ctypes.c_double), ("global_var", ctypes.c_double),
        ("mean0", ctypes.c_double), ("mean1", ctypes.c_double),
        ("var0", ctypes.c_double), ("var1", ctypes.c_double),
        ("w0", ctypes.c_double), ("w1", ctypes.c_double),
        ("iters", ctypes.c_int), ("jumped", ctypes.c_int),
        ("selected", ctypes.c_int64), ("min_margin", ctypes.c_double),
    ]


class Config(ctypes.Structure):
    _fields_ = [
        ("max_iter", ctypes.c_int), ("rel_tol", ctypes.c_double), ("abs_tol", ctypes.c_double),
        ("block_size", ctypes.c_int), ("do_hybrid_precond", ctypes.c_int),
        ("select_bad_dofs_from_l1_norm", ctypes.c_int), ("decompose_subdomains", ctypes.c_int),
        ("expand_subdomains", ctypes.c_int), ("min_subdomain_size", ctypes.c_int64),
        ("max_subdomain_size", ctypes.c_int64), ("gmm_jump_threshold", ctypes.c_double),
        ("gmm_tol", ctypes.c_double), ("max_gmm_iterations", ctypes.c_int),
        ("conditioning_threshold", ctypes.c_double), ("base_precond", ctypes.c_int),
        ("threads", ctypes.c_int), ("damping", ctypes.c_double), ("composition", ctypes.c_int),
        ("solver_mode", ctypes.c_int), ("use_bad_dof", ctypes.c_int), ("bad_dof_strategy", ctypes.c_int),
        ("mgmm_jump", ctypes.c_double), ("mgmm_tol", ctypes.c_double), ("mgmm_max_iter", ctypes.c_int),
        ("mgmm_expand", ctypes.c_int), ("km_fraction", ctypes.c_double), ("km_jump", ctypes.c_double),
        ("km_max_iter", ctypes.c_int), ("km_expand", ctypes.c_int), ("true_residual_period", ctypes.c_int),
        ("precond_norm", ctypes.c_int),
    ]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liblss_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run __graft_entry__.build() (or make -C oracle)")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        L.orc_canonical_sum.restype = ctypes.c_double
        L.orc_canonical_sum.argtypes = [ctypes.c_int64, P]
        L.orc_row_l1.argtypes = [ctypes.c_int32, P, P, P]
        L.orc_gmm_select.argtypes = [ctypes.c_int32, P, ctypes.c_double, ctypes.c_int, ctypes.c_double, P, P,
                                     ctypes.POINTER(GmmInfo)]
        L.orc_filter.argtypes = [ctypes.c_int32, P, P, P, P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                 ctypes.c_double, P, P, P]
        L.orc_expand.argtypes = [ctypes.c_int32, P, P, P, ctypes.c_int64, P, P]
        L.orc_decompose.argtypes = [ctypes.c_int32, P, P, P, ctypes.c_int64, ctypes.c_int64, P, P, P]
        L.orc_partition.argtypes = [ctypes.c_int32, P, P, ctypes.c_int, P, P]
        L.orc_spmv.argtypes = [ctypes.c_int32, P, P, P, P, P]
        L.orc_mas_create.restype = P
        L.orc_mas_create.argtypes = [ctypes.c_int32, P, P, P, ctypes.c_int, P, P, ctypes.c_int32]
        L.orc_mas_destroy.argtypes = [P]
        L.orc_mas_info.argtypes = [P, P, P, P]
        L.orc_mas_map.restype = ctypes.c_int32
        L.orc_mas_map.argtypes = [P, ctypes.c_int32, ctypes.c_int]
        L.orc_mas_inverse.argtypes = [P, ctypes.c_int, ctypes.c_int32, P]
        L.orc_mas_apply.argtypes = [P, P, P]
        L.orc_hybrid_create.restype = P
        L.orc_hybrid_create.argtypes = [ctypes.POINTER(Config)]
        L.orc_hybrid_destroy.argtypes = [P]
        L.orc_hybrid_set_problematic.argtypes = [P, P, ctypes.c_int64]
        L.orc_hybrid_factorize.argtypes = [P, ctypes.c_int32, P, P, P]
        L.orc_hybrid_solve.argtypes = [P, P, P]
        L.orc_hybrid_info.argtypes = [P, P, P, P, P, P, P]
        L.orc_hybrid_stiff.restype = ctypes.c_int64
        L.orc_hybrid_stiff.argtypes = [P, P, ctypes.c_int64, P, ctypes.c_int64]
        L.orc_hybrid_precond.argtypes = [P, P, P]
        L.orc_hybrid_set_partition.argtypes = [P, P, ctypes.c_int32, P, ctypes.c_int32]
        L.orc_mas_indicator.argtypes = [ctypes.c_int, ctypes.c_int32, P, P, P, ctypes.c_int, P, P]
        L.orc_mas_gmm_positions.restype = ctypes.c_int64
        L.orc_mas_gmm_positions.argtypes = [ctypes.c_int64, P, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_int, P, P]
        L.orc_mas_elbow_positions.restype = ctypes.c_int64
        L.orc_mas_elbow_positions.argtypes = [ctypes.c_int64, P, P, P]
        L.orc_mas_kmeans_positions.restype = ctypes.c_int64
        L.orc_mas_kmeans_positions.argtypes = [ctypes.c_int64, P, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                               ctypes.c_int, P, P]
        L.orc_hybrid_mas_status.restype = ctypes.c_int
        L.orc_hybrid_mas_status.argtypes = [P, P]
        L.orc_hybrid_omega.restype = ctypes.c_double
        L.orc_hybrid_omega.argtypes = [P]
        L.orc_hybrid_set_omega.restype = None
        L.orc_hybrid_set_omega.argtypes = [P, ctypes.c_double]
        L.orc_hybrid_damping_retry.argtypes = [P]
        L.orc_hybrid_set_run.restype = None
        L.orc_hybrid_set_run.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.orc_hybrid_partition.restype = ctypes.c_int32
        L.orc_hybrid_partition.argtypes = [P, P, P]
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def canonical_sum(v) -> float:
    v = _f64(v)
    return lib().orc_canonical_sum(v.size, _p(v))


def row_l1(ptr, vals) -> np.ndarray:
    ptr = _i32(ptr)
    vals = _f64(vals)
    out = np.empty(ptr.size - 1, np.float64)
    lib().orc_row_l1(ptr.size - 1, _p(ptr), _p(vals), _p(out))
    return out


def gmm_select(norms, tol=1e-3, max_iter=20, jump=10.0, want_gamma=False):
    norms = _f64(norms)
    n = norms.size
    flags = np.zeros(n, np.uint8)
    gamma = np.empty(2 * n, np.float64) if want_gamma else None
    info = GmmInfo()
    lib().orc_gmm_select(n, _p(norms), tol, max_iter, jump, _p(flags), _p(gamma), ctypes.byref(info))
    res = {k: getattr(info, k) for k, _ in GmmInfo._fields_}
    res["set"] = np.flatnonzero(flags).astype(np.int32)
    if want_gamma:
        res["gamma0"], res["gamma1"] = gamma[:n], gamma[n:]
    return res


def filter_set(ptr, idx, vals, s, min_size=4, max_size=10_000_000, cond=100.0):
    ptr, idx, vals, s = _i32(ptr), _i32(idx), _f64(vals), _i32(s)
    out = np.empty(max(s.size, 1), np.int32)
    nout = ctypes.c_int64()
    stats = np.zeros(3, np.int64)
    lib().orc_filter(ptr.size - 1, _p(ptr), _p(idx), _p(vals), _p(s), s.size, min_size, max_size, cond, _p(out),
                     ctypes.byref(nout), _p(stats))
    return out[: nout.value].copy(), stats


def expand_set(ptr, idx, s):
    ptr, idx, s = _i32(ptr), _i32(idx), _i32(s)
    out = np.empty(ptr.size - 1, np.int32)
    nout = ctypes.c_int64()
    lib().orc_expand(ptr.size - 1, _p(ptr), _p(idx), _p(s), s.size, _p(out), ctypes.byref(nout))
    return out[: nout.value].copy()


def decompose_set(ptr, idx, s, max_size=10_000_000):
    ptr, idx, s = _i32(ptr), _i32(idx), _i32(s)
    dofs = np.empty(max(s.size, 1), np.int32)
    off = np.empty(s.size + 1, np.int64)
    nc = ctypes.c_int64()
    lib().orc_decompose(ptr.size - 1, _p(ptr), _p(idx), _p(s), s.size, max_size, _p(dofs), _p(off), ctypes.byref(nc))
    k = nc.value
    return [dofs[off[c]:off[c + 1]].copy() for c in range(k)]


def partition(bptr, bidx, max_part=32):
    bptr = np.ascontiguousarray(bptr, np.int64)
    bidx = _i32(bidx)
    nb = bptr.size - 1
    part = np.empty(nb, np.int32)
    np_ = ctypes.c_int32()
    lib().orc_partition(nb, _p(bptr), _p(bidx), max_part, _p(part), ctypes.byref(np_))
    return part, np_.value


def spmv(ptr, idx, vals, x):
    ptr, idx, vals, x = _i32(ptr), _i32(idx), _f64(vals), _f64(x)
    y = np.empty(ptr.size - 1, np.float64)
    lib().orc_spmv(ptr.size - 1, _p(ptr), _p(idx), _p(vals), _p(x), _p(y))
    return y


class Mas:
    def __init__(self, ptr, idx, vals, bs, perm, part_off):
        self._keep = [_i32(ptr), _i32(idx), _f64(vals), _i32(perm), _i32(part_off)]
        ptr, idx, vals, perm, part_off = self._keep
        self.n = ptr.size - 1
        self.bs = bs
        h = lib().orc_mas_create(self.n, _p(ptr), _p(idx), _p(vals), bs, _p(perm), _p(part_off), part_off.size - 1)
        if not h:
            raise RuntimeError("oracle MAS construction failed (non-SPD bank matrix)")
        self._h = ctypes.c_void_p(h)
        lv = ctypes.c_int()
        counts = np.zeros(8, np.int32)
        cco = np.zeros(8, np.int32)
        lib().orc_mas_info(self._h, ctypes.byref(lv), _p(counts), _p(cco))
        self.levels = lv.value
        self.counts = counts[: self.levels].copy()
        self.cco = cco[: self.levels - 1].copy()
        self.dim = 32 * bs

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.orc_mas_destroy(self._h)

    def map(self, node, level):
        return lib().orc_mas_map(self._h, node, level)

    def inverse(self, level, bank):
        out = np.empty((self.dim, self.dim), np.float64)
        if lib().orc_mas_inverse(self._h, level, bank, _p(out)) != 0:
            raise IndexError("bank out of range")
        return out

    def apply(self, r):
        r = _f64(r)
        z = np.empty(self.n, np.float64)
        lib().orc_mas_apply(self._h, _p(r), _p(z))
        return z


def config(block_size=1, rel_tol=1e-8, abs_tol=0.0, max_iter=10000, base="mas", threads=0, **kw):
    c = Config()
    c.max_iter = max_iter
    c.rel_tol = rel_tol
    c.abs_tol = abs_tol
    c.block_size = block_size
    c.do_hybrid_precond = int(kw.get("do_hybrid_precond", True))
    c.select_bad_dofs_from_l1_norm = int(kw.get("select_bad_dofs_from_l1_norm", True))
    c.decompose_subdomains = int(kw.get("decompose_subdomains", True))
    c.expand_subdomains = int(kw.get("expand_subdomains", True))
    c.min_subdomain_size = kw.get("min_subdomain_size", 4)
    c.max_subdomain_size = kw.get("max_subdomain_size", 10_000_000)
    c.gmm_jump_threshold = kw.get("gmm_jump_threshold", 10.0)
    c.gmm_tol = kw.get("gmm_tol", 1e-3)
    c.max_gmm_iterations = kw.get("max_gmm_iterations", 20)
    c.conditioning_threshold = kw.get("conditioning_threshold", 100.0)
    c.base_precond = {"mas": 0, "jacobi": 1, "identity": 2}[base]
    c.threads = threads
    c.damping = kw.get("damping", 1.8)
    c.composition = {"BPB": 0, "PBP": 1, "additive": 2}[kw.get("composition", "BPB")]
    # the standalone MAS solver (MASSolver.cu) and its stiff-DOF strategies
    c.solver_mode = {"GPUHybrid": 0, "MAS": 1}[kw.get("solver", "GPUHybrid")]
    c.use_bad_dof = int(kw.get("use_bad_dof_precond", False))
    c.bad_dof_strategy = {"elbow": 0, "kmean": 1, "gmm": 2}[kw.get("bad_dof_strategy", "elbow")]
    c.mgmm_jump = kw.get("bad_dof_gmm_jump_threshold", 10.0)
    c.mgmm_tol = kw.get("bad_dof_gmm_tol", 1e-3)
    c.mgmm_max_iter = kw.get("bad_dof_gmm_max_iterations", 5)
    c.mgmm_expand = kw.get("bad_dof_gmm_expand_neighbors", 32)
    c.km_fraction = kw.get("bad_dof_kmeans_search_fraction", 0.02)
    c.km_jump = kw.get("bad_dof_kmeans_jump_threshold", 10.0)
    c.km_max_iter = kw.get("bad_dof_kmeans_max_iterations", 32)
    c.km_expand = kw.get("bad_dof_kmeans_expand_neighbors", 0)
    c.true_residual_period = kw.get("true_residual_period", 4)
    c.precond_norm = int(kw.get("use_preconditioned_residual_norm", False))
    return c


def mas_indicator(mode, A_colptr, A_rowidx, A_vals, bs, perm):
    """Per permuted scalar row: mode 0 sum |a_ij|, 1 |a_ii|, 2 max |a_ij| (j != i)."""
    ptr, idx, vals, perm = _i32(A_colptr), _i32(A_rowidx), _f64(A_vals), _i32(perm)
    n = ptr.size - 1
    out = np.empty(perm.size * bs, np.float64)
    lib().orc_mas_indicator(mode, n, _p(ptr), _p(idx), _p(vals), bs, _p(perm), _p(out))
    return out


def mas_gmm_positions(x, tol=1e-3, max_iter=5, jump=10.0, expand=32):
    x = _f64(x)
    out = np.empty(max(x.size, 1), np.int64)
    st = np.zeros(8)
    k = lib().orc_mas_gmm_positions(x.size, _p(x), tol, max_iter, jump, expand, _p(out), _p(st))
    keys = ("mean0", "mean1", "var0", "var1", "global_mean", "global_var", "ratio", "iters")
    return out[:k].copy(), dict(zip(keys, st))


def mas_elbow_positions(x):
    x = _f64(x)
    out = np.empty(max(x.size, 1), np.int64)
    cut = np.zeros(1, np.int64)
    k = lib().orc_mas_elbow_positions(x.size, _p(x), _p(out), _p(cut))
    return out[:k].copy(), int(cut[0])


def mas_kmeans_positions(x, fraction=0.02, jump=10.0, max_iter=32, expand=0):
    x = _f64(x)
    out = np.empty(max(x.size, 1), np.int64)
    st = np.zeros(4)
    k = lib().orc_mas_kmeans_positions(x.size, _p(x), fraction, jump, max_iter, expand, _p(out), _p(st))
    return out[:k].copy(), dict(zip(("c0", "c1", "ratio", "iters"), st))


class Hybrid:
    """CPU restatement of GPUHybrid factorize()/solve() with B = MAS (or Jacobi)."""

    def __init__(self, cfg: Config):
        self._cfg = cfg
        self._h = ctypes.c_void_p(lib().orc_hybrid_create(ctypes.byref(cfg)))
        self._keep = None

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.orc_hybrid_destroy(self._h)

    def set_partition(self, perm, part_off):
        perm, part_off = _i32(perm), _i32(part_off)
        self._keep = (perm, part_off)
        lib().orc_hybrid_set_partition(self._h, _p(perm), perm.size, _p(part_off), part_off.size - 1)

    def set_problematic_dofs(self, dofs):
        d = _i32(dofs)
        lib().orc_hybrid_set_problematic(self._h, _p(d), d.size)

    def factorize(self, ptr, idx, vals):
        ptr, idx, vals = _i32(ptr), _i32(idx), _f64(vals)
        rc = lib().orc_hybrid_factorize(self._h, ptr.size - 1, _p(ptr), _p(idx), _p(vals))
        if rc != 0:
            raise RuntimeError(f"oracle factorize failed ({rc})")
        self.n = ptr.size - 1

    def solve(self, b, x0=None):
        b = _f64(b)
        x = np.zeros(self.n) if x0 is None else _f64(x0).copy()
        lib().orc_hybrid_solve(self._h, _p(b), _p(x))
        return x

    def precond(self, r):
        r = _f64(r)
        z = np.empty_like(r)
        lib().orc_hybrid_precond(self._h, _p(r), _p(z))
        return z

    def set_omega(self, omega):
        """Damping of B inside the multiplicative sweep (replaces factorize's estimate)."""
        lib().orc_hybrid_set_omega(self._h, float(omega))

    def set_run(self, threads, max_iter):
        """Thread count and iteration cap of the following solves."""
        lib().orc_hybrid_set_run(self._h, int(threads), int(max_iter))

    def info(self):
        it = ctypes.c_int()
        res, ts, tv = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        ns, nc = ctypes.c_int64(), ctypes.c_int64()
        lib().orc_hybrid_info(self._h, ctypes.byref(it), ctypes.byref(res), ctypes.byref(ts), ctypes.byref(tv),
                              ctypes.byref(ns), ctypes.byref(nc))
        return {"num_iterations": it.value, "final_res_norm": res.value, "t_setup": ts.value,
                "t_solve": tv.value, "n_stiff": ns.value, "n_components": nc.value,
                "omega": lib().orc_hybrid_omega(self._h),
                "mas_status": lib().orc_hybrid_mas_status(self._h, None),
                "damping_retry": bool(lib().orc_hybrid_damping_retry(self._h))}

    def stiff_components(self):
        inf = self.info()
        dofs = np.empty(max(inf["n_stiff"], 1), np.int32)
        off = np.empty(inf["n_components"] + 1, np.int64)
        lib().orc_hybrid_stiff(self._h, _p(dofs), dofs.size, _p(off), off.size)
        return [dofs[off[c]:off[c + 1]].copy() for c in range(inf["n_components"])]

    def partition(self):
        nparts = lib().orc_hybrid_partition(self._h, None, None)
        if nparts < 0:
            return None, None
        nb = (self.n + self._cfg.block_size - 1) // self._cfg.block_size
        perm = np.empty(nb, np.int32)
        off = np.empty(nparts + 1, np.int32)
        lib().orc_hybrid_partition(self._h, _p(perm), _p(off))
        return perm, off
