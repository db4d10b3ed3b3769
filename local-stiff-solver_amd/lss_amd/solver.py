"""Python mirror of the polysolve::linear::Solver surface for the GPU hybrid solver.

Reference interface: src/polysolve/linear/Solver.hpp:32-149 (create,
set_parameters, get_info, analyze_pattern, factorize, solve, name,
set_problematic_dofs) as implemented by GPUHybrid (GPUHybrid.hpp:27-192).
Errors raise LssError (the reference throws std::runtime_error for invalid
configurations, Solver.cpp:157-165, and exits on device errors).
"""
from __future__ import annotations

import ctypes
import json as _json

import numpy as np

from ._lib import LssError, load

_P = ctypes.c_void_p


def _ptr(a: np.ndarray):
    return _P(a.ctypes.data)


class Solver:
    """GPUHybrid drop-in: ``Solver.create({"solver": "GPUHybrid", "GPUHybrid": {...}})``."""

    def __init__(self, handle, lib):
        self._h = handle
        self._lib = lib
        self._n = None

    # ---------------------------------------------------------------- factory
    @classmethod
    def create(cls, params: dict | str | None = None) -> "Solver":
        lib = load()
        text = params if isinstance(params, str) or params is None else _json.dumps(params)
        h = _P()
        rc = lib.lss_create(text.encode() if text else None, ctypes.byref(h))
        if rc != 0:
            msg = lib.lss_last_error(h).decode() if h.value else "lss_create failed"
            if h.value:
                lib.lss_destroy(h)
            raise LssError(rc, msg)
        return cls(h, lib)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.lss_destroy(h)
            self._h = None

    def _check(self, rc: int):
        if rc != 0:
            raise LssError(rc, self._lib.lss_last_error(self._h).decode())

    # ---------------------------------------------------------------- Solver surface
    def set_parameters(self, params: dict | str):
        text = params if isinstance(params, str) else _json.dumps(params)
        self._check(self._lib.lss_set_parameters(self._h, text.encode()))

    def get_info(self) -> dict:
        need = ctypes.c_size_t()
        self._check(self._lib.lss_get_info(self._h, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value + 1)
        self._check(self._lib.lss_get_info(self._h, buf, need.value + 1, None))
        return _json.loads(buf.value.decode())

    def name(self) -> str:
        return self._lib.lss_name(self._h).decode()

    def set_problematic_dofs(self, dofs):
        d = np.ascontiguousarray(np.fromiter(dofs, dtype=np.int32) if not isinstance(dofs, np.ndarray) else dofs,
                                 dtype=np.int32)
        self._check(self._lib.lss_set_problematic_dofs(self._h, _ptr(d), d.size))

    def analyze_pattern(self, A, precond_num: int = 0):
        self._check(self._lib.lss_analyze_pattern(self._h, 0, 0, None, None))

    def factorize(self, A):
        """A: scipy.sparse matrix or (n, colptr, rowidx, vals) CSC arrays (host)."""
        n, colptr, rowidx, vals = _csc_arrays(A)
        self._check(self._lib.lss_factorize_csc(self._h, n, int(colptr[-1]), _ptr(colptr), _ptr(rowidx), _ptr(vals)))
        self._n = n

    def factorize_device(self, n: int, colptr, rowidx, vals, stream=None):
        """Device CSC tensors (torch, on the GPU)."""
        nnz = int(rowidx.numel())
        self._check(self._lib.lss_factorize_csc_device(self._h, n, nnz, _P(colptr.data_ptr()),
                                                       _P(rowidx.data_ptr()), _P(vals.data_ptr()),
                                                       _stream(stream)))
        self._n = n

    def solve(self, b, x=None) -> np.ndarray:
        """Solve A x = b; x (optional) is the initial guess (GPUHybrid.cu:342-348)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b) if x is None else np.array(x, dtype=np.float64, copy=True)
        self._check(self._lib.lss_solve(self._h, _ptr(b), _ptr(x)))
        return x

    def solve_device(self, b, x, stream=None):
        """b, x: contiguous float64 torch tensors on the GPU; x is updated in place."""
        self._check(self._lib.lss_solve_device(self._h, _P(b.data_ptr()), _P(x.data_ptr()), _stream(stream)))
        return x

    def solve_multi_device(self, B, X, stream=None):
        """B, X: (k, n) contiguous float64 device tensors (one right-hand side per row)."""
        k = int(B.shape[0])
        self._check(self._lib.lss_solve_multi_device(self._h, k, _P(B.data_ptr()), _P(X.data_ptr()),
                                                     _stream(stream)))
        return X

    # ---------------------------------------------------------------- inspection
    def stiff_components(self) -> list[np.ndarray]:
        nd, nc = ctypes.c_int64(), ctypes.c_int64()
        self._check(self._lib.lss_get_stiff_dofs(self._h, None, 0, ctypes.byref(nd), None, 0, ctypes.byref(nc)))
        dofs = np.empty(max(nd.value, 1), np.int32)
        off = np.empty(nc.value + 1, np.int64)
        self._check(self._lib.lss_get_stiff_dofs(self._h, _ptr(dofs), dofs.size, ctypes.byref(nd), _ptr(off),
                                                 off.size, ctypes.byref(nc)))
        return [dofs[off[c]:off[c + 1]].copy() for c in range(nc.value)]

    def row_norms(self) -> np.ndarray:
        out = np.empty(self._n, np.float64)
        self._check(self._lib.lss_debug_row_norms(self._h, _ptr(out), self._n))
        return out

    def gmm(self) -> dict:
        params = np.zeros(8, np.float64)
        it, jumped, sel = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        self._check(self._lib.lss_debug_gmm(self._h, _ptr(params), ctypes.byref(it), ctypes.byref(jumped),
                                            ctypes.byref(sel)))
        keys = ["global_mean", "global_var", "mean0", "mean1", "var0", "var1", "w0", "w1"]
        d = dict(zip(keys, params.tolist()))
        d.update(iters=it.value, jumped=jumped.value, selected=sel.value)
        return d

    def partition(self, nb: int):
        nparts = ctypes.c_int32()
        self._check(self._lib.lss_debug_partition(self._h, None, 0, None, 0, ctypes.byref(nparts)))
        perm = np.empty(nb, np.int32)
        off = np.empty(nparts.value + 1, np.int32)
        self._check(self._lib.lss_debug_partition(self._h, _ptr(perm), nb, _ptr(off), off.size,
                                                  ctypes.byref(nparts)))
        return perm, off

    def spmv(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        self._check(self._lib.lss_debug_spmv(self._h, _ptr(x), _ptr(y)))
        return y

    def base_apply(self, r) -> np.ndarray:
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty_like(r)
        self._check(self._lib.lss_debug_base_apply(self._h, _ptr(r), _ptr(z)))
        return z

    def precond_apply(self, r) -> np.ndarray:
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty_like(r)
        self._check(self._lib.lss_debug_precond_apply(self._h, _ptr(r), _ptr(z)))
        return z


def _stream(stream):
    if stream is None:
        return None
    h = getattr(stream, "cuda_stream", stream)
    return _P(int(h))


def _csc_arrays(A):
    if isinstance(A, tuple):
        n, colptr, rowidx, vals = A
    elif hasattr(A, "colptr"):
        n, colptr, rowidx, vals = A.n, A.colptr, A.rowidx, A.vals
    else:
        import scipy.sparse as sp

        C = sp.csc_matrix(A)
        C.sort_indices()
        n, colptr, rowidx, vals = C.shape[0], C.indptr, C.indices, C.data
    return (int(n), np.ascontiguousarray(colptr, np.int32), np.ascontiguousarray(rowidx, np.int32),
            np.ascontiguousarray(vals, np.float64))
