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
LSS_ERR_SIZE, LSS_ERR_NOT_FACTORIZED, LSS_ERR_ARG = 2, 3, 8


def _ptr(a: np.ndarray):
    return _P(a.ctypes.data)


def require_device(t, dtype, name: str, numel: int | None = None, shape: tuple | None = None):
    """Check a tensor handed to the C ABI as a raw device pointer: a contiguous GPU
    tensor of `dtype` with `numel` elements (or `shape`).  The C side reads the
    pointer as that type and length, so anything else would be read out of bounds."""
    import torch

    if not isinstance(t, torch.Tensor):
        raise LssError(LSS_ERR_ARG, f"{name} must be a torch.Tensor")
    if not t.is_cuda:
        raise LssError(LSS_ERR_ARG, f"{name} must be a GPU tensor (got device {t.device})")
    if t.dtype != dtype:
        raise LssError(LSS_ERR_ARG, f"{name} must be {dtype} (got {t.dtype})")
    if not t.is_contiguous():
        raise LssError(LSS_ERR_ARG, f"{name} must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise LssError(LSS_ERR_SIZE, f"{name} must have shape {tuple(shape)} (got {tuple(t.shape)})")
    if numel is not None and t.numel() != numel:
        raise LssError(LSS_ERR_SIZE, f"{name} must have {numel} elements (got {t.numel()})")
    return t


def _host_vector(a, n: int, name: str) -> np.ndarray:
    v = np.ascontiguousarray(a, dtype=np.float64)
    if v.ndim != 1 or v.size != n:
        raise LssError(LSS_ERR_SIZE, f"{name} must be a vector of {n} values (got shape {v.shape})")
    return v


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
        if colptr.size != n + 1:
            raise LssError(LSS_ERR_SIZE, f"colptr must have n + 1 = {n + 1} entries (got {colptr.size})")
        nnz = int(colptr[-1])
        if rowidx.size < nnz or vals.size < nnz:
            raise LssError(LSS_ERR_SIZE, f"rowidx and vals must hold colptr[n] = {nnz} entries")
        self._check(self._lib.lss_factorize_csc(self._h, n, int(colptr[-1]), _ptr(colptr), _ptr(rowidx), _ptr(vals)))
        self._n = n

    def factorize_device(self, n: int, colptr, rowidx, vals, stream=None):
        """Device CSC tensors (torch, on the GPU): colptr, rowidx int32; vals float64."""
        import torch

        require_device(colptr, torch.int32, "colptr", numel=n + 1)
        require_device(rowidx, torch.int32, "rowidx")
        nnz = int(rowidx.numel())
        require_device(vals, torch.float64, "vals", numel=nnz)
        self._check(self._lib.lss_factorize_csc_device(self._h, n, nnz, _P(colptr.data_ptr()),
                                                       _P(rowidx.data_ptr()), _P(vals.data_ptr()),
                                                       _stream(stream)))
        self._n = n

    def solve(self, b, x=None) -> np.ndarray:
        """Solve A x = b; x (optional) is the initial guess (GPUHybrid.cu:342-348)."""
        n = self._factorized_size()
        b = _host_vector(b, n, "b")
        x = np.zeros_like(b) if x is None else _host_vector(x, n, "x").copy()
        self._check(self._lib.lss_solve(self._h, _ptr(b), _ptr(x)))
        return x

    def solve_device(self, b, x, stream=None):
        """b, x: contiguous float64 torch tensors on the GPU; x is updated in place."""
        import torch

        n = self._factorized_size()
        require_device(b, torch.float64, "b", numel=n)
        require_device(x, torch.float64, "x", numel=n)
        self._check(self._lib.lss_solve_device(self._h, _P(b.data_ptr()), _P(x.data_ptr()), _stream(stream)))
        return x

    def solve_multi_device(self, B, X, stream=None):
        """B, X: (k, n) contiguous float64 device tensors (one right-hand side per row)."""
        import torch

        n = self._factorized_size()
        require_device(B, torch.float64, "B")
        if B.dim() != 2:
            raise LssError(LSS_ERR_SIZE, f"B must be (k, n) (got shape {tuple(B.shape)})")
        k = int(B.shape[0])
        require_device(B, torch.float64, "B", shape=(k, n))
        require_device(X, torch.float64, "X", shape=(k, n))
        self._check(self._lib.lss_solve_multi_device(self._h, k, _P(B.data_ptr()), _P(X.data_ptr()),
                                                     _stream(stream)))
        return X

    def _factorized_size(self) -> int:
        if self._n is None:
            raise LssError(LSS_ERR_NOT_FACTORIZED, "solve called before factorize")
        return self._n

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
        out = np.empty(self._factorized_size(), np.float64)
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

    def amg_levels(self) -> list[dict]:
        """The AMG hierarchy (base_preconditioner "AMG"): per level n, cf, A (levels >= 1)
        and P as scipy CSR matrices."""
        import scipy.sparse as sp

        c_i32, c_i64 = ctypes.c_int32, ctypes.c_int64
        nl = c_i32()
        self._check(self._lib.lss_debug_amg(self._h, -1, ctypes.byref(nl), None, None, None, None, None, None, None,
                                            None, None, None))
        out = []
        for lv in range(nl.value):
            n, na, npp = c_i32(), c_i64(), c_i64()
            self._check(self._lib.lss_debug_amg(self._h, lv, None, ctypes.byref(n), ctypes.byref(na),
                                                ctypes.byref(npp), None, None, None, None, None, None, None))
            last = lv == nl.value - 1
            cf = np.zeros(n.value if not last else 1, np.int8)
            ap = np.zeros(n.value + 1, np.int64)
            ac = np.zeros(max(na.value, 1), np.int32)
            av = np.zeros(max(na.value, 1), np.float64)
            pp = np.zeros(n.value + 1, np.int64)
            pc = np.zeros(max(npp.value, 1), np.int32)
            pv = np.zeros(max(npp.value, 1), np.float64)
            self._check(self._lib.lss_debug_amg(self._h, lv, None, None, None, None,
                                                None if last else _ptr(cf), _ptr(ap) if lv else None,
                                                _ptr(ac) if lv else None, _ptr(av) if lv else None,
                                                None if last else _ptr(pp), None if last else _ptr(pc),
                                                None if last else _ptr(pv)))
            d = {"n": n.value, "cf": None if last else cf}
            if lv:
                d["A"] = sp.csr_matrix((av[:na.value], ac[:na.value], ap), shape=(n.value, n.value))
            if not last:
                d["P"] = (pp, pc[:npp.value], pv[:npp.value])
            out.append(d)
        for lv in range(len(out) - 1):
            pp, pc, pv = out[lv]["P"]
            out[lv]["P"] = sp.csr_matrix((pv, pc, pp), shape=(out[lv]["n"], out[lv + 1]["n"]))
        return out

    def spmv(self, x) -> np.ndarray:
        x = _host_vector(x, self._factorized_size(), "x")
        y = np.empty_like(x)
        self._check(self._lib.lss_debug_spmv(self._h, _ptr(x), _ptr(y)))
        return y

    def base_apply(self, r) -> np.ndarray:
        r = _host_vector(r, self._factorized_size(), "r")
        z = np.empty_like(r)
        self._check(self._lib.lss_debug_base_apply(self._h, _ptr(r), _ptr(z)))
        return z

    def precond_apply(self, r) -> np.ndarray:
        r = _host_vector(r, self._factorized_size(), "r")
        z = np.empty_like(r)
        self._check(self._lib.lss_debug_precond_apply(self._h, _ptr(r), _ptr(z)))
        return z


def residual_norm(colptr, rowidx, vals, b, x, stream=None) -> tuple[float, float]:
    """(||b - A x||_2, ||b||_2) in double-double arithmetic on the GPU (lss_residual_norm);
    device CSC tensors (int32, int32, float64) read as operator rows, float64 b and x."""
    import torch

    require_device(colptr, torch.int32, "colptr")
    n = int(colptr.numel()) - 1
    require_device(rowidx, torch.int32, "rowidx")
    nnz = int(rowidx.numel())
    require_device(vals, torch.float64, "vals", numel=nnz)
    require_device(b, torch.float64, "b", numel=n)
    require_device(x, torch.float64, "x", numel=n)
    out = torch.empty(2, dtype=torch.float64, device=b.device)
    lib = load()
    st = _stream(stream) if stream is not None else _P(int(torch.cuda.current_stream().cuda_stream))
    rc = lib.lss_residual_norm(n, nnz, _P(colptr.data_ptr()), _P(rowidx.data_ptr()), _P(vals.data_ptr()),
                               _P(b.data_ptr()), _P(x.data_ptr()), _P(out.data_ptr()), st)
    if rc != 0:
        raise LssError(rc, lib.lss_last_error(None).decode())
    h = out.cpu()
    return float(h[0]), float(h[1])


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


def nested_dissection(adj_ptr, adj_idx, leaf: int = 128) -> dict:
    """Ordering and front structure the sparse stiff factorization builds for a
    symmetric graph (lss_debug_nested_dissection; host only, no GPU needed).
    adj_ptr/adj_idx: CSR without self loops.  Returns order (position -> vertex) and,
    per front in postorder, p0 / np / parent and its update positions struct_ptr /
    struct_idx."""
    lib = load()
    ap = np.ascontiguousarray(adj_ptr, dtype=np.int32)
    ai = np.ascontiguousarray(adj_idx, dtype=np.int32)
    m = ap.size - 1
    nf = ctypes.c_int32()
    ns = ctypes.c_int64()

    def call(*arrays):
        rc = lib.lss_debug_nested_dissection(m, _ptr(ap), _ptr(ai) if ai.size else None, int(leaf), *arrays[:1],
                                             ctypes.byref(nf), ctypes.byref(ns), *arrays[1:])
        if rc:
            raise LssError(rc, lib.lss_last_error(None).decode())

    call(None, None, None, None, None, None)
    out = {"order": np.empty(m, np.int32), "p0": np.empty(nf.value, np.int32), "np": np.empty(nf.value, np.int32),
           "parent": np.empty(nf.value, np.int32), "struct_ptr": np.empty(nf.value + 1, np.int64),
           "struct_idx": np.empty(max(ns.value, 1), np.int32)}
    call(_ptr(out["order"]), _ptr(out["p0"]), _ptr(out["np"]), _ptr(out["parent"]), _ptr(out["struct_ptr"]),
         _ptr(out["struct_idx"]))
    out["struct_idx"] = out["struct_idx"][: ns.value]
    return out
