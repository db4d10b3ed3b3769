"""The C-ABI library without a GPU: it loads, exports every entry point that
include/lss.h declares, and reports configuration and argument errors with
status codes rather than exiting (the reference exits, GPUHybrid.cu:34-63)."""
import ctypes
import os
import re

import pytest

from lss_amd import _lib
from lss_amd.solver import Solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "lss.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?(lss_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = _declared()
    assert len(names) >= 20
    assert sorted(_lib.EXPORTS) == names
    for name in names:
        assert hasattr(L, name), name


def _create(doc):
    L = _lib.load()
    h = ctypes.c_void_p()
    rc = L.lss_create(doc.encode(), ctypes.byref(h))
    msg = L.lss_last_error(h).decode() if h.value else ""
    if h.value:
        L.lss_destroy(h)
    return rc, msg


@pytest.mark.parametrize("doc,frag", [
    ('{"solver": "Eigen::SimplicialLDLT"}', "GPUHybrid"),
    ('{"solver": "GPUHybrid", "GPUHybrid": {"no_such_key": 1}}', "no_such_key"),
    ('{"solver": "GPUHybrid", "GPUHybrid": {"block_size": 4}}', "block_size"),
    ('{"solver": "GPUHybrid", "GPUHybrid": {"precision": "int8"}}', "precision"),
    ('{"solver": "GPUHybrid", "GPUHybrid": {"max_iter": "many"}}', "max_iter"),
    ('{"solver": "GPUHybrid", ', "JSON"),
    ('{"solver": "MAS", "MAS": {"true_residual_period": -1}}', "true_residual_period"),
    ('{"solver": "GPUHybrid", "GPUHybrid": {"base_preconditioner": "ILU"}}', "base_preconditioner"),
    ('{"solver": "GPUHybrid", "GPUHybrid": {"bank_apply": "wmma"}}', "bank_apply"),
    ('{"solver": "GPUHybrid", "GPUHybrid": {"bank_apply": "mfma", "mfma_flush": 6}}', "mfma_flush"),
    ('{"solver": "GPUHybrid", "GPUHybrid": {"sparse_batch_rows": -1}}', "sparse_batch_rows"),
    ('{"solver": "MAS", "MAS": {"sparse_batch_bytes": -5}}', "sparse_batch_bytes"),
])
def test_bad_config_is_rejected(doc, frag):
    rc, msg = _create(doc)
    assert rc == 1, (rc, msg)  # LSS_ERR_CONFIG
    assert frag.lower() in msg.lower()


def test_null_arguments():
    L = _lib.load()
    assert L.lss_create(b"{}", None) == 8  # LSS_ERR_ARG
    assert L.lss_solve(None, None, None) == 8
    L.lss_destroy(None)


def test_no_device_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(_lib.LssError) as e:
        Solver.create({"solver": "GPUHybrid"})
    assert e.value.code == 4  # LSS_ERR_HIP: no silent CPU path


# ---------------------------------------------------------------- wrapper argument checks
# The Python wrappers hand raw device pointers to the C ABI, which reads them as the
# documented type and length; every mismatch must raise before a native call.  A Solver
# without a native handle shows that the checks run first (no GPU is needed).
def _unbound(n=None):
    s = Solver.__new__(Solver)
    s._h, s._lib, s._n = None, None, n
    return s


def test_solve_before_factorize_raises():
    import numpy as np

    with pytest.raises(_lib.LssError) as e:
        _unbound().solve(np.zeros(3))
    assert e.value.code == 3  # LSS_ERR_NOT_FACTORIZED


@pytest.mark.parametrize("b,code", [([1.0, 2.0], 2), (list(range(5)), 2), ([[1.0, 2.0, 3.0]], 2)])
def test_host_vector_size_checked(b, code):
    with pytest.raises(_lib.LssError) as e:
        _unbound(3).solve(b)
    assert e.value.code == code


def test_host_matrix_arrays_checked():
    import numpy as np

    s = _unbound()
    with pytest.raises(_lib.LssError) as e:  # colptr of n entries instead of n + 1
        s.factorize((3, np.array([0, 1, 2], np.int32), np.array([0, 1, 2], np.int32), np.ones(3)))
    assert e.value.code == 2
    with pytest.raises(_lib.LssError) as e:  # colptr[n] beyond the value array
        s.factorize((2, np.array([0, 1, 4], np.int32), np.array([0, 1], np.int32), np.ones(2)))
    assert e.value.code == 2


def test_device_tensor_checks():
    import torch

    from lss_amd.solver import require_device

    cpu = torch.zeros(4, dtype=torch.float64)
    with pytest.raises(_lib.LssError) as e:
        require_device(cpu, torch.float64, "b", numel=4)
    assert e.value.code == 8 and "GPU" in str(e.value)
    with pytest.raises(_lib.LssError) as e:
        require_device([0.0] * 4, torch.float64, "b")
    assert e.value.code == 8
    # the device wrappers check their tensors before any native call: CPU tensors,
    # int64 indices and float32 values are all rejected with LSS_ERR_ARG
    s = _unbound(4)
    with pytest.raises(_lib.LssError) as e:
        s.solve_device(cpu, cpu)
    assert e.value.code == 8
    with pytest.raises(_lib.LssError) as e:
        s.factorize_device(4, torch.zeros(5, dtype=torch.int64), torch.zeros(4, dtype=torch.int64), cpu)
    assert e.value.code == 8
    with pytest.raises(_lib.LssError) as e:
        s.solve_multi_device(torch.zeros(2, 4, dtype=torch.float32), cpu)
    assert e.value.code == 8
    from lss_amd import autograd

    with pytest.raises(_lib.LssError) as e:
        autograd.pattern_grad(torch.zeros(5, dtype=torch.int64), torch.zeros(4, dtype=torch.int32), cpu, cpu)
    assert e.value.code == 8
