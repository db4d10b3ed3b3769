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
