"""Build and run the C++ AMG restatement (oracle/amg_ref.cpp) through
tests/amg_setup_shim.cpp: the whole hierarchy of a CSR matrix as scipy matrices.
Test infrastructure only."""
import ctypes
import os
import shutil
import struct
import subprocess

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(HERE, "..", "oracle")


def build(workdir: str):
    """Compile the shim into workdir; returns run(A, block_size, theta) -> levels."""
    cxx = shutil.which("g++")
    if cxx is None:
        raise FileNotFoundError("g++")
    so = os.path.join(workdir, "amgref.so")
    subprocess.run([cxx, "-O2", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-Wno-unknown-pragmas", "-shared",
                    "-fPIC", "-I", ORACLE, os.path.join(HERE, "amg_setup_shim.cpp"),
                    os.path.join(ORACLE, "amg_ref.cpp"), "-o", so], check=True)
    lib = ctypes.CDLL(so)
    P = ctypes.c_void_p
    lib.amg_shim_setup.argtypes = [ctypes.c_int32, P, P, P, ctypes.c_int32, ctypes.c_double, ctypes.c_char_p]
    lib.amg_shim_setup.restype = ctypes.c_int

    def run(A, bs, theta):
        """levels: dicts with cf (int8), A (scipy CSR) and P (None on the coarsest)."""
        A = sp.csr_matrix(A)
        A.sort_indices()
        ptr = A.indptr.astype(np.int64)
        col = A.indices.astype(np.int32)
        val = np.ascontiguousarray(A.data, dtype=np.float64)
        out = os.path.join(workdir, "levels.bin")
        rc = lib.amg_shim_setup(A.shape[0], ptr.ctypes.data, col.ctypes.data, val.ctypes.data, bs, theta,
                                out.encode())
        assert rc == 0
        levels = []
        with open(out, "rb") as f:
            def vec(dt):
                k = struct.unpack("q", f.read(8))[0]
                return np.frombuffer(f.read(k * np.dtype(dt).itemsize), dt)

            nl = struct.unpack("q", f.read(8))[0]
            for _ in range(nl):
                cf, ap, ac, av = vec(np.int8), vec(np.int64), vec(np.int32), vec(np.float64)
                pp, pc, pv = vec(np.int64), vec(np.int32), vec(np.float64)
                m = struct.unpack("q", f.read(8))[0]
                n = ap.size - 1
                levels.append({"cf": cf, "A": sp.csr_matrix((av, ac, ap), shape=(n, n)),
                               "P": sp.csr_matrix((pv, pc, pp), shape=(n, m)) if pp.size else None})
        os.remove(out)
        return levels

    return run
