"""linear_solve end to end on the GPU: the reference's example system from Matrix Market files."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "local-stiff-solver_amd", "bin", "linear_solve")
GOLD = os.path.join(ROOT, "tests", "golden")


def _fixture(tmp_path):
    d = np.load(os.path.join(GOLD, "A_contact.npz"))
    n, ptr, idx, vals = int(d["n"]), d["colptr"], d["rowidx"], d["vals"]
    f = tmp_path / "A_contact.mtx"
    cols = np.repeat(np.arange(n), np.diff(ptr))
    with open(f, "w") as h:
        h.write("%%MatrixMarket matrix coordinate real general\n")
        h.write(f"{n} {n} {vals.size}\n")
        h.writelines(f"{int(r) + 1} {int(c) + 1} {float(v)!r}\n" for r, c, v in zip(idx, cols, vals))
    return f, d


def _spec(tmp_path, **kw):
    g = {"block_size": 3, "relative_tolerance": 1e-8, "gmm_jump_threshold": 10.0, "min_subdomain_size": 4}
    g.update(kw)
    f = tmp_path / "spec.json"
    f.write_text(json.dumps({"solver": "GPUHybrid", "GPUHybrid": g}))
    return f


def _read_vector(path):
    with open(path) as h:
        lines = [s for s in h.read().splitlines() if s and not s.startswith("%")]
    n = int(lines[0].split()[0])
    return np.array([float(s) for s in lines[1:1 + n]])


def test_linear_solve_matrix_market_run(gpu, tmp_path):
    A, d = _fixture(tmp_path)
    spec = _spec(tmp_path)
    xo = tmp_path / "x.mtx"
    p = subprocess.run([CLI, "-A", str(A), "-j", str(spec), "-w", "1", "-r", "2", "-o", str(xo), "--info"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    out = p.stdout.splitlines()
    res = [s for s in out if "[residual=" in s]
    assert len(res) == 2  # the warm-up repetition prints nothing
    assert all(s.startswith("[GPUHybrid]") for s in res)
    rel = float(res[-1].split("[relative_residual=")[1].split("]")[0])
    # operator rows reach 6e11: the FP64 floor of the true residual is near 1e-6
    assert rel <= 2e-6
    infos = [json.loads(s) for s in out if s.startswith("{")]
    assert len(infos) == 2 and infos[-1]["stop_reason"] == "converged_rel"
    x = _read_vector(xo)
    g = np.load(os.path.join(GOLD, "A_contact_golden.npz"))
    assert np.linalg.norm(x - g["x_direct"]) <= 1e-6 * np.linalg.norm(g["x_direct"])


def test_linear_solve_problem_info_gives_user_stiff_set(gpu, tmp_path):
    A, d = _fixture(tmp_path)
    spec = _spec(tmp_path, select_bad_dofs_from_l1_norm=False)
    ind = O.row_l1(d["colptr"], d["vals"])
    fi = tmp_path / "info.mtx"
    with open(fi, "w") as h:
        h.write(f"%%MatrixMarket matrix array real general\n{ind.size} 1\n")
        h.writelines(f"{float(v)!r}\n" for v in ind)
    thr = float(np.quantile(ind, 0.97))
    p = subprocess.run([CLI, "-A", str(A), "-j", str(spec), "-w", "0", "-i", str(fi), "-t", repr(thr), "--info"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    res = [s for s in p.stdout.splitlines() if "[residual=" in s]
    assert len(res) == 1 and res[0].startswith("[GPUAMGF]")
    info = json.loads([s for s in p.stdout.splitlines() if s.startswith("{")][0])
    assert info["stop_reason"] == "converged_rel"
    assert info["stiff"]["n_selected"] == int((ind > thr).sum())


def test_linear_solve_rejects_other_solvers(gpu, tmp_path):
    A, _ = _fixture(tmp_path)
    f = tmp_path / "ldlt.json"
    f.write_text(json.dumps({"solver": "Eigen::SimplicialLDLT"}))
    p = subprocess.run([CLI, "-A", str(A), "-j", str(f)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1 and "configuration" in p.stderr
