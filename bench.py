"""Benchmark: time-to-1e-8 residual of the hybrid solve on the 128^3 contact system.

One step = factorize + solve (GPUHybrid factorize()/solve()) of the 6.29M-DOF
contact-stiffened elasticity system from device-resident CSC arrays to a
relative residual of 1e-8.  The mixed-precision path (FP32 MAS inverses and
preconditioner-internal SpMV, FP64 outer PCG and stiff Cholesky) is `value`;
the plain FP64 path is timed beside it.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "local-stiff-solver_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def spmv_bytes(info: dict, value_bytes: int = 8) -> float:
    """Algorithmic bytes of one BSR SpMV: nnzb (b^2 s_v + 4) + (nb + 1) 4 + n s_x + n s_y."""
    bs = info["block_size"]
    nnzb = info["nnz_blocks"]
    n = info["n"]
    nb = (n + bs - 1) // bs
    return nnzb * (bs * bs * value_bytes + 4) + (nb + 1) * 4 + n * 8 + n * 8


def load_traffic(workload: str):
    """HBM bytes per SpMV launch from the committed rocprofv3 --pmc summary, if present."""
    path = os.path.join(ROOT, "profiles", f"spmv_{workload}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("traffic_bytes_per_launch")


def host_cores():
    """(threads this process may run on, logical CPUs of the host)."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return aff, os.cpu_count() or aff


def cpu_baseline(prob, A, b, gpu_iters: int, sample_iters: int, threads: int, full: bool):
    """Oracle (CPU FP64 restatement) factorize + a bounded PCG sample, extrapolated to the
    GPU's iteration count; with full=True the oracle solves to convergence instead."""
    from oracle import oracle as O

    aff, ncpu = host_cores()
    cfg = O.config(block_size=prob.block_size, rel_tol=1e-8, max_iter=(10000 if full else sample_iters),
                   threads=threads, gmm_jump_threshold=10.0, min_subdomain_size=4)
    H = O.Hybrid(cfg)
    t0 = time.perf_counter()
    H.factorize(A.colptr, A.rowidx, A.vals)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    H.solve(b)
    t_loop = time.perf_counter() - t0
    its = H.info()["num_iterations"]
    t_iter = t_loop / max(1, its)
    value = t_setup + (t_loop if full else t_iter * gpu_iters)
    # the same PCG iteration on one thread (a few iterations, same factorization)
    H.set_run(1, 3)
    t0 = time.perf_counter()
    H.solve(b)
    t_iter1 = (time.perf_counter() - t0) / max(1, H.info()["num_iterations"])
    if full:
        sample = (f"oracle factorize ({t_setup:.2f} s) + PCG to convergence ({its} iterations, "
                  f"{t_iter * 1e3:.1f} ms each) on {threads} threads, measured")
    else:
        sample = (f"oracle factorize ({t_setup:.2f} s) + {its} PCG iterations ({t_iter * 1e3:.1f} ms each) "
                  f"on {threads} threads on the same system, extrapolated to the {gpu_iters} iterations of "
                  f"the FP64 GPU path; 3 iterations on 1 thread: {t_iter1 * 1e3:.0f} ms each")
    return {"value": round(value, 3), "unit": "s", "cores": threads, "kind": "port",
            "host_threads_available": aff, "host_logical_cpus": ncpu,
            "factorize_s": round(t_setup, 3), "iterations_run": its, "measured_to_convergence": full,
            "pcg_iteration_s": {str(threads): round(t_iter, 4), "1": round(t_iter1, 4)}, "sample": sample}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="elastic_contact")
    ap.add_argument("--N", type=int, default=None, help="grid nodes per axis (default: the workload's)")
    ap.add_argument("--cpu-sample-iters", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (default: all this process may use)")
    ap.add_argument("--cpu-full", action="store_true",
                    help="run the CPU baseline to convergence instead of a bounded sample")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward+adjoint timing")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--max-iter", type=int, default=20000,
                    help="PCG iteration cap (profiling runs cut it; such a line is marked invalid)")
    ap.add_argument("--mixed-precision", default="fp32", choices=["fp32", "bf16", "fp16"],
                    help="bank storage of the mixed path")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=JSON",
                    help="extra GPUHybrid configuration entry for both paths (experiments)")
    ap.add_argument("--direct-max", type=int, default=1_000_000,
                    help="largest system (DOF) whose accuracy reference is an FP64 sparse-direct GPU solve")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device (HBM counter calibration)")
    ap.add_argument("--multi-rhs", type=int, default=0, metavar="K",
                    help="also time one fused K-right-hand-side solve of the mixed path against K single solves")
    ap.add_argument("--amg", action="store_true",
                    help="also time the AMG base preconditioner (BoomerAMG restated): hybrid with base AMG and "
                         "AMG alone (GPUAMG), FP64")
    ap.add_argument("--mas-solver", action="store_true",
                    help="also time the standalone MAS solver (its defaults, rtol 1e-6, at most 3000 iterations) "
                         "without and with each stiff-DOF strategy")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the solver has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from lss_amd import problems
    from lss_amd.autograd import pattern_grad
    from lss_amd.solver import Solver, residual_norm

    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    if args.copy_probe_mib:
        # a plain copy of known size: FETCH_SIZE / WRITE_SIZE per dispatch of this
        # kernel calibrate the counters against exact byte counts
        src = torch.full((args.copy_probe_mib << 17,), 1.0, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)  # noqa: E731
    prob = problems.make(args.workload, N=args.N)
    log("workload", args.workload, "N", prob.N)
    A = prob.csc()
    b = prob.rhs(0)
    n = A.n
    d_ptr = torch.from_numpy(A.colptr).to(dev)
    d_idx = torch.from_numpy(A.rowidx).to(dev)
    d_val = torch.from_numpy(A.vals).to(dev)
    d_b = torch.from_numpy(b).to(dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    log("system on the device:", n, "DOF,", A.nnz, "stored entries")
    bnorm = float(np.linalg.norm(b))
    stream = torch.cuda.current_stream()
    base = {"block_size": prob.block_size, "relative_tolerance": 1e-8, "gmm_jump_threshold": 10.0,
            "min_subdomain_size": 4, "max_iter": args.max_iter}
    for kv in args.set:
        k, v = kv.split("=", 1)
        base[k] = json.loads(v)
    paths = {"fp64": dict(base, precision="fp64"), "mixed": dict(base, precision=args.mixed_precision)}
    res = {}
    cp_d, ri_d, va_d = d_ptr, d_idx, d_val
    # forward+adjoint: L = 0.5 ||x - x_t||^2, dL/dx = x - x_t, A^T lam = dL/dx (A symmetric:
    # the same factorization), dL/dA = -lam x^T on the stored pattern
    x_t = torch.from_numpy(problems.random_vector(n, 1)).to(dev)
    lam = torch.zeros_like(x)
    for name, g in paths.items():
        s = Solver.create({"solver": "GPUHybrid", "GPUHybrid": g})

        def step():
            s.factorize_device(n, d_ptr, d_idx, d_val, stream=stream)
            x.zero_()
            s.solve_device(d_b, x, stream=stream)

        def adj_step():
            step()
            gx = x - x_t
            lam.zero_()
            s.solve_device(gx, lam, stream=stream)
            return pattern_grad(cp_d, ri_d, lam, x)

        for _ in range(args.warmup):
            step()
            log(name, "warm-up step done")
        times = []
        for _ in range(args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            step()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        info = s.get_info()
        rel = info["final_res_norm"] / bnorm
        r_dd, _ = residual_norm(cp_d, ri_d, va_d, d_b, x)
        res[name] = {"s": float(np.median(times)), "all": times, "iters": info["num_iterations"], "rel_res": rel,
                     "rel_res_dd": r_dd / bnorm, "info": info, "x": x.cpu().numpy().copy(),
                     "stop": info["stop_reason"]}
        log(name, f"{res[name]['s']:.3f} s (median of {args.steps})", info["num_iterations"], "iterations",
            info["stop_reason"], f"true rel residual (double-double) {r_dd / bnorm:.3e}")
        if not args.no_adjoint:
            for _ in range(args.warmup):
                adj_step()
            at = []
            for _ in range(args.steps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                grad = adj_step()
                torch.cuda.synchronize()
                at.append(time.perf_counter() - t0)
            ainfo = s.get_info()
            res[name]["adjoint"] = {"s": float(np.median(at)), "all": at, "adjoint_iters": ainfo["num_iterations"],
                                    "adjoint_stop": ainfo["stop_reason"],
                                    "grad_norm": float(torch.linalg.norm(grad))}
            log(name, f"forward+adjoint {res[name]['adjoint']['s']:.3f} s")
        del s
    # accuracy: the reference solution is, for systems of up to 40 000 DOF, scipy's sparse
    # LU; up to --direct-max DOF, an FP64 sparse-direct solve on the GPU (the multifrontal
    # Cholesky of the stiff-unit path with every DOF in one unit, so the first PCG iterate is
    # the direct solution); above, the FP64 GPU path solved far past the benchmark tolerance
    ref_kind = "FP64 GPU path, recurrence residual 1e-13"
    s = Solver.create({"solver": "GPUHybrid", "GPUHybrid": dict(paths["fp64"], relative_tolerance=1e-13)})
    s.factorize_device(n, d_ptr, d_idx, d_val, stream=stream)
    x.zero_()
    s.solve_device(d_b, x, stream=stream)
    ref_info = s.get_info()
    x_deep = x.cpu().numpy().copy()
    x_ref = x_deep
    r_ref_dd, _ = residual_norm(cp_d, ri_d, va_d, d_b, x)
    ref_iters = ref_info["num_iterations"]
    del s
    if n <= 40_000:
        import scipy.sparse.linalg as spla

        x_ref = spla.spsolve(A.to_scipy().tocsc(), b)
        ref_kind = "scipy sparse LU (SuperLU), FP64"
        ref_iters = None
        r_ref_dd, _ = residual_norm(cp_d, ri_d, va_d, d_b, torch.from_numpy(x_ref).to(dev))
    elif n <= args.direct_max:
        g = {"block_size": prob.block_size, "relative_tolerance": 1e-13, "max_iter": 10,
             "select_bad_dofs_from_l1_norm": False, "decompose_subdomains": False,
             "stiff_factorization": "sparse"}
        s = Solver.create({"solver": "GPUHybrid", "GPUHybrid": g})
        s.set_problematic_dofs(np.arange(n, dtype=np.int32))
        t0 = time.perf_counter()
        s.factorize_device(n, d_ptr, d_idx, d_val, stream=stream)
        x.zero_()
        s.solve_device(d_b, x, stream=stream)
        torch.cuda.synchronize()
        t_direct = time.perf_counter() - t0
        dinfo = s.get_info()
        sp = dinfo["stiff"]["sparse"]
        x_ref = x.cpu().numpy().copy()
        r_ref_dd, _ = residual_norm(cp_d, ri_d, va_d, d_b, x)
        ref_iters = dinfo["num_iterations"]
        ref_kind = (f"FP64 sparse direct on the GPU (multifrontal Cholesky, every DOF one unit: "
                    f"{sp['fronts'][0]} fronts, largest {sp['max_front'][0]} rows, {sp['factor_entries'][0]:.3e} "
                    f"factor entries, {t_direct:.2f} s; PCG iterations {dinfo['num_iterations']})")
        log("direct reference", ref_kind)
        del s
    accuracy = {"reference": ref_kind, "reference_rel_residual_dd": r_ref_dd / bnorm,
                "reference_iterations": ref_iters}
    for name in ("mixed", "fp64"):
        accuracy[f"x_{name}_vs_ref"] = float(np.linalg.norm(res[name]["x"] - x_ref) / np.linalg.norm(x_ref))
    if x_ref is not x_deep:  # the FP64 path iterated to 1e-13, against the direct reference
        accuracy["x_fp64_deep_vs_ref"] = float(np.linalg.norm(x_deep - x_ref) / np.linalg.norm(x_ref))
        accuracy["fp64_deep_iterations"] = ref_info["num_iterations"]
    multi = None
    if args.multi_rhs > 1:
        # solve phase only: one factorization, K right-hand sides fused vs one by one
        K = args.multi_rhs
        s = Solver.create({"solver": "GPUHybrid", "GPUHybrid": paths["mixed"]})
        s.factorize_device(n, d_ptr, d_idx, d_val, stream=stream)
        dB = torch.stack([torch.from_numpy(prob.rhs(c)).to(dev) for c in range(K)])
        dX = torch.zeros_like(dB)
        s.solve_multi_device(dB, dX, stream=stream)  # warm-up
        dX.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.solve_multi_device(dB, dX, stream=stream)
        torch.cuda.synchronize()
        t_fused = time.perf_counter() - t0
        minfo = s.get_info()
        xs = torch.zeros_like(dB)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for c in range(K):
            s.solve_device(dB[c], xs[c], stream=stream)
        torch.cuda.synchronize()
        t_seq = time.perf_counter() - t0
        bn = torch.linalg.norm(dB, dim=1).cpu().numpy()
        multi = {"k": K, "fused_solve_s": round(t_fused, 4), "single_solves_s": round(t_seq, 4),
                 "speedup": round(t_seq / t_fused, 3),
                 "bit_identical_to_single": bool(torch.equal(dX, xs)),
                 "iterations": [c["num_iterations"] for c in minfo["columns"]],
                 "rel_residual": [c["final_res_norm"] / float(bn[i]) for i, c in enumerate(minfo["columns"])]}
        log("multi-rhs", K, f"fused {t_fused:.3f} s", f"single {t_seq:.3f} s")
        del s, dB, dX, xs
    mas_runs = None
    if args.mas_solver:
        mas_runs = {}
        variants = {"mas": {}, "mas+gmm": {"use_bad_dof_precond": True, "bad_dof_strategy": "gmm"},
                    "mas+elbow": {"use_bad_dof_precond": True, "bad_dof_strategy": "elbow"},
                    "mas+kmean": {"use_bad_dof_precond": True, "bad_dof_strategy": "kmean"},
                    # the recurrence residual kept (no b - A x every 4 iterations)
                    "mas/keep_r": {"true_residual_period": 0},
                    "mas+gmm/keep_r": {"use_bad_dof_precond": True, "bad_dof_strategy": "gmm",
                                       "true_residual_period": 0}}
        for vname, extra in variants.items():
            # MAS alone converges slowly on contact stiffness: cap the iterations
            m = {"block_dim": prob.block_size, "precision": args.mixed_precision, "max_iter": 3000}
            m.update(extra)
            log(vname, "start")
            s = Solver.create({"solver": "MAS", "MAS": m})

            def mstep():
                s.factorize_device(n, d_ptr, d_idx, d_val, stream=stream)
                x.zero_()
                s.solve_device(d_b, x, stream=stream)

            try:
                mstep()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                mstep()
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                mi = s.get_info()
                mas_runs[vname] = {"s": round(dt, 4), "factorize_s": round(mi["timings"]["factorize"], 4),
                                   "iterations": mi["solver_iter"], "status": mi["solver_status"],
                                   "rel_residual": mi["final_res_norm"] / bnorm,
                                   "stiff_dofs": mi["stiff"]["n_stiff_dofs"],
                                   "components": mi["stiff"]["n_components"],
                                   "max_component": mi["stiff"]["max_component"]}
                sp = mi["stiff"]["sparse"]
                if sp["units"]:  # units above max_dense_component: the multifrontal factor
                    mas_runs[vname]["sparse"] = {k: sp[k] for k in ("dofs", "fronts", "levels", "max_front",
                                                                     "factor_entries", "t_symbolic", "t_numeric")}
            except Exception as e:  # e.g. a front beyond the sparse factor's dense workspace limit
                mas_runs[vname] = {"error": str(e)}
            log(vname, mas_runs[vname])
            del s
    amg_runs = None
    if args.amg:
        amg_runs = {}
        # AMG alone has no treatment of the contact stiffness: its iterations are capped
        for vname, extra in {"hybrid_amg": {}, "amg_only": {"do_hybrid_precond": False, "max_iter": 1000}}.items():
            g = dict(paths["fp64"], base_preconditioner="AMG", **extra)
            log(vname, "start")
            s = Solver.create({"solver": "GPUHybrid", "GPUHybrid": g})

            def astep():
                s.factorize_device(n, d_ptr, d_idx, d_val, stream=stream)
                x.zero_()
                s.solve_device(d_b, x, stream=stream)

            astep()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            astep()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            ai = s.get_info()
            rdd = residual_norm(cp_d, ri_d, va_d, d_b, x, stream=stream)[0]
            amg_runs[vname] = {"solver": ai["solver"], "s": round(dt, 4), "phases_s": ai["timings"],
                               "amg_setup_s": ai["timings"]["mas"], "iterations": ai["num_iterations"],
                               "stop_reason": ai["stop_reason"], "true_rel_residual_dd": rdd / bnorm,
                               "levels": ai["amg"]["rows"], "level_nnz": ai["amg"]["nnz"],
                               "setup_phases_s": ai["amg"]["setup_s"]}
            log(vname, amg_runs[vname])
            del s
    # roofline of the outer FP64 SpMV, timed with HIP events on the solver's stream
    g = dict(paths["mixed"], profile_kernels=True)
    s = Solver.create({"solver": "GPUHybrid", "GPUHybrid": g})
    s.factorize_device(n, d_ptr, d_idx, d_val, stream=stream)
    x.zero_()
    s.solve_device(d_b, x, stream=stream)
    pinfo = s.get_info()
    del s
    spmv_ms = pinfo["kernel_ms"]["spmv_outer"]
    by = spmv_bytes(pinfo)
    achieved = by / (spmv_ms * 1e-3) / 1e9
    workload = f"{args.workload}_{prob.N}"
    traffic = load_traffic(workload)
    mixed, fp64 = res["mixed"], res["fp64"]
    # the reference stops on the recurrence residual (GPUHybrid.cu:1484-1496), as both
    # paths here do; beside it the true residual of each solution in double-double
    # arithmetic and eps || |A| |x| ||, below which no vector of doubles gets on stiff rows
    valid = mixed["stop"] == "converged_rel" and fp64["stop"] == "converged_rel"
    S = A.to_scipy().tocsr()
    absx = np.abs(fp64["x"])
    floor = float(np.linalg.norm(abs(S) @ absx) * np.finfo(np.float64).eps / bnorm)
    x_diff = float(np.linalg.norm(mixed["x"] - fp64["x"]) / np.linalg.norm(fp64["x"]))
    cpu = None
    if not args.no_cpu:
        log("cpu baseline", "to convergence" if args.cpu_full else "sample")
        threads = args.cpu_threads or host_cores()[0]
        cpu = cpu_baseline(prob, A, b, fp64["iters"], args.cpu_sample_iters, threads, args.cpu_full)
    line = {
        "metric": "time-to-1e-8 residual (s), mixed vs FP64 GPU, 6.3M-DOF; SpMV HBM GB/s",
        "value": round(mixed["s"], 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["s"] * 1e3, 2),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64 outer PCG / f32 MAS",
        "data": "synthetic",
        "config": {"workload": workload, "n_dof": n, "nnz": A.nnz, "block_size": prob.block_size,
                   "contacts": prob.n_contacts, "rel_tol": 1e-8},
        "fp64_path_s": round(fp64["s"], 4),
        "speedup_mixed_vs_fp64": round(fp64["s"] / mixed["s"], 3),
        "iterations": {"mixed": mixed["iters"], "fp64": fp64["iters"]},
        "rel_residual": {"mixed": mixed["rel_res"], "fp64": fp64["rel_res"]},
        "true_rel_residual_dd": {"mixed": mixed["rel_res_dd"], "fp64": fp64["rel_res_dd"]},
        "stop_reason": {"mixed": mixed["stop"], "fp64": fp64["stop"]},
        "rel_residual_fp64_floor": floor,
        "x_mixed_vs_fp64": x_diff,
        "accuracy": accuracy,
        "valid": valid,
        "step_s": {"mixed": [round(t, 4) for t in mixed["all"]], "fp64": [round(t, 4) for t in fp64["all"]]},
        "phases_mixed_s": mixed["info"]["timings"],
        "stiff": {"n_stiff_dofs": mixed["info"]["stiff"]["n_stiff_dofs"],
                  "n_components": mixed["info"]["stiff"]["n_components"],
                  "max_component": mixed["info"]["stiff"]["max_component"]},
        "mas": mixed["info"]["mas"],
        "sell_padding": round(mixed["info"]["sell_slots"] / mixed["info"]["nnz_blocks"] - 1.0, 4),
        "roofline": {"kernel": "spmv_sell_bs3_fp64", "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "bytes_per_launch": by, "avg_launch_ms": spmv_ms,
                     "launches_timed": pinfo["kernel_ms"]["spmv_outer_n"]},
        "kernel_ms": pinfo["kernel_ms"],
        "cpu_baseline": cpu,
    }
    if "adjoint" in mixed:
        ma, fa = mixed["adjoint"], fp64["adjoint"]
        line["forward_adjoint"] = {
            "mixed_s": round(ma["s"], 4), "fp64_s": round(fa["s"], 4), "speedup_mixed_vs_fp64": round(fa["s"] / ma["s"], 3),
            "adjoint_iterations": {"mixed": ma["adjoint_iters"], "fp64": fa["adjoint_iters"]},
            "adjoint_stop_reason": {"mixed": ma["adjoint_stop"], "fp64": fa["adjoint_stop"]},
            "what": "factorize + forward solve + adjoint solve A^T lam = x - x_t + dL/dA on the pattern (median)"}
    if multi is not None:
        line["multi_rhs"] = multi
    if mas_runs is not None:
        line["mas_solver"] = mas_runs
    if amg_runs is not None:
        line["amg"] = amg_runs
    print(json.dumps(line))


if __name__ == "__main__":
    main()
