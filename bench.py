"""Benchmark: wall-clock to a converged temperature field on one MI355X.

A step is one converged source iteration (one pass of the hot path over one synthetic problem):
StaticBTESolver::solve on a seeded 2D Poisson-disk/Delaunay mesh of a 1e-7 m square (~20k
triangles), 10 seeded bands, DM=2, t=8, p=8 (32 directions), left wall 301 K, right wall 299 K,
top/bottom diffuse, stop at margin <= 1e-5.  All 320 band x direction systems of each outer
iteration run in one batched launch.  Inputs are generated and uploaded before the timed region.

Prints one JSON line (see the driver contract) with a roofline object for the dominant kernel (the
batched inner solve, timed with hipEvents on the library's stream) and a cpu_baseline measured on a
bounded sample with the CPU oracle.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "wall-clock to converged T field, learned+batched vs plain FP64 @≤1e-4 rel err"
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # name: (kind, size, n_band, DM, ntheta, nphi, L, seed)
    "2d-20k": ("square", 20000, 10, 2, 8, 8, 1e-7, 2),
    "2d-ref12k": ("file", None, None, 2, 8, 8, 1e-8, None),
    "1d-100": ("1d", 100, 3, 2, 8, 1, 1e-8, 1),
    "3d-200k": ("cube", 32, 20, 3, 4, 4, 1e-7, 3),
}


def make_case(name, workdir):
    from staticbtesolver_amd import capi, meshgen

    kind, size, nb, DM, nt, npp, L, seed = CONFIGS[name]
    if kind == "square":
        p = meshgen.make_square_case(workdir, size, seed, nb, seed, L=L)
        kw = dict(geometry=p["mesh"], band_file=p["band"], bc_file=p["bc"], Lx=L, Ly=L)
    elif kind == "cube":
        p = meshgen.make_cube_case(workdir, size, seed, nb, seed, L=L)
        kw = dict(geometry=p["mesh"], band_file=p["band"], bc_file=p["bc"], Lx=L, Ly=L, Lz=L)
    elif kind == "1d":
        band = os.path.join(workdir, "band1d.dat")
        meshgen.write_bands(band, meshgen.random_bands(nb, seed))
        bc = os.path.join(workdir, "bc1d.dat")
        meshgen.write_bcs(bc, [(0, 1, 301.0), (1, 1, 299.0)])
        kw = dict(geometry=str(size), band_file=band, bc_file=bc, Lx=L)
    else:
        d = os.path.join(ROOT, "tests", "golden", "ref", "2Dcase2")
        kw = dict(geometry=os.path.join(d, "2D_12000.mphtxt"), band_file=os.path.join(d, "Input-dispersion-relation-fp.dat"),
                  bc_file=os.path.join(ROOT, "tests", "golden", "2Dcase2_bc31.dat"), Lx=L, Ly=L)
    kw.update(DM=DM, ntheta=nt, nphi=npp, T_ref=300.0)
    return kw, capi.Case(**kw)


def spmv_bytes(case, w=8):
    """Survey-canonical bytes of one CSR product of every system: nnz*(w+4) + 4(n+1) + 2wn, with
    nnz = n + incoming interior faces of the system's direction (band independent)."""
    geo = case.geometry()
    _, dirv, _ = case.quadrature()
    n = case.N_cell
    dn = np.einsum("cfk,dk->dcf", geo["normal"], dirv)
    inc = ((dn < 0) & (geo["nbr"][None] >= 0)).sum(axis=(1, 2))  # per direction
    nnz = n + inc
    per_dir = nnz * (w + 4) + 4 * (n + 1) + 2 * w * n
    return float(per_dir.sum() * case.N_band)


def cpu_baseline(kw, gpu_iters, budget_s=15.0):
    """PETSc-default restatement (GMRES(30) + ILU(0), rtol 1e-9) timed on a bounded sample: the
    first outer iteration's systems, taken in order until budget_s of CPU time, then extrapolated
    to all systems and to the GPU's outer-iteration count."""
    from oracle import oracle as O

    o = O.OracleCase(**kw)
    o.set_inner(O.INNER_PETSC)
    S = o.N_band * o.N_dir
    done, spent = 0, 0.0
    while done < S and spent < budget_s:
        spent += o.time_systems(done, done + 1)
        done += 1
    per_sys = spent / done
    est = per_sys * S * gpu_iters
    return dict(value=est, unit="s (extrapolated)", cores=1, kind="port",
                sample=f"first outer iteration, {done}/{S} systems in {spent:.1f}s on 1 core; "
                       f"x{S} systems x{gpu_iters} outer iterations")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="2d-20k", choices=sorted(CONFIGS))
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-run", action="store_true", help="one untimed solve without graphs (for rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only"}))
        sys.exit(2)
    from staticbtesolver_amd import capi

    work = tempfile.mkdtemp(prefix="bte_bench_")
    t0 = time.time()
    kw, case = make_case(args.config, work)
    print(f"# case {args.config}: N_cell={case.N_cell} N_dir={case.N_dir} N_band={case.N_band} "
          f"(inputs {time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    prec = capi.PREC_MIXED if args.precision == "mixed" else capi.PREC_FP64
    if args.profile_run:
        r = case.solve(precision=prec, use_graph=0, max_iter=10000)
        print(f"# profile run: {r['iters']} outer iterations, {r['solve_seconds']:.3f}s", file=sys.stderr)
        return
    for i in range(args.warmup):
        r = case.solve(precision=prec)
        print(f"# warmup {i}: {r['iters']} it {r['solve_seconds']:.3f}s setup {r['setup_seconds']:.3f}s",
              file=sys.stderr, flush=True)
    times, iters = [], []
    for i in range(args.steps):
        r = case.solve(precision=prec)
        times.append(r["solve_seconds"])
        iters.append(r["iters"])
        print(f"# step {i}: {r['iters']} it {r['solve_seconds']:.4f}s", file=sys.stderr, flush=True)
    T_run = r["T"].copy()
    # roofline of the dominant kernel: a separate solve with per-kernel hipEvents on the library stream
    rt = case.solve(precision=prec, timing=1)
    st = case.stats()
    solve_ms = st["ms"]["solve"] / max(1, st["calls"]["solve"])
    # one product + one sweep per system per launch when the sweep is exact (inner_iters == 1 each)
    per_launch_sys_iters = rt["inner_iters"] / max(1, rt["iters"]) / (case.N_band * case.N_dir)
    b_alg = 2.0 * spmv_bytes(case, 8 if prec == capi.PREC_FP64 else 4) * max(1.0, per_launch_sys_iters)
    achieved = b_alg / (solve_ms * 1e-3) / 1e9
    # accuracy of this run against the plain FP64 form
    if prec != capi.PREC_FP64:
        ref = case.solve(precision=capi.PREC_FP64)
        rel = float(np.max(np.abs(T_run - ref["T"]) / np.abs(ref["T"])))
    else:
        rel = 0.0
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(kw, int(statistics.median(iters)))
    t_med = statistics.median(times)
    line = {
        "metric": METRIC,
        "value": round(t_med, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_med, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == capi.PREC_FP64 else "f32+f64",
        "data": "synthetic (seeded mesh and bands)",
        "config": {"workload": args.config, "n_cell": case.N_cell, "n_dir": case.N_dir, "n_band": case.N_band,
                   "systems_per_launch": case.N_band * case.N_dir, "form": args.precision + "+sweep",
                   "outer_iterations": int(statistics.median(iters)),
                   "ms_per_outer_iteration": round(1e3 * t_med / statistics.median(iters), 4),
                   "sweep_levels_max": st["max_levels"], "lagged_edges": st["lagged_edges"],
                   "block": st["block"], "sweep_variant": st["variant"],
                   "max_rel_err_T_vs_fp64": rel},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "kernel": "k_solve",
                     "kernel_ms": round(solve_ms, 5), "alg_bytes_per_launch": b_alg},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
