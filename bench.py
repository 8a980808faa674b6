"""Benchmark: wall-clock to a converged temperature field on one MI355X, learned + batched form
against the plain FP64 form.

Default workload (near-diffusive 2D): seeded Poisson-disk/Delaunay triangle mesh of a 1e-5 m square
(~50k cells), 10 seeded bands (v ~ U[1e3, 6e3] m/s, tau ~ logU[1e-12, 1e-10] s, C ~ U[1e5, 1e6]),
DM=2, t=8, p=8 (32 directions), left wall 301 K, right wall 299 K, top/bottom diffuse at 300 K.
A step is one solve of the form, outer iteration by outer iteration, all 320 band x direction
systems of each outer iteration in one batched launch, each outer iteration one hipGraph; the library
calls are timed (bte_begin's per-solve set-up included) and inputs are resident before the timed
region.  value = the median time
until the temperature field is within 1e-4 relative (max over cells) of a tightly converged FP64
field (Anderson to margin 1e-12); the check runs outside the timed calls.  Each step continues to
the reference's stop rule (margin <= 1e-5, StaticBTESolver.cpp:798), whose time and error are
reported too.

Forms: learned (Anderson mixing + residual-judged network corrections), anderson (mixing only),
plain (the reference's source iteration).  The same run measures the plain FP64 form once, both to
the 1e-4 target and to the stop rule.  --config uneven-64 runs the 64-problem uneven batch instead.

Prints one JSON line with a roofline object for the dominant kernel (the batched inner solve, timed
with hipEvents on the library's stream in a separate solve) and a cpu_baseline measured on a bounded
sample with the CPU oracle's PETSc-default restatement, extrapolated.
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
    "2d-nd-50k": ("square", 50000, 10, 2, 8, 8, 1e-5, 4),
    "2d-20k": ("square", 20000, 10, 2, 8, 8, 1e-7, 2),
    "2d-ref12k": ("file", None, None, 2, 8, 8, 1e-8, None),
    "1d-100": ("1d", 100, 3, 2, 8, 1, 1e-8, 1),
    # 3D: a seeded jittered Kuhn lattice of 32^3 cubes (196,608 tetrahedra).  Delaunay tetrahedra of
    # seeded points (meshgen.cube_delaunay_mesh) hold slivers on which the reference's source
    # iteration diverges (the oracle's too): there is no plain-form baseline on them.
    "3d-200k": ("cube", 32, 20, 3, 4, 4, 1e-7, 3),
}


def say(*a):
    print("#", *a, file=sys.stderr, flush=True)


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


UNEVEN = dict(n=64, seed=5, cells=(2e3, 6e4), L=(1e-8, 1e-5), bands=(5, 10), dT=(1.0, 5.0))


def _make_uneven_one(spec):
    from staticbtesolver_amd import meshgen

    d, cells, seed, nb, L, th, tc = spec
    p = meshgen.make_square_case(d, cells, seed, nb, seed, L=L,
                                 bc_rows=[(0, 1, th), (1, 2, 300.0), (2, 2, 300.0), (3, 1, tc)])
    p.update(L=L)
    return p


def make_uneven(workdir):
    """64 seeded squares: cells logU[2k, 60k], side logU[1e-8, 1e-5] m, 5-10 seeded bands, isothermal
    walls at 300 + U[1, 5] K and 300 - U[1, 5] K, diffuse walls between.  Meshes are generated by a
    process pool forked before this process touches the GPU."""
    import multiprocessing as mp

    u = UNEVEN
    rng = np.random.default_rng(u["seed"])
    specs = []
    for k in range(u["n"]):
        cells = int(round(10 ** rng.uniform(np.log10(u["cells"][0]), np.log10(u["cells"][1]))))
        L = float(10 ** rng.uniform(np.log10(u["L"][0]), np.log10(u["L"][1])))
        nb = int(rng.integers(u["bands"][0], u["bands"][1] + 1))
        th, tc = 300.0 + rng.uniform(*u["dT"]), 300.0 - rng.uniform(*u["dT"])
        specs.append((os.path.join(workdir, f"u{k}"), cells, 5000 + k, nb, L, th, tc))
    with mp.get_context("fork").Pool(min(16, os.cpu_count() or 1)) as pool:
        return pool.map(_make_uneven_one, specs)


def bench_uneven(args):
    """Uneven batch: all problems in one bte_batch_solve versus one solve per problem."""
    from staticbtesolver_amd import capi

    work = tempfile.mkdtemp(prefix="bte_bench_")
    t0 = time.time()
    paths = make_uneven(work)
    cases = [capi.Case(p["mesh"], p["band"], p["bc"], 2, 8, 8, 300.0, Lx=p["L"], Ly=p["L"]) for p in paths]
    n_cell = sum(c.N_cell for c in cases)
    n_sys = sum(c.N_band * c.N_dir for c in cases)
    say(f"uneven batch: {len(cases)} problems, {n_cell} cells, {n_sys} systems (inputs {time.time() - t0:.1f}s)")
    form = dict(max_iter=100000, accel=capi.ACCEL_ANDERSON, accel_memory=args.accel_memory)
    net = None
    if args.form == "learned":
        from staticbtesolver_amd.gnn import Gnn
        from staticbtesolver_amd.train_gnn import DEFAULT_WEIGHTS
        net = Gnn(layers=4, seed=0, device=0)
        net.set_params(np.load(args.weights or DEFAULT_WEIGHTS))
        for c in cases:
            c.attach_gnn(net)
        form.update(learned=1)
    elif args.form == "plain":
        form.update(accel=capi.ACCEL_NONE)
    batch = capi.Batch(cases)
    if args.profile_run:
        out = batch.solve(**form)
        say(f"profile run: outer iterations {min(r['iters'] for r in out)}..{max(r['iters'] for r in out)}")
        return
    for i in range(args.warmup):
        t = time.perf_counter()
        out = batch.solve(**form)
        say(f"warmup {i}: {time.perf_counter() - t:.3f}s, outer iterations {min(r['iters'] for r in out)}.."
            f"{max(r['iters'] for r in out)}")
    times = []
    for i in range(args.steps):
        t = time.perf_counter()
        out = batch.solve(**form)
        times.append(time.perf_counter() - t)
        say(f"step {i}: {times[-1]:.4f}s")
    # the same problems one at a time (same form), and agreement with the batch
    t_seq, worst = 0.0, 0.0
    for c, r in zip(cases, out):
        t = time.perf_counter()
        s = c.solve(**form)
        t_seq += time.perf_counter() - t
        worst = max(worst, float(np.max(np.abs(s["T"] - r["T"]) / np.abs(s["T"]))))
    t_med = statistics.median(times)
    say(f"one at a time: {t_seq:.3f}s; batch {t_med:.3f}s")
    # roofline of the batched sweep: a timed solve (hipEvents around each launch); member k sweeps
    # its B_shared bytes (each direction's operator once, right-hand sides in, intensities out) in
    # each of its outer iterations
    if net is not None:
        for c in cases:
            c.attach_gnn(net)
    timed = batch.solve(timing=1, **form)
    st = batch.stats()
    sweep_ms = st["ms"]["solve"] / max(1, st["calls"]["solve"])
    member_bytes = [alg_bytes(c, 8, True)[0] for c in cases]
    total_bytes = sum(b * r["iters"] for b, r in zip(member_bytes, timed))
    achieved = total_bytes / (st["ms"]["solve"] * 1e-3) / 1e9
    prof = profiled_traffic("uneven-64", args.form, "fp64")
    # mixed precision on three members alone (FP32 BiCGSTAB + FP64 refinement, the sweep as the
    # preconditioner) against the FP64 direct sweep: outer iteration time
    by_size = sorted(range(len(cases)), key=lambda k: cases[k].N_cell)
    mixed = []
    for k in (by_size[0], by_size[len(by_size) // 2], by_size[-1]):
        c = cases[k]
        c.attach_gnn(None)
        f64 = c.solve(max_iter=4, stop_margin=0.0)
        mx = c.solve(max_iter=4, stop_margin=0.0, precision=capi.PREC_MIXED)
        mixed.append(dict(n_cell=c.N_cell, fp64_ms_per_iteration=round(1e3 * f64["solve_seconds"] / 4, 3),
                          mixed_ms_per_iteration=round(1e3 * mx["solve_seconds"] / 4, 3)))
        say(f"member {c.N_cell} cells: fp64 {mixed[-1]['fp64_ms_per_iteration']} ms, mixed "
            f"{mixed[-1]['mixed_ms_per_iteration']} ms per outer iteration")
    cpu = None if args.no_cpu else cpu_baseline_uneven(paths, cases, [r["iters"] for r in timed])
    line = {
        "metric": METRIC, "value": round(t_med, 6), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * t_med, 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded meshes and bands)",
        "config": {"workload": "uneven-64", "problems": len(cases), "n_cell_total": n_cell, "systems": n_sys,
                   "form": args.form, "precision": "fp64"},
        "comparison": {"one_at_a_time_s": round(t_seq, 4), "speedup_vs_one_at_a_time": round(t_seq / t_med, 2),
                       "max_rel_diff_T_vs_one_at_a_time": worst,
                       "outer_iterations": [r["iters"] for r in out],
                       "learned_accepted": sum(r["learned_accepted"] for r in out),
                       "learned_proposed": sum(r["learned_proposed"] for r in out),
                       "mixed_precision_members": mixed},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": prof["bytes"] if prof else None,
                     "kernel": "k_sweep_batch", "kernel_ms": round(sweep_ms, 5),
                     "rhs_ms": round(st["ms"]["rhs"] / max(1, st["calls"]["rhs"]), 5),
                     "alg_bytes_per_launch": total_bytes / max(1, st["calls"]["solve"]),
                     "alg_bytes_formula": "sum over the problems still iterating of B_shared",
                     "traffic_source": prof},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline_uneven(paths, cases, iters, budget_s=15.0):
    """The oracle's PETSc-default restatement on the uneven batch: the first outer iteration's
    systems of the smallest, the median and the largest problem timed on all usable cores (at most
    16, directions round-robin), within budget_s, give a CPU time per system and cell; extrapolated
    to every problem's systems and cells times the outer iterations the batch form took for it (the
    oracle's iteration is the plain one; the mixing steps are not counted)."""
    from oracle import oracle as O

    model, n_online, n_usable = cpu_model()
    cap = max(1, min(16, n_usable or 1))
    order = sorted(range(len(cases)), key=lambda k: cases[k].N_cell)
    pick = [order[0], order[len(order) // 2], order[-1]]
    spent, cell_systems = 0.0, 0
    W = 1
    for k in pick:
        p, c = paths[k], cases[k]
        o = O.OracleCase(p["mesh"], p["band"], p["bc"], 2, 8, 8, 300.0, Lx=p["L"], Ly=p["L"])
        o.set_inner(O.INNER_PETSC)
        D = o.N_dir
        W = max(w for w in range(1, D + 1) if D % w == 0 and w <= cap)
        o.set_threads(W)
        spent += o.time_systems(0, D)  # every direction of the first band
        cell_systems += D * o.N_cell
        if spent > budget_s:
            break
    per_cell_system = spent / cell_systems
    total = sum(per_cell_system * c.N_band * c.N_dir * c.N_cell * it for c, it in zip(cases, iters))
    return dict(value=total, unit="s (extrapolated)", cores=W, kind="port",
                sample=f"first outer iteration, band 0 of {len(pick)} problems ({cell_systems} system-cells) in "
                       f"{spent:.2f}s on {W} worker threads (GMRES(30)+ILU(0) restatement); x every problem's "
                       f"systems and cells x its outer iterations in the batch form",
                cpu_model=model, nproc=n_online, cpus_usable=n_usable)


def spmv_bytes(case, w=8):
    """Canonical bytes of one CSR product of every system: nnz*(w+4) + 4(n+1) + 2wn, with nnz = n +
    incoming interior faces of the system's direction (band independent)."""
    geo = case.geometry()
    _, dirv, _ = case.quadrature()
    n = case.N_cell
    dn = np.einsum("cfk,dk->dcf", geo["normal"], dirv)
    inc = ((dn < 0) & (geo["nbr"][None] >= 0)).sum(axis=(1, 2))  # per direction
    nnz = n + inc
    per_dir = nnz * (w + 4) + 4 * (n + 1) + 2 * w * n
    return float(per_dir.sum() * case.N_band)


def alg_bytes(case, w, direct, per_sys=1.0, jacobi=False):
    """Algorithmic bytes of one inner-solve launch (all systems), and the per-system canonical count.

    The stored operator of a direction serves all its bands (the band enters as the scalar
    v_b tau_b), so a product of all systems needs each direction's operator once:
    B_shared = sum_d [nnz_d (w+4) + 4 (n+1)] + sum_(b,d) 2 w n.  An exact sweep moves one product's
    worth (operator, right-hand side in, intensities out): B_shared.  A BiCGSTAB iteration: 2 B_shared
    + 21 w n per system (the fused vector passes' lower bound) + the Jacobi preconditioner's per-band
    diagonal reads 2 (w + 8) n, times the mean iterations per system.  The canonical count charges
    every system its own operator (B_spmv per system)."""
    n, S = case.N_cell, case.N_band * case.N_dir
    canon = spmv_bytes(case, w)  # sum over systems of B_spmv
    vec = 2.0 * w * n * S
    op_dir = canon / case.N_band - 2.0 * w * n * case.N_dir  # every direction's operator once
    shared = op_dir + vec
    if direct:
        return shared, canon
    jac = 2.0 * (w + 8) * n * S if jacobi else 0.0
    return ((2.0 * shared + 21.0 * w * n * S + jac) * per_sys, (2.0 * canon + 21.0 * w * n * S) * per_sys)


def cpu_model():
    """(model name, logical CPUs online, CPUs this process may use) of the host."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return model, os.cpu_count(), usable


def cpu_baseline(kw, outer_iters, budget_s=15.0):
    """The oracle's PETSc-default restatement (GMRES(30) + ILU(0), rtol 1e-9, one system at a time
    like the reference's _solve_matrix) timed on bounded samples of the first outer iteration and
    extrapolated to all systems and to the plain form's outer-iteration count.

    - 1 core: systems in order until budget_s of CPU time.
    - all cores: W workers, W the largest divisor of N_dir not above the usable CPUs (at most 16),
      each taking the directions d = w, w + W, ... of every band, as the reference splits
      directions over MPI ranks (StaticBTESolver.cpp:725); sampled over whole bands (every
      direction of the first k bands) to about budget_s of wall time."""
    from oracle import oracle as O

    o = O.OracleCase(**kw)
    o.set_inner(O.INNER_PETSC)
    D, B = o.N_dir, o.N_band
    S = B * D
    o.set_threads(1)
    done, spent = 0, 0.0
    while done < S and spent < budget_s:
        spent += o.time_systems(done, done + 1)
        done += 1
    per_sys = spent / done
    one = dict(value=per_sys * S * outer_iters, unit="s (extrapolated)", cores=1,
               sample=f"first outer iteration, {done}/{S} systems in {spent:.2f}s")
    model, n_online, n_usable = cpu_model()
    cap = max(1, min(16, n_usable or 1))
    W = max(w for w in range(1, D + 1) if D % w == 0 and w <= cap)
    o.set_threads(W)
    k = int(max(1, min(B, budget_s * W / max(per_sys * D, 1e-9))))
    wall = o.time_systems(0, k * D)
    per_sys_all = wall / (k * D)
    return dict(value=per_sys_all * S * outer_iters, unit="s (extrapolated)", cores=W, kind="port",
                sample=f"first outer iteration, {k * D}/{S} systems (bands 0..{k - 1}, all directions) in "
                       f"{wall:.1f}s on {W} worker threads, directions round-robin (GMRES(30)+ILU(0) "
                       f"restatement); x{S} systems x{outer_iters} outer iterations of the plain form",
                cpu_model=model, nproc=n_online, cpus_usable=n_usable, one_core=one)


def roofline_only(case, form, k, prec):
    """K outer iterations with per-kernel hipEvents; roofline of the inner solve (all kernels of
    the solve step: right-hand sides, inner solves and, in the mixed form, FP64 re-solves), with
    the algorithmic bytes of alg_bytes (FP64 re-solve iterations of the mixed form counted at
    w = 4: a lower bound).  achieved_canonical charges every system its own operator; with the
    operator shared through the caches it can exceed the HBM traffic actually moved."""
    from staticbtesolver_amd import capi

    f = dict(form, max_iter=k, stop_margin=0.0, timing=1)
    r = case.solve(**f)
    st = case.stats()
    launches = max(1, st["calls"]["solve"])
    solve_ms = st["ms"]["solve"] / launches
    rhs_ms = st["ms"]["rhs"] / max(1, st["calls"]["rhs"])
    S = case.N_band * case.N_dir
    per_sys = r["inner_iters"] / launches / S
    w = 8 if prec == capi.PREC_FP64 else 4
    direct = st["variant"] > 0
    G = max(1, st.get("band_group", 1))
    b_alg, b_canon = alg_bytes(case, w, direct, per_sys, form.get("precond") == capi.PC_JACOBI)
    achieved = b_alg / (solve_ms * 1e-3) / 1e9
    return {"roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 5), "kernel_ms": round(solve_ms, 4),
                         "rhs_ms": round(rhs_ms, 4),
                         "alg_bytes_per_launch": b_alg,
                         "achieved_canonical": round(b_canon / (solve_ms * 1e-3) / 1e9, 2),
                         "inner_iterations_per_system": round(per_sys, 2),
                         "inner_limit_hits": r["inner_failures"],
                         "form": "direct sweep" if direct else "BiCGSTAB"},
            "config": {"n_cell": case.N_cell, "systems": S, "outer_iterations": r["iters"],
                       "sweep_variant": st["variant"], "block": st["block"], "band_group": G}}


def profiled_traffic(workload, form, precision):
    """HBM bytes per dispatch of the dominant kernel from the committed rocprofv3 summary of this
    workload and form (profiles/<kernel>_<workload>_<form>_<precision>.json, written by
    staticbtesolver_amd/prof_summary.py from FETCH_SIZE / WRITE_SIZE passes scaled by a copy kernel
    of known size), or None."""
    import glob

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", f"*_{workload}_{form}_{precision}.json"))):
        with open(path) as f:
            res = json.load(f)
        if res.get("traffic_bytes"):
            return dict(bytes=res["traffic_bytes"], kernel=res.get("dominant"), avg_us=res.get("avg_us"),
                        source=os.path.relpath(path, ROOT))
    return None


def errors(T, ref, T_ref=300.0):
    dT = np.abs(T - ref)
    return dict(max_rel=float(np.max(dT / np.abs(ref))), max_dT_rel=float(np.max(dT) / np.max(np.abs(ref - T_ref))))


def track(case, form, T_ref_field, tol, coarse_after=None):
    """Runs one solve of `form` outer iteration by outer iteration to its stop rule, timing the
    library's work: bte_begin (per-solve set-up: state reset, mixing history, the network's static
    node features, graph capture when the form changed) and the bte_step calls.  Returns the time
    and iterations until the temperature field first meets max |T - T*| / |T*| <= tol (the check
    runs outside the timed calls) with the error there, and the same until the stop rule.  With
    coarse_after, the check runs every 10 iterations after that many (the reported time to target
    is then at most 9 iterations late)."""
    from staticbtesolver_amd import capi

    t = time.perf_counter()
    case.begin(capi.default_opts(**form))
    t_begin = time.perf_counter() - t
    t_solve, k, conv = t_begin, 0, False
    tta, tta_k, err_tta = None, None, None
    while not conv:
        n = 1 if (tta is not None or coarse_after is None or k < coarse_after) else 10
        if tta is not None:
            n = 1000
        t = time.perf_counter()
        done, conv = case.step(n)
        t_solve += time.perf_counter() - t
        k += done
        if done == 0:
            break
        if tta is None:
            T = case.temperature()
            if np.max(np.abs(T - T_ref_field) / np.abs(T_ref_field)) <= tol:
                tta, tta_k, err_tta = t_solve, k, errors(T, T_ref_field)
    r = case.finish()
    return dict(tta_s=tta, tta_iters=tta_k, err_tta=err_tta, stop_s=t_solve, begin_s=t_begin, iters=r["iters"],
                err_stop=errors(r["T"], T_ref_field), accepted=r["learned_accepted"],
                proposed=r["learned_proposed"], converged=r["converged"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="2d-nd-50k", choices=sorted(CONFIGS) + ["uneven-64"])
    ap.add_argument("--form", default="learned", choices=["learned", "anderson", "plain"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--tol", type=float, default=1e-4, help="target max relative error of T")
    ap.add_argument("--accel-memory", type=int, default=50)
    ap.add_argument("--weights", default=None, help="network parameters (.npy); default: the packaged weights")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 comparison solve")
    ap.add_argument("--profile-run", action="store_true", help="one untimed solve of the form without graphs")
    ap.add_argument("--precond", default="sweep", choices=["sweep", "jacobi"],
                    help="jacobi: the reference GPU solver's right-Jacobi BiCGSTAB (Krylov inner solves)")
    ap.add_argument("--roofline-only", type=int, default=0, metavar="K",
                    help="only time K outer iterations of the form's inner solve and print its roofline")
    ap.add_argument("--block", type=int, default=0, choices=[0, 64, 128, 256, 512, 1024],
                    help="work-group size of the inner solve (0: autotuned)")
    ap.add_argument("--variant", type=int, default=0, choices=[0, 1, 2, 3, 4, 5, 6, 7, 8],
                    help="sweep variant of the direct form (0: autotuned)")
    ap.add_argument("--band-group", type=int, default=0, choices=[0, 1, 2, 4],
                    help="jacobi: bands per work-group of the batched BiCGSTAB (0: autotuned)")
    ap.add_argument("--tol32", type=float, default=0.0, help="mixed: FP32 inner tolerance (0: library default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU (the box has one); multi-GPU sharding is "
                                   "design-only -- run with --gpus 1"}))
        sys.exit(2)
    if args.config == "uneven-64":
        return bench_uneven(args)
    from staticbtesolver_amd import capi

    work = tempfile.mkdtemp(prefix="bte_bench_")
    t0 = time.time()
    kw, case = make_case(args.config, work)
    say(f"case {args.config}: N_cell={case.N_cell} N_dir={case.N_dir} N_band={case.N_band} "
        f"(inputs {time.time() - t0:.1f}s)")
    prec = capi.PREC_MIXED if args.precision == "mixed" else capi.PREC_FP64
    form = dict(precision=prec, max_iter=100000, precond=capi.PC_JACOBI if args.precond == "jacobi" else capi.PC_SWEEP)
    if args.block:
        form.update(block=args.block)
    if args.variant:
        form.update(sweep_variant=args.variant)
    if args.band_group:
        form.update(band_group=args.band_group)
    if args.tol32:
        form.update(inner_tol32=args.tol32)
    net = None
    if args.form in ("learned", "anderson"):
        form.update(accel=capi.ACCEL_ANDERSON, accel_memory=args.accel_memory)
    if args.form == "learned":
        from staticbtesolver_amd.gnn import Gnn
        from staticbtesolver_amd.train_gnn import DEFAULT_WEIGHTS
        wpath = args.weights or DEFAULT_WEIGHTS
        net = Gnn(layers=4, seed=0, device=0)
        net.set_params(np.load(wpath))
        case.attach_gnn(net)
        form.update(learned=1)
    if args.roofline_only:
        print(json.dumps(roofline_only(case, form, args.roofline_only, prec)), flush=True)
        return
    if args.profile_run:
        r = case.solve(use_graph=0, **form)
        say(f"profile run: {r['iters']} outer iterations, {r['solve_seconds']:.3f}s")
        return
    # the converged field every accuracy figure refers to: FP64, Anderson to margin 1e-12
    ref = case.solve(accel=capi.ACCEL_ANDERSON, stop_margin=1e-12, max_iter=100000)
    for i in range(args.warmup):
        r = track(case, form, ref["T"], args.tol)
        say(f"warmup {i}: to {args.tol:g}: {r['tta_s']:.4f}s ({r['tta_iters']} it); to the stop rule: "
            f"{r['stop_s']:.4f}s ({r['iters']} it)")
    runs = []
    for i in range(args.steps):
        r = track(case, form, ref["T"], args.tol)
        runs.append(r)
        say(f"step {i}: to {args.tol:g}: {r['tta_s']:.4f}s ({r['tta_iters']} it); to the stop rule: {r['stop_s']:.4f}s "
            f"({r['iters']} it, err {r['err_stop']['max_rel']:.1e}); accepted {r['accepted']}/{r['proposed']}")
    tta = statistics.median(r["tta_s"] for r in runs)
    t_stop = statistics.median(r["stop_s"] for r in runs)
    last = runs[-1]
    # roofline of the dominant kernel (the inner solve: one launch sweeps all systems): a separate
    # solve with hipEvents on the library stream around that kernel alone, and around the
    # right-hand-side pass before it
    rt = case.solve(timing=1, **form)
    st = case.stats()
    solve_ms = st["ms"]["solve"] / max(1, st["calls"]["solve"])
    rhs_ms = st["ms"]["rhs"] / max(1, st["calls"]["rhs"])
    # per launch: one exact sweep per system (FP64 direct form) or the refinement loop's products
    per_sys = rt["inner_iters"] / max(1, st["calls"]["solve"]) / (case.N_band * case.N_dir)
    w = 8 if prec == capi.PREC_FP64 else 4
    b_alg, b_canon = alg_bytes(case, w, st["variant"] > 0, per_sys, form.get("precond") == capi.PC_JACOBI)
    achieved = b_alg / (solve_ms * 1e-3) / 1e9
    cmp = dict(form=args.form, precision=args.precision, target_rel_err=args.tol,
               time_to_target_s=round(tta, 5), iterations_to_target=last["tta_iters"],
               err_at_target=last["err_tta"], begin_s=round(statistics.median(r["begin_s"] for r in runs), 6),
               time_to_stop_rule_s=round(t_stop, 5), iterations_to_stop_rule=last["iters"],
               err_at_stop_rule=last["err_stop"], learned_accepted=last["accepted"],
               learned_proposed=last["proposed"], reference_field="FP64, Anderson mixing to margin 1e-12")
    if args.form == "learned":
        # the same mixing without the network: what the learned corrections add
        case.attach_gnn(None)
        an = track(case, {k: v for k, v in form.items() if k != "learned"}, ref["T"], args.tol)
        case.attach_gnn(net)
        cmp.update(anderson_fp64_time_to_target_s=round(an["tta_s"], 5) if an["tta_s"] else None,
                   anderson_fp64_iterations_to_target=an["tta_iters"],
                   anderson_fp64_time_to_stop_rule_s=round(an["stop_s"], 5),
                   anderson_fp64_iterations_to_stop_rule=an["iters"])
        say(f"anderson only: to {args.tol:g}: {an['tta_s']:.4f}s ({an['tta_iters']} it); to the stop rule "
            f"{an['stop_s']:.4f}s ({an['iters']} it)")
    plain_iters = None
    if not args.no_plain:
        if net is not None:
            case.attach_gnn(None)
        pl = track(case, dict(max_iter=100000), ref["T"], args.tol, coarse_after=100)
        plain_iters = pl["iters"]
        cmp.update(plain_fp64_time_to_target_s=round(pl["tta_s"], 4), plain_fp64_iterations_to_target=pl["tta_iters"],
                   plain_fp64_time_to_stop_rule_s=round(pl["stop_s"], 4), plain_fp64_iterations_to_stop_rule=pl["iters"],
                   plain_fp64_err_at_stop_rule=pl["err_stop"],
                   speedup_to_target_vs_plain_fp64=round(pl["tta_s"] / tta, 1) if pl["tta_s"] else None,
                   speedup_to_stop_rule_vs_plain_fp64=round(pl["stop_s"] / t_stop, 2))
        say(f"plain fp64: to {args.tol:g}: {pl['tta_s']:.3f}s ({pl['tta_iters']} it); to the stop rule "
            f"{pl['stop_s']:.3f}s ({pl['iters']} it, err {pl['err_stop']['max_rel']:.1e})")
    if net is not None:
        case.attach_gnn(None)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(kw, plain_iters or last["iters"])
    prof = profiled_traffic(args.config, args.form, args.precision)
    line = {
        "metric": METRIC,
        "value": round(tta, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * tta, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == capi.PREC_FP64 else "f32+f64",
        "data": "synthetic (seeded mesh and bands)",
        "config": {"workload": args.config, "n_cell": case.N_cell, "n_dir": case.N_dir, "n_band": case.N_band,
                   "systems_per_launch": case.N_band * case.N_dir, "form": args.form, "precision": args.precision,
                   "accel_memory": form.get("accel_memory"),
                   "ms_per_outer_iteration": round(1e3 * t_stop / max(1, last["iters"]), 4),
                   "sweep_levels_max": st["max_levels"], "lagged_edges": st["lagged_edges"],
                   "block": st["block"], "sweep_variant": st["variant"]},
        "comparison": cmp,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": prof["bytes"] if prof else None,
                     "kernel": {4: "k_sweep_ring", 5: "k_sweep_wave", 6: "k_sweep_ring_bands",
                                7: "k_sweep_ring_bands", 8: "k_sweep_ring"}.get(st["variant"], "inner solve"),
                     "kernel_ms": round(solve_ms, 5), "rhs_ms": round(rhs_ms, 5), "alg_bytes_per_launch": b_alg,
                     "alg_bytes_formula": "B_shared: each direction's operator once + 2 w n per system",
                     "achieved_canonical": round(b_canon / (solve_ms * 1e-3) / 1e9, 2),
                     "traffic_source": prof},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
