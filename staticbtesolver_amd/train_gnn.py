"""Training of the learned outer-iteration correction on pairs the solver generates itself.

Family: seeded Poisson-disk/Delaunay triangle meshes of a square of side 1e-5 m (near-diffusive:
Knudsen numbers ~1e-4-0.1 for the seeded band sets), 10 seeded bands, DM=2, t=8, p=8, walls 0/3
isothermal at 301/299 K and 1/2 diffuse at 300 K.  For every training geometry the
Anderson-accelerated solver runs to a tight stop for the converged field T*; the network inputs
(bte_learn_features) are exported at states where the solver would propose a correction, from
Anderson-mixed and from plain source iterations; the target is (T* - x_next) / T_span.  A second
stage adds states reached after the stage-one network's own accepted corrections (the states the
solver meets when the network is in use) and trains on the union.  The network is trained on the
GPU by the library's own BF16-MFMA forward/backward kernels and Adam (bte_gnn_backward /
bte_gnn_adam).

    python -m staticbtesolver_amd.train_gnn --out weights.npy --report report.json

The evaluation part solves held-out geometries with and without the network (same settings) and
reports iterations, wall time and the error against a tightly converged solution.
"""
from __future__ import annotations

import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

from staticbtesolver_amd import meshgen  # noqa: E402

FAMILY = dict(L=1e-5, n_band=10, DM=2, ntheta=8, nphi=8, T_ref=300.0)
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "gnn_square_nd_L4.npy")
# export points (state steps): Anderson-mixed and plain iterations without the network, and the
# same with the stage-one network proposing (every 8 / every 16 steps from step 1)
REC_AA, REC_SI = (1, 9, 17), (1, 33, 129)
REC_AA_L, REC_SI_L = (5, 13), (9, 41)
EVERY_AA, EVERY_SI = 8, 16
STATIC0 = 6  # first static input (see k_learn_features)


def cells_for(seed: int, lo: int, hi: int) -> int:
    rng = np.random.default_rng(seed + 7919)
    return int(round(10 ** rng.uniform(math.log10(lo), math.log10(hi))))


def _make(args):
    directory, seed, n_cells = args
    p = meshgen.make_square_case(directory, n_cells, seed, FAMILY["n_band"], seed, L=FAMILY["L"], name=f"g{seed}")
    p["seed"] = seed
    return p


def open_case(p):
    from staticbtesolver_amd import capi
    L = FAMILY["L"]
    return capi.Case(p["mesh"], p["band"], p["bc"], FAMILY["DM"], FAMILY["ntheta"], FAMILY["nphi"], FAMILY["T_ref"],
                     Lx=L, Ly=L)


def solver_opts(**kw):
    from staticbtesolver_amd import capi
    base = dict(accel=capi.ACCEL_ANDERSON, accel_memory=20, max_iter=4000)
    base.update(kw)
    return capi.default_opts(**base)


def record(case, opts, steps):
    """Runs the iteration of `opts` up to the last of `steps`, exporting the network inputs after
    each; returns [(X, x_next, span)]."""
    case.begin(opts)
    recs, k = [], 0
    for s in steps:
        done, conv = case.step(s - k)
        k += done
        if conv:
            break
        if k == s:
            recs.append(case.learn_features())
    case.finish()
    return recs


def loss_weight(X):
    """Per-node loss weights: errors within a few mean free paths of an isothermal wall dominate the
    residual that judges a correction (the iteration damps smooth interior errors only slowly, so
    they hardly show in the residual), so those nodes weigh up to 10x more.  Static input 7 is
    log10(1 + distance in mean free paths) / 2."""
    dm = 10.0 ** (2.0 * X[..., STATIC0 + 7].astype(np.float64)) - 1.0
    return (1.0 + 9.0 * np.exp(-dm / 3.0)).astype(np.float32)


def _pairs(recs, T):
    return [(r[0], ((T - r[1]) / r[2]).astype(np.float32)) for r in recs]


def generate(seeds, lo, hi, stop, workdir, workers, log=print):
    """Stage-one data: list of dict(seed, files, nbr, T, samples=[(X, y)])."""
    from staticbtesolver_amd import capi
    jobs = [(os.path.join(workdir, f"g{s}"), s, cells_for(s, lo, hi)) for s in seeds]
    data = []
    t0 = time.time()
    ctx = mp.get_context("fork")
    with ctx.Pool(workers) as pool:
        for i, p in enumerate(pool.imap(_make, jobs)):
            case = open_case(p)
            ref = case.solve(solver_opts(stop_margin=stop))
            T = ref["T"]
            recs = record(case, solver_opts(), REC_AA)
            recs += record(case, solver_opts(accel=capi.ACCEL_NONE, learned=1, learned_every=0), REC_SI)
            nbr = case.geometry()["nbr"].astype(np.int32)
            case.close()
            data.append(dict(seed=p["seed"], files=p, nbr=nbr, T=T, samples=_pairs(recs, T), iters=ref["iters"]))
            log(f"[data] {i + 1}/{len(jobs)} seed {p['seed']} cells {len(nbr)} iters {ref['iters']} "
                f"|y| {[round(float(np.sqrt(np.mean(y ** 2))), 4) for _, y in data[-1]['samples']]} "
                f"({time.time() - t0:.0f} s)")
    return data


def augment(net, data, log=print):
    """Stage-two data: states reached with the network proposing corrections."""
    from staticbtesolver_amd import capi
    t0 = time.time()
    for i, d in enumerate(data):
        case = open_case(d["files"])
        case.attach_gnn(net)
        recs = record(case, solver_opts(learned=1, learned_every=EVERY_AA), REC_AA_L)
        recs += record(case, solver_opts(accel=capi.ACCEL_NONE, learned=1, learned_every=EVERY_SI), REC_SI_L)
        case.attach_gnn(None)
        case.close()
        d["samples"] += _pairs(recs, d["T"])
        if (i + 1) % 32 == 0:
            log(f"[augment] {i + 1}/{len(data)} ({time.time() - t0:.0f} s)")


def train(data, net=None, layers=4, epochs=30, lr=2e-3, lr_min=1e-4, seed=0, val=None, log=print):
    from staticbtesolver_amd.gnn import Gnn
    if net is None:
        net = Gnn(layers=layers, seed=seed, device=0)
    rng = np.random.default_rng(seed)
    n_steps = epochs * sum(len(d["samples"]) for d in data)
    step = 0
    hist = []
    t0 = time.time()
    for ep in range(epochs):
        tot, cnt = 0.0, 0
        for gi in rng.permutation(len(data)):
            d = data[gi]
            net.set_graph(d["nbr"])
            for si in rng.permutation(len(d["samples"])):
                X, y = d["samples"][si]
                a = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * step / max(1, n_steps)))
                loss = net.backward(X, y, loss_weight(X))
                net.adam(a, weight_decay=0.0)
                tot += loss
                cnt += 1
                step += 1
        msg = dict(epoch=ep, train=tot / max(cnt, 1))
        if val:
            msg["val_rel"] = relative_error(net, val)
        hist.append(msg)
        log(f"[train] epoch {ep} loss {msg['train']:.4e}" +
            (f" val rel err {msg['val_rel']:.3f}" if val else "") + f" ({time.time() - t0:.0f} s)")
    return net, hist


def relative_error(net, data):
    """Mean over samples of |y_pred - y| / |y| (1 = no better than no correction)."""
    r = []
    for d in data:
        net.set_graph(d["nbr"])
        for X, y in d["samples"]:
            yp = net.forward(X)
            r.append(np.linalg.norm(yp - y) / max(np.linalg.norm(y), 1e-30))
    return float(np.mean(r))


def evaluate(net, seeds, lo, hi, workdir, stop, stop_ref, budget=200, log=print):
    """Held-out geometries: Anderson with and without the network (to the stop margin), and the
    plain iteration with and without the network (error after `budget` outer iterations)."""
    from staticbtesolver_amd import capi
    out = []
    for s in seeds:
        p = _make((os.path.join(workdir, f"e{s}"), s, cells_for(s, lo, hi)))
        case = open_case(p)
        case.attach_gnn(net)
        ref = case.solve(solver_opts(stop_margin=stop_ref))["T"]
        row = dict(seed=s, cells=case.N_cell)
        for name, kw in (("anderson", {}), ("anderson+learned", dict(learned=1, learned_every=EVERY_AA)),
                         ("plain", dict(accel=capi.ACCEL_NONE, max_iter=budget)),
                         ("plain+learned", dict(accel=capi.ACCEL_NONE, learned=1, learned_every=EVERY_SI,
                                                max_iter=budget))):
            t = time.perf_counter()
            r = case.solve(solver_opts(**dict(dict(stop_margin=stop), **kw)))
            dt = time.perf_counter() - t
            dT = np.abs(r["T"] - ref)
            row[name] = dict(iters=r["iters"], seconds=round(dt, 4), converged=r["converged"],
                             rel_err=float(np.max(dT / np.abs(ref))),
                             rms_err_K=float(np.sqrt(np.mean(dT ** 2))),
                             accepted=r["learned_accepted"], proposed=r["learned_proposed"])
        case.attach_gnn(None)
        case.close()
        log(f"[eval] seed {s} cells {row['cells']} | " + " | ".join(
            f"{k}: {v['iters']} it {v['seconds']:.3f}s err {v['rel_err']:.1e} acc {v['accepted']}/{v['proposed']}"
            for k, v in row.items() if isinstance(v, dict)))
        out.append(row)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--train-seeds", default="1000:1256")
    ap.add_argument("--val-seeds", default="1900:1908")
    ap.add_argument("--eval-seeds", default="2000:2004")
    ap.add_argument("--cells", default="5000:50000", help="lo:hi of the log-uniform cell count")
    ap.add_argument("--eval-cells", default="50000:50001")
    ap.add_argument("--stop", type=float, default=1e-5)
    ap.add_argument("--stop-ref", type=float, default=1e-12)
    ap.add_argument("--collect-stop", type=float, default=1e-10)
    ap.add_argument("--layers", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--epochs2", type=int, default=10)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--workers", type=int, default=12)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)

    def rng_of(s):
        lo, hi = (int(x) for x in s.split(":"))
        return list(range(lo, hi))

    log = lambda m: print(m, flush=True)  # noqa: E731
    lo, hi = (int(x) for x in a.cells.split(":"))
    elo, ehi = (int(x) for x in a.eval_cells.split(":"))
    with tempfile.TemporaryDirectory(prefix="bte_train_") as wd:
        t0 = time.time()
        # one worker pool, forked before this process touches the GPU
        tr, va = rng_of(a.train_seeds), rng_of(a.val_seeds)
        data = generate(tr + va, lo, hi, a.collect_stop, wd, a.workers, log)
        train_data = [d for d in data if d["seed"] in set(tr)]
        val_data = [d for d in data if d["seed"] in set(va)]
        t_data = time.time() - t0
        net, hist = train(train_data, layers=a.layers, epochs=a.epochs, lr=a.lr, val=val_data, log=log)
        augment(net, data, log)
        net, hist2 = train(train_data, net=net, epochs=a.epochs2, lr=a.lr / 4, val=val_data, seed=1, log=log)
        t_train = time.time() - t0 - t_data
        params = net.get_params()
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        np.save(a.out, params)
        log(f"[save] {a.out} ({params.size} parameters)")
        ev = evaluate(net, rng_of(a.eval_seeds), elo, ehi, wd, a.stop, a.stop_ref, log=log)
    rep = dict(family=FAMILY, train_seeds=a.train_seeds, val_seeds=a.val_seeds, eval_seeds=a.eval_seeds,
               cells=a.cells, layers=a.layers, epochs=[a.epochs, a.epochs2], lr=a.lr, seconds_data=t_data,
               seconds_train=t_train, history=hist + hist2, eval=ev)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    return rep


if __name__ == "__main__":
    main()
