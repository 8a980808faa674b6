"""Training of the learned outer-iteration correction on pairs the solver generates itself.

Family: seeded Poisson-disk/Delaunay triangle meshes of a square of side 1e-5 m (near-diffusive:
Knudsen numbers ~0.01-0.1 for the seeded band sets), 10 seeded bands, DM=2, t=8, p=8, walls
0/3 isothermal at 301/299 K and 1/2 diffuse at 300 K.  For every training geometry the
Anderson-accelerated solver runs to a tight stop; at the state steps where the solver would
propose a correction (1, 1 + every, ...) it exports the network inputs (bte_learn_features); the
target is (T_converged - x_next) / T_span.  The network is trained on the GPU by the library's own
BF16-MFMA forward/backward kernels and Adam (bte_gnn_backward / bte_gnn_adam).

    python -m staticbtesolver_amd.train_gnn --out weights.npy --report report.json

The evaluation part solves held-out geometries with and without the network (same Anderson
settings) and reports iterations, wall time and the error against a tightly converged solution.
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


def collect(case, record_steps, stop):
    """Runs the Anderson-accelerated iteration to `stop`; returns the exported inputs at the record
    steps, the converged T and the iteration count."""
    case.begin(solver_opts(stop_margin=stop))
    recs, k, conv = [], 0, False
    for s in sorted(record_steps):
        if conv:
            break
        done, conv = case.step(s - k)
        k += done
        if not conv and k == s:
            X, x_next, span = case.learn_features()
            recs.append((X, x_next, span))
    while not conv:
        done, conv = case.step(500)
        k += done
        if done == 0:
            break
    r = case.finish()
    return recs, r["T"], r["iters"]


def loss_weight(X):
    """Per-node loss weights: errors within a few mean free paths of an isothermal wall dominate the
    residual that judges a correction (the iteration damps smooth interior errors only slowly, so
    they hardly show in the residual), so those nodes weigh up to 10x more.  Input 15 is
    log10(1 + distance in mean free paths) / 2."""
    dm = 10.0 ** (2.0 * X[..., 15].astype(np.float64)) - 1.0
    return (1.0 + 9.0 * np.exp(-dm / 3.0)).astype(np.float32)


def generate(seeds, lo, hi, record_steps, stop, workdir, workers, log=print):
    """Dataset: list of dict(seed, nbr, X [S, N, 16], y [S, N])."""
    jobs = [(os.path.join(workdir, f"g{s}"), s, cells_for(s, lo, hi)) for s in seeds]
    data = []
    t0 = time.time()
    ctx = mp.get_context("fork")
    with ctx.Pool(workers) as pool:
        for i, p in enumerate(pool.imap(_make, jobs)):
            case = open_case(p)
            recs, T, iters = collect(case, record_steps, stop)
            nbr = case.geometry()["nbr"].astype(np.int32)
            case.close()
            for f in (p["mesh"], p["band"], p["bc"]):
                os.remove(f)
            if not recs:
                continue
            X = np.stack([r[0] for r in recs])
            y = np.stack([((T - r[1]) / r[2]).astype(np.float32) for r in recs])
            data.append(dict(seed=p["seed"], nbr=nbr, X=X, y=y, w=loss_weight(X), iters=iters))
            log(f"[data] {i + 1}/{len(jobs)} seed {p['seed']} cells {len(nbr)} iters {iters} "
                f"|y| {np.sqrt(np.mean(y ** 2, axis=1)).round(4).tolist()} ({time.time() - t0:.0f} s)")
    return data


def train(data, layers=4, epochs=30, lr=2e-3, lr_min=1e-4, seed=0, val=None, log=print):
    from staticbtesolver_amd.gnn import Gnn
    net = Gnn(layers=layers, seed=seed, device=0)
    rng = np.random.default_rng(seed)
    n_steps = epochs * sum(d["X"].shape[0] for d in data)
    step = 0
    hist = []
    t0 = time.time()
    for ep in range(epochs):
        tot, cnt = 0.0, 0
        for gi in rng.permutation(len(data)):
            d = data[gi]
            net.set_graph(d["nbr"])
            for si in rng.permutation(d["X"].shape[0]):
                a = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * step / max(1, n_steps)))
                loss = net.backward(d["X"][si], d["y"][si], d["w"][si])
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
        for si in range(d["X"].shape[0]):
            yp = net.forward(d["X"][si])
            r.append(np.linalg.norm(yp - d["y"][si]) / max(np.linalg.norm(d["y"][si]), 1e-30))
    return float(np.mean(r))


def evaluate(net, seeds, lo, hi, workdir, stop, stop_ref, learned_every=8, precision=None, log=print):
    """Held-out geometries: Anderson with and without the network (same settings)."""
    from staticbtesolver_amd import capi
    prec = capi.PREC_FP64 if precision is None else precision
    out = []
    for s in seeds:
        p = _make((os.path.join(workdir, f"e{s}"), s, cells_for(s, lo, hi)))
        case = open_case(p)
        case.attach_gnn(net)
        ref = case.solve(solver_opts(stop_margin=stop_ref))
        row = dict(seed=s, cells=case.N_cell)
        for name, lrn in (("anderson", 0), ("learned", 1)):
            r = case.solve(solver_opts(stop_margin=stop, learned=lrn, learned_every=learned_every, precision=prec))
            dT = np.abs(r["T"] - ref["T"])
            row[name] = dict(iters=r["iters"], seconds=r["solve_seconds"],
                             rel_err=float(np.max(dT / np.abs(ref["T"]))),
                             dT_rel=float(np.max(dT) / np.max(np.abs(ref["T"] - FAMILY["T_ref"]))),
                             accepted=r["learned_accepted"], proposed=r["learned_proposed"])
        case.attach_gnn(None)
        case.close()
        log(f"[eval] seed {s} cells {row['cells']} anderson {row['anderson']['iters']} it "
            f"{row['anderson']['seconds']:.3f} s | learned {row['learned']['iters']} it "
            f"{row['learned']['seconds']:.3f} s acc {row['learned']['accepted']}/{row['learned']['proposed']} "
            f"dT_rel {row['learned']['dT_rel']:.2e}")
        out.append(row)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--train-seeds", default="1000:1256")
    ap.add_argument("--val-seeds", default="1900:1908")
    ap.add_argument("--eval-seeds", default="2000:2004")
    ap.add_argument("--cells", default="5000:50000", help="lo:hi of the log-uniform cell count")
    ap.add_argument("--eval-cells", default="50000:50001")
    ap.add_argument("--record", default="1,9,17,25")
    ap.add_argument("--stop", type=float, default=1e-5)
    ap.add_argument("--stop-ref", type=float, default=1e-12)
    ap.add_argument("--collect-stop", type=float, default=1e-10)
    ap.add_argument("--layers", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=30)
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
    rec = [int(x) for x in a.record.split(",")]
    with tempfile.TemporaryDirectory(prefix="bte_train_") as wd:
        t0 = time.time()
        # one worker pool, forked before this process touches the GPU
        tr, va = rng_of(a.train_seeds), rng_of(a.val_seeds)
        data = generate(tr + va, lo, hi, rec, a.collect_stop, wd, a.workers, log)
        train_data = [d for d in data if d["seed"] in set(tr)]
        val_data = [d for d in data if d["seed"] in set(va)]
        t_data = time.time() - t0
        net, hist = train(train_data, layers=a.layers, epochs=a.epochs, lr=a.lr, val=val_data, log=log)
        t_train = time.time() - t0 - t_data
        params = net.get_params()
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        np.save(a.out, params)
        log(f"[save] {a.out} ({params.size} parameters)")
        ev = evaluate(net, rng_of(a.eval_seeds), elo, ehi, wd, a.stop, a.stop_ref, log=log)
    rep = dict(family=FAMILY, train_seeds=a.train_seeds, val_seeds=a.val_seeds, eval_seeds=a.eval_seeds,
               cells=a.cells, record=rec, layers=a.layers, epochs=a.epochs, lr=a.lr, seconds_data=t_data,
               seconds_train=t_train, history=hist, eval=ev)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    return rep


if __name__ == "__main__":
    main()
