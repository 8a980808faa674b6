"""GPU: learned correction of the outer iteration -- network inputs, the residual judge, and what
the trained network gains over Anderson mixing alone on held-out near-diffusive geometries."""
import os

import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import dijkstra

from staticbtesolver_amd import capi, meshgen
from staticbtesolver_amd.gnn import Gnn
from staticbtesolver_amd.train_gnn import DEFAULT_WEIGHTS

pytestmark = pytest.mark.gpu

L_ND = 1e-5  # near-diffusive square side


def say(*a):
    print(*a, flush=True)


def nd_case(directory, n_cells, seed):
    p = meshgen.make_square_case(str(directory), n_cells, seed, 10, seed, L=L_ND)
    return capi.Case(p["mesh"], p["band"], p["bc"], 2, 8, 8, 300.0, Lx=L_ND, Ly=L_ND)


def aa(**kw):
    return dict(accel=capi.ACCEL_ANDERSON, max_iter=20000, **kw)


@pytest.fixture(scope="module")
def case3k(tmp_path_factory):
    c = nd_case(tmp_path_factory.mktemp("nd3k"), 3000, 11)
    yield c
    c.close()


def hop_distance(nbr, codes):
    """Graph distance in interior faces from the cells touching boundary entities `codes`."""
    N, F = nbr.shape
    r, f = np.nonzero(nbr >= 0)
    G = csr_matrix((np.ones(len(r)), (r, nbr[r, f])), shape=(N, N))
    seeds = np.nonzero(np.isin(nbr, codes).any(axis=1))[0]
    return dijkstra(G, directed=False, indices=seeds, unweighted=True, min_only=True)


def test_learn_features_match_their_definition(case3k):
    """Dynamic inputs against the exported state; static inputs against graph distances computed
    independently (scipy's Dijkstra) from the face table."""
    say("learned: features")
    c = case3k
    c.begin(capi.default_opts(**aa()))
    c.step(3)
    X, x_next, span = c.learn_features()
    c.finish()
    N = c.N_cell
    assert X.shape == (N, 16) and np.all(np.isfinite(X))
    assert span == 2.0  # walls at 301 and 299 K
    # 0: residual over its RMS -> unit RMS
    assert abs(np.sqrt(np.mean(X[:, 0].astype(np.float64) ** 2)) - 1.0) < 1e-5
    np.testing.assert_allclose(X[:, 3], (x_next - 300.0) / 2.0, atol=1e-6)
    assert np.all(X[:, 4] == X[0, 4])
    nbr = c.geometry()["nbr"]
    # boundary entity k is stored as neighbour -k-1; 0 and 3 isothermal (301 / 299 K), 1 and 2 diffuse
    d_hot, d_cold, d_refl = hop_distance(nbr, [-1]), hop_distance(nbr, [-4]), hop_distance(nbr, [-2, -3])
    d_iso = np.minimum(d_hot, d_cold)
    st = X[:, 6:]  # static inputs
    np.testing.assert_array_equal(st[:, 0], (d_iso == 0).astype(np.float32))
    np.testing.assert_array_equal(st[:, 1], (d_refl == 0).astype(np.float32))
    w_h, w_c = 1.0 / (d_hot + 0.5), 1.0 / (d_cold + 0.5)
    T_int = (301.0 * w_h + 299.0 * w_c) / (w_h + w_c)
    np.testing.assert_allclose(st[:, 3], (T_int - 300.0) / 2.0, atol=1e-6)
    np.testing.assert_allclose(st[:, 4], d_iso / (d_iso + d_refl + 1e-30), atol=1e-6)
    np.testing.assert_allclose(st[:, 5], np.log10(1 + 0.5 * (d_hot + d_cold)) / 3, atol=1e-6)
    np.testing.assert_allclose(X[:, 5], X[:, 3] - st[:, 3], atol=1e-6)
    # 7: distance from the nearest isothermal wall in mean free paths (mfp / h from static 2)
    kn = 10.0 ** (2 * st[:, 2].astype(np.float64))
    np.testing.assert_allclose(st[:, 7], np.log10(1 + d_iso / kn) / 2, atol=2e-5)
    # 8: conductivity fraction of bands with mean free path beyond (d_iso + 1/2) h; 9: spread (uniform)
    assert np.all((st[:, 8] >= 0) & (st[:, 8] <= 1 + 1e-6))
    assert st[:, 8][d_iso == 0].mean() >= st[:, 8][d_iso > 5].mean()
    assert np.all(st[:, 9] == st[0, 9]) and st[0, 9] > 0


def test_rejected_corrections_leave_the_iteration_unchanged(case3k):
    """A network that proposes a 2000 K jump everywhere: every proposal raises the residual, is
    rejected, and the iteration runs as without the network (bit for bit); proposals stop after two
    rejections in a row."""
    say("learned: rejection")
    c = case3k
    g = Gnn(layers=4, seed=1, device=0)
    p = np.zeros(g.n_params, np.float32)
    p[-1] = 1000.0  # output bias: y = 1000 -> correction 2000 K
    g.set_params(p)
    c.attach_gnn(g)
    base = c.solve(**aa())
    r = c.solve(**aa(learned=1, learned_every=4))
    c.attach_gnn(None)
    say("iters", base["iters"], r["iters"], "proposed", r["learned_proposed"], "accepted", r["learned_accepted"])
    assert r["learned_proposed"] == 2 and r["learned_accepted"] == 0
    assert r["iters"] == base["iters"] and r["converged"]
    np.testing.assert_array_equal(r["T"], base["T"])
    np.testing.assert_array_equal(r["margins"], base["margins"])


def test_null_correction_changes_nothing(case3k):
    """An all-zero network proposes zero corrections: whether accepted or not, the iterates are
    those of Anderson mixing alone."""
    say("learned: null network")
    c = case3k
    g = Gnn(layers=4, seed=2, device=0)
    g.set_params(np.zeros(g.n_params, np.float32))
    c.attach_gnn(g)
    base = c.solve(**aa())
    r = c.solve(**aa(learned=1, learned_every=3))
    c.attach_gnn(None)
    assert r["learned_proposed"] >= 1
    assert r["iters"] == base["iters"]
    np.testing.assert_array_equal(r["T"], base["T"])


def test_proposals_on_a_fixed_point_are_not_made(tmp_path):
    """Every wall at T_ref = 0 K: the initial state is the exact fixed point and every sum that forms
    the next one is exactly 0 (at 300 K the band-weighted cell temperature would round away from
    300 by an ulp), so each proposal (launched before its norms are known) finds no residual to
    correct, applies nothing and counts as not made; the run equals Anderson mixing alone bit for
    bit."""
    say("learned: proposals at a fixed point")
    p = meshgen.make_square_case(str(tmp_path), 2000, 12, 5, 12, L=L_ND)
    bc = str(tmp_path / "bc_equal.dat")
    meshgen.write_bcs(bc, [(0, 1, 0.0), (1, 1, 0.0), (2, 1, 0.0), (3, 1, 0.0)])
    c = capi.Case(p["mesh"], p["band"], bc, 2, 8, 8, 0.0, Lx=L_ND, Ly=L_ND)
    g = Gnn(layers=4, seed=0, device=0)
    g.set_params(np.load(DEFAULT_WEIGHTS))
    c.attach_gnn(g)
    opts = dict(accel=capi.ACCEL_ANDERSON, stop_margin=-1.0, max_iter=6)
    base = c.solve(**opts)
    r = c.solve(learned=1, learned_every=1, **opts)
    c.attach_gnn(None)
    c.close()
    assert r["learned_proposed"] == 0 and r["learned_accepted"] == 0 and r["iters"] == base["iters"] == 6
    assert np.all(r["T"] == 0.0)
    np.testing.assert_array_equal(r["T"], base["T"])
    np.testing.assert_array_equal(r["margins"], base["margins"])


def test_learned_needs_an_attached_network(case3k):
    c = case3k
    c.attach_gnn(None)
    with pytest.raises(capi.BTEError):
        c.solve(**aa(learned=1))


def test_learned_with_plain_iteration_reaches_the_same_field(case3k):
    """learned with the reference's plain source iteration (no mixing): a random network's
    proposals are judged; the result still meets the plain iteration's fixed point."""
    say("learned: plain iteration")
    c = case3k
    ref = c.solve(**aa(stop_margin=1e-12))
    g = Gnn(layers=4, seed=3, device=0)
    c.attach_gnn(g)
    r = c.solve(max_iter=60000, stop_margin=1e-9, learned=1, learned_every=50)
    c.attach_gnn(None)
    assert r["converged"] and r["learned_proposed"] >= 1
    assert np.max(np.abs(r["T"] - ref["T"]) / ref["T"]) < 1e-6


def iterations_to(c, opts, ref, tol, budget):
    """Outer iterations until max |T - T*| / T* <= tol (None if not within budget) and the error at
    the end of the budget (or at the stop rule, if earlier)."""
    c.begin(capi.default_opts(**opts))
    hit, k, conv, err = None, 0, False, None
    while k < budget and not conv:
        done, conv = c.step(1)
        k += done
        err = np.max(np.abs(c.temperature() - ref) / ref)
        if hit is None and err <= tol:
            hit = k
    r = c.finish()
    return hit, err, r


@pytest.mark.skipif(not os.path.exists(DEFAULT_WEIGHTS), reason="trained weights not present")
@pytest.mark.parametrize("seed,n_cells", [(2001, 20000), (2002, 50000)])
def test_trained_network_gains(tmp_path, seed, n_cells):
    """Held-out geometries of the training family (seeds outside 1000-1255).  What the trained
    network gains: with it, the field is within 1e-4 of the converged one after a handful of outer
    iterations, where Anderson mixing alone needs tens to hundreds and the plain source iteration
    thousands; at a budget of 60 outer iterations the plain iteration with the network is >= 10x
    more accurate than without.  Every judged correction obeys the residual rule, and the solves
    still run to the same fixed point (<= 1e-4 at the stop rule)."""
    say("learned: trained network", seed, n_cells)
    c = nd_case(tmp_path, n_cells, seed)
    g = Gnn(layers=4, seed=0, device=0)
    g.set_params(np.load(DEFAULT_WEIGHTS))
    c.attach_gnn(g)
    ref = c.solve(**aa(stop_margin=1e-12))["T"]
    k_aa, _, r_aa = iterations_to(c, aa(), ref, 1e-4, 20000)
    k_l, _, r_l = iterations_to(c, aa(learned=1), ref, 1e-4, 20000)
    _, e_pl, _ = iterations_to(c, dict(max_iter=60), ref, 1e-4, 60)
    _, e_pll, r_pll = iterations_to(c, dict(max_iter=60, learned=1, learned_every=16), ref, 1e-4, 60)
    c.attach_gnn(None)
    c.close()
    say(f"iterations to 1e-4: anderson {k_aa}, anderson+network {k_l} (accepted "
        f"{r_l['learned_accepted']}/{r_l['learned_proposed']}); error after 60 plain iterations {e_pl:.2e}, "
        f"with the network {e_pll:.2e}")
    assert r_l["learned_accepted"] >= 1 and r_pll["learned_accepted"] >= 1
    assert k_l is not None and k_aa is not None and k_l * 5 <= k_aa and k_l <= 10
    assert e_pll * 10 <= e_pl
    for r in (r_aa, r_l):
        assert r["converged"] and np.max(np.abs(r["T"] - ref) / ref) <= 1e-4
