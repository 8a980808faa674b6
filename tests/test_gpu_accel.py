"""GPU: outer-iteration acceleration (Anderson mixing) converges to the plain iteration's fixed point."""
import numpy as np
import pytest

from oracle import oracle as O
from staticbtesolver_amd import capi, meshgen
from tests.cases import case_kwargs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name", ["2DM2DG", "3DM3DG", "3DM2DG", "2DM1DG"])
def test_anderson_reaches_the_oracle_fixed_point(name):
    say("anderson", name)
    kw = case_kwargs(name)
    o = O.OracleCase(**kw)
    o.set_inner(O.INNER_EXACT)
    o.set_stop(1e-12)
    ref = o.solve(100000)
    c = capi.Case(**kw)
    r = c.solve(accel=capi.ACCEL_ANDERSON, stop_margin=1e-12)
    p = c.solve(stop_margin=1e-12)
    assert r["converged"] and r["iters"] <= p["iters"]
    np.testing.assert_allclose(r["T"], ref["T"], rtol=1e-10)
    qmax = np.abs(ref["flux"]).max()
    assert np.all(np.abs(r["flux"] - ref["flux"]) <= 1e-7 * qmax)


def test_anderson_cuts_near_diffusive_iterations(tmp_path):
    """1e-5 m square, ~2k triangles, 10 bands: plain source iteration needs thousands of outer
    iterations; Anderson mixing reaches the same field (1e-6 relative) in a small fraction."""
    say("anderson near-diffusive")
    p = meshgen.make_square_case(str(tmp_path), 2000, 4, 10, 4, L=1e-5)
    c = capi.Case(p["mesh"], p["band"], p["bc"], 2, 8, 8, 300.0, Lx=1e-5, Ly=1e-5)
    ref = c.solve(max_iter=60000, stop_margin=1e-9)
    plain = c.solve(max_iter=60000)
    aa = c.solve(max_iter=60000, accel=capi.ACCEL_ANDERSON)
    say("plain", plain["iters"], "anderson", aa["iters"])
    assert aa["converged"] and aa["iters"] * 10 <= plain["iters"]
    assert np.max(np.abs(aa["T"] - ref["T"]) / ref["T"]) < 1e-6


def test_anderson_with_iteration_graphs_matches_eager_launches(tmp_path):
    """Anderson mixing between graph-launched outer iterations gives the iterates of the same solve
    launched one kernel at a time (a 5k-cell mesh: the cells run in the library's internal order)."""
    say("anderson: graph against eager")
    p = meshgen.make_square_case(str(tmp_path), 5000, 4, 10, 4, L=1e-5)
    c = capi.Case(p["mesh"], p["band"], p["bc"], 2, 8, 8, 300.0, Lx=1e-5, Ly=1e-5)
    form = dict(max_iter=20000, accel=capi.ACCEL_ANDERSON, accel_memory=30, block=128, sweep_variant=4)
    g = c.solve(**form)
    e = c.solve(use_graph=0, **form)
    say("iterations", g["iters"], e["iters"])
    assert g["converged"] and g["iters"] == e["iters"]
    np.testing.assert_array_equal(g["T"], e["T"])
    np.testing.assert_array_equal(g["margins"], e["margins"])
    np.testing.assert_array_equal(g["flux"], e["flux"])
