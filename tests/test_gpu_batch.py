"""GPU: uneven batch of problems -- one sweep launch per outer iteration over the systems of several
meshes; every problem's result equals its own single-problem solve."""
import numpy as np
import pytest

from staticbtesolver_amd import capi, meshgen
from tests.cases import case_kwargs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def same(a, b):
    assert a["iters"] == b["iters"] and a["converged"] == b["converged"]
    np.testing.assert_array_equal(a["T"], b["T"])
    np.testing.assert_array_equal(a["margins"], b["margins"])
    np.testing.assert_array_equal(a["flux"], b["flux"])


def test_batch_of_reference_cases_matches_single_solves():
    """2D and 3D golden problems of different sizes, band counts and boundary types in one batch."""
    say("batch: reference cases")
    cases = [capi.Case(**case_kwargs(n)) for n in ("2DM2DG", "3DM3DG", "2DM1DG", "3DM2DG")]
    single = [c.solve(sweep_variant=2, block=256, stop_margin=1e-6) for c in cases]
    batch = capi.Batch(cases)
    out = batch.solve(stop_margin=1e-6)
    for s, b in zip(single, out):
        same(s, b)
    # a second solve of the same batch starts from scratch
    for s, b in zip(single, batch.solve(stop_margin=1e-6)):
        same(s, b)
    batch.close()


def test_batch_with_anderson_on_uneven_meshes(tmp_path):
    """Near-diffusive and ballistic squares of 1k-6k cells with Anderson mixing: problems converge at
    different outer iterations and leave the batch one by one."""
    say("batch: uneven meshes")
    cases = []
    for k, (n, L, nb) in enumerate([(1500, 1e-5, 5), (6000, 1e-7, 8), (3000, 1e-6, 10)]):
        p = meshgen.make_square_case(str(tmp_path / f"m{k}"), n, 30 + k, nb, 30 + k, L=L)
        cases.append(capi.Case(p["mesh"], p["band"], p["bc"], 2, 8, 8, 300.0, Lx=L, Ly=L))
    opts = dict(accel=capi.ACCEL_ANDERSON)
    single = [c.solve(sweep_variant=2, block=256, **opts) for c in cases]
    assert len({s["iters"] for s in single}) == 3
    out = capi.Batch(cases).solve(**opts)
    for s, b in zip(single, out):
        same(s, b)


def test_large_batch_matches_single_solves(tmp_path):
    """Eight squares of 10 bands (2,560 systems, more work-groups than the device holds at once)
    with Anderson mixing: every problem matches its single solve bit for bit."""
    say("batch: large")
    cases = []
    for k in range(8):
        n, L = 1200 + 400 * k, 10 ** (-5 - 0.3 * k)
        p = meshgen.make_square_case(str(tmp_path / f"m{k}"), n, 40 + k, 10, 40 + k, L=L)
        cases.append(capi.Case(p["mesh"], p["band"], p["bc"], 2, 8, 8, 300.0, Lx=L, Ly=L))
    assert sum(c.N_band * c.N_dir for c in cases) >= 2048
    opts = dict(accel=capi.ACCEL_ANDERSON)
    single = [c.solve(sweep_variant=2, block=256, **opts) for c in cases]
    assert len({s["iters"] for s in single}) > 1
    out = capi.Batch(cases).solve(**opts)
    for s, b in zip(single, out):
        same(s, b)


def test_batch_rejects_krylov_forms():
    c = capi.Case(**case_kwargs("2DM2DG"))
    with pytest.raises(capi.BTEError):
        capi.Batch([c]).solve(precision=capi.PREC_MIXED)


def test_batch_with_learned_correction(tmp_path):
    """Learned corrections (one trained network shared by the problems) proposed, judged and undone
    per problem inside the batch: every problem's result equals its own single-problem solve."""
    from staticbtesolver_amd.gnn import Gnn
    from staticbtesolver_amd.train_gnn import DEFAULT_WEIGHTS

    say("batch: learned")
    net = Gnn(layers=4, seed=0, device=0)
    net.set_params(np.load(DEFAULT_WEIGHTS))
    cases = []
    for k, (n, nb) in enumerate([(1500, 5), (4500, 10), (2500, 8)]):
        p = meshgen.make_square_case(str(tmp_path / f"m{k}"), n, 40 + k, nb, 40 + k, L=1e-5)
        cases.append(capi.Case(p["mesh"], p["band"], p["bc"], 2, 8, 8, 300.0, Lx=1e-5, Ly=1e-5))
    opts = dict(accel=capi.ACCEL_ANDERSON, learned=1, accel_memory=20, max_iter=20000)
    single = []
    for c in cases:
        c.attach_gnn(net)
        single.append(c.solve(sweep_variant=2, block=256, **opts))
    assert sum(s["learned_proposed"] for s in single) > 0
    out = capi.Batch(cases).solve(**opts)
    for s, b in zip(single, out):
        same(s, b)
        assert s["learned_accepted"] == b["learned_accepted"] and s["learned_proposed"] == b["learned_proposed"]
    for c in cases:
        c.attach_gnn(None)
