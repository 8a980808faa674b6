"""GPU: the intensity layouts of the pipelined sweep.  Band-major intensities (the reference's layout,
ee[b][d][c]) and direction-major ones with each cell's bands interleaved (ee[d][c][b], bands padded
to a multiple of 4) hold the same values: every kernel does the same arithmetic in the same order,
except the band-temperature pass's margin sums, which it groups by cells instead of bands (rounding
only).  Compared: temperatures, fluxes and the exported intensities bit for bit, iteration counts,
margins to 1e-12."""
import numpy as np
import pytest

from staticbtesolver_amd import capi, meshgen
from staticbtesolver_amd.gnn import Gnn
from staticbtesolver_amd.train_gnn import DEFAULT_WEIGHTS

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def same(a, b):
    assert a["iters"] == b["iters"] and a["converged"] == b["converged"]
    np.testing.assert_array_equal(a["T"], b["T"])
    np.testing.assert_array_equal(a["flux"], b["flux"])
    np.testing.assert_allclose(b["margins"], a["margins"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("accel", [capi.ACCEL_NONE, capi.ACCEL_ANDERSON])
def test_cube_layouts_agree(tmp_path, accel):
    """3D, 20 bands x 128 directions on a 10,368-tet lattice: the interleaved layout is the default
    there (no padding: 20 bands fill five 32-byte sectors per cell)."""
    say("layouts: cube, accel", accel)
    p = meshgen.make_cube_case(str(tmp_path), 12, 3, 20, 3, L=1e-7)
    c = capi.Case(p["mesh"], p["band"], p["bc"], 3, 4, 4, 300.0, Lx=1e-7, Ly=1e-7, Lz=1e-7)
    kw = dict(sweep_variant=4, block=256, accel=accel, max_iter=400)
    band = c.solve(intensity_layout=1, **kw)
    assert c.stats()["intensity_layout"] == 1
    e_band = c.intensity()
    auto = c.solve(**kw)
    assert c.stats()["intensity_layout"] == 2 and c.stats()["variant"] == 4
    e_auto = c.intensity()
    same(band, auto)
    np.testing.assert_array_equal(e_band, e_auto)
    c.close()


def test_square_layouts_agree_with_padding_and_learned_corrections(tmp_path):
    """2D, 10 bands (padded to 12 in the interleaved layout) with Anderson mixing and network
    corrections, which read and write boundary intensities through the layout: identical runs."""
    say("layouts: square, learned")
    p = meshgen.make_square_case(str(tmp_path), 3000, 21, 10, 21, L=1e-5)
    c = capi.Case(p["mesh"], p["band"], p["bc"], 2, 8, 8, 300.0, Lx=1e-5, Ly=1e-5)
    g = Gnn(layers=4, seed=0, device=0)
    g.set_params(np.load(DEFAULT_WEIGHTS))
    c.attach_gnn(g)
    kw = dict(sweep_variant=4, block=128, accel=capi.ACCEL_ANDERSON, learned=1, learned_every=4, max_iter=20000)
    band = c.solve(**kw)
    assert c.stats()["intensity_layout"] == 1  # the default in 2D
    inter = c.solve(intensity_layout=2, **kw)
    assert c.stats()["intensity_layout"] == 2
    assert band["learned_proposed"] > 0
    assert (band["learned_proposed"], band["learned_accepted"]) == (inter["learned_proposed"], inter["learned_accepted"])
    same(band, inter)
    c.attach_gnn(None)
    c.close()


@pytest.mark.parametrize("layout", [1, 2])
def test_wide_ring_with_mixing_and_corrections_matches_variant_4(tmp_path, layout):
    """Variant 8 forms the interior right-hand sides inside the sweep from the source temperature
    that every writer of the emission density records (cell temperatures, Anderson updates and
    gathers, network corrections and their undo); variant 4 reads them from the separate pass.
    With Anderson mixing and network corrections: identical runs in either layout."""
    say("wide ring, learned, layout", layout)
    p = meshgen.make_square_case(str(tmp_path), 3000, 21, 10, 21, L=1e-5)
    c = capi.Case(p["mesh"], p["band"], p["bc"], 2, 8, 8, 300.0, Lx=1e-5, Ly=1e-5)
    g = Gnn(layers=4, seed=0, device=0)
    g.set_params(np.load(DEFAULT_WEIGHTS))
    c.attach_gnn(g)
    kw = dict(block=128, accel=capi.ACCEL_ANDERSON, learned=1, learned_every=4, max_iter=20000,
              intensity_layout=layout)
    a = c.solve(sweep_variant=4, **kw)
    assert c.stats()["variant"] == 4
    b = c.solve(sweep_variant=8, **kw)
    assert c.stats()["variant"] == 8 and c.stats()["intensity_layout"] == layout
    assert a["learned_proposed"] > 0
    assert (a["learned_proposed"], a["learned_accepted"]) == (b["learned_proposed"], b["learned_accepted"])
    assert a["iters"] == b["iters"]
    np.testing.assert_array_equal(a["T"], b["T"])
    np.testing.assert_array_equal(a["margins"], b["margins"])
    np.testing.assert_array_equal(a["flux"], b["flux"])
    c.attach_gnn(None)
    c.close()


def test_other_forms_stay_band_major(tmp_path):
    """The interleaved layout is only for the pipelined sweep: asked for with another form, the
    solve runs band-major."""
    p = meshgen.make_square_case(str(tmp_path), 1500, 22, 4, 22, L=1e-7)
    c = capi.Case(p["mesh"], p["band"], p["bc"], 2, 8, 8, 300.0, Lx=1e-7, Ly=1e-7)
    r2 = c.solve(intensity_layout=2, sweep_variant=2, block=128)
    assert c.stats()["intensity_layout"] == 1 and c.stats()["variant"] == 2
    rj = c.solve(intensity_layout=2, precond=capi.PC_JACOBI)
    assert c.stats()["intensity_layout"] == 1
    assert r2["converged"] and rj["converged"]
    c.close()
