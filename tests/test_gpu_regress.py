"""Regression tests of fixed defects in the C ABI's state handling.

- Captured iteration graphs bake in the inner-solve options: a second solve on the same context
  that changes only inner_maxit must not replay the old graph.
- Buffers sized by the boundary-condition count and the band count: after bte_set_bcs adds a
  boundary condition or bte_set_bands adds bands on a context that has solved, the next solve must
  match a fresh context with the same inputs.
- Band temperatures over more (band, cell) pairs than one launch's threads: every pair is updated.
"""
import numpy as np
import pytest

from staticbtesolver_amd import capi, meshgen
from tests.cases import case_kwargs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_graph_follows_inner_options():
    say("graph options 2DM2DG")
    c = capi.Case(**case_kwargs("2DM2DG"))
    c.solve(precond=capi.PC_JACOBI, inner_maxit=2, max_iter=3)  # captured with inner_maxit 2
    # the facade re-uploads before each solve; begin/step/finish keep the context's state
    c.begin(precond=capi.PC_JACOBI, inner_maxit=2, max_iter=3, use_graph=1)
    c.step(3)
    c.finish()
    c.begin(precond=capi.PC_JACOBI, inner_maxit=1000, max_iter=5, use_graph=1)
    c.step(5)
    g = c.finish()
    c.begin(precond=capi.PC_JACOBI, inner_maxit=1000, max_iter=5, use_graph=0)
    c.step(5)
    e = c.finish()
    assert np.array_equal(g["T"], e["T"]) and np.array_equal(g["margins"], e["margins"])
    assert g["inner_iters"] == e["inner_iters"] > 5 * 2 * c.N_dir


def test_setters_grow_buffers(tmp_path):
    say("setters grow buffers")
    p = meshgen.make_square_case(str(tmp_path), 600, 5, 1, 5, L=1e-7)
    a = capi.Case(p["mesh"], p["band"], p["bc"], 2, 8, 8, 300.0, Lx=1e-7, Ly=1e-7)
    ctx = a.ctx()
    N = a.N_cell
    first = capi.ctx_solve(ctx, N, 1, a.n_bc, max_iter=50, block=256, sweep_variant=4)
    assert first["iters"] > 1
    # three bands and one more (unused) boundary condition on the context that has solved
    bands = meshgen.random_bands(3, 9)
    capi.ctx_set_bands(ctx, *bands.T)
    rows = [(0, 1, 301), (1, 2, 300), (2, 2, 300), (3, 1, 299), (7, 1, 300)]
    capi.ctx_set_bcs(ctx, [-k - 1 for k, _, _ in rows], [t for _, t, _ in rows], [T for _, _, T in rows])
    grown = capi.ctx_solve(ctx, N, 3, len(rows), max_iter=50, block=256, sweep_variant=4)
    bf = tmp_path / "band3.dat"
    meshgen.write_bands(str(bf), bands)
    bc = tmp_path / "bc5.dat"
    meshgen.write_bcs(str(bc), rows)
    fresh = capi.Case(p["mesh"], str(bf), str(bc), 2, 8, 8, 300.0, Lx=1e-7, Ly=1e-7)
    ref = capi.ctx_solve(fresh.ctx(), N, 3, len(rows), max_iter=50, block=256, sweep_variant=4)
    assert grown["iters"] == ref["iters"]
    assert np.array_equal(grown["T"], ref["T"])
    np.testing.assert_array_equal(grown["band_flux"][:4], ref["band_flux"][:4])


def _strip_geometry(nx, ny, L):
    """Right triangles of an nx x ny grid on [0, L] x [0, L/nx*ny], built directly for the C ABI:
    volumes, face areas, outward normals and neighbours (face f opposite vertex f)."""
    h = L / nx
    N = 2 * nx * ny
    vol = np.full(N, 0.5 * h * h)
    area = np.zeros((N, 3))
    normal = np.zeros((N, 3, 3))
    nbr = np.zeros((N, 3), dtype=np.int32)
    i = np.repeat(np.arange(nx), ny)
    j = np.tile(np.arange(ny), nx)
    lo = 2 * (i * ny + j)  # vertices (i,j), (i+1,j), (i,j+1): faces hypotenuse, left, bottom
    hi = lo + 1            # vertices (i+1,j+1), (i,j+1), (i+1,j): faces hypotenuse, right, top
    s2 = np.sqrt(0.5)
    area[lo] = [h * np.sqrt(2), h, h]
    normal[lo] = [[s2, s2, 0], [-1, 0, 0], [0, -1, 0]]
    area[hi] = [h * np.sqrt(2), h, h]
    normal[hi] = [[-s2, -s2, 0], [1, 0, 0], [0, 1, 0]]
    nbr[lo, 0] = hi
    nbr[hi, 0] = lo
    nbr[lo, 1] = np.where(i > 0, 2 * ((i - 1) * ny + j) + 1, -1)   # left wall: entity 0
    nbr[lo, 2] = np.where(j > 0, 2 * (i * ny + j - 1) + 1, -2)     # bottom: entity 1
    nbr[hi, 1] = np.where(i < nx - 1, 2 * ((i + 1) * ny + j), -4)  # right wall: entity 3
    nbr[hi, 2] = np.where(j < ny - 1, 2 * (i * ny + j + 1), -3)    # top: entity 2
    return vol, area.ravel(), normal.ravel(), nbr.ravel()


@pytest.mark.parametrize("layout,bands", [(1, 5), (2, 10)])
def test_band_temperatures_beyond_one_launch(layout, bands):
    """Band-major: 5 bands x 3.4M cells = 17M (band, cell) pairs, more than the 65536 x 256 threads
    of a capped launch.  Bands interleaved (two pairs per thread): 10 bands, 34M pairs, more than
    65536 x 512.  After two iterations every band temperature lies between the wall temperatures."""
    nx, ny = 1300, 1300
    say("band temperatures over", bands * 2 * nx * ny, "pairs, layout", layout)
    L = 1e-6
    vol, area, normal, nbr = _strip_geometry(nx, ny, L)
    N = len(vol)
    assert bands * N > 65536 * 256 * layout
    ref = capi.Case(**case_kwargs("2DM2DG"))  # DM 2, t = p = 8 quadrature of the same kind
    om, d, S = ref.quadrature()
    L_ = capi.lib()
    ctx = capi.C.c_void_p()
    assert L_.bte_ctx_create(0, capi.C.byref(ctx)) == 0
    try:
        assert L_.bte_set_geometry(ctx, 2, N, 3, capi._dp(vol), capi._dp(area), capi._dp(normal), capi._ip(nbr)) == 0
        om, d, S = [np.ascontiguousarray(a) for a in (om, d.ravel(), S.ravel())]
        assert L_.bte_set_quadrature(ctx, 2, len(om), capi._dp(om), capi._dp(d), capi._dp(S), 8, 8,
                                     float(ref.solid_angle), 300.0) == 0
        capi.ctx_set_bands(ctx, *meshgen.random_bands(bands, 4).T)
        capi.ctx_set_bcs(ctx, [-1, -2, -3, -4], [1, 2, 2, 1], [301.0, 300.0, 300.0, 299.0])
        r = capi.ctx_solve(ctx, N, bands, 4, max_iter=2, stop_margin=0.0, block=256, sweep_variant=4,
                           intensity_layout=layout)
        assert r["iters"] == 2
        Tb = r["T_band"]
        assert np.all(np.isfinite(Tb))
        assert Tb.min() >= 299.0 - 1e-9 and Tb.max() <= 301.0 + 1e-9
        # the first iteration's margin counts every pair: positive, finite
        assert np.all(np.isfinite(r["margins"])) and r["margins"][0] > 0
    finally:
        L_.bte_ctx_destroy(ctx)
