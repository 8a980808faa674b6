"""bte_solve_noalloc against bte_solve through the C ABI.

On one context: a solve with bte_solve, then bte_solve_noalloc of the same problem, then the
context's inputs replaced by a smaller problem (fewer cells, bands and boundary faces) and
bte_solve_noalloc again, checked against bte_solve of that problem on a fresh context.  The two
entry points run the same kernels in the same order, with the reductions in different memory, so
results must agree bit for bit (plain and Anderson forms, use_graph 0 and a fixed launch shape).
A context that never solved has no device buffers, and the entry point refuses it.
"""
import numpy as np
import pytest

from staticbtesolver_amd import capi, meshgen

pytestmark = pytest.mark.gpu

FORM = dict(block=256, sweep_variant=4, use_graph=0, max_iter=300)
ROWS = [(0, 1, 301.0), (1, 2, 300.0), (2, 2, 300.0), (3, 1, 299.0)]


def say(*a):
    print(*a, flush=True)


def _same(a, b):
    assert a["iters"] == b["iters"]
    assert np.array_equal(a["T"], b["T"])
    assert np.array_equal(a["T_band"], b["T_band"])
    assert np.array_equal(a["margins"], b["margins"])
    assert np.array_equal(a["band_flux"], b["band_flux"])


def _load(ctx, case, bands):
    """Puts case's geometry, quadrature, bands and boundary conditions into ctx through the setters."""
    L = capi.lib()
    g = case.geometry()
    vol = np.ascontiguousarray(g["vol"])
    area = np.ascontiguousarray(g["area"].ravel())
    nrm = np.ascontiguousarray(g["normal"].ravel())
    nbr = np.ascontiguousarray(g["nbr"].ravel(), dtype=np.int32)
    assert L.bte_set_geometry(ctx, case.dim, case.N_cell, case.N_face, capi._dp(vol), capi._dp(area), capi._dp(nrm),
                              capi._ip(nbr)) == 0
    om, d, S = case.quadrature()
    om, d, S = np.ascontiguousarray(om), np.ascontiguousarray(d.ravel()), np.ascontiguousarray(S.ravel())
    assert L.bte_set_quadrature(ctx, 2, case.N_dir, capi._dp(om), capi._dp(d), capi._dp(S), 8, 8, case.solid_angle,
                                case.T_ref) == 0
    capi.ctx_set_bands(ctx, *bands.T)
    capi.ctx_set_bcs(ctx, [-k - 1 for k, _, _ in ROWS], [t for _, t, _ in ROWS], [T for _, _, T in ROWS])


def test_noalloc_matches_solve_twice_second_smaller(tmp_path):
    pa = meshgen.make_square_case(str(tmp_path / "a"), 2400, 11, 5, 11, L=1e-7, bc_rows=ROWS)
    pb = meshgen.make_square_case(str(tmp_path / "b"), 700, 12, 3, 12, L=3e-8, bc_rows=ROWS)
    a = capi.Case(pa["mesh"], pa["band"], pa["bc"], 2, 8, 8, 300.0, Lx=1e-7, Ly=1e-7)
    b = capi.Case(pb["mesh"], pb["band"], pb["bc"], 2, 8, 8, 300.0, Lx=3e-8, Ly=3e-8)
    assert b.N_cell < a.N_cell and b.N_band < a.N_band
    ctx = a.ctx()
    for form in (dict(FORM), dict(FORM, accel=capi.ACCEL_ANDERSON, accel_memory=20)):
        say("problem a", a.N_cell, "cells", form.get("accel", 0))
        ref = capi.ctx_solve(ctx, a.N_cell, a.N_band, a.n_bc, **form)
        got = capi.ctx_solve(ctx, a.N_cell, a.N_band, a.n_bc, noalloc=True, **form)
        _same(got, ref)
    # the smaller problem through the setters of the same context, then the allocation-free solve
    _load(ctx, b, np.loadtxt(pb["band"]))
    ref_b = capi.ctx_solve(b.ctx(), b.N_cell, b.N_band, b.n_bc, **FORM)
    for form in (dict(FORM), dict(FORM, accel=capi.ACCEL_ANDERSON, accel_memory=20)):
        say("problem b", b.N_cell, "cells", form.get("accel", 0))
        got_b = capi.ctx_solve(ctx, b.N_cell, b.N_band, b.n_bc, noalloc=True, **form)
        want = ref_b if not form.get("accel") else capi.ctx_solve(b.ctx(), b.N_cell, b.N_band, b.n_bc, **form)
        _same(got_b, want)


def test_noalloc_refuses_a_context_without_buffers(tmp_path):
    say("fresh context")
    p = meshgen.make_square_case(str(tmp_path), 500, 13, 2, 13, L=1e-7, bc_rows=ROWS)
    c = capi.Case(p["mesh"], p["band"], p["bc"], 2, 8, 8, 300.0, Lx=1e-7, Ly=1e-7)
    ctx = capi.C.c_void_p()
    L = capi.lib()
    assert L.bte_ctx_create(0, capi.C.byref(ctx)) == 0
    try:
        _load(ctx, c, np.loadtxt(p["band"]))
        with pytest.raises(capi.BTEError) as e:
            capi.ctx_solve(ctx, c.N_cell, c.N_band, c.n_bc, noalloc=True, **FORM)
        assert e.value.code == capi.BTE_ERR_STATE
        # the refusal leaves the context usable: a normal solve allocates, then the variant runs
        ref = capi.ctx_solve(ctx, c.N_cell, c.N_band, c.n_bc, **FORM)
        got = capi.ctx_solve(ctx, c.N_cell, c.N_band, c.n_bc, noalloc=True, **FORM)
        _same(got, ref)
    finally:
        L.bte_ctx_destroy(ctx)
