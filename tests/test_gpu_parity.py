"""GPU parity: the HIP path (through the C ABI) against the CPU oracle and the reference goldens.

Tolerances (FP64 plain form): assembly and RHS to 1e-13 relative (only the multiplication order
differs); inner solves to 1e-11 relative to the exact sweep; converged T to 1e-10 and main heat
fluxes to 1e-8 relative to the exact oracle; adiabatic-wall fluxes to 1e-8 of the largest flux.
Mixed precision: T within 1e-4 relative of the FP64 oracle (target of the mixed form).
"""
import numpy as np
import pytest

from oracle import oracle as O
from staticbtesolver_amd import capi, meshgen
from tests.cases import GOLDEN_CASES, case_kwargs, check_against_golden, golden_sol

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def assert_margins_close(m, m_ref, E, rel, floor):
    """Margins are mean squared changes between iterates; their square roots (RMS change) agree to
    rel, down to an absolute floor of floor * E, E the scale of the intensities (rounding of the
    iterates themselves: e ~ 1e8 in 1D, where T_ref = 0)."""
    a, b = np.sqrt(m), np.sqrt(m_ref)
    assert np.all(np.abs(a - b) <= rel * b + floor * E), np.max(np.abs(a - b) / (rel * b + floor * E))


def csr_to_faces(rp, col, val, nbr):
    N, F = nbr.shape
    diag, off = np.zeros(N), np.zeros((N, F))
    for c in range(N):
        for k in range(rp[c], rp[c + 1]):
            j = col[k]
            if j == c:
                diag[c] = val[k]
            else:
                f = int(np.nonzero(nbr[c] == j)[0][0])
                off[c, f] = val[k]
    return diag, off


@pytest.fixture(scope="module", params=sorted(GOLDEN_CASES))
def pair(request):
    name = request.param
    kw = case_kwargs(name)
    return name, O.OracleCase(**kw), capi.Case(**kw)


def test_assembly_matches_oracle(pair):
    name, o, c = pair
    say("assembly", name)
    nbr = c.geometry()["nbr"]
    for b in range(c.N_band):
        for d in range(c.N_dir):
            rp, col, val = o.assemble(b, d)
            od, oo = csr_to_faces(rp, col, val, nbr)
            gd, go = c.debug_assemble(b, d)
            np.testing.assert_allclose(gd, od, rtol=1e-13, err_msg=f"diag b={b} d={d}")
            np.testing.assert_allclose(go, oo, rtol=1e-13, atol=1e-13 * np.abs(oo).max(), err_msg=f"off b={b} d={d}")
            assert np.array_equal(go != 0, oo != 0)


def _random_state(c, seed):
    rng = np.random.default_rng(seed)
    scale = 1e5 if c.dim > 1 else 1e7
    ee = rng.uniform(0.5, 1.5, (c.N_band, c.N_dir, c.N_cell)) * scale
    dens = rng.uniform(-1, 1, (c.N_cell, c.N_band)) * scale
    return ee, dens


def test_rhs_matches_oracle(pair):
    name, o, c = pair
    say("rhs", name)
    ee, dens = _random_state(c, 1)
    g = c.debug_rhs(ee, dens)
    for b in range(c.N_band):
        for d in range(c.N_dir):
            r = o.rhs(b, d, ee, dens)
            np.testing.assert_allclose(g[b, d], r, rtol=1e-12, atol=1e-13 * np.abs(r).max(), err_msg=f"b={b} d={d}")


@pytest.mark.parametrize("direct", [1, 0])
def test_inner_solve_matches_exact_sweep(pair, direct):
    name, o, c = pair
    say("inner", name, "direct sweep" if direct else "bicgstab+sweep")
    ee, dens = _random_state(c, 2)
    x, its = c.debug_inner(ee, dens, direct_sweep=direct)
    for b in range(c.N_band):
        for d in range(c.N_dir):
            rp, col, val = o.assemble(b, d)
            rhs = o.rhs(b, d, ee, dens)
            xe, _ = O.inner_solve(O.INNER_EXACT, rp, col, val, rhs)
            np.testing.assert_allclose(x[b, d], xe, rtol=1e-11, atol=1e-12 * np.abs(xe).max(), err_msg=f"b={b} d={d}")
    assert its.max() <= 2  # the sweep is exact on these acyclic upwind graphs


def test_full_solve_matches_oracle_and_golden(pair):
    name, o, c = pair
    say("solve", name)
    o.set_inner(O.INNER_EXACT)
    o.set_stop(1e-4)
    ref = o.solve(10000)
    r = c.solve(stop_margin=1e-4, max_iter=10000)
    assert r["status"] == capi.BTE_OK
    assert r["iters"] == ref["iters"]
    assert_margins_close(r["margins"], ref["margins"], np.abs(ref["ee"]).max(), 1e-9, 1e-13)
    np.testing.assert_allclose(r["T"], ref["T"], rtol=1e-10)
    np.testing.assert_allclose(r["T_band"], ref["T_band"], rtol=1e-10)
    qmax = np.abs(ref["flux"]).max()
    main = np.abs(ref["flux"]) > 1e-3 * qmax
    assert np.all(np.abs(r["flux"] - ref["flux"])[main] <= 1e-8 * np.abs(ref["flux"])[main])
    assert np.all(np.abs(r["flux"] - ref["flux"]) <= 1e-8 * qmax)
    gm, gf = golden_sol(name)
    check_against_golden(r["margins"], r["flux"], gm, gf)


@pytest.mark.parametrize("name", ["2DM2DG", "3DM3DG", "2DM1DG"])
def test_sweep_variants_and_block_sizes_agree(name):
    """Every launch shape the autotuner can pick solves the same systems (FP64, same operations up
    to the order of a 3-term sum)."""
    say("variants", name)
    c = capi.Case(**case_kwargs(name))
    base = c.solve(sweep_variant=1, block=256, stop_margin=1e-4)
    for variant in (2, 3, 4, 5, 6, 7, 8):
        for block in (64, 128, 256, 512):
            r = c.solve(sweep_variant=variant, block=block, stop_margin=1e-4)
            st = c.stats()
            assert st["block"] == block and st["variant"] in (variant, 2)
            assert r["iters"] == base["iters"]
            np.testing.assert_allclose(r["T"], base["T"], rtol=1e-13)
            np.testing.assert_allclose(r["flux"], base["flux"], rtol=1e-9, atol=1e-12 * np.abs(base["flux"]).max())
    r = c.solve(stop_margin=1e-4)  # autotuned
    st = c.stats()
    assert st["block"] in (64, 128, 256, 512) and st["variant"] in (2, 3, 4, 5, 6, 7, 8)
    np.testing.assert_allclose(r["T"], base["T"], rtol=1e-13)


@pytest.mark.parametrize("variant,block,n_cells", [(4, 64, 20000), (4, 256, 20000), (5, 64, 20000),
                                                   (4, 256, 50000), (5, 64, 50000), (6, 128, 20000),
                                                   (6, 256, 50000), (7, 256, 20000), (7, 512, 50000),
                                                   (8, 128, 20000), (8, 256, 50000)])
def test_ring_sweep_matches_level_sweep(tmp_path, variant, block, n_cells):
    """The pipelined LDS-ring sweeps (variant 4: work-group per system; 5: one wave per system;
    6 / 7: 2 / 4 bands of a direction per work-group, 10 bands so the last group of 4 is short;
    8: the wide ring of 16384 values) perform the level-ordered sweep's arithmetic in the same
    order: identical iterates, including levels wider than the pipelined lanes (finished
    unpipelined) and, at 50k cells, upwind values older than the ring (read back from HBM)."""
    say("ring sweep", variant, block, n_cells)
    p = meshgen.make_square_case(str(tmp_path), n_cells, 2, 10, 2, L=1e-7)
    c = capi.Case(p["mesh"], p["band"], p["bc"], 2, 8, 8, 300.0, Lx=1e-7, Ly=1e-7)
    a = c.solve(sweep_variant=2, block=block)
    b = c.solve(sweep_variant=variant, block=block)
    assert c.stats()["variant"] == variant
    assert a["iters"] == b["iters"]
    np.testing.assert_array_equal(a["T"], b["T"])
    np.testing.assert_array_equal(a["margins"], b["margins"])


def test_wide_ring_sweep_matches_level_sweep_in_3d(tmp_path):
    """Variant 8 (the ring of 16384 values) on a 24,576-tet cube with 4 bands: a few upwind values
    are older than even the wide ring, so its own level flags and HBM read-back run.  Band-major,
    identical iterates to the level sweep and to variant 4; with the interleaved layout (the 3D
    default) the same up to the margins' summation order."""
    say("wide ring 3D")
    p = meshgen.make_cube_case(str(tmp_path), 16, 3, 4, 3, L=1e-7)
    c = capi.Case(p["mesh"], p["band"], p["bc"], 3, 4, 4, 300.0, Lx=1e-7, Ly=1e-7, Lz=1e-7)
    kw = dict(block=256, max_iter=25, intensity_layout=1)
    a = c.solve(sweep_variant=2, **kw)
    for v in (4, 8):
        b = c.solve(sweep_variant=v, **kw)
        assert c.stats()["variant"] == v
        assert a["iters"] == b["iters"]
        np.testing.assert_array_equal(a["T"], b["T"])
        np.testing.assert_array_equal(a["margins"], b["margins"])
        np.testing.assert_array_equal(a["flux"], b["flux"])
    i = c.solve(sweep_variant=8, block=512, max_iter=25)
    assert c.stats()["variant"] == 8 and c.stats()["intensity_layout"] == 2
    np.testing.assert_array_equal(a["T"], i["T"])
    np.testing.assert_allclose(i["margins"], a["margins"], rtol=1e-12, atol=0)
    c.close()


def test_default_stop_and_repeat_determinism():
    kw = case_kwargs("2DM2DG")
    c = capi.Case(**kw)
    say("determinism 2DM2DG")
    r1 = c.solve()
    r2 = c.solve()
    assert r1["margins"][-1] <= 1e-5 and r1["iters"] == r2["iters"]
    assert np.array_equal(r1["T"], r2["T"]) and np.array_equal(r1["flux"], r2["flux"])
    r3 = c.solve(use_graph=0)
    assert np.array_equal(r1["T"], r3["T"])


def test_mixed_precision_within_1e4_of_fp64_oracle():
    kw = case_kwargs("2DM2DG")
    say("mixed 2DM2DG")
    o = O.OracleCase(**kw)
    o.set_inner(O.INNER_EXACT)
    ref = o.solve(10000)
    c = capi.Case(**kw)
    r = c.solve(precision=capi.PREC_MIXED)
    assert r["status"] == capi.BTE_OK
    assert np.max(np.abs(r["T"] - ref["T"]) / np.abs(ref["T"])) <= 1e-4
    dT, dTr = r["T"] - o.T_ref, ref["T"] - o.T_ref
    assert np.max(np.abs(dT - dTr)) <= 1e-4 * np.max(np.abs(dTr))


def test_jacobi_path_matches_reference_bicgstab_restatement():
    kw = case_kwargs("2DM2DG")
    say("jacobi 2DM2DG")
    o = O.OracleCase(**kw)
    o.set_inner(O.INNER_BICGSTAB)
    o.set_stop(1e-4)
    ref = o.solve(10000)
    c = capi.Case(**kw)
    r = c.solve(precond=capi.PC_JACOBI, stop_margin=1e-4)
    assert r["iters"] == ref["iters"]
    np.testing.assert_allclose(r["T"], ref["T"], rtol=1e-9)
    # both stop their inner solves at a 1e-9 recursive residual: iterates agree to ~1e-9 of e
    assert_margins_close(r["margins"], ref["margins"], np.abs(ref["ee"]).max(), 1e-6, 1e-8)


@pytest.mark.parametrize("group", [2, 4])
def test_band_groups_match_per_system_kernel(tmp_path, group):
    """Band-batched FP64 BiCGSTAB (one work-group per direction and group of bands, each operator
    row loaded once for the group, per-band masks) does each band's arithmetic of the
    one-system-per-work-group kernel.  5 bands: the last group is short (masked bands).  The
    compiler may contract the two kernels' products differently, so iterates agree to rounding,
    not bit for bit.  The mixed form always runs one system per work-group."""
    say("band group", group)
    p = meshgen.make_square_case(str(tmp_path), 3000, 4, 5, 4, L=1e-6)
    c = capi.Case(p["mesh"], p["band"], p["bc"], 2, 8, 8, 300.0, Lx=1e-6, Ly=1e-6)
    form = dict(precond=capi.PC_JACOBI, stop_margin=1e-4, block=256)
    a = c.solve(band_group=1, **form)
    assert c.stats()["band_group"] == 1
    b = c.solve(band_group=group, **form)
    assert c.stats()["band_group"] == group
    say("iterations", a["iters"], b["iters"], "inner", a["inner_iters"], b["inner_iters"])
    assert a["inner_failures"] == b["inner_failures"] == 0
    assert a["iters"] == b["iters"]
    np.testing.assert_allclose(b["T"], a["T"], rtol=1e-11)
    # late margins (~1e-4) sit at the inner solves' 1e-9 noise: absolute floor 1e-6
    np.testing.assert_allclose(b["margins"], a["margins"], rtol=1e-6, atol=1e-6)
    assert abs(b["inner_iters"] - a["inner_iters"]) <= 0.02 * a["inner_iters"]
    c.solve(precision=capi.PREC_MIXED, band_group=group, max_iter=2, **form)
    assert c.stats()["band_group"] == 1


def test_mixed_jacobi_inner_solves_match_fp64_on_transport_systems(tmp_path):
    """Mixed precision with the Jacobi preconditioner on a 3D tetrahedral cube from the zero state
    (right-hand sides only at the walls, 20 bands x 128 directions): FP32 BiCGSTAB stalls on some
    of these far-from-normal systems; those are re-solved in FP64, so every system's solution
    matches the FP64 inner solve to the refinement's tolerance."""
    say("mixed jacobi 3D cube")
    p = meshgen.make_cube_case(str(tmp_path), 12, 3, 20, 3, L=1e-7)
    c = capi.Case(geometry=p["mesh"], band_file=p["band"], bc_file=p["bc"], DM=3, ntheta=4, nphi=4, T_ref=300.0,
                  Lx=1e-7, Ly=1e-7, Lz=1e-7)
    ee = np.zeros((c.N_band, c.N_dir, c.N_cell))
    dens = np.zeros((c.N_cell, c.N_band))
    x64, _ = c.debug_inner(ee, dens, precond=capi.PC_JACOBI)
    x32, its = c.debug_inner(ee, dens, precond=capi.PC_JACOBI, precision=capi.PREC_MIXED)
    scale = np.maximum(np.abs(x64).max(axis=2), 1e-300)
    err = np.abs(x32 - x64).max(axis=2) / scale
    say("max rel err", err.max(), "inner its mean", its.mean(), "max", its.max())
    assert err.max() <= 1e-6
    r = c.solve(precision=capi.PREC_MIXED, precond=capi.PC_JACOBI, max_iter=2)
    assert r["inner_failures"] == 0


def test_equal_wall_temperature_is_exact_fixed_point(tmp_path):
    say("fixed point")
    (tmp_path / "bc.dat").write_text("4\n\n#\n0 1 305\n1 2 0\n2 31 0\n3 1 305\n")
    kw = case_kwargs("2DM2DG")
    kw["bc_file"] = str(tmp_path / "bc.dat")
    c = capi.Case(**kw)
    r = c.solve()
    # converged to margin <= 1e-5 from e = 0: residual error ~1e-3 of e ~ 6e5
    np.testing.assert_allclose(r["T"], 305.0, rtol=1e-9)
    qmax = 1.627e6 / (4 * np.pi) * 5 * 2677  # scale of a 5 K flux
    assert np.all(np.abs(r["flux"]) < 1e-7 * qmax)


@pytest.mark.parametrize("bad", ["33", "7"])
def test_unsupported_boundary_type_is_an_error(tmp_path, bad):
    say("bad bc", bad)
    (tmp_path / "bc.dat").write_text(f"4\n\n#\n0 1 301\n1 {bad} 300\n2 31 300\n3 1 300\n")
    kw = case_kwargs("2DM2DG")
    kw["bc_file"] = str(tmp_path / "bc.dat")
    c = capi.Case(**kw)
    with pytest.raises(capi.BTEError) as e:
        c.solve(max_iter=2)
    assert e.value.code == capi.BTE_ERR_BC
