"""CPU-side checks of the native libraries: exported symbols, the C++ facade's readers and
preprocessing against the oracle and against the reference's own reader build."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from staticbtesolver_amd import capi, meshgen
from tests.cases import GOLDEN_CASES, case_kwargs
from tests.conftest import ROOT


def declared_symbols(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bte_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    hip = C.CDLL(capi.HIP_LIB)
    for s in declared_symbols("bte.h"):
        assert hasattr(hip, s), s
    host = capi.lib()
    for s in declared_symbols("bte_case.h"):
        assert hasattr(host, s), s
    assert set(declared_symbols("bte.h")) == set(capi.HIP_SYMBOLS)
    assert set(declared_symbols("bte_case.h")) == set(capi.CASE_SYMBOLS)


def test_no_device_is_reported_not_crashed():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    c = capi.Case(**case_kwargs("2DM2DG"))
    with pytest.raises(capi.BTEError) as e:
        c.solve(max_iter=1)
    assert e.value.code in (capi.BTE_ERR_NODEVICE, capi.BTE_ERR_HIP)


@pytest.mark.parametrize("name", sorted(GOLDEN_CASES))
def test_facade_preprocess_matches_oracle(name):
    kw = case_kwargs(name)
    o = O.OracleCase(**kw)
    c = capi.Case(**kw)
    assert (c.N_cell, c.N_dir, c.N_band, c.N_face, c.n_bc, c.dim) == (o.N_cell, o.N_dir, o.N_band, o.N_face, o.n_bc,
                                                                      o.dim)
    assert c.T_ref == o.T_ref and c.solid_angle == o.solid_angle
    og, cg = o.geometry(), c.geometry()
    for k in ("vol", "centre", "area", "normal"):
        np.testing.assert_allclose(cg[k], og[k], rtol=1e-14, atol=1e-30, err_msg=k)
    np.testing.assert_array_equal(cg["nbr"], og["nbr"])
    for a, b in zip(c.quadrature(), o.quadrature()):
        np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 16, 33])
def test_gauss_legendre_matches_numpy(n):
    x, w = np.polynomial.legendre.leggauss(n)
    g = O.gauss(-1.0, 1.0, n)
    np.testing.assert_allclose(g[:, 0], x, atol=2e-15)
    np.testing.assert_allclose(g[:, 1], w, rtol=5e-13)
    # mapped interval, as used for the 1D angular quadrature on [0, pi]
    g2 = O.gauss(0.0, np.pi, n)
    np.testing.assert_allclose(g2[:, 0], (x + 1) * np.pi / 2, atol=5e-15)
    np.testing.assert_allclose(g2[:, 1], w * np.pi / 2, rtol=5e-13)


@pytest.mark.parametrize("n", [1, 4, 8, 12])
def test_gauss_legendre_matches_reference_build(n):
    R = O.ref_lib()
    if R is None:
        pytest.skip("reference sources not built (oracle/_ref)")
    out = np.zeros(2 * n)
    R.ref_gauss(0.0, np.pi, n, out.ctypes.data_as(C.POINTER(C.c_double)))
    np.testing.assert_allclose(O.gauss(0.0, np.pi, n).ravel(), out, rtol=1e-14, atol=1e-15)


def _ref_read_mesh(path, L=(1.0, 1.0, 1.0)):
    R = O.ref_lib()
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double if a.dtype == np.float64 else C.c_int))  # noqa: E731
    cnt = np.zeros(5, dtype=np.int32)
    assert R.ref_read_mesh(str(path).encode(), *L, P(cnt), None, None, None, None, None, None) == 0
    dim, npts, nseg, ntri, ntet = [int(v) for v in cnt]
    pts = np.zeros(3 * npts)
    seg, se = np.zeros(2 * max(nseg, 1), dtype=np.int32), np.zeros(max(nseg, 1), dtype=np.int32)
    tri, te = np.zeros(3 * max(ntri, 1), dtype=np.int32), np.zeros(max(ntri, 1), dtype=np.int32)
    tet = np.zeros(4 * max(ntet, 1), dtype=np.int32)
    R.ref_read_mesh(str(path).encode(), *L, P(cnt), P(pts), P(seg), P(se), P(tri), P(te), P(tet))
    return dict(dim=dim, pts=pts.reshape(-1, 3), seg=seg[:2 * nseg].reshape(-1, 2), seg_ent=se[:nseg],
                tri=tri[:3 * ntri].reshape(-1, 3), tri_ent=te[:ntri], tet=tet[:4 * ntet].reshape(-1, 4))


def test_generated_meshes_read_by_reference_reader_and_oracle(tmp_path):
    if O.ref_lib() is None:
        pytest.skip("reference sources not built (oracle/_ref)")
    pts, tri, e, en = meshgen.square_mesh(400, 7)
    meshgen.write_mphtxt(tmp_path / "sq.mphtxt", pts, tri, e, en)
    r = _ref_read_mesh(tmp_path / "sq.mphtxt")
    assert r["dim"] == 2 and len(r["tri"]) == len(tri) and len(r["seg"]) == len(e)
    np.testing.assert_array_equal(r["tri"], tri)
    np.testing.assert_array_equal(r["seg"], e)
    np.testing.assert_array_equal(r["seg_ent"], en)
    np.testing.assert_allclose(r["pts"][:, :2], pts, rtol=0, atol=0)
    p3, tet, f3, e3 = meshgen.cube_mesh(4, 3)
    meshgen.write_mphtxt(tmp_path / "cu.mphtxt", p3, tet, f3, e3)
    r = _ref_read_mesh(tmp_path / "cu.mphtxt")
    assert r["dim"] == 3
    np.testing.assert_array_equal(r["tet"], tet)
    np.testing.assert_array_equal(r["tri"], f3)
    np.testing.assert_array_equal(r["tri_ent"], e3)


def _write_case(tmp_path, kind):
    if kind == "square":
        pts, cells, bf, be = meshgen.square_mesh(600, 11)
        rows = [(0, 1, 301), (1, 2, 300), (2, 31, 300), (3, 1, 299)]
        kw = dict(DM=2, ntheta=8, nphi=4, Lx=1e-7, Ly=1e-7)
    else:
        pts, cells, bf, be = meshgen.cube_mesh(4, 5)
        rows = [(0, 1, 301), (1, 32, 300), (2, 33, 300), (3, 2, 300), (4, 32, 300), (5, 1, 299)]
        kw = dict(DM=3, ntheta=2, nphi=2, Lx=1e-7, Ly=1e-7, Lz=1e-7)
    meshgen.write_mphtxt(tmp_path / "m.mphtxt", pts, cells, bf, be)
    meshgen.write_bands(tmp_path / "b.dat", meshgen.random_bands(3, 1))
    meshgen.write_bcs(tmp_path / "c.dat", rows)
    return dict(geometry=str(tmp_path / "m.mphtxt"), band_file=str(tmp_path / "b.dat"),
                bc_file=str(tmp_path / "c.dat"), **kw)


@pytest.mark.parametrize("kind", ["square", "cube"])
def test_facade_preprocess_matches_oracle_on_generated_meshes(tmp_path, kind):
    kw = _write_case(tmp_path, kind)
    o, c = O.OracleCase(**kw), capi.Case(**kw)
    og, cg = o.geometry(), c.geometry()
    for k in ("vol", "area", "normal"):
        np.testing.assert_allclose(cg[k], og[k], rtol=1e-14, atol=1e-30, err_msg=k)
    np.testing.assert_array_equal(cg["nbr"], og["nbr"])
    # interior faces come in symmetric pairs with opposite normals; boundary faces name an entity
    nbr = cg["nbr"]
    cells = np.repeat(np.arange(c.N_cell)[:, None], c.N_face, axis=1)
    nint = nbr >= 0
    pairs = set(zip(cells[nint].tolist(), nbr[nint].tolist()))
    assert len(pairs) == nint.sum() and all((b, a) in pairs for a, b in pairs)
    assert set((-nbr[~nint] - 1).tolist()) <= set(range(6))


def test_bands_and_bc_file_formats(tmp_path):
    b = meshgen.random_bands(20, 4)
    assert np.all((b[:, 0] >= 1e3) & (b[:, 0] <= 6e3))
    assert np.all((b[:, 1] >= 1e-12) & (b[:, 1] <= 1e-10))
    assert abs(b[:, 3].sum() - 1) < 1e-14
    meshgen.write_bands(tmp_path / "b.dat", b)
    R = O.ref_lib()
    if R is None:
        pytest.skip("reference sources not built (oracle/_ref)")
    out = np.zeros(4 * 32)
    n = R.ref_read_bands(str(tmp_path / "b.dat").encode(), out.ctypes.data_as(C.POINTER(C.c_double)))
    assert n == 20
    np.testing.assert_array_equal(out[:80].reshape(20, 4), b)
    meshgen.write_bcs(tmp_path / "c.dat", [(0, 1, 301.5), (3, 31, 300)])
    idx, typ, T = np.zeros(4, dtype=np.int32), np.zeros(4, dtype=np.int32), np.zeros(4)
    n = R.ref_read_bcs(str(tmp_path / "c.dat").encode(), idx.ctypes.data_as(C.POINTER(C.c_int)),
                       typ.ctypes.data_as(C.POINTER(C.c_int)), T.ctypes.data_as(C.POINTER(C.c_double)))
    assert n == 2 and list(idx[:2]) == [-1, -4] and list(typ[:2]) == [1, 31] and list(T[:2]) == [301.5, 300]


def test_btecmd_rejects_bad_input(tmp_path):
    r = subprocess.run([capi.BTECMD, "-g", "missing.mphtxt", "-m", "x", "-b", "y"], capture_output=True, text=True)
    assert r.returncode == 1 and "not open" in r.stdout
    r = subprocess.run([capi.BTECMD, "-g", "mesh.stl", "-m", "x", "-b", "y"], capture_output=True, text=True)
    assert r.returncode == 1 and "not supported" in r.stdout
