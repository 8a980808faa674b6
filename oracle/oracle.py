"""ctypes binding of the CPU oracle (oracle/bte_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker.  The product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libbte_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_util.so")

INNER_EXACT = 0
INNER_PETSC = 1
INNER_BICGSTAB = 2

_lib = None


def build(ref: bool = False) -> None:
    targets = ["all"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    L.orc_create.restype = C.c_void_p
    L.orc_create.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_char_p, C.c_char_p,
                             C.c_int, C.c_int, C.c_int, C.c_double, C.c_char_p, C.c_int]
    L.orc_destroy.argtypes = [C.c_void_p]
    L.orc_sizes.argtypes = [C.c_void_p, ip]
    L.orc_params.argtypes = [C.c_void_p, dp]
    L.orc_set_inner.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
    L.orc_set_stop.argtypes = [C.c_void_p, C.c_double]
    L.orc_set_threads.argtypes = [C.c_void_p, C.c_int]
    L.orc_quadrature.argtypes = [C.c_void_p, dp, dp, dp]
    L.orc_geometry.argtypes = [C.c_void_p, dp, dp, dp, dp, ip]
    L.orc_mesh.argtypes = [C.c_void_p, dp, ip]
    L.orc_assemble.restype = C.c_int
    L.orc_assemble.argtypes = [C.c_void_p, C.c_int, C.c_int, ip, ip, dp]
    L.orc_rhs.restype = C.c_int
    L.orc_rhs.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, dp, dp]
    L.orc_inner_solve.restype = C.c_int
    L.orc_inner_solve.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, ip, ip, dp, dp, dp]
    L.orc_sweep_levels.restype = C.c_int
    L.orc_sweep_levels.argtypes = [C.c_int, ip, ip, ip]
    L.orc_solve.restype = C.c_int
    L.orc_solve.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.orc_time_systems.restype = C.c_double
    L.orc_time_systems.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.orc_results.argtypes = [C.c_void_p, dp, dp, dp, dp, dp]
    L.orc_intensity.argtypes = [C.c_void_p, dp]
    L.orc_iterations.restype = C.c_int
    L.orc_iterations.argtypes = [C.c_void_p]
    L.orc_worst_resid.restype = C.c_double
    L.orc_worst_resid.argtypes = [C.c_void_p]
    L.orc_inner_iters.restype = C.c_longlong
    L.orc_inner_iters.argtypes = [C.c_void_p]
    L.orc_gauss.argtypes = [C.c_double, C.c_double, C.c_int, dp]
    L.orc_error.restype = C.c_char_p
    L.orc_error.argtypes = [C.c_void_p]
    L.orc_begin.restype = C.c_int
    L.orc_begin.argtypes = [C.c_void_p]
    L.orc_step.restype = C.c_double
    L.orc_step.argtypes = [C.c_void_p, C.c_int]
    L.orc_get_T.argtypes = [C.c_void_p, dp]
    L.orc_set_T.argtypes = [C.c_void_p, dp]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def gauss(a: float, b: float, n: int) -> np.ndarray:
    out = np.zeros(2 * n)
    lib().orc_gauss(a, b, n, _dp(out))
    return out.reshape(n, 2)


def inner_solve(mode, rowptr, col, val, b, tol=-1.0, maxit=-1):
    n = len(b)
    x = np.zeros(n)
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    it = lib().orc_inner_solve(mode, tol, maxit, n, _ip(rowptr), _ip(col), _dp(val), _dp(b), _dp(x))
    return x, it


def sweep_levels(rowptr, col):
    n = len(rowptr) - 1
    lv = np.zeros(n, dtype=np.int32)
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    nl = lib().orc_sweep_levels(n, _ip(rowptr), _ip(col), _ip(lv))
    return lv, nl


class OracleCase:
    """One reference case: geometry + bands + BCs + (DM, ntheta, nphi, T_ref)."""

    def __init__(self, geometry, band_file, bc_file, DM, ntheta, nphi, T_ref=300.0,
                 Lx=0.0, Ly=0.0, Lz=0.0):
        L = lib()
        err = C.create_string_buffer(256)
        h = L.orc_create(str(geometry).encode(), Lx, Ly, Lz, str(band_file).encode(),
                         str(bc_file).encode(), DM, ntheta, nphi, T_ref, err, 256)
        if not h:
            raise RuntimeError("oracle: " + err.value.decode())
        self.h = h
        s = np.zeros(8, dtype=np.int32)
        L.orc_sizes(h, _ip(s))
        (self.N_cell, self.N_dir, self.N_band, self.N_face, self.n_bc, self.dim,
         self.n_pts, self.n_bnd) = [int(v) for v in s]
        p = np.zeros(3)
        L.orc_params(h, _dp(p))
        self.T_ref, self.solid_angle, self.WFACTOR = [float(v) for v in p]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_inner(self, mode, tol=-1.0, maxit=-1):
        lib().orc_set_inner(self.h, mode, tol, maxit)

    def set_stop(self, stop):
        lib().orc_set_stop(self.h, stop)

    def set_threads(self, n):
        """Workers of an outer iteration's systems, directions split round-robin (1: serial)."""
        lib().orc_set_threads(self.h, int(n))

    def quadrature(self):
        D = self.N_dir
        om, d, S = np.zeros(D), np.zeros(3 * D), np.zeros(3 * D)
        lib().orc_quadrature(self.h, _dp(om), _dp(d), _dp(S))
        return om, d.reshape(D, 3), S.reshape(D, 3)

    def geometry(self):
        N, F = self.N_cell, self.N_face
        vol, cen = np.zeros(N), np.zeros(3 * N)
        area, nrm = np.zeros(N * F), np.zeros(3 * N * F)
        nbr = np.zeros(N * F, dtype=np.int32)
        lib().orc_geometry(self.h, _dp(vol), _dp(cen), _dp(area), _dp(nrm), _ip(nbr))
        return dict(vol=vol, centre=cen.reshape(N, 3), area=area.reshape(N, F),
                    normal=nrm.reshape(N, F, 3), nbr=nbr.reshape(N, F))

    def mesh(self):
        pts = np.zeros(3 * self.n_pts)
        cells = np.zeros(self.N_cell * (self.dim + 1), dtype=np.int32)
        lib().orc_mesh(self.h, _dp(pts), _ip(cells))
        return pts.reshape(-1, 3), cells.reshape(self.N_cell, self.dim + 1)

    def assemble(self, b, d):
        N, F = self.N_cell, self.N_face
        rp = np.zeros(N + 1, dtype=np.int32)
        col = np.zeros(N * F + 1, dtype=np.int32)
        val = np.zeros(N * F + 1)
        nnz = lib().orc_assemble(self.h, b, d, _ip(rp), _ip(col), _dp(val))
        return rp, col[:nnz].copy(), val[:nnz].copy()

    def rhs(self, b, d, ee_prev, dens):
        Re = np.zeros(self.N_cell)
        ee_prev = np.ascontiguousarray(ee_prev, dtype=np.float64)
        dens = np.ascontiguousarray(dens, dtype=np.float64)
        if lib().orc_rhs(self.h, b, d, _dp(ee_prev), _dp(dens), _dp(Re)) != 0:
            raise RuntimeError("oracle rhs: " + lib().orc_error(self.h).decode())
        return Re

    def solve(self, max_iter=10000, verbose=False):
        it = lib().orc_solve(self.h, max_iter, 1 if verbose else 0)
        if it < 0:
            raise RuntimeError("oracle solve: " + lib().orc_error(self.h).decode())
        N, B = self.N_cell, self.N_band
        T, Tb = np.zeros(N), np.zeros(N * B)
        fl, bfl = np.zeros(self.n_bc), np.zeros(self.n_bc * B)
        m = np.zeros(it)
        lib().orc_results(self.h, _dp(T), _dp(Tb), _dp(fl), _dp(bfl), _dp(m))
        ee = np.zeros(B * self.N_dir * N)
        lib().orc_intensity(self.h, _dp(ee))
        return dict(T=T, T_band=Tb.reshape(N, B), flux=fl, band_flux=bfl.reshape(self.n_bc, B),
                    margins=m, iters=it, ee=ee.reshape(B, self.N_dir, N),
                    worst_resid=lib().orc_worst_resid(self.h),
                    inner_iters=int(lib().orc_inner_iters(self.h)))

    # stepping form (outer-iteration correction studies)
    def begin(self):
        if lib().orc_begin(self.h) != 0:
            raise RuntimeError("oracle begin: " + lib().orc_error(self.h).decode())

    def step(self, n=1):
        return lib().orc_step(self.h, n)

    def get_T(self):
        T = np.zeros(self.N_cell)
        lib().orc_get_T(self.h, _dp(T))
        return T

    def set_T(self, T):
        T = np.ascontiguousarray(T, dtype=np.float64)
        lib().orc_set_T(self.h, _dp(T))

    def time_systems(self, begin, end):
        return lib().orc_time_systems(self.h, begin, end)


# ---- reference reader/utility build (oracle/_ref), only where /root/reference exists ----
_ref = None


def ref_lib():
    """The reference's own reader/utility sources built into oracle/_ref, or None."""
    global _ref
    if _ref is not None:
        return _ref
    if not os.path.exists(REF_PATH):
        return None
    L = C.CDLL(REF_PATH)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    L.ref_gauss.argtypes = [C.c_double, C.c_double, C.c_int, dp]
    L.ref_read_mesh.restype = C.c_int
    L.ref_read_mesh.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_double, ip, dp, ip, ip, ip, ip, ip]
    L.ref_read_bands.restype = C.c_int
    L.ref_read_bands.argtypes = [C.c_char_p, dp]
    L.ref_read_bcs.restype = C.c_int
    L.ref_read_bcs.argtypes = [C.c_char_p, ip, ip, dp]
    L.ref_tri_area.restype = C.c_double
    L.ref_tri_area.argtypes = [dp]
    L.ref_tet_volume.restype = C.c_double
    L.ref_tet_volume.argtypes = [dp]
    _ref = L
    return L
