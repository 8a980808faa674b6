"""ctypes binding of the library's C entry points (include/bte.h, include/bte_case.h).

The shared libraries are built in-tree (staticbtesolver_amd/lib/); loading fails loudly when they
are missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
HIP_LIB = os.path.join(LIB_DIR, "libbte_hip.so")
HOST_LIB = os.path.join(LIB_DIR, "libStaticBTESolver.so")
BTECMD = os.path.join(HERE, "bin", "BTEcmd")

BTE_OK = 0
BTE_ERR_ARG = -1
BTE_ERR_HIP = -2
BTE_ERR_STATE = -3
BTE_ERR_BC = -4
BTE_ERR_MESH = -5
BTE_ERR_NOCONV = -6
BTE_ERR_NODEVICE = -7

PREC_FP64 = 0
PREC_MIXED = 1
PC_SWEEP = 0
PC_JACOBI = 1
ACCEL_NONE = 0
ACCEL_ANDERSON = 1


class BTEError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (status {code})")
        self.code = code


class Opts(C.Structure):
    _fields_ = [("precision", C.c_int), ("precond", C.c_int), ("max_iter", C.c_int),
                ("stop_margin", C.c_double), ("inner_tol", C.c_double), ("inner_maxit", C.c_int),
                ("refine_steps", C.c_int), ("block", C.c_int), ("use_graph", C.c_int), ("verbose", C.c_int),
                ("timing", C.c_int), ("learned", C.c_int), ("learned_every", C.c_int),
                ("direct_sweep", C.c_int), ("sweep_variant", C.c_int), ("accel", C.c_int),
                ("accel_memory", C.c_int), ("band_group", C.c_int),
                ("inner_tol32", C.c_double), ("intensity_layout", C.c_int)]


class Result(C.Structure):
    _fields_ = [("T", C.POINTER(C.c_double)), ("T_band", C.POINTER(C.c_double)),
                ("bc_flux", C.POINTER(C.c_double)), ("bc_band_flux", C.POINTER(C.c_double)),
                ("margins", C.POINTER(C.c_double)), ("margins_capacity", C.c_int),
                ("iterations", C.c_int), ("converged", C.c_int), ("inner_iters", C.c_longlong),
                ("inner_failures", C.c_int), ("solve_seconds", C.c_double), ("setup_seconds", C.c_double),
                ("learned_accepted", C.c_int), ("learned_proposed", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("n_kernels", C.c_int), ("names", C.c_char_p * 16), ("ms", C.c_double * 16),
                ("calls", C.c_longlong * 16), ("spmv_bytes", C.c_double), ("spmv_rows", C.c_longlong),
                ("ell_width", C.c_int), ("max_levels", C.c_int), ("lagged_edges", C.c_int), ("block", C.c_int),
                ("variant", C.c_int), ("band_group", C.c_int), ("intensity_layout", C.c_int)]


# symbols declared in include/bte.h and include/bte_case.h
HIP_SYMBOLS = ["bte_default_opts", "bte_ctx_create", "bte_ctx_destroy", "bte_last_error", "bte_set_geometry",
               "bte_set_quadrature", "bte_set_bands", "bte_set_bcs", "bte_solve", "bte_solve_noalloc", "bte_get_intensity",
               "bte_get_stats", "bte_debug_assemble", "bte_debug_rhs", "bte_debug_inner", "bte_begin", "bte_step",
               "bte_get_temperature", "bte_set_temperature", "bte_finish", "bte_gnn_create", "bte_gnn_destroy",
               "bte_gnn_last_error", "bte_gnn_num_params", "bte_gnn_get_params", "bte_gnn_set_params",
               "bte_gnn_get_grads", "bte_gnn_set_graph", "bte_gnn_forward", "bte_gnn_backward", "bte_gnn_adam",
               "bte_attach_gnn", "bte_learn_features", "bte_batch_create", "bte_batch_destroy",
               "bte_batch_last_error", "bte_batch_solve", "bte_batch_get_stats"]
CASE_SYMBOLS = ["bte_case_open", "bte_case_close", "bte_case_sizes", "bte_case_params", "bte_case_solve",
                "bte_case_geometry", "bte_case_quadrature", "bte_case_ctx", "bte_case_error"]

_lib = None


def lib():
    """Loads libStaticBTESolver.so (which links libbte_hip.so).  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not (os.path.exists(HOST_LIB) and os.path.exists(HIP_LIB)):
        raise RuntimeError(f"staticbtesolver_amd native libraries are not built ({LIB_DIR}); run "
                           "`python -c 'import __graft_entry__ as g; g.build()'` or `make -C staticbtesolver_amd`")
    C.CDLL(HIP_LIB, mode=C.RTLD_GLOBAL)
    L = C.CDLL(HOST_LIB)
    dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
    L.bte_default_opts.argtypes = [C.POINTER(Opts)]
    L.bte_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.bte_ctx_destroy.argtypes = [vp]
    L.bte_last_error.restype = C.c_char_p
    L.bte_last_error.argtypes = [vp]
    L.bte_set_geometry.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp, dp, ip]
    L.bte_set_quadrature.argtypes = [vp, C.c_int, C.c_int, dp, dp, dp, C.c_int, C.c_int, C.c_double, C.c_double]
    L.bte_set_bands.argtypes = [vp, C.c_int, dp, dp, dp, dp]
    L.bte_set_bcs.argtypes = [vp, C.c_int, ip, ip, dp]
    L.bte_solve.argtypes = [vp, C.POINTER(Opts), C.POINTER(Result)]
    L.bte_solve_noalloc.argtypes = [vp, C.POINTER(Opts), C.POINTER(Result)]
    L.bte_get_intensity.argtypes = [vp, dp]
    L.bte_get_stats.argtypes = [vp, C.POINTER(Stats)]
    L.bte_debug_assemble.argtypes = [vp, C.c_int, C.c_int, dp, dp]
    L.bte_debug_rhs.argtypes = [vp, dp, dp, dp]
    L.bte_debug_inner.argtypes = [vp, C.POINTER(Opts), dp, dp, dp, ip]
    L.bte_begin.argtypes = [vp, C.POINTER(Opts)]
    L.bte_step.argtypes = [vp, C.c_int, ip, ip]
    L.bte_get_temperature.argtypes = [vp, dp, dp]
    L.bte_set_temperature.argtypes = [vp, dp]
    L.bte_finish.argtypes = [vp, C.POINTER(Result)]
    L.bte_attach_gnn.argtypes = [vp, vp]
    L.bte_learn_features.argtypes = [vp, C.POINTER(C.c_float), dp, dp]
    L.bte_batch_create.argtypes = [C.c_int, C.POINTER(vp), C.POINTER(vp)]
    L.bte_batch_destroy.argtypes = [vp]
    L.bte_batch_last_error.restype = C.c_char_p
    L.bte_batch_last_error.argtypes = [vp]
    L.bte_batch_solve.argtypes = [vp, C.POINTER(Opts), C.POINTER(Result)]
    L.bte_batch_get_stats.argtypes = [vp, C.POINTER(Stats)]
    L.bte_case_open.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_char_p, C.c_char_p, C.c_int,
                                C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(vp), C.c_char_p, C.c_int]
    L.bte_case_close.argtypes = [vp]
    L.bte_case_sizes.argtypes = [vp, ip]
    L.bte_case_params.argtypes = [vp, dp]
    L.bte_case_solve.argtypes = [vp, C.POINTER(Opts), C.c_int, C.POINTER(Result)]
    L.bte_case_geometry.argtypes = [vp, dp, dp, dp, dp, ip]
    L.bte_case_quadrature.argtypes = [vp, dp, dp, dp]
    L.bte_case_ctx.restype = vp
    L.bte_case_ctx.argtypes = [vp]
    L.bte_case_error.restype = C.c_char_p
    L.bte_case_error.argtypes = [vp]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def default_opts(**kw) -> Opts:
    o = Opts()
    lib().bte_default_opts(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


class Case:
    """A problem loaded from the reference's input files through the C++ facade
    (StaticBTESolver + BTEMesh + BTEBand + BTEBoundaryCondition)."""

    def __init__(self, geometry, band_file, bc_file, DM, ntheta, nphi, T_ref=300.0, Lx=0.0, Ly=0.0, Lz=0.0,
                 device=-1):
        L = lib()
        err = C.create_string_buffer(512)
        h = C.c_void_p()
        rc = L.bte_case_open(str(geometry).encode(), Lx, Ly, Lz, str(band_file).encode(), str(bc_file).encode(),
                             DM, ntheta, nphi, T_ref, device, C.byref(h), err, 512)
        if rc != BTE_OK:
            raise BTEError(rc, err.value.decode())
        self.h = h
        s = np.zeros(6, dtype=np.int32)
        L.bte_case_sizes(h, _ip(s))
        self.N_cell, self.N_dir, self.N_band, self.N_face, self.n_bc, self.dim = [int(x) for x in s]
        p = np.zeros(2)
        L.bte_case_params(h, _dp(p))
        self.T_ref, self.solid_angle = float(p[0]), float(p[1])

    def close(self):
        if getattr(self, "h", None):
            lib().bte_case_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def geometry(self):
        N, F = self.N_cell, self.N_face
        vol, cen = np.zeros(N), np.zeros(3 * N)
        area, nrm = np.zeros(N * F), np.zeros(3 * N * F)
        nbr = np.zeros(N * F, dtype=np.int32)
        lib().bte_case_geometry(self.h, _dp(vol), _dp(cen), _dp(area), _dp(nrm), _ip(nbr))
        return dict(vol=vol, centre=cen.reshape(N, 3), area=area.reshape(N, F), normal=nrm.reshape(N, F, 3),
                    nbr=nbr.reshape(N, F))

    def quadrature(self):
        D = self.N_dir
        om, d, S = np.zeros(D), np.zeros(3 * D), np.zeros(3 * D)
        lib().bte_case_quadrature(self.h, _dp(om), _dp(d), _dp(S))
        return om, d.reshape(D, 3), S.reshape(D, 3)

    def solve(self, opts: Opts | None = None, verbose=False, margins_capacity=None, **kw):
        if opts is None:
            opts = default_opts(**kw)
        N, B, nbc = self.N_cell, self.N_band, self.n_bc
        cap = margins_capacity or max(1, opts.max_iter)
        T, Tb, fl, bfl, m = np.zeros(N), np.zeros(N * B), np.zeros(max(nbc, 1)), np.zeros(max(nbc, 1) * B), np.zeros(cap)
        r = Result()
        r.T, r.T_band, r.bc_flux, r.bc_band_flux, r.margins = _dp(T), _dp(Tb), _dp(fl), _dp(bfl), _dp(m)
        r.margins_capacity = cap
        rc = lib().bte_case_solve(self.h, C.byref(opts), 1 if verbose else 0, C.byref(r))
        if rc not in (BTE_OK, BTE_ERR_NOCONV):
            raise BTEError(rc, lib().bte_case_error(self.h).decode())
        it = r.iterations
        return dict(T=T, T_band=Tb.reshape(N, B), flux=fl[:nbc], band_flux=bfl[:nbc * B].reshape(nbc, B),
                    margins=m[:min(it, cap)].copy(), iters=it, converged=bool(r.converged),
                    inner_iters=int(r.inner_iters), inner_failures=int(r.inner_failures),
                    solve_seconds=r.solve_seconds, setup_seconds=r.setup_seconds, status=rc,
                    learned_accepted=r.learned_accepted, learned_proposed=r.learned_proposed)

    def ctx(self):
        c = lib().bte_case_ctx(self.h)
        if not c:
            raise BTEError(BTE_ERR_NODEVICE, lib().bte_case_error(self.h).decode())
        return c

    # ---- piecewise entry points on the loaded context ----
    def _check(self, rc, ctx):
        if rc not in (BTE_OK, BTE_ERR_NOCONV):
            raise BTEError(rc, lib().bte_last_error(ctx).decode())
        return rc

    def debug_assemble(self, band, d):
        ctx = self.ctx()
        diag, off = np.zeros(self.N_cell), np.zeros(self.N_cell * self.N_face)
        self._check(lib().bte_debug_assemble(ctx, band, d, _dp(diag), _dp(off)), ctx)
        return diag, off.reshape(self.N_cell, self.N_face)

    def debug_rhs(self, ee_prev, dens):
        ctx = self.ctx()
        out = np.zeros(self.N_band * self.N_dir * self.N_cell)
        ee_prev = np.ascontiguousarray(ee_prev, dtype=np.float64)
        dens = np.ascontiguousarray(dens, dtype=np.float64)
        self._check(lib().bte_debug_rhs(ctx, _dp(ee_prev), _dp(dens), _dp(out)), ctx)
        return out.reshape(self.N_band, self.N_dir, self.N_cell)

    def debug_inner(self, ee_prev, dens, opts: Opts | None = None, **kw):
        ctx = self.ctx()
        if opts is None:
            opts = default_opts(**kw)
        out = np.zeros(self.N_band * self.N_dir * self.N_cell)
        its = np.zeros(self.N_band * self.N_dir, dtype=np.int32)
        ee_prev = np.ascontiguousarray(ee_prev, dtype=np.float64)
        dens = np.ascontiguousarray(dens, dtype=np.float64)
        self._check(lib().bte_debug_inner(ctx, C.byref(opts), _dp(ee_prev), _dp(dens), _dp(out), _ip(its)), ctx)
        return out.reshape(self.N_band, self.N_dir, self.N_cell), its

    # ---- stepping (bte_begin / bte_step / bte_get_temperature / bte_set_temperature / bte_finish) ----
    def begin(self, opts: Opts | None = None, **kw):
        ctx = self.ctx()
        self._run_opts = opts if opts is not None else default_opts(**kw)
        self._check(lib().bte_begin(ctx, C.byref(self._run_opts)), ctx)

    def step(self, n):
        ctx = self.ctx()
        done, conv = C.c_int(), C.c_int()
        self._check(lib().bte_step(ctx, n, C.byref(done), C.byref(conv)), ctx)
        return done.value, bool(conv.value)

    def temperature(self, with_bands=False):
        ctx = self.ctx()
        T = np.zeros(self.N_cell)
        Tb = np.zeros(self.N_cell * self.N_band) if with_bands else None
        self._check(lib().bte_get_temperature(ctx, _dp(T), _dp(Tb) if with_bands else None), ctx)
        return (T, Tb.reshape(self.N_cell, self.N_band)) if with_bands else T

    def set_temperature(self, T):
        ctx = self.ctx()
        T = np.ascontiguousarray(T, dtype=np.float64)
        self._check(lib().bte_set_temperature(ctx, _dp(T)), ctx)

    def finish(self):
        ctx = self.ctx()
        N, B, nbc = self.N_cell, self.N_band, self.n_bc
        cap = max(1, self._run_opts.max_iter)
        T, Tb, fl, bfl, m = np.zeros(N), np.zeros(N * B), np.zeros(max(nbc, 1)), np.zeros(max(nbc, 1) * B), np.zeros(cap)
        r = Result()
        r.T, r.T_band, r.bc_flux, r.bc_band_flux, r.margins = _dp(T), _dp(Tb), _dp(fl), _dp(bfl), _dp(m)
        r.margins_capacity = cap
        rc = self._check(lib().bte_finish(ctx, C.byref(r)), ctx)
        it = r.iterations
        return dict(T=T, T_band=Tb.reshape(N, B), flux=fl[:nbc], band_flux=bfl[:nbc * B].reshape(nbc, B),
                    margins=m[:min(it, cap)].copy(), iters=it, converged=bool(r.converged),
                    inner_iters=int(r.inner_iters), solve_seconds=r.solve_seconds, status=rc,
                    learned_accepted=r.learned_accepted, learned_proposed=r.learned_proposed)

    # ---- learned correction ----
    def attach_gnn(self, gnn):
        """Network used by opts.learned (a staticbtesolver_amd.gnn.Gnn, or None to detach).  The
        caller keeps the Gnn alive while it is attached."""
        ctx = self.ctx()
        self._gnn = gnn
        self._check(lib().bte_attach_gnn(ctx, None if gnn is None else gnn.h), ctx)

    def learn_features(self):
        """(X [N_cell, 16] float32, x_next [N_cell], span) of the current iteration state."""
        ctx = self.ctx()
        X = np.zeros((self.N_cell, 16), dtype=np.float32)
        x = np.zeros(self.N_cell)
        span = C.c_double()
        self._check(lib().bte_learn_features(ctx, X.ctypes.data_as(C.POINTER(C.c_float)), _dp(x), C.byref(span)),
                    ctx)
        return X, x, span.value

    def intensity(self):
        ctx = self.ctx()
        out = np.zeros(self.N_band * self.N_dir * self.N_cell)
        self._check(lib().bte_get_intensity(ctx, _dp(out)), ctx)
        return out.reshape(self.N_band, self.N_dir, self.N_cell)

    def stats(self):
        ctx = self.ctx()
        s = Stats()
        self._check(lib().bte_get_stats(ctx, C.byref(s)), ctx)
        names = [s.names[i].decode() for i in range(s.n_kernels)]
        return dict(ms={n: s.ms[i] for i, n in enumerate(names)}, calls={n: s.calls[i] for i, n in enumerate(names)},
                    spmv_rows=s.spmv_rows, ell_width=s.ell_width, max_levels=s.max_levels,
                    lagged_edges=s.lagged_edges, block=s.block, variant=s.variant,
                    band_group=s.band_group, intensity_layout=s.intensity_layout)


def ctx_set_bands(ctx, v, tau, Cv, Lr):
    """bte_set_bands on a context (the C-ABI setter, without the facade's re-upload)."""
    v, tau, Cv, Lr = [np.ascontiguousarray(a, dtype=np.float64) for a in (v, tau, Cv, Lr)]
    rc = lib().bte_set_bands(ctx, len(v), _dp(v), _dp(tau), _dp(Cv), _dp(Lr))
    if rc != BTE_OK:
        raise BTEError(rc, lib().bte_last_error(ctx).decode())


def ctx_set_bcs(ctx, index, types, temps):
    """bte_set_bcs on a context (index = -k-1 per the reference's convention)."""
    index = np.ascontiguousarray(index, dtype=np.int32)
    types = np.ascontiguousarray(types, dtype=np.int32)
    temps = np.ascontiguousarray(temps, dtype=np.float64)
    rc = lib().bte_set_bcs(ctx, len(index), _ip(index), _ip(types), _dp(temps))
    if rc != BTE_OK:
        raise BTEError(rc, lib().bte_last_error(ctx).decode())


def ctx_solve(ctx, N, B, nbc, opts: Opts | None = None, noalloc=False, **kw):
    """bte_solve (bte_solve_noalloc with noalloc) on a context with the inputs it already holds."""
    if opts is None:
        opts = default_opts(**kw)
    cap = max(1, opts.max_iter)
    T, Tb, fl, bfl, m = np.zeros(N), np.zeros(N * B), np.zeros(max(nbc, 1)), np.zeros(max(nbc, 1) * B), np.zeros(cap)
    r = Result()
    r.T, r.T_band, r.bc_flux, r.bc_band_flux, r.margins = _dp(T), _dp(Tb), _dp(fl), _dp(bfl), _dp(m)
    r.margins_capacity = cap
    fn = lib().bte_solve_noalloc if noalloc else lib().bte_solve
    rc = fn(ctx, C.byref(opts), C.byref(r))
    if rc not in (BTE_OK, BTE_ERR_NOCONV):
        raise BTEError(rc, lib().bte_last_error(ctx).decode())
    it = r.iterations
    return dict(T=T, T_band=Tb.reshape(N, B), flux=fl[:nbc], band_flux=bfl[:nbc * B].reshape(nbc, B),
                margins=m[:min(it, cap)].copy(), iters=it, converged=bool(r.converged),
                inner_iters=int(r.inner_iters), inner_failures=int(r.inner_failures), status=rc)


class Batch:
    """Several loaded cases solved together (bte_batch_*): each outer iteration sweeps the systems of
    all still-iterating cases in one launch.  The cases must stay open while the batch exists."""

    def __init__(self, cases):
        L = lib()
        self.cases = list(cases)
        arr = (C.c_void_p * len(self.cases))(*[c.ctx() for c in self.cases])
        h = C.c_void_p()
        rc = L.bte_batch_create(len(self.cases), arr, C.byref(h))
        if rc != BTE_OK:
            raise BTEError(rc, "bte_batch_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().bte_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def solve(self, opts: Opts | None = None, **kw):
        if opts is None:
            opts = default_opts(**kw)
        cap = max(1, opts.max_iter)
        res = (Result * len(self.cases))()
        bufs = []
        for c, r in zip(self.cases, res):
            N, B, nbc = c.N_cell, c.N_band, c.n_bc
            T, Tb, fl, bfl, m = np.zeros(N), np.zeros(N * B), np.zeros(max(nbc, 1)), np.zeros(max(nbc, 1) * B), \
                np.zeros(cap)
            r.T, r.T_band, r.bc_flux, r.bc_band_flux, r.margins = _dp(T), _dp(Tb), _dp(fl), _dp(bfl), _dp(m)
            r.margins_capacity = cap
            bufs.append((T, Tb, fl, bfl, m))
        rc = lib().bte_batch_solve(self.h, C.byref(opts), res)
        if rc not in (BTE_OK, BTE_ERR_NOCONV):
            raise BTEError(rc, lib().bte_batch_last_error(self.h).decode())
        out = []
        for c, r, (T, Tb, fl, bfl, m) in zip(self.cases, res, bufs):
            nbc, B, it = c.n_bc, c.N_band, r.iterations
            out.append(dict(T=T, T_band=Tb.reshape(c.N_cell, B), flux=fl[:nbc], band_flux=bfl[:nbc * B].reshape(nbc, B),
                            margins=m[:min(it, cap)].copy(), iters=it, converged=bool(r.converged),
                            inner_iters=int(r.inner_iters), solve_seconds=r.solve_seconds,
                            learned_accepted=r.learned_accepted, learned_proposed=r.learned_proposed))
        return out

    def stats(self):
        """hipEvent timing of the last solve run with timing=1: ms and calls of "rhs" and "solve"."""
        st = Stats()
        rc = lib().bte_batch_get_stats(self.h, C.byref(st))
        if rc != BTE_OK:
            raise BTEError(rc, lib().bte_batch_last_error(self.h).decode())
        names = [st.names[i].decode() for i in range(st.n_kernels)]
        return dict(ms={n: st.ms[i] for i, n in enumerate(names)}, calls={n: st.calls[i] for i, n in enumerate(names)})
