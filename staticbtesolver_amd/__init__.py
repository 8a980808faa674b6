"""MI355X-native steady-state phonon BTE source iteration (drop-in for StaticBTESolver's solve).

Native code: staticbtesolver_amd/lib/libbte_hip.so (HIP kernels + C ABI, include/bte.h) and
staticbtesolver_amd/lib/libStaticBTESolver.so (C++ facade, include/StaticBTESolver/*.h).
"""
from .capi import Case, Opts, BTEError, default_opts, lib  # noqa: F401
from .capi import PREC_FP64, PREC_MIXED, PC_SWEEP, PC_JACOBI  # noqa: F401
