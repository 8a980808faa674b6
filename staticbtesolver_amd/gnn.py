"""ctypes binding of the mesh graph network entry points (bte_gnn_* in include/bte.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .capi import BTE_OK, BTEError, lib

NIN = 16
HID = 64

_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    vp, fp = C.c_void_p, C.POINTER(C.c_float)
    L.bte_gnn_create.argtypes = [C.c_int, C.c_int, C.c_ulonglong, C.POINTER(vp)]
    L.bte_gnn_destroy.argtypes = [vp]
    L.bte_gnn_last_error.restype = C.c_char_p
    L.bte_gnn_last_error.argtypes = [vp]
    L.bte_gnn_num_params.argtypes = [vp]
    L.bte_gnn_get_params.argtypes = [vp, fp]
    L.bte_gnn_set_params.argtypes = [vp, fp]
    L.bte_gnn_get_grads.argtypes = [vp, fp]
    L.bte_gnn_set_graph.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int)]
    L.bte_gnn_forward.argtypes = [vp, fp, fp]
    L.bte_gnn_backward.argtypes = [vp, fp, fp, fp, C.POINTER(C.c_double)]
    L.bte_gnn_adam.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]
    _bound = True
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Gnn:
    """The network on one device (parameters, Adam state, graph, activations)."""

    def __init__(self, layers=4, seed=0, device=-1):
        L = _bind()
        h = C.c_void_p()
        rc = L.bte_gnn_create(device, layers, seed, C.byref(h))
        if rc != BTE_OK:
            raise BTEError(rc, "bte_gnn_create failed")
        self.h, self.layers = h, layers
        self.n_params = L.bte_gnn_num_params(h)
        self.n_node = 0

    def _check(self, rc):
        if rc != BTE_OK:
            raise BTEError(rc, lib().bte_gnn_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().bte_gnn_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def get_params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        self._check(lib().bte_gnn_get_params(self.h, _fp(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        self._check(lib().bte_gnn_set_params(self.h, _fp(p)))

    def grads(self):
        g = np.zeros(self.n_params, dtype=np.float32)
        self._check(lib().bte_gnn_get_grads(self.h, _fp(g)))
        return g

    def set_graph(self, nbr):
        nbr = np.ascontiguousarray(nbr, dtype=np.int32)
        self._check(lib().bte_gnn_set_graph(self.h, nbr.shape[0], nbr.shape[1],
                                            nbr.ctypes.data_as(C.POINTER(C.c_int))))
        self.n_node = nbr.shape[0]

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        assert X.shape == (self.n_node, NIN)
        y = np.zeros(self.n_node, dtype=np.float32)
        self._check(lib().bte_gnn_forward(self.h, _fp(X), _fp(y)))
        return y

    def backward(self, X, target, weight=None):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float32)
        loss = C.c_double()
        self._check(lib().bte_gnn_backward(self.h, _fp(X), _fp(t), None if w is None else _fp(w), C.byref(loss)))
        return loss.value

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        self._check(lib().bte_gnn_adam(self.h, lr, beta1, beta2, eps, weight_decay))
