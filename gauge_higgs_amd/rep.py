"""Representation query (host-side construction inside libghx.so; no device needed)."""
import ctypes as C

import numpy as np

from . import _lib


def rep_query(N, spec):
    """Return dict(d, nality, real, dimT, W, T) for SU(N) rep `spec` (fund|adj|rows[:real])."""
    L = _lib.load()
    d, nal, real, dimT = C.c_int(), C.c_int(), C.c_int(), C.c_long()
    rc = L.ghx_rep_query(N, spec.encode(), C.byref(d), C.byref(nal), C.byref(real), C.byref(dimT), None, None)
    if rc != 0:
        raise ValueError(f"invalid representation '{spec}' for SU({N}) ({rc})")
    W = np.zeros(d.value * dimT.value)
    T = np.zeros((N * N - 1) * d.value * d.value * 2)
    dp = C.POINTER(C.c_double)
    L.ghx_rep_query(N, spec.encode(), None, None, None, None, W.ctypes.data_as(dp), T.ctypes.data_as(dp))
    return dict(d=d.value, nality=nal.value, real=bool(real.value), dimT=dimT.value,
                W=W.reshape(d.value, dimT.value), T=T.view(np.complex128).reshape(N * N - 1, d.value, d.value))
