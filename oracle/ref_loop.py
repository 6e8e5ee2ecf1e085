"""ctypes wrapper of the oracle's C loop nest -- TEST INFRASTRUCTURE ONLY.

apply_measurement(T, m, mode, x) runs forward_quick_cpu / adjoint_quick_cpu
(operator_born.m:332-596) for one measurement of a BornTables-like object.
"""
import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = HERE / "libborn_ref_loop.so"
        if not path.exists():
            from . import build_ref
            build_ref.build()
        L = C.CDLL(str(path))
        ip = C.POINTER(C.c_int)
        dp = C.POINTER(C.c_double)
        L.ref_apply.argtypes = [C.c_int, C.c_int, C.c_int, dp, ip, C.c_int, C.c_int, ip, dp, C.c_int, ip, ip, ip, ip,
                                dp, dp, dp, C.c_int, C.c_int]
        L.ref_apply.restype = C.c_int
        L.ref_face_sum3.argtypes = [C.c_long, dp, C.c_int, dp, C.c_int, dp, C.c_double, C.c_double, C.c_double, dp]
        L.ref_face_sum3.restype = None
        _LIB = L
    return _LIB


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Prepared:
    """Column-major FP64 buffers of the tables one measurement needs."""

    def __init__(self, T, m, mixes=None):
        ms = T.meas[m]
        self.h = np.ascontiguousarray(np.asarray(T.h_ref, dtype=np.complex128).T)      # f fastest
        self.shift = np.ascontiguousarray((np.asarray(T.shift) - 1).T.astype(np.int32))  # [e][p]
        self.p = np.ascontiguousarray(np.asarray(ms.p_incident, dtype=np.complex128).T)
        self.fg = (np.asarray(ms.idx_f_to_global) - 1).astype(np.int32)
        sel = ms.mixes if mixes is None else [ms.mixes[k] for k in mixes]
        self.Nf = np.array([len(mx.idx_f_to_meas) for mx in sel], dtype=np.int32)
        self.fmeas = np.concatenate([np.asarray(mx.idx_f_to_meas) - 1 for mx in sel]).astype(np.int32)
        self.Na = np.array([len(mx.idx_active) for mx in sel], dtype=np.int32)
        self.act = np.concatenate([np.asarray(mx.idx_active) - 1 for mx in sel]).astype(np.int32)
        self.pref = np.concatenate([np.ascontiguousarray(np.asarray(mx.prefactors, dtype=np.complex128).T).reshape(-1)
                                    for mx in sel])
        self.N_pts = T.shift.shape[0]
        self.N_fu = np.asarray(T.h_ref).shape[0]
        self.N_e = T.shift.shape[1]
        self.n_obs = int(self.Nf.sum())


def apply_measurement(P: Prepared, mode: int, x: np.ndarray, fused: bool = False, threads: int = 0) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    if mode == 1:
        out = np.zeros(P.n_obs, dtype=np.complex128)
    else:
        out = np.zeros(P.N_pts, dtype=np.complex128)
    rc = lib().ref_apply(mode, P.N_pts, P.N_fu, _dp(P.h.view(np.float64)), _ip(P.shift), P.N_e,
                         len(P.fg), _ip(P.fg), _dp(P.p.view(np.float64)), len(P.Nf), _ip(P.Nf), _ip(P.fmeas),
                         _ip(P.Na), _ip(P.act), _dp(P.pref.view(np.float64)), _dp(x.view(np.float64)),
                         _dp(out.view(np.float64)), int(fused), int(threads))
    if rc != 0:
        raise MemoryError("ref_apply failed to allocate")
    return out


def face_sum3(positions: np.ndarray, centres: np.ndarray, offsets: np.ndarray, dA: float, k: complex) -> np.ndarray:
    """(n_pts, n_e) complex: the C face sum ref_face_sum3 for one wavenumber."""
    pos = np.ascontiguousarray(positions, dtype=np.float64)
    ctr = np.ascontiguousarray(centres, dtype=np.float64)
    off = np.ascontiguousarray(offsets, dtype=np.float64)
    out = np.empty((pos.shape[0], ctr.shape[0]), dtype=np.complex128)
    lib().ref_face_sum3(pos.shape[0], _dp(pos), ctr.shape[0], _dp(ctr), off.shape[0], _dp(off), float(dA),
                        float(np.real(k)), float(np.imag(k)), _dp(out.view(np.float64)))
    return out
