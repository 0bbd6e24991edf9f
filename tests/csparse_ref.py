"""ctypes access to oracle/_ref: the reference's vendored CSparse
(external_packages/CSparse) and its header-only casadi_ldl runtime, both
compiled from the reference sources by oracle/Makefile.  Test infrastructure."""
import ctypes as C
import os

import numpy as np

from oracle_ref import REF_CSPARSE, REF_LDL

AVAILABLE = os.path.exists(REF_CSPARSE) and os.path.exists(REF_LDL)


class cs(C.Structure):
    _fields_ = [("nzmax", C.c_int), ("m", C.c_int), ("n", C.c_int), ("p", C.POINTER(C.c_int)),
                ("i", C.POINTER(C.c_int)), ("x", C.POINTER(C.c_double)), ("nz", C.c_int)]


class css(C.Structure):
    _fields_ = [("pinv", C.POINTER(C.c_int)), ("q", C.POINTER(C.c_int)), ("parent", C.POINTER(C.c_int)),
                ("cp", C.POINTER(C.c_int)), ("leftmost", C.POINTER(C.c_int)), ("m2", C.c_int),
                ("lnz", C.c_double), ("unz", C.c_double)]


class csn(C.Structure):
    _fields_ = [("L", C.POINTER(cs)), ("U", C.POINTER(cs)), ("pinv", C.POINTER(C.c_int)),
                ("B", C.POINTER(C.c_double))]


_lib = None
_ldl = None


def lib():
    global _lib, _ldl
    if _lib is None:
        _lib = C.CDLL(REF_CSPARSE)
        _lib.cs_amd.restype = C.POINTER(C.c_int)
        _lib.cs_amd.argtypes = [C.c_int, C.POINTER(cs)]
        _lib.cs_etree.restype = C.POINTER(C.c_int)
        _lib.cs_etree.argtypes = [C.POINTER(cs), C.c_int]
        _lib.cs_schol.restype = C.POINTER(css)
        _lib.cs_schol.argtypes = [C.c_int, C.POINTER(cs)]
        _lib.cs_chol.restype = C.POINTER(csn)
        _lib.cs_chol.argtypes = [C.POINTER(cs), C.POINTER(css)]
        _lib.cs_sqr.restype = C.POINTER(css)
        _lib.cs_sqr.argtypes = [C.c_int, C.POINTER(cs), C.c_int]
        _lib.cs_qr.restype = C.POINTER(csn)
        _lib.cs_qr.argtypes = [C.POINTER(cs), C.POINTER(css)]
        _lib.cs_free.restype = C.c_void_p
        _lib.cs_free.argtypes = [C.c_void_p]
        _ldl = C.CDLL(REF_LDL)
        PI = C.POINTER(C.c_longlong)
        PD = C.POINTER(C.c_double)
        _ldl.ref_ldl.argtypes = [PI, PD, PI, PD, PD, PI, PD]
        _ldl.ref_ldl_solve.argtypes = [PD, C.c_longlong, PI, PD, PD, PI, PD]
        _ldl.ref_det.restype = C.c_double
        _ldl.ref_det.argtypes = [PI, PD, PI, PD, PD]
    return _lib


class Mat:
    """Keeps the numpy buffers of a cs struct alive."""

    def __init__(self, sp, values=None):
        self.p = np.ascontiguousarray(sp.colind, np.int32)
        self.i = np.ascontiguousarray(sp.row, np.int32)
        self.x = np.ascontiguousarray(values if values is not None else np.ones(sp.nnz()), np.float64)
        self.s = cs(sp.nnz(), sp.nrow, sp.ncol, self.p.ctypes.data_as(C.POINTER(C.c_int)),
                    self.i.ctypes.data_as(C.POINTER(C.c_int)), self.x.ctypes.data_as(C.POINTER(C.c_double)), -1)

    @property
    def ptr(self):
        return C.pointer(self.s)


def _arr(ptr, n, dtype=np.int32):
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(np.int64 if dtype == np.int32 else dtype).copy()


def amd(sp):
    m = Mat(sp)
    P = lib().cs_amd(1, m.ptr)
    out = _arr(P, sp.ncol)
    lib().cs_free(C.cast(P, C.c_void_p))
    return out


def etree(sp, ata=False):
    m = Mat(sp)
    P = lib().cs_etree(m.ptr, int(ata))
    out = _arr(P, sp.ncol)
    lib().cs_free(C.cast(P, C.c_void_p))
    return out


def chol_pattern(sp, values):
    """Pattern of the Cholesky factor L (natural order) of an SPD matrix."""
    m = Mat(sp, values)
    S = lib().cs_schol(0, m.ptr)
    N = lib().cs_chol(m.ptr, S)
    Lm = N.contents.L.contents
    colind = _arr(Lm.p, Lm.n + 1)
    row = _arr(Lm.i, int(colind[-1]))
    return colind, row


def qr_patterns(sp):
    """(V colind, V rows, R colind, R rows, pinv, m2) of CSparse's QR with natural order."""
    m = Mat(sp, np.arange(1, sp.nnz() + 1, dtype=float))
    S = lib().cs_sqr(0, m.ptr, 1)
    N = lib().cs_qr(m.ptr, S)
    V, R = N.contents.L.contents, N.contents.U.contents
    vc = _arr(V.p, V.n + 1)
    vr = _arr(V.i, int(vc[-1]))
    rc = _arr(R.p, R.n + 1)
    rr = _arr(R.i, int(rc[-1]))
    m2 = S.contents.m2
    pinv = _arr(S.contents.pinv, m2)
    return vc, vr, rc, rr, pinv, m2


def ref_ldl(sp_a, a, sp_lt, p):
    lib()
    sp_a = np.ascontiguousarray(sp_a, np.int64)
    sp_lt = np.ascontiguousarray(sp_lt, np.int64)
    p = np.ascontiguousarray(p, np.int64)
    n = int(sp_lt[1])
    lt = np.zeros(int(sp_lt[2 + n]))
    d = np.zeros(n)
    w = np.zeros(n)
    PI = C.POINTER(C.c_longlong)
    PD = C.POINTER(C.c_double)
    a = np.ascontiguousarray(a, float)
    _ldl.ref_ldl(sp_a.ctypes.data_as(PI), a.ctypes.data_as(PD), sp_lt.ctypes.data_as(PI), lt.ctypes.data_as(PD),
                 d.ctypes.data_as(PD), p.ctypes.data_as(PI), w.ctypes.data_as(PD))
    return lt, d


def ref_ldl_solve(b, sp_lt, lt, d, p):
    lib()
    x = np.array(b, dtype=float)
    sp_lt = np.ascontiguousarray(sp_lt, np.int64)
    p = np.ascontiguousarray(p, np.int64)
    n = int(sp_lt[1])
    w = np.zeros(n)
    PI = C.POINTER(C.c_longlong)
    PD = C.POINTER(C.c_double)
    _ldl.ref_ldl_solve(x.ctypes.data_as(PD), x.size // n, sp_lt.ctypes.data_as(PI), lt.ctypes.data_as(PD),
                       d.ctypes.data_as(PD), p.ctypes.data_as(PI), w.ctypes.data_as(PD))
    return x


def ref_det(sp_v, v, sp_r, r, beta):
    """The reference's casadi_det (casadi/core/runtime/casadi_det.hpp), without
    the permutation sign LinsolQr::det multiplies in."""
    lib()
    PI = C.POINTER(C.c_longlong)
    PD = C.POINTER(C.c_double)
    sp_v, sp_r = np.ascontiguousarray(sp_v, np.int64), np.ascontiguousarray(sp_r, np.int64)
    v, r, beta = (np.ascontiguousarray(x, float) for x in (v, r, beta))
    return _ldl.ref_det(sp_v.ctypes.data_as(PI), v.ctypes.data_as(PD), sp_r.ctypes.data_as(PI), r.ctypes.data_as(PD),
                        beta.ctypes.data_as(PD))
