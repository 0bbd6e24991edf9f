"""ctypes access to the C oracle (oracle/liboracle.so) and to oracle/_ref.

Test infrastructure only: the oracle is the checker, never the product.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF_LDL = os.path.join(ROOT, "oracle", "_ref", "libref_ldl.so")
REF_CSPARSE = os.path.join(ROOT, "oracle", "_ref", "libref_csparse.so")

P = C.c_void_p
I64 = C.c_int64
PD = C.POINTER(C.c_double)
PI64 = C.POINTER(C.c_int64)


def _load():
    if not os.path.exists(ORACLE):
        import subprocess
        subprocess.run(["make", "-C", os.path.dirname(ORACLE), "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    lib = C.CDLL(ORACLE)
    lib.or_fun.restype = C.c_double
    lib.or_fun.argtypes = [C.c_int, C.c_double, C.c_double]
    lib.or_der.restype = None
    lib.or_der.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, PD]
    lib.or_eval.restype = C.c_int
    lib.or_eval.argtypes = [P, I64, P, P, PD]
    lib.or_map_eval.restype = C.c_int
    lib.or_map_eval.argtypes = [P, I64, I64, C.c_int, PI64, C.c_int, PI64, I64, P, P, C.c_int]
    lib.or_ldl.restype = None
    lib.or_ldl.argtypes = [PI64, PD, PI64, PD, PD, PI64, PD]
    lib.or_ldl_solve.restype = None
    lib.or_ldl_solve.argtypes = [PD, I64, PI64, PD, PD, PI64, PD]
    lib.or_qr.restype = None
    lib.or_qr.argtypes = [PI64, PD, PD, PI64, PD, PI64, PD, PD, PI64, PI64]
    lib.or_qr_solve.restype = None
    lib.or_qr_solve.argtypes = [PD, I64, C.c_int, PI64, PD, PI64, PD, PD, PI64, PI64, PD]
    lib.or_qr_singular.restype = I64
    lib.or_qr_singular.argtypes = [PD, PI64, PD, PI64, PI64, C.c_double]
    lib.or_set_perturb.restype = None
    lib.or_set_perturb.argtypes = [C.c_int]
    lib.or_adj_map.restype = C.c_int
    lib.or_adj_map.argtypes = [P, I64, I64, C.c_int, PI64, C.c_int, PI64, I64, P, P, P, P, C.c_int]
    lib.or_eval_calls.restype = C.c_int
    lib.or_eval_calls.argtypes = [P, I64, P, P, PD, P, PI64, PI64, P, P, P, P, P, P]
    lib.or_sp_forward.restype = C.c_int
    lib.or_sp_forward.argtypes = [P, I64, P, P, P]
    lib.or_sp_reverse.restype = C.c_int
    lib.or_sp_reverse.argtypes = [P, I64, I64, P, P, P]
    lib.or_qr_det.restype = C.c_double
    lib.or_qr_det.argtypes = [PI64, PD, PI64, PD, PD, PI64, PI64]
    lib.or_ldl_batch.restype = I64
    lib.or_ldl_batch.argtypes = [PI64, PI64, PI64, I64, PD, PD, PD, C.c_int]
    return lib


lib = _load()


def _pd(a):
    return a.ctypes.data_as(PD)


def _pi(a):
    return a.ctypes.data_as(PI64)


def _ptrs(arrs):
    out = (C.c_void_p * max(len(arrs), 1))()
    for i, a in enumerate(arrs):
        out[i] = a.ctypes.data if a is not None and a.size else None
    return out


def fun(op, x, y=0.0):
    return lib.or_fun(op, x, y)


def der(op, x, y, f):
    d = np.zeros(2)
    lib.or_der(op, x, y, f, _pd(d))
    return d


def map_eval(f, n, args_aos, n_threads=1):
    """Evaluate Function f over n instances on the oracle VM (Map serial / OpenMP).
    args_aos[i]: (n, nnz_in(i)) array or None; returns list of (n, nnz_out(i))."""
    instr = f.instructions()
    nnz_in = np.array([f.nnz_in(i) for i in range(f.n_in())], np.int64)
    nnz_out = np.array([f.nnz_out(i) for i in range(f.n_out())], np.int64)
    arg = [np.ascontiguousarray(a, dtype=float) if a is not None else None for a in args_aos]
    res = [np.zeros((n, k)) for k in nnz_out]
    flag = lib.or_map_eval(instr.ctypes.data, instr.size, f.sz_w(), len(nnz_in), _pi(nnz_in),
                           len(nnz_out), _pi(nnz_out), n, _ptrs(arg), _ptrs(res), n_threads)
    assert flag == 0, "oracle VM rejected the tape"
    return res


def eval_calls(instr, ws, nnz_in, nnz_out, calls, args):
    """A tape with OP_CALL records on the oracle, instance by instance.
    calls[k] = (callee Function, argument slots, result slots); args[i]:
    (n, nnz_in[i]) arrays.  Returns list of (n, nnz_out[j])."""
    n = args[0].shape[0]
    cf = [c[0] for c in calls]
    ci = [f.instructions() for f in cf]
    c_n = np.array([a.size for a in ci], np.int64)
    c_ws = np.array([f.sz_w() for f in cf], np.int64)
    c_nin = (C.c_int * max(len(cf), 1))(*[f.n_in() for f in cf])
    c_nout = (C.c_int * max(len(cf), 1))(*[f.n_out() for f in cf])
    nin = [np.array([f.nnz_in(j) for j in range(f.n_in())], np.int64) for f in cf]
    nout = [np.array([f.nnz_out(j) for j in range(f.n_out())], np.int64) for f in cf]
    dep = [np.ascontiguousarray(c[1], np.int32) for c in calls]
    res_slot = [np.ascontiguousarray(c[2], np.int32) for c in calls]
    instr = np.ascontiguousarray(instr)
    out = [np.zeros((n, k)) for k in nnz_out]
    w = np.zeros(max(int(ws), 1))
    for k in range(n):
        a = [np.ascontiguousarray(x[k], dtype=float) for x in args]
        r = [np.zeros(m) for m in nnz_out]
        flag = lib.or_eval_calls(instr.ctypes.data, instr.size, _ptrs(a), _ptrs(r), w.ctypes.data_as(PD),
                                 _ptrs(ci), c_n.ctypes.data_as(PI64), c_ws.ctypes.data_as(PI64), c_nin,
                                 _ptrs(nin), c_nout, _ptrs(nout), _ptrs(dep), _ptrs(res_slot))
        assert flag == 0, "oracle rejected the tape"
        for j in range(len(nnz_out)):
            out[j][k] = r[j]
    return out


def sp_forward(f, seeds):
    """SXFunction::sp_forward on the oracle, word by word.  seeds[i]:
    (nnz_in(i), n_words) uint64 or None; returns (nnz_out(j), n_words) uint64."""
    instr = f.instructions()
    nw = next(s.shape[1] for s in seeds if s is not None)
    out = [np.zeros((f.nnz_out(j), nw), np.uint64) for j in range(f.n_out())]
    w = np.zeros(max(f.sz_w(), 1), np.uint64)
    for k in range(nw):
        a = [None if s is None else np.ascontiguousarray(s[:, k]) for s in seeds]
        r = [np.zeros(f.nnz_out(j), np.uint64) for j in range(f.n_out())]
        assert lib.or_sp_forward(instr.ctypes.data, instr.size, _ptrs(a), _ptrs(r), w.ctypes.data) == 0
        for j in range(f.n_out()):
            out[j][:, k] = r[j]
    return out


def sp_reverse(f, seeds):
    """SXFunction::sp_reverse on the oracle, word by word.  seeds[j]:
    (nnz_out(j), n_words) uint64 or None; returns (nnz_in(i), n_words) uint64."""
    instr = f.instructions()
    nw = next(s.shape[1] for s in seeds if s is not None)
    out = [np.zeros((f.nnz_in(i), nw), np.uint64) for i in range(f.n_in())]
    w = np.zeros(max(f.sz_w(), 1), np.uint64)
    for k in range(nw):
        r = [None if s is None else np.ascontiguousarray(s[:, k]) for s in seeds]
        a = [np.zeros(f.nnz_in(i), np.uint64) for i in range(f.n_in())]
        assert lib.or_sp_reverse(instr.ctypes.data, instr.size, f.sz_w(), _ptrs(a), _ptrs(r), w.ctypes.data) == 0
        assert all(x is None or not x.any() for x in r)   # seeds are consumed
        for i in range(f.n_in()):
            out[i][:, k] = a[i]
    return out


def adj_map(f, n, args_aos, seeds_aos, n_threads=1):
    """Numeric reverse sweep (ad_reverse rules on numbers) over n instances.
    Returns (res, asens) lists of (n, nnz) arrays."""
    instr = f.instructions()
    nnz_in = np.array([f.nnz_in(i) for i in range(f.n_in())], np.int64)
    nnz_out = np.array([f.nnz_out(i) for i in range(f.n_out())], np.int64)
    arg = [np.ascontiguousarray(a, dtype=float) if a is not None else None for a in args_aos]
    seed = [np.ascontiguousarray(a, dtype=float) if a is not None else None for a in seeds_aos]
    res = [np.zeros((n, k)) for k in nnz_out]
    sens = [np.zeros((n, k)) for k in nnz_in]
    flag = lib.or_adj_map(instr.ctypes.data, instr.size, f.sz_w(), len(nnz_in), _pi(nnz_in), len(nnz_out),
                          _pi(nnz_out), n, _ptrs(arg), _ptrs(seed), _ptrs(res), _ptrs(sens), n_threads)
    assert flag == 0, "oracle reverse sweep rejected the tape"
    return res, sens


def mapsum_eval(f, n, args, reduce_in, reduce_out):
    """MapSum::eval_gen (casadi/core/mapsum.cpp:155-186): a reduced input is
    one block read by every instance, a reduced output is the sum over the
    instances, added instance by instance in order (casadi_add).
    args[i]: (nnz,) for reduced inputs, (n, nnz) otherwise; returns (nnz,) for
    reduced outputs, (n, nnz) otherwise."""
    aos = [None if a is None else
           np.tile(np.asarray(a, float).reshape(1, -1), (n, 1)) if reduce_in[i] else a
           for i, a in enumerate(args)]
    res = map_eval(f, n, aos)
    return [_serial_sum(r) if reduce_out[j] else r for j, r in enumerate(res)]


def mapsum_adj(f, n, args, seeds, reduce_in, reduce_out):
    """The reverse of a MapSum (MapSum::get_reverse,
    casadi/core/mapsum.cpp:354-415: the MapSum of f.reverse(1) whose seeds of
    reduced outputs are shared and whose sensitivities of reduced inputs are
    summed).  Returns (outputs, sensitivities) shaped as in mapsum_eval."""
    aos = [None if a is None else
           np.tile(np.asarray(a, float).reshape(1, -1), (n, 1)) if reduce_in[i] else a
           for i, a in enumerate(args)]
    sd = [None if a is None else
          np.tile(np.asarray(a, float).reshape(1, -1), (n, 1)) if reduce_out[j] else a
          for j, a in enumerate(seeds)]
    res, sens = adj_map(f, n, aos, sd)
    return ([_serial_sum(r) if reduce_out[j] else r for j, r in enumerate(res)],
            [_serial_sum(g) if reduce_in[i] else g for i, g in enumerate(sens)])


def _serial_sum(rows):
    """sum_k rows[k], accumulated from zero in instance order."""
    acc = np.zeros(rows.shape[1])
    for k in range(rows.shape[0]):
        acc = acc + rows[k]
    return acc


def ldl(sp_a, a, sp_lt, p):
    sp_a = np.ascontiguousarray(sp_a, np.int64)
    sp_lt = np.ascontiguousarray(sp_lt, np.int64)
    p = np.ascontiguousarray(p, np.int64)
    n = int(sp_lt[1])
    lt = np.zeros(int(sp_lt[2 + n]))
    d = np.zeros(n)
    w = np.zeros(n)
    lib.or_ldl(_pi(sp_a), _pd(np.ascontiguousarray(a, float)), _pi(sp_lt), _pd(lt), _pd(d), _pi(p), _pd(w))
    return lt, d


def ldl_solve(x, sp_lt, lt, d, p):
    x = np.array(x, dtype=float)
    n = int(sp_lt[1])
    w = np.zeros(n)
    lib.or_ldl_solve(_pd(x), x.size // n, _pi(np.ascontiguousarray(sp_lt, np.int64)), _pd(lt), _pd(d),
                     _pi(np.ascontiguousarray(p, np.int64)), _pd(w))
    return x


def qr(sp_a, a, sp_v, sp_r, prinv, pc):
    sp_a, sp_v, sp_r = (np.ascontiguousarray(s, np.int64) for s in (sp_a, sp_v, sp_r))
    prinv, pc = np.ascontiguousarray(prinv, np.int64), np.ascontiguousarray(pc, np.int64)
    ncol = int(sp_a[1])
    nrow_ext = int(sp_v[0])
    v = np.zeros(int(sp_v[2 + ncol]))
    r = np.zeros(int(sp_r[2 + ncol]))
    beta = np.zeros(ncol)
    x = np.zeros(nrow_ext)
    lib.or_qr(_pi(sp_a), _pd(np.ascontiguousarray(a, float)), _pd(x), _pi(sp_v), _pd(v), _pi(sp_r), _pd(r),
              _pd(beta), _pi(prinv), _pi(pc))
    return v, r, beta


def qr_solve(x, tr, sp_v, v, sp_r, r, beta, prinv, pc):
    x = np.array(x, dtype=float)
    sp_v = np.ascontiguousarray(sp_v, np.int64)
    w = np.zeros(max(int(sp_v[0]), int(sp_v[1])))
    ncol = int(sp_v[1])
    lib.or_qr_solve(_pd(x), x.size // ncol, int(tr), _pi(sp_v), _pd(v), _pi(np.ascontiguousarray(sp_r, np.int64)),
                    _pd(r), _pd(beta), _pi(np.ascontiguousarray(prinv, np.int64)),
                    _pi(np.ascontiguousarray(pc, np.int64)), _pd(w))
    return x


def qr_singular(r, sp_r, pc, eps):
    rmin = C.c_double()
    irmin = C.c_int64()
    nul = lib.or_qr_singular(C.byref(rmin), C.byref(irmin), _pd(np.ascontiguousarray(r, float)),
                             _pi(np.ascontiguousarray(sp_r, np.int64)), _pi(np.ascontiguousarray(pc, np.int64)),
                             eps)
    return nul, rmin.value, irmin.value


def qr_det(sp_v, v, sp_r, r, beta, prinv, pc):
    return lib.or_qr_det(_pi(np.ascontiguousarray(sp_v, np.int64)), _pd(np.ascontiguousarray(v, float)),
                         _pi(np.ascontiguousarray(sp_r, np.int64)), _pd(np.ascontiguousarray(r, float)),
                         _pd(np.ascontiguousarray(beta, float)), _pi(np.ascontiguousarray(prinv, np.int64)),
                         _pi(np.ascontiguousarray(pc, np.int64)))


class perturbed_transcendentals:
    """Context manager: every transcendental result of the oracle moves by one
    ulp (alternating direction) -- the rounding envelope of a workload."""

    def __enter__(self):
        lib.or_set_perturb(1)
        return self

    def __exit__(self, *exc):
        lib.or_set_perturb(0)
        return False


def rel_err(y, yref):
    """max |y - y_ref| / max(|y_ref|, 1) -- the 1e-13 acceptance metric."""
    y = np.asarray(y, float)
    yref = np.asarray(yref, float)
    return float(np.max(np.abs(y - yref) / np.maximum(np.abs(yref), 1.0))) if y.size else 0.0
