"""Sparsity and SX: CasADi's pattern and symbolic-matrix types, Python side.

``SX`` mirrors ``casadi.SX`` (casadi/core/sx_instantiator.cpp,
casadi/core/generic_matrix.hpp): a compressed-column pattern plus one scalar
expression node per nonzero.  Scalar nodes, their simplification and all
graph algorithms live in the native library (casadi_amd/csrc/sx_node.cpp);
this module only arranges nonzeros.
"""
import numbers

import numpy as np

from . import _lib as L
from .ops import OP

__all__ = ["Sparsity", "SX", "vertcat", "horzcat", "veccat", "jacobian", "gradient",
           "hessian", "if_else", "sumsqr", "dot", "sum1", "sum2", "mtimes", "fmin", "fmax",
           "atan2", "sqrt", "sin", "cos", "tan", "exp", "log", "sq", "fabs", "tanh", "sinh",
           "cosh", "asin", "acos", "atan", "erf", "erfinv", "floor", "ceil", "sign", "power",
           "constpow", "logic_and", "logic_or", "logic_not", "if_else_zero", "log1p", "expm1",
           "hypot", "copysign", "fmod", "asinh", "acosh", "atanh", "cross", "diag", "repmat",
           "densify", "norm_2"]


class Sparsity:
    """Compressed-column sparsity pattern (casadi::Sparsity)."""

    def __init__(self, nrow, ncol, colind=None, row=None):
        self.nrow = int(nrow)
        self.ncol = int(ncol)
        self.colind = np.zeros(self.ncol + 1, np.int64) if colind is None else np.asarray(colind, np.int64)
        self.row = np.zeros(0, np.int64) if row is None else np.asarray(row, np.int64)

    # -- constructors
    @staticmethod
    def dense(n, m=1):
        colind = np.arange(m + 1, dtype=np.int64) * n
        row = np.tile(np.arange(n, dtype=np.int64), m)
        return Sparsity(n, m, colind, row)

    @staticmethod
    def diag(n):
        return Sparsity(n, n, np.arange(n + 1), np.arange(n))

    @staticmethod
    def triplet(nrow, ncol, r, c):
        r = np.asarray(r, np.int64)
        c = np.asarray(c, np.int64)
        key = np.unique(c * max(nrow, 1) + r)
        cc, rr = np.divmod(key, max(nrow, 1)) if nrow else (key, key)
        colind = np.zeros(ncol + 1, np.int64)
        np.add.at(colind, cc + 1, 1)
        return Sparsity(nrow, ncol, np.cumsum(colind), rr)

    @staticmethod
    def from_compressed(sp):
        sp = np.asarray(sp, np.int64)
        nr, nc = int(sp[0]), int(sp[1])
        colind = sp[2:3 + nc]
        row = sp[3 + nc:3 + nc + colind[-1]]
        return Sparsity(nr, nc, colind.copy(), row.copy())

    # -- queries
    def nnz(self):
        return int(self.colind[-1])

    def numel(self):
        return self.nrow * self.ncol

    def size1(self):
        return self.nrow

    def size2(self):
        return self.ncol

    @property
    def shape(self):
        return (self.nrow, self.ncol)

    def is_dense(self):
        return self.nnz() == self.numel()

    def is_scalar(self):
        return self.nrow == 1 and self.ncol == 1

    def is_symmetric(self):
        return self.nrow == self.ncol and self == self.T

    def get_col(self):
        return np.repeat(np.arange(self.ncol, dtype=np.int64), np.diff(self.colind))

    def find(self):
        """Linear (column-major) indices of the nonzeros."""
        return self.row + self.get_col() * self.nrow

    def compressed(self):
        return np.concatenate([[self.nrow, self.ncol], self.colind, self.row]).astype(np.int64)

    @property
    def T(self):
        return Sparsity.triplet(self.ncol, self.nrow, self.get_col(), self.row)

    def __eq__(self, other):
        return (isinstance(other, Sparsity) and self.shape == other.shape
                and np.array_equal(self.colind, other.colind) and np.array_equal(self.row, other.row))

    def __ne__(self, other):
        return not self == other

    def __hash__(self):
        return hash((self.shape, self.colind.tobytes(), self.row.tobytes()))

    def __repr__(self):
        return f"Sparsity({self.nrow}x{self.ncol},{self.nnz()}nz)"

    def dense_mask(self):
        m = np.zeros((self.nrow, self.ncol), bool)
        m[self.row, self.get_col()] = True
        return m

    # -- graph algorithms (native, casadi/core/sparsity_internal.cpp)
    def _call(self, fn, *extra):
        sp = self.compressed()
        h = L.P()
        L.check(fn(sp.ctypes.data_as(L.PI64), *extra, L.C.byref(h)))
        return L.take_buf(h)

    def star_coloring(self, ordering=1):
        return Sparsity.from_compressed(self._call(L.cxs_sp_star_coloring, ordering))

    def uni_coloring(self, cutoff=0):
        """Colour the columns of this pattern (structurally orthogonal groups)."""
        out = self._call(L.cxs_sp_uni_coloring, cutoff)
        return Sparsity.from_compressed(out) if out.size else None

    def largest_first(self):
        return self._call(L.cxs_sp_largest_first)

    def amd(self):
        return self._call(L.cxs_sp_amd)

    def etree(self, ata=False):
        return self._call(L.cxs_sp_etree, int(ata))

    def ldl(self, amd=True):
        """Pattern of L^T (strictly upper) and the permutation p (Sparsity::ldl)."""
        sp = self.compressed()
        h1, h2 = L.P(), L.P()
        L.check(L.cxs_sp_ldl(sp.ctypes.data_as(L.PI64), int(amd), L.C.byref(h1), L.C.byref(h2)))
        return Sparsity.from_compressed(L.take_buf(h1)), L.take_buf(h2)

    def qr_sparse(self, amd=True):
        """(V, R, prinv, pc) of the Householder QR symbolic factorisation."""
        sp = self.compressed()
        hs = [L.P() for _ in range(4)]
        L.check(L.cxs_sp_qr(sp.ctypes.data_as(L.PI64), int(amd), *[L.C.byref(h) for h in hs]))
        v, r, prinv, pc = [L.take_buf(h) for h in hs]
        return Sparsity.from_compressed(v), Sparsity.from_compressed(r), prinv, pc

    def mtimes(self, other):
        a, b = self.compressed(), other.compressed()
        h = L.P()
        L.check(L.cxs_sp_mtimes(a.ctypes.data_as(L.PI64), b.ctypes.data_as(L.PI64), L.C.byref(h)))
        return Sparsity.from_compressed(L.take_buf(h))


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def _const_node(v):
    return int(L.cxs_const(float(v)))


class SX:
    """Sparse matrix of scalar symbolic expressions (casadi::SX)."""

    __array_priority__ = 100

    def __init__(self, sp=None, nz=None):
        if sp is None:
            sp = Sparsity(1, 1)
            nz = np.zeros(0, np.uint64)
        if isinstance(sp, numbers.Number):
            v = sp
            sp = Sparsity.dense(1, 1)
            nz = np.array([_const_node(v)], np.uint64)
        self.sparsity = sp
        self.nz = _u64(nz if nz is not None else np.full(sp.nnz(), _const_node(0), np.uint64))
        assert self.nz.size == sp.nnz()

    # -- constructors
    @staticmethod
    def sym(name, n=1, m=1):
        sp = n if isinstance(n, Sparsity) else Sparsity.dense(n, m)
        k = sp.nnz()
        if k == 1:
            nodes = [L.cxs_sym(name.encode())]
        else:
            nodes = [L.cxs_sym(f"{name}_{i}".encode()) for i in range(k)]
        return SX(sp, np.array(nodes, np.uint64))

    @staticmethod
    def zeros(n=1, m=1):
        sp = n if isinstance(n, Sparsity) else Sparsity.dense(n, m)
        return SX(sp, np.full(sp.nnz(), _const_node(0), np.uint64))

    @staticmethod
    def ones(n=1, m=1):
        sp = n if isinstance(n, Sparsity) else Sparsity.dense(n, m)
        return SX(sp, np.full(sp.nnz(), _const_node(1), np.uint64))

    @staticmethod
    def from_value(v):
        if isinstance(v, SX):
            return v
        a = np.atleast_2d(np.asarray(v, dtype=float))
        if np.asarray(v).ndim == 1:
            a = a.T
        sp = Sparsity.dense(*a.shape)
        nodes = np.array([_const_node(x) for x in a.flatten(order="F")], np.uint64)
        return SX(sp, nodes)

    # -- shape
    @property
    def shape(self):
        return self.sparsity.shape

    def size1(self):
        return self.sparsity.nrow

    def size2(self):
        return self.sparsity.ncol

    def numel(self):
        return self.sparsity.numel()

    def nnz(self):
        return self.sparsity.nnz()

    def is_scalar(self):
        return self.shape == (1, 1)

    def __len__(self):
        return self.size1()

    def nonzeros(self):
        return list(self.nz)

    def densify(self):
        if self.sparsity.is_dense():
            return self
        full = np.full(self.numel(), _const_node(0), np.uint64)
        full[self.sparsity.find()] = self.nz
        return SX(Sparsity.dense(*self.shape), full)

    def dense_nodes(self):
        """(nrow, ncol) object array of node handles (structural zeros as 0)."""
        d = self.densify()
        return d.nz.reshape(self.shape[::-1]).T

    @staticmethod
    def from_dense_nodes(a):
        a = np.asarray(a, np.uint64)
        return SX(Sparsity.dense(*a.shape), a.T.reshape(-1))

    @property
    def T(self):
        d = self.dense_nodes()
        return SX.from_dense_nodes(d.T)

    # -- elementwise arithmetic
    def _binary(self, op, other, swap=False):
        other = other if isinstance(other, SX) else SX.from_value(other)
        a, b = (other, self) if swap else (self, other)
        if a.is_scalar() and not b.is_scalar():
            b = b.densify() if b.sparsity.is_dense() else b.densify()
            out = np.empty(b.nnz(), np.uint64)
            L.check(L.cxs_binary_n(op, b.nnz(), a.densify().nz.ctypes.data_as(L.PU64), 0,
                                   b.nz.ctypes.data_as(L.PU64), 1, out.ctypes.data_as(L.PU64)))
            return SX(b.sparsity, out)
        if b.is_scalar() and not a.is_scalar():
            a = a.densify()
            out = np.empty(a.nnz(), np.uint64)
            L.check(L.cxs_binary_n(op, a.nnz(), a.nz.ctypes.data_as(L.PU64), 1,
                                   b.densify().nz.ctypes.data_as(L.PU64), 0, out.ctypes.data_as(L.PU64)))
            return SX(a.sparsity, out)
        if a.shape != b.shape:
            raise ValueError(f"Dimension mismatch: {a.shape} vs {b.shape}")
        if a.sparsity != b.sparsity:
            a, b = a.densify(), b.densify()
        out = np.empty(a.nnz(), np.uint64)
        L.check(L.cxs_binary_n(op, a.nnz(), a.nz.ctypes.data_as(L.PU64), 1, b.nz.ctypes.data_as(L.PU64), 1,
                               out.ctypes.data_as(L.PU64)))
        return SX(a.sparsity, out)

    def _unary(self, op):
        out = np.empty(self.nnz(), np.uint64)
        L.check(L.cxs_unary_n(op, self.nnz(), self.nz.ctypes.data_as(L.PU64), out.ctypes.data_as(L.PU64)))
        return SX(self.sparsity, out)

    def __add__(self, o): return self._binary(OP.ADD, o)
    def __radd__(self, o): return self._binary(OP.ADD, o, True)
    def __sub__(self, o): return self._binary(OP.SUB, o)
    def __rsub__(self, o): return self._binary(OP.SUB, o, True)
    def __mul__(self, o): return self._binary(OP.MUL, o)
    def __rmul__(self, o): return self._binary(OP.MUL, o, True)
    def __truediv__(self, o): return self._binary(OP.DIV, o)
    def __rtruediv__(self, o): return self._binary(OP.DIV, o, True)
    def __pow__(self, o): return self._binary(OP.POW, o)
    def __rpow__(self, o): return self._binary(OP.POW, o, True)
    def __lt__(self, o): return self._binary(OP.LT, o)
    def __le__(self, o): return self._binary(OP.LE, o)
    def __gt__(self, o): return self._binary(OP.LT, o, True)
    def __ge__(self, o): return self._binary(OP.LE, o, True)
    def __neg__(self):
        out = np.array([L.cxs_neg(int(x)) for x in self.nz], np.uint64)
        return SX(self.sparsity, out)
    def __pos__(self): return self
    def __matmul__(self, o): return mtimes(self, o)
    def __rmatmul__(self, o): return mtimes(SX.from_value(o), self)

    def eq(self, o): return self._binary(OP.EQ, o)
    def ne(self, o): return self._binary(OP.NE, o)

    # -- indexing (dense semantics)
    def __getitem__(self, key):
        d = self.dense_nodes()
        if not isinstance(key, tuple):
            if self.size2() == 1:
                key = (key, slice(None))
            elif self.size1() == 1:
                key = (slice(None), key)
            else:
                flat = d.T.reshape(-1)[key]
                flat = np.atleast_1d(flat)
                return SX.from_dense_nodes(flat.reshape(-1, 1))
        r, c = key
        sub = d[r, c]
        if np.ndim(sub) == 0:
            return SX.from_dense_nodes(np.array([[sub]], np.uint64))
        if np.ndim(sub) == 1:
            sub = sub.reshape(-1, 1) if not isinstance(r, numbers.Integral) else sub.reshape(1, -1)
        return SX.from_dense_nodes(sub)

    def __setitem__(self, key, value):
        d = self.dense_nodes().copy()
        v = value if isinstance(value, SX) else SX.from_value(value)
        vd = v.dense_nodes()
        if not isinstance(key, tuple):
            key = (key, slice(None)) if self.size2() == 1 else (slice(None), key)
        tgt = d[key]
        d[key] = vd.reshape(np.shape(tgt)) if np.ndim(tgt) else vd.reshape(-1)[0]
        new = SX.from_dense_nodes(d)
        self.sparsity, self.nz = new.sparsity, new.nz

    def __iter__(self):
        for i in range(self.numel()):
            yield self[i]

    # -- scalar node inspection
    def is_constant(self):
        return all(_node_op(x) == OP.CONST for x in self.nz)

    def to_double(self):
        """Value of a constant expression (dense numpy array)."""
        vals = np.zeros(self.numel())
        for k, x in zip(self.sparsity.find(), self.nz):
            op, v = _node_op(x, value=True)
            if op != OP.CONST:
                raise ValueError("expression is not constant")
            vals[k] = v
        return vals.reshape(self.shape[::-1]).T

    def __float__(self):
        return float(self.to_double().reshape(-1)[0])

    def __repr__(self):
        return f"SX({self.size1()}x{self.size2()}, {self.nnz()} nz)"


def _node_op(h, value=False):
    op = L.I32()
    v = L.D()
    L.check(L.cxs_node_info(int(h), L.C.byref(op), L.C.byref(v), None, None))
    return (op.value, v.value) if value else op.value


def _as_sx(x):
    return x if isinstance(x, SX) else SX.from_value(x)


# -- concatenation ---------------------------------------------------------
def vertcat(*xs):
    xs = [_as_sx(x) for x in xs if not (isinstance(x, SX) and x.numel() == 0)]
    if not xs:
        return SX(Sparsity(0, 1), np.zeros(0, np.uint64))
    return SX.from_dense_nodes(np.vstack([x.dense_nodes() for x in xs]))


def horzcat(*xs):
    xs = [_as_sx(x) for x in xs if not (isinstance(x, SX) and x.numel() == 0)]
    if not xs:
        return SX(Sparsity(1, 0), np.zeros(0, np.uint64))
    return SX.from_dense_nodes(np.hstack([x.dense_nodes() for x in xs]))


def veccat(*xs):
    return vertcat(*[SX.from_dense_nodes(_as_sx(x).dense_nodes().T.reshape(-1, 1)) for x in xs])


def repmat(x, n, m=1):
    return SX.from_dense_nodes(np.tile(_as_sx(x).dense_nodes(), (n, m)))


def diag(x):
    x = _as_sx(x)
    d = x.dense_nodes()
    if x.size2() == 1:
        n = x.size1()
        out = np.full((n, n), _const_node(0), np.uint64)
        out[np.arange(n), np.arange(n)] = d[:, 0]
        return SX.from_dense_nodes(out)
    return SX.from_dense_nodes(np.diag(d).reshape(-1, 1))


def densify(x):
    return _as_sx(x).densify()


# -- elementwise functions --------------------------------------------------
def _un(op):
    def f(x):
        if isinstance(x, SX):
            return x._unary(op)
        return SX.from_value(x)._unary(op)
    return f


sin = _un(OP.SIN)
cos = _un(OP.COS)
tan = _un(OP.TAN)
asin = _un(OP.ASIN)
acos = _un(OP.ACOS)
atan = _un(OP.ATAN)
exp = _un(OP.EXP)
log = _un(OP.LOG)
sqrt = _un(OP.SQRT)
sq = _un(OP.SQ)
fabs = _un(OP.FABS)
sinh = _un(OP.SINH)
cosh = _un(OP.COSH)
tanh = _un(OP.TANH)
asinh = _un(OP.ASINH)
acosh = _un(OP.ACOSH)
atanh = _un(OP.ATANH)
erf = _un(OP.ERF)
erfinv = _un(OP.ERFINV)
floor = _un(OP.FLOOR)
ceil = _un(OP.CEIL)
sign = _un(OP.SIGN)
log1p = _un(OP.LOG1P)
expm1 = _un(OP.EXPM1)
logic_not = _un(OP.NOT)


def _bin(op):
    def f(x, y):
        return _as_sx(x)._binary(op, y)
    return f


fmin = _bin(OP.FMIN)
fmax = _bin(OP.FMAX)
atan2 = _bin(OP.ATAN2)
power = _bin(OP.POW)
constpow = _bin(OP.CONSTPOW)
logic_and = _bin(OP.AND)
logic_or = _bin(OP.OR)
if_else_zero = _bin(OP.IF_ELSE_ZERO)
hypot = _bin(OP.HYPOT)
copysign = _bin(OP.COPYSIGN)
fmod = _bin(OP.FMOD)


def if_else(c, a, b):
    """if_else(c, a, b) = if_else_zero(c, a) + if_else_zero(!c, b) (casadi/core/sx_elem.hpp:288-290)."""
    c, a, b = _as_sx(c), _as_sx(a), _as_sx(b)
    n = max(c.numel(), a.numel(), b.numel())
    cd, ad, bd = (x.densify() for x in (c, a, b))
    pick = lambda x, i: int(x.nz[0 if x.numel() == 1 else i])
    shape = next(x.shape for x in (c, a, b) if x.numel() == n)
    out = np.array([L.cxs_if_else(pick(cd, i), pick(ad, i), pick(bd, i)) for i in range(n)], np.uint64)
    return SX(Sparsity.dense(*shape), out)


# -- reductions / linear algebra --------------------------------------------
def mtimes(a, b):
    a, b = _as_sx(a), _as_sx(b)
    if a.size2() != b.size1():
        raise ValueError(f"mtimes: dimension mismatch {a.shape} x {b.shape}")
    A, B = a.dense_nodes(), b.dense_nodes()
    zero = _const_node(0)
    out = np.empty((a.size1(), b.size2()), np.uint64)
    for i in range(a.size1()):
        for j in range(b.size2()):
            acc = zero
            for k in range(a.size2()):
                acc = L.cxs_binary(OP.ADD, acc, L.cxs_binary(OP.MUL, int(A[i, k]), int(B[k, j])))
            out[i, j] = acc
    return SX.from_dense_nodes(out)


def dot(a, b):
    a, b = _as_sx(a).densify(), _as_sx(b).densify()
    acc = _const_node(0)
    for x, y in zip(a.nz, b.nz):
        acc = L.cxs_binary(OP.ADD, acc, L.cxs_binary(OP.MUL, int(x), int(y)))
    return SX(Sparsity.dense(1, 1), np.array([acc], np.uint64))


def sumsqr(a):
    a = _as_sx(a)
    acc = _const_node(0)
    for x in a.nz:
        acc = L.cxs_binary(OP.ADD, acc, L.cxs_unary(OP.SQ, int(x)))
    return SX(Sparsity.dense(1, 1), np.array([acc], np.uint64))


def sum1(a):
    d = _as_sx(a).dense_nodes()
    out = []
    for j in range(d.shape[1]):
        acc = _const_node(0)
        for i in range(d.shape[0]):
            acc = L.cxs_binary(OP.ADD, acc, int(d[i, j]))
        out.append(acc)
    return SX.from_dense_nodes(np.array([out], np.uint64))


def sum2(a):
    return sum1(_as_sx(a).T).T


def norm_2(a):
    return sqrt(sumsqr(a))


def cross(a, b):
    a, b = _as_sx(a), _as_sx(b)
    return vertcat(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


# -- calculus (casadi/core/sx_instantiator.cpp:1155-1177) -----------------------
def _calc(fn, f, x, *extra):
    f, x = _as_sx(f), _as_sx(x)
    spf, spx = f.sparsity.compressed(), x.sparsity.compressed()
    h1, h2 = L.P(), L.P()
    L.check(fn(spf.ctypes.data_as(L.PI64), f.nz.ctypes.data_as(L.PU64), spx.ctypes.data_as(L.PI64),
               x.nz.ctypes.data_as(L.PU64), *extra, L.C.byref(h1), L.C.byref(h2)))
    sp = Sparsity.from_compressed(L.take_buf(h1))
    return SX(sp, L.take_buf(h2, "uint64"))


def jacobian(f, x, symmetric=False):
    return _calc(L.cxs_jacobian, f, x, int(symmetric))


def gradient(f, x):
    return _calc(L.cxs_gradient, f, x)


def hessian(f, x):
    """Hessian and gradient of scalar f: (H, g) as casadi.hessian."""
    H = _calc(L.cxs_hessian, f, x)
    return H, gradient(f, x)
