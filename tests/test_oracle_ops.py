"""Pin the oracle's scalar semantics (calculus.hpp restatement) and the host VM.

Values are checked against numpy / scipy / math closed forms, partial
derivatives against central finite differences, plus CasADi's documented
special cases.  A dropped term, flipped sign or swapped operand in or_fun /
or_der fails one of these.
"""
import math

import numpy as np
import pytest
import scipy.special as sps

import oracle_ref as O
from casadi_amd.ops import OP, UNARY, BINARY

VALUES = {
    OP.ADD: lambda x, y: x + y, OP.SUB: lambda x, y: x - y, OP.MUL: lambda x, y: x * y,
    OP.DIV: lambda x, y: x / y, OP.NEG: lambda x, y: -x, OP.EXP: lambda x, y: np.exp(x),
    OP.LOG: lambda x, y: np.log(x), OP.POW: lambda x, y: x ** y, OP.CONSTPOW: lambda x, y: x ** y,
    OP.SQRT: lambda x, y: np.sqrt(x), OP.SQ: lambda x, y: x * x, OP.TWICE: lambda x, y: 2 * x,
    OP.SIN: lambda x, y: np.sin(x), OP.COS: lambda x, y: np.cos(x), OP.TAN: lambda x, y: np.tan(x),
    OP.ASIN: lambda x, y: np.arcsin(x), OP.ACOS: lambda x, y: np.arccos(x),
    OP.ATAN: lambda x, y: np.arctan(x), OP.LT: lambda x, y: float(x < y), OP.LE: lambda x, y: float(x <= y),
    OP.EQ: lambda x, y: float(x == y), OP.NE: lambda x, y: float(x != y), OP.NOT: lambda x, y: float(not x),
    OP.AND: lambda x, y: float(bool(x) and bool(y)), OP.OR: lambda x, y: float(bool(x) or bool(y)),
    OP.FLOOR: lambda x, y: np.floor(x), OP.CEIL: lambda x, y: np.ceil(x), OP.FMOD: lambda x, y: np.fmod(x, y),
    OP.REMAINDER: lambda x, y: math.remainder(x, y), OP.FABS: lambda x, y: abs(x),
    OP.SIGN: lambda x, y: np.sign(x), OP.COPYSIGN: lambda x, y: np.copysign(x, y),
    OP.IF_ELSE_ZERO: lambda x, y: 0.0 if x == 0 else y, OP.ERF: lambda x, y: sps.erf(x),
    OP.FMIN: lambda x, y: min(x, y), OP.FMAX: lambda x, y: max(x, y), OP.INV: lambda x, y: 1 / x,
    OP.SINH: lambda x, y: np.sinh(x), OP.COSH: lambda x, y: np.cosh(x), OP.TANH: lambda x, y: np.tanh(x),
    OP.ASINH: lambda x, y: np.arcsinh(x), OP.ACOSH: lambda x, y: np.arccosh(x),
    OP.ATANH: lambda x, y: np.arctanh(x), OP.ATAN2: lambda x, y: np.arctan2(x, y),
    OP.ERFINV: lambda x, y: sps.erfinv(x), OP.LOG1P: lambda x, y: np.log1p(x),
    OP.EXPM1: lambda x, y: np.expm1(x), OP.HYPOT: lambda x, y: np.hypot(x, y),
    OP.ASSIGN: lambda x, y: x, OP.PRINTME: lambda x, y: x, OP.LIFT: lambda x, y: x,
}

# points inside every op's domain
PTS = [(0.3, 0.7), (0.55, 1.9), (0.9, 0.4)]
NONSMOOTH = {OP.LT, OP.LE, OP.EQ, OP.NE, OP.NOT, OP.AND, OP.OR, OP.FLOOR, OP.CEIL, OP.SIGN}


def _pt(op, x, y):
    if op == OP.ACOSH:
        return 1.0 + x + y, y
    return x, y


@pytest.mark.parametrize("op", sorted(VALUES))
def test_values_match_closed_forms(op):
    for x, y in PTS:
        x, y = _pt(op, x, y)
        ref = VALUES[op](x, y)
        got = O.fun(op, x, y)
        tol = 1e-5 if op == OP.ERFINV else 4e-16 * max(1.0, abs(ref))
        assert abs(got - ref) <= tol, (op, x, y, got, ref)


@pytest.mark.parametrize("op", sorted(set(VALUES) - NONSMOOTH - {OP.IF_ELSE_ZERO}))
def test_partials_match_finite_differences(op):
    for x, y in PTS:
        x, y = _pt(op, x, y)
        f = O.fun(op, x, y)
        d = O.der(op, x, y, f)
        h = 1e-6
        dx = (O.fun(op, x + h, y) - O.fun(op, x - h, y)) / (2 * h)
        assert abs(d[0] - dx) <= 1e-6 * max(1, abs(dx)), (op, "d/dx", d[0], dx)
        if op in BINARY:
            dy = (O.fun(op, x, y + h) - O.fun(op, x, y - h)) / (2 * h)
            if op == OP.CONSTPOW:
                dy = 0.0  # the exponent of constpow is a constant
            assert abs(d[1] - dy) <= 1e-6 * max(1, abs(dy)), (op, "d/dy", d[1], dy)
        else:
            assert d[1] == 0


def test_special_cases():
    nan = float("nan")
    assert math.isnan(O.fun(OP.SIGN, nan))                # sign(nan) == nan
    assert O.fun(OP.IF_ELSE_ZERO, 0.0, nan) == 0.0         # if_else_zero(0, nan) = 0
    assert O.fun(OP.IF_ELSE_ZERO, 2.0, 5.0) == 5.0
    assert list(O.der(OP.FMIN, 1.0, 1.0, 1.0)) == [0.5, 0.5]  # tie splits evenly
    assert list(O.der(OP.FMAX, 2.0, 2.0, 2.0)) == [0.5, 0.5]
    assert list(O.der(OP.FMIN, 1.0, 2.0, 1.0)) == [1.0, 0.0]
    assert list(O.der(OP.IF_ELSE_ZERO, 3.0, 4.0, 4.0)) == [0.0, 3.0]
    # pow derivative w.r.t. the exponent is log(x) * f
    d = O.der(OP.POW, 2.0, 3.0, 8.0)
    assert d[0] == 3.0 * 2.0 ** 2 and d[1] == math.log(2.0) * 8.0
    # erfinv at the domain edges
    assert O.fun(OP.ERFINV, 1.0) == math.inf and O.fun(OP.ERFINV, -1.0) == -math.inf
    assert math.isnan(O.fun(OP.ERFINV, 1.5))
    # CasADi's rational approximant polished by two Newton steps: the first
    # guess is poor near |x| = 0.5 .. 0.7, so the result is only 1e-11 .. 1e-6
    # accurate there ...
    for v in (-0.5, 0.1, 0.69, 0.71, 0.99):
        assert abs(O.fun(OP.ERFINV, v) - sps.erfinv(v)) < 1e-5
    assert abs(O.fun(OP.ERFINV, 0.1) - sps.erfinv(0.1)) < 1e-15
    # ... but CasADi returns the unpolished rational approximant for x < -0.7
    # (early return, casadi/core/calculus.hpp:306-308): only ~1e-7 accurate
    for v in (-0.95, -0.75):
        err = abs(O.fun(OP.ERFINV, v) - sps.erfinv(v))
        assert 1e-12 < err < 1e-6


def test_every_operation_covered():
    assert set(UNARY + BINARY) <= set(VALUES)


def test_mutation_is_caught():
    """A planted sign error in a partial is caught by the FD check above."""
    x, y = 0.3, 0.7
    f = O.fun(OP.ATAN2, x, y)
    d = O.der(OP.ATAN2, x, y, f)
    wrong = -d[1]
    h = 1e-6
    dy = (O.fun(OP.ATAN2, x, y + h) - O.fun(OP.ATAN2, x, y - h)) / (2 * h)
    assert abs(wrong - dy) > 1e-3 and abs(d[1] - dy) < 1e-8
