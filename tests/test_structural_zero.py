"""Reverse sweeps skip what CasADi's symbolic reverse drops.

SXFunction::ad_reverse (casadi/core/sx_function.cpp:1454-1560) builds the
adjoint symbolically: a constant-zero partial (floor, ceil, sign, not, the
comparisons, the exponent of constpow, ...; casadi/core/calculus.hpp) and an
adjoint no output reaches multiply as the SXElem zero, and x * 0 -> 0,
w + 0 -> w remove them.  Evaluated numerically, CasADi's reverse(1) Function
therefore never forms 0 * inf.  The numeric sweeps here (the oracle's, the
interpreter's, the generated kernels') skip the same contributions, so an
infinite partial behind a floor or a comparison gives the reference's finite
adjoint, not NaN.
"""
import numpy as np

import casadi_amd as ca
import oracle_ref as O


def edge_functions():
    x, y = ca.SX.sym("x"), ca.SX.sym("y")
    return [
        # d/dx floor(sqrt x) at 0: the floor's zero partial hides sqrt's infinite one
        (ca.Function("f1", [x, y], [ca.floor(ca.sqrt(x)) + x * y]), [0.0, 2.0]),
        # sign(log x) at 0: log's partial is infinite, sign's partial zero
        (ca.Function("f2", [x, y], [ca.sign(ca.log(x)) * y + y]), [0.0, 3.0]),
        # a comparison of an infinite-slope value
        (ca.Function("f3", [x, y], [ca.if_else_zero(ca.sqrt(x) < y, y) + x]), [0.0, 1.0]),
        # constpow's exponent and copysign's sign operand
        (ca.Function("f4", [x, y], [ca.copysign(x, ca.log(y)) + ca.constpow(x, 2.5)]), [1.5, 0.0]),
    ]


def test_oracle_sweep_equals_reverse_function_on_edge_inputs():
    for f, pt in edge_functions():
        args = [np.array([[v]]) for v in pt]
        seeds = [np.ones((1, 1))]
        _, g = O.adj_map(f, 1, args, seeds)
        R = f.reverse(1)
        gr = O.map_eval(R, 1, args + [None] + seeds)
        for a, b in zip(g, gr):
            if b.size == 0:   # reverse(1) drops a structurally zero sensitivity's sparsity (#4345)
                assert np.all(a == 0), (f.name(), a)
            else:
                assert np.array_equal(a, b, equal_nan=True), (f.name(), a, b)
            assert np.all(np.isfinite(a)), (f.name(), a)
