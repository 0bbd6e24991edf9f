"""Edge-case inputs through every operation, on the GPU, against the oracle.

One Function evaluates every CasADi operation on two inputs (x, y), and the
batch is the grid of special-value pairs: NaN, +-Inf, +-0, +-1, halves,
twos, the largest and smallest normal and subnormal magnitudes, and values
where fmod / remainder / floor / ceil / comparisons tie.  Every pair with
x == y is a tie for fmin / fmax / comparisons, (0, -0) a signed-zero tie,
if_else_zero(0, NaN) is in the grid.

* Forward values: the interpreter and the specialised kernels give the same
  bits (NaN where NaN).  Against the oracle (casadi_math<double>::fun,
  casadi/core/calculus.hpp:598-1018): operations that IEEE 754 rounds
  exactly (arithmetic, sqrt, comparisons, floor/ceil/fabs/sign/copysign,
  fmin/fmax, fmod/remainder, if_else_zero) bit for bit, the sign of zero
  included; transcendental ones with the same NaN / infinity pattern and
  finite values within 1e-13 * max(|ref|, 1).
* Adjoints (the fused forward + reverse sweep, every output seeded): the two
  engines give the same bits; against the oracle's numeric reverse sweep the
  same classes apply to the partials casadi_math<double>::der gives
  (calculus.hpp:1302-1439).
"""
import numpy as np
import pytest

import oracle_ref as O
from conftest import seed_of

pytestmark = pytest.mark.gpu
TOL = 1e-13

SPECIAL = [np.nan, np.inf, -np.inf, 0.0, -0.0, 1.0, -1.0, 0.5, -0.5, 2.0, -2.0, 3.0, 0.75, 10.0,
           1e308, -1e308, 2.2250738585072014e-308, 5e-324, -5e-324, 1e-300, 700.0, -745.0]


def _edge_function():
    import casadi_amd as ca
    x, y = ca.SX.sym("x"), ca.SX.sym("y")
    ops = [
        ("add", x + y, True), ("sub", x - y, True), ("mul", x * y, True), ("div", x / y, True),
        ("neg", -x, True), ("exp", ca.exp(x), False), ("log", ca.log(x), False), ("pow", x ** y, False),
        ("constpow", ca.constpow(x, 3.0), False), ("sqrt", ca.sqrt(x), True), ("sq", ca.sq(x), True),
        ("twice", 2 * x, True), ("sin", ca.sin(x), False), ("cos", ca.cos(x), False), ("tan", ca.tan(x), False),
        ("asin", ca.asin(x), False), ("acos", ca.acos(x), False), ("atan", ca.atan(x), False),
        ("lt", x < y, True), ("le", x <= y, True), ("eq", x.eq(y), True), ("ne", x.ne(y), True),
        ("not", ca.logic_not(x), True), ("and", ca.logic_and(x, y), True), ("or", ca.logic_or(x, y), True),
        ("floor", ca.floor(x), True), ("ceil", ca.ceil(x), True), ("fmod", ca.fmod(x, y), True),
        ("fabs", ca.fabs(x), True), ("sign", ca.sign(x), True), ("copysign", ca.copysign(x, y), True),
        ("if_else_zero", ca.if_else_zero(x, y), True), ("erf", ca.erf(x), False),
        ("fmin", ca.fmin(x, y), True), ("fmax", ca.fmax(x, y), True), ("inv", 1 / x, True),
        ("sinh", ca.sinh(x), False), ("cosh", ca.cosh(x), False), ("tanh", ca.tanh(x), False),
        ("asinh", ca.asinh(x), False), ("acosh", ca.acosh(x), False), ("atanh", ca.atanh(x), False),
        ("atan2", ca.atan2(x, y), False), ("erfinv", ca.erfinv(x), False), ("log1p", ca.log1p(x), False),
        ("expm1", ca.expm1(x), False), ("hypot", ca.hypot(x, y), False),
    ]
    f = ca.Function("edge", [x, y], [ca.vertcat(*[e for _, e, _ in ops])])
    return f, [n for n, _, _ in ops], np.array([ex for _, _, ex in ops])


def _grid():
    v = np.array(SPECIAL)
    xs, ys = np.meshgrid(v, v, indexing="ij")
    return xs.reshape(1, -1), ys.reshape(1, -1)


def _dev(*arrs):
    import torch
    return [torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda") for a in arrs]


def _same_bits(a, b):
    """Equal, NaN where NaN, and the same sign where both are zero."""
    eq = (a == b) | (np.isnan(a) & np.isnan(b))
    return eq & ((a != 0) | (np.signbit(a) == np.signbit(b)))


def _close(a, b):
    """The same NaN / infinity pattern, finite values within the bar."""
    same = (np.isnan(a) & np.isnan(b)) | (a == b)
    fin = np.isfinite(a) & np.isfinite(b)
    near = np.zeros_like(same)
    near[fin] = np.abs(a[fin] - b[fin]) <= TOL * np.maximum(np.abs(b[fin]), 1.0)
    return same | near


def _report(name_of_row, ok, a, b, xs, ys):
    bad = np.argwhere(~ok)
    for r, k in bad[:12]:
        print(f"  {name_of_row[r]}({xs[0, k]!r}, {ys[0, k]!r}): gpu {a[r, k]!r} oracle {b[r, k]!r}", flush=True)
    return len(bad)


def test_edge_values_forward_and_adjoint(cuda):
    from casadi_amd.hip import HipTape
    f, names, exact = _edge_function()
    xs, ys = _grid()
    n = xs.shape[1]
    seed = np.random.default_rng(seed_of(12)).standard_normal((len(names), n))
    print(f"[case] edge values: {len(names)} operations x {n} special pairs", flush=True)
    ti, tj = HipTape(f), HipTape(f, engine="jit")
    (yi,) = [t.cpu().numpy() for t in ti.eval(_dev(xs, ys))]
    (yj,) = [t.cpu().numpy() for t in tj.eval(_dev(xs, ys))]
    assert _same_bits(yi, yj).all()
    (yo,) = O.map_eval(f, n, [xs.T, ys.T])
    yo = yo.T
    ok = np.where(exact[:, None], _same_bits(yj, yo), _close(yj, yo))
    assert _report(names, ok, yj, yo, xs, ys) == 0
    # adjoints: every output seeded
    (oi,), gi = ti.fwd_adj(_dev(xs, ys), _dev(seed))
    (oj,), gj = tj.fwd_adj(_dev(xs, ys), _dev(seed))
    gi = [g.cpu().numpy() for g in gi]
    gj = [g.cpu().numpy() for g in gj]
    for a, b in zip(gi, gj):
        assert _same_bits(a, b).all()
    _, go = O.adj_map(f, n, [xs.T, ys.T], [seed.T])
    for a, b, which in zip(gj, go, ("x", "y")):
        b = b.T
        ok = _close(a, b)
        assert _report([f"d/d{which}"], ok, a, b, xs, ys) == 0


def test_structural_zero_contributions_skipped(cuda):
    """An infinite partial behind a constant-zero one (floor, sign, a
    comparison, constpow's exponent, copysign's sign operand): both engines
    skip the contribution as CasADi's symbolic reverse does, and give the
    oracle's finite adjoint bit for bit (tests/test_structural_zero.py pins
    the oracle against reverse(1))."""
    from casadi_amd.hip import HipTape
    from test_structural_zero import edge_functions
    for f, pt in edge_functions():
        n = 130
        ins = [np.full((1, n), v) for v in pt]
        seed = np.ones((1, n))
        print(f"[case] structural zero {f.name()} at {pt}", flush=True)
        _, go = O.adj_map(f, n, [a.T for a in ins], [seed.T])
        for engine in ("interp", "jit"):
            _, g = HipTape(f, engine=engine).fwd_adj(_dev(*ins), _dev(seed))
            for a, b in zip(g, go):
                a = a.cpu().numpy()
                assert np.all(np.isfinite(a)) and np.array_equal(a, b.T), (f.name(), engine)
