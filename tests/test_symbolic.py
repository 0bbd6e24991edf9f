"""Host symbolic layer: simplification, tape compilation, derivative Functions.

Pinned against closed forms (numpy re-implementations of the model equations),
finite differences, and CasADi's documented rewrite rules
(casadi/core/calculus.hpp:1812-2139, casadi/core/sx_elem.cpp:189-263).
"""
import numpy as np
import pytest

import casadi_amd as ca
import oracle_ref as O
from casadi_amd import workloads as W
from casadi_amd.ops import OP
from casadi_amd.sx import _node_op


def node(e):
    return int(e.nz[0])


def op_of(e):
    return _node_op(node(e))


def test_simplification_rules():
    x, y = ca.SX.sym("x"), ca.SX.sym("y")
    assert node(x + 0) == node(x)
    assert node(0 + x) == node(x)
    assert node(x * 1) == node(x)
    assert op_of(x - x) == OP.CONST and float(x - x) == 0
    assert op_of(x * 2) == OP.TWICE                 # 2*x -> twice(x)
    assert op_of(x + x) == OP.TWICE                 # x+x -> 2*x -> twice(x)
    assert op_of(x * x) == OP.SQ                    # x*x -> sq(x)
    assert node(-(-x)) == node(x)
    assert node(ca.sqrt(x) * ca.sqrt(x)) == node(x)   # x*x -> sq(x), sq(sqrt(x)) -> x
    assert node(ca.sq(ca.sqrt(x))) == node(x)       # sqrt(x)^2 = x
    assert node(ca.log(ca.exp(x))) == node(x)
    assert op_of(x ** 0.5) == OP.SQRT
    assert op_of(x ** 3) == OP.MUL                  # x*x^2
    assert op_of(x ** 2.5) == OP.CONSTPOW
    assert float(ca.sin(x) ** 2 + ca.cos(x) ** 2) == 1.0
    assert op_of(1 / x) == OP.INV
    assert float(ca.SX(3.0) * 4.0) == 12.0          # constant folding
    # x * if_else_zero(c, y) -> if_else_zero(c, x*y) (NaN-safe rule 1)
    c = ca.SX.sym("c")
    e = x * ca.if_else_zero(c, y)
    assert op_of(e) == OP.IF_ELSE_ZERO


def _deps(e):
    """(op, constant value, dep0 node, dep1 node) of the node of scalar e."""
    from casadi_amd import _lib as L
    op, v, d0, d1 = L.I32(), L.D(), L.U64(), L.U64()
    L.check(L.cxs_node_info(node(e), L.C.byref(op), L.C.byref(v), L.C.byref(d0), L.C.byref(d1)))
    return op.value, v.value, d0.value, d1.value


def test_more_rewrite_rules():
    """Rules of common_simp_unary / common_simp_binary
    (casadi/core/calculus.hpp:1812-2139) and SXElem::binary / unary
    (casadi/core/sx_elem.cpp:189-263), each checked on the node it must
    build: the expression the reference returns, with its operation and
    operands.  A rule applied in the wrong branch order or not at all builds
    a different node."""
    x, y, z, c = (ca.SX.sym(s) for s in "xyzc")
    nx, ny, nz = node(x), node(y), node(z)
    # unary
    assert _deps(ca.sq(-x))[0] == OP.SQ and _deps(ca.sq(-x))[2] == nx                  # (-x)^2 -> x^2
    assert _deps(ca.sq(ca.fabs(x)))[0] == OP.SQ and _deps(ca.sq(ca.fabs(x)))[2] == nx   # |x|^2 -> x^2
    assert _deps(ca.fabs(-x))[0] == OP.FABS and _deps(ca.fabs(-x))[2] == nx             # fabs(-x) -> fabs(x)
    # fabs of a nonnegative unary result is that result (sqrt, sq, exp, not,
    # fabs); the binary ones of the checker table (lt, hypot, ...) are not
    # taken, since is_nonnegative asks is_unary first (sx_elem.cpp:366-374)
    for e in (ca.sq(x), ca.sqrt(x), ca.exp(x), ca.logic_not(x)):
        assert node(ca.fabs(e)) == node(e)
    assert _deps(ca.fabs(x < y))[0] == OP.FABS and _deps(ca.fabs(ca.hypot(x, y)))[0] == OP.FABS
    assert node(1 / (1 / x)) == nx                           # 1/(1/x) -> x
    assert _deps(ca.sqrt(ca.sq(x)))[0] == OP.FABS            # sqrt(x^2) -> fabs(x)
    assert _deps(ca.cos(-x))[0] == OP.COS and _deps(ca.cos(-x))[2] == nx                # cos(-x) -> cos(x)
    assert _deps(ca.cos(ca.fabs(x)))[2] == nx                # cos(|x|) -> cos(x)
    assert node(2 * (0.5 * x)) == nx                         # twice(0.5 x) -> x
    # addition
    assert _deps(x + (-y)) [0] == OP.SUB and _deps(x + (-y))[2:] == (nx, ny)             # x + (-y) -> x - y
    assert _deps((-x) + y)[0] == OP.SUB and _deps((-x) + y)[2:] == (ny, nx)              # (-x) + y -> y - x
    assert node((x - y) + y) == nx                           # (x - y) + y -> x
    assert node(y + (x - y)) == nx                           # y + (x - y) -> x
    assert float(ca.sq(ca.cos(x)) + ca.sq(ca.sin(x))) == 1.0  # cos^2 + sin^2 -> 1
    op, _, d0, d1 = _deps(2 * x + x)                          # twice(x) + x -> 3*x (is_doubled)
    assert op == OP.MUL and _node_op(d0, value=True) == (OP.CONST, 3.0) and d1 == nx
    # le / lt decided when the difference is a nonnegative unary result:
    # (x + z^2) - x simplifies to z^2 first
    assert float(x <= x + ca.sq(z)) == 1.0 and float(x + ca.sq(z) < x) == 0.0
    # subtraction
    assert _deps(0 - x)[0] == OP.NEG                          # 0 - y -> -y
    assert _deps(x - (-y))[0] == OP.ADD and _deps(x - (-y))[2:] == (nx, ny)              # x - (-y) -> x + y
    assert node((x + y) - y) == nx and node((x + y) - x) == ny   # (x + y) - y -> x, (x + y) - x -> y
    assert _deps(y - (x + y))[0] == OP.NEG and _deps(y - (x + y))[2] == nx             # y - (x + y) -> -x
    op, _, d0, _ = _deps((-x) - y)                           # (-x) - y -> -(x + y)
    assert op == OP.NEG
    assert node(2 * x - x) == nx                             # 2x - x -> x (is_doubled)
    # multiplication
    assert _deps(x * 3.0)[0] == OP.MUL and _deps(x * 3.0)[3] == nx                      # x*c -> c*x
    assert _deps(x * 0)[:2] == (OP.CONST, 0.0)
    assert _deps(x * -1)[0] == OP.NEG                         # x*(-1) -> -x
    assert _deps(x * (1 / y))[0] == OP.DIV and _deps(x * (1 / y))[2:] == (nx, ny)        # x * inv(y) -> x / y
    assert node(5.0 * (0.2 * x)) == nx                        # 5*(0.2 x) -> x
    assert node(0.5 * (2 * x)) == nx                          # 0.5 * twice(x) -> x
    assert node(5.0 * (x / 5.0)) == nx                        # 5*(x/5) -> x
    assert node((y / x) * x) == ny                            # (y/x)*x -> y
    op, _, d0, _ = _deps((-x) * y)                            # (-x)*y -> -(x*y)
    assert op == OP.NEG and _node_op(d0) == OP.MUL
    op, v, d0, d1 = _deps(3.0 * (2.5 * x))                    # a*(b*x) -> (a*b)*x
    assert op == OP.MUL and _node_op(d0, value=True) == (OP.CONST, 7.5) and d1 == nx
    # division
    assert float(x / x) == 1.0
    assert _deps(x / 0.5)[0] == OP.TWICE                      # x / 0.5 -> 2x
    assert node((2 * x) / 2.0) == nx                          # twice(x)/2 -> x
    assert float((2 * x) / x) == 2.0                          # twice(x)/x -> 2
    assert node((x * y) / x) == ny and node((x * y) / y) == nx
    assert _deps(1 / x)[0] == OP.INV                          # 1/y -> inv(y)
    assert float((-x) / x) == -1.0 and float(x / (-x)) == -1.0
    op, _, d0, _ = _deps((-x) / y)                            # (-x)/y -> -(x/y)
    assert op == OP.NEG and _node_op(d0) == OP.DIV
    # powers: integer exponents become products, negative ones an inverse
    op, _, d0, d1 = _deps(x ** 4)                             # even: (x^2)*(x^2) -> sq rule: sq(x^2)
    assert op == OP.SQ and _node_op(d0) == OP.SQ
    assert _deps(x ** -2)[0] == OP.INV                        # 1/pow(x, 2)
    assert float(x ** 0) == 1.0
    assert _deps(x ** 101)[0] == OP.CONSTPOW                  # |n| > 100: constpow
    # comparisons, fmin / fmax, if_else_zero
    assert float(x.eq(x)) == 1.0 and float(x.ne(x)) == 0.0
    assert node(ca.fmin(x, x)) == nx and node(ca.fmax(x, x)) == nx
    assert node(ca.fmin(x, float("inf"))) == nx and node(ca.fmax(float("-inf"), x)) == nx
    assert float(ca.fmin(x, float("-inf"))) == float("-inf") and float(ca.fmax(x, float("inf"))) == float("inf")
    assert float(ca.if_else_zero(c, 0)) == 0.0 and node(ca.if_else_zero(1.0, y)) == ny
    assert float(ca.if_else_zero(0.0, y)) == 0.0
    # if_else_zero(c, x) + if_else_zero(c, y) -> if_else_zero(c, x + y) (rule 2)
    op, _, d0, d1 = _deps(ca.if_else_zero(c, x) + ca.if_else_zero(c, y))
    assert op == OP.IF_ELSE_ZERO and d0 == node(c) and _node_op(d1) == OP.ADD
    # logic: not(not(x)) -> x; twice of twice -> 4x
    assert node(ca.logic_not(ca.logic_not(x))) == nx
    op, _, d0, d1 = _deps(2 * (2 * x))
    assert op == OP.MUL and _node_op(d0, value=True) == (OP.CONST, 4.0) and d1 == nx


def _np_cartpole_ode(s, u):
    mc, mp, l, g = W.CP_MC, W.CP_MP, W.CP_L, W.CP_G
    th, xd, thd = s[1], s[2], s[3]
    sn, cs = np.sin(th), np.cos(th)
    temp = (u + mp * l * thd * thd * sn) / (mc + mp)
    thacc = (g * sn - cs * temp) / (l * (4.0 / 3.0 - mp * cs * cs / (mc + mp)))
    xacc = temp - mp * l * thacc * cs / (mc + mp)
    return np.array([xd, thd, xacc, thacc])


def _np_rk4(s, u, dt):
    k1 = _np_cartpole_ode(s, u)
    k2 = _np_cartpole_ode(s + dt / 2 * k1, u)
    k3 = _np_cartpole_ode(s + dt / 2 * k2, u)
    k4 = _np_cartpole_ode(s + dt * k3, u)
    return s + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)


def test_cartpole_tape_matches_closed_form():
    f = W.cartpole_step()
    rng = np.random.default_rng(1)
    n = 200
    x = rng.uniform(-1, 1, (n, 4))
    u = rng.uniform(-10, 10, (n, 1))
    (y,) = O.map_eval(f, n, [x, u])
    ref = np.array([_np_rk4(x[k], u[k, 0], 0.05) for k in range(n)])
    assert O.rel_err(y, ref) < 1e-13


def test_rollout_matches_closed_form():
    f = W.cartpole_rollout(20)
    rng = np.random.default_rng(2)
    x0 = rng.uniform(-1, 1, 4)
    us = rng.uniform(-10, 10, 20)
    s, J = x0.copy(), 0.0
    for t in range(20):
        J += s @ s + 0.01 * us[t] ** 2
        s = _np_rk4(s, us[t], 0.05)
    J += s @ s
    assert abs(float(f(x0, us)) - J) <= 1e-13 * max(1, abs(J))


def test_host_vm_equals_oracle_vm_bitwise():
    f = W.quadrotor_step()
    rng = np.random.default_rng(3)
    n = 64
    x = rng.uniform(-1, 1, (n, 12))
    u = rng.uniform(0, 2, (n, 4))
    (ref,) = O.map_eval(f, n, [x, u])
    got = f.map(n, "serial")(x.T, u.T)
    assert np.array_equal(got, ref.T)
    got_omp = f.map(n, "openmp", 4)(x.T, u.T)
    assert np.array_equal(got_omp, ref.T)


def test_map_equals_loop():
    f = W.arm_dynamics()
    rng = np.random.default_rng(4)
    n = 10
    z = rng.uniform(-1, 1, (6, n))
    F = f.map(n, "serial")(z)
    for k in range(n):
        assert np.array_equal(F[:, k], f(z[:, k]).reshape(-1))


def test_tape_structure():
    f = W.cartpole_step()
    a = f.instructions()
    assert np.sum(a["op"] == OP.INPUT) == 5
    assert np.sum(a["op"] == OP.OUTPUT) == 4
    assert f.sz_w() < f.n_instructions()          # live-variable slot reuse
    assert a["i0"][a["op"] != OP.OUTPUT].max() < f.sz_w()
    # unary instructions carry i2 == i1
    un = np.isin(a["op"], [OP.SIN, OP.COS, OP.SQ, OP.NEG, OP.TWICE])
    assert np.array_equal(a["i1"][un], a["i2"][un])


def _tape(f):
    inv = {v: k for k, v in vars(OP).items() if not k.startswith("_")}
    return [(inv[int(e["op"])], int(e["i0"]), int(e["i1"]), int(e["i2"])) for e in f.instructions()], f.sz_w()


def test_tape_hand_traced():
    """The tape SXFunction::init emits (casadi/core/sx_function.cpp:476-841),
    traced by hand for two small Functions.

    Nodes come in the depth-first order of sort_depth_first
    (casadi/core/x_function.hpp:412-438: dependencies left to right, each
    node once), one OUTPUT per output nonzero.  Work-vector slots are handed
    out in that order from a LIFO stack of freed slots: an instruction first
    releases the slots of the operands it reads for the last time, the last
    operand first (so the first operand's slot is on top), then takes the top
    of the stack, or a new slot when the stack is empty (lines 689-753).
    Symbolic nodes become INPUT records after the slots are placed.

    f(x, y) = [(x + y) sin(x) - y, x + y]: nodes x, y, a = x+y, b = sin x,
    c = a b, d = c - y.  b reads x for the last time: x's slot 0 is freed and
    reused by b; c frees b's slot 0 and takes it; d frees y's slot 1, then
    c's slot 0 (on top) and takes 0; a keeps slot 2 until the second output.

    g(x[2], y) = [(y + 3) x1]: x0 is never read (no INPUT record); the ADD
    frees the constant's slot 1 and then y's slot 0 and takes 0; x1 takes the
    freed slot 1."""
    x, y = ca.SX.sym("x"), ca.SX.sym("y")
    a = x + y
    f = ca.Function("f", [x, y], [a * ca.sin(x) - y, a])
    assert _tape(f) == ([("INPUT", 0, 0, 0), ("INPUT", 1, 1, 0), ("ADD", 2, 0, 1), ("SIN", 0, 0, 0),
                         ("MUL", 0, 2, 0), ("SUB", 0, 0, 1), ("OUTPUT", 0, 0, 0), ("OUTPUT", 1, 2, 0)], 3)
    x, y = ca.SX.sym("x", 2), ca.SX.sym("y")
    g = ca.Function("g", [x, y], [(y + 3.0) * x[1]])
    three = np.array([3.0]).view(np.int32)       # the CONST record's i1, i2 hold the double's two words
    assert _tape(g) == ([("INPUT", 0, 1, 0), ("CONST", 1, int(three[0]), int(three[1])), ("ADD", 0, 0, 1),
                         ("INPUT", 1, 0, 1), ("MUL", 0, 0, 1), ("OUTPUT", 0, 0, 0)], 2)


def test_free_variable_error():
    x, y = ca.SX.sym("x"), ca.SX.sym("y")
    with pytest.raises(ca.CasadiHipError, match="free"):
        ca.Function("f", [x], [x * y])
    f = ca.Function("f", [x], [x * y], opts={"allow_free": True})
    assert f.has_free()


def _fd_jac(fun, x, h=1e-6):
    cols = []
    for i in range(x.size):
        e = np.zeros_like(x)
        e.flat[i] = h
        cols.append((fun(x + e) - fun(x - e)) / (2 * h))
    return np.stack(cols, axis=1)


def test_jacobian_function_matches_fd():
    f = W.cartpole_step()
    J = f.jacobian()
    assert J.name() == "jac_cartpole"
    assert J.name_out() == ["jac_xf_x", "jac_xf_u"]
    x0, u0 = np.array([0.1, -0.4, 0.3, 0.2]), 2.0
    jx, ju = J(x0, u0, 0)
    fdx = _fd_jac(lambda v: f(v, u0).reshape(-1), x0)
    fdu = _fd_jac(lambda v: f(x0, v[0]).reshape(-1), np.array([u0]))
    assert np.allclose(jx, fdx, atol=1e-7)
    assert np.allclose(ju, fdu, atol=1e-7)
    # structural zeros: the cart position does not enter the dynamics
    sp = J.sparsity_out(0)
    assert not sp.dense_mask()[1, 0] and sp.dense_mask()[0, 0]


def test_forward_and_reverse_functions():
    f = W.arm_dynamics()
    z0 = np.array([0.3, -0.2, 0.5, 0.1, 1.0, -0.5])
    fd = _fd_jac(lambda v: f(v).reshape(-1), z0)
    v = np.array([0.2, -1.0, 0.3, 0.0, 0.7, 0.1])
    Ff = f.forward(1)
    assert Ff.name_in() == ["z", "out_zdot", "fwd_z"]
    assert np.allclose(Ff(z0, 0, v).reshape(-1), fd @ v, atol=1e-7)
    w = np.array([1.0, -2.0, 0.5, 0.25])
    Fr = f.reverse(1)
    assert Fr.name_in() == ["z", "out_zdot", "adj_zdot"] and Fr.name_out() == ["adj_z"]
    assert np.allclose(Fr(z0, 0, w).reshape(-1), fd.T @ w, atol=1e-7)


def test_hessian_symmetric_and_correct():
    x = ca.SX.sym("x", 3)
    e = ca.sin(x[0] * x[1]) + x[2] ** 3 * x[0] + ca.exp(x[1])
    H, g = ca.hessian(e, x)
    assert H.sparsity.is_symmetric()
    fH = ca.Function("h", [x], [H])
    x0 = np.array([0.3, 0.6, -0.2])
    Hv = fH(x0)
    a, b, c = x0
    ref = np.array([[-b * b * np.sin(a * b), np.cos(a * b) - a * b * np.sin(a * b), 3 * c * c],
                    [np.cos(a * b) - a * b * np.sin(a * b), -a * a * np.sin(a * b) + np.exp(b), 0],
                    [3 * c * c, 0, 6 * c * a]])
    assert np.allclose(Hv, ref, atol=1e-14)


def test_symbolic_call_inlines():
    f = W.cartpole_step()
    x, u = ca.SX.sym("x", 4), ca.SX.sym("u")
    y = f(f(x, u), u)
    g = ca.Function("g", [x, u], [y])
    x0 = np.array([0.1, 0.2, 0.3, 0.4])
    assert np.array_equal(g(x0, 1.0), f(f(x0, 1.0), 1.0))


def test_hessian_pattern_drops_one_sided_dependencies():
    # The gradient of L = x0*x1 + x0*floor(x2) is [x1 + floor(x2), x0, 0]: its
    # first entry depends on x2 structurally (floor passes the dependency on),
    # but the derivative with respect to x2 carries no x0 term.  The Jacobian
    # of the gradient therefore has (0,2) and not (2,0); the symmetric
    # pattern is the intersection with the transpose
    # (casadi/core/function_internal.cpp:2051-2062), here {(0,1), (1,0)}.
    x = ca.SX.sym("x", 3)
    L = x[0] * x[1] + x[0] * ca.floor(x[2])
    _, g = ca.hessian(L, x)
    J = ca.jacobian(g, x)
    assert J.sparsity.dense_mask()[0, 2] and not J.sparsity.dense_mask()[2, 0]
    H, _ = ca.hessian(L, x)
    assert np.array_equal(H.sparsity.dense_mask(), np.array([[0, 1, 0], [1, 0, 0], [0, 0, 0]], bool))
    fH = ca.Function("h", [x], [H])
    for x0 in (np.array([0.3, 0.7, 1.5]), np.array([-2.0, 0.1, -0.4])):
        assert np.array_equal(fH(x0), np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 0.0]]))
    # A second case with sign and fmin: L = x0*x1*x2 + x1*sign(x0) + fmin(x2, 1)*x0
    L2 = x[0] * x[1] * x[2] + x[1] * ca.sign(x[0]) + ca.fmin(x[2], 1.0) * x[0]
    H2, _ = ca.hessian(L2, x)
    assert H2.sparsity.is_symmetric()
    a, b, c = 0.3, -0.6, 0.4          # fmin(c, 1) = c: its partial is 1
    ref = np.array([[0.0, c, b + 1.0], [c, 0.0, a], [b + 1.0, a, 0.0]])
    assert np.array_equal(ca.Function("h2", [x], [H2])(np.array([a, b, c])), ref)


def test_symmetric_jac_sparsity_full_and_compact():
    # the intersection is taken in the full space; the compact pattern is the
    # full one stripped back to the nonzeros (the same here: dense blocks)
    x = ca.SX.sym("x", 3)
    L = x[0] * x[1] + x[0] * ca.floor(x[2]) + ca.sin(x[2]) * x[2]
    f = ca.Function("g", [x], [ca.gradient(L, x)])
    full = f.jac_sparsity(0, 0, compact=False, symmetric=True)
    assert np.array_equal(full.dense_mask(), np.array([[0, 1, 0], [1, 0, 0], [0, 0, 1]], bool))
    compact = f.jac_sparsity(0, 0, compact=True, symmetric=True)
    assert compact == full
