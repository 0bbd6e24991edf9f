"""Map semantics on the host (no GPU): every CPU parallelization equals the
loop of single calls (as test/python/function.py:659-760 checks for CasADi),
implicit parallel evaluation of horizontally stacked arguments
(FunctionInternal::check_mat, casadi/core/function_internal.cpp:3649-3670),
and the map of a derivative equals the derivative per instance
(Map::get_forward / get_reverse, casadi/core/map.cpp:219-325)."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def fun():
    import casadi_amd as ca
    x = ca.SX.sym("x")
    y = ca.SX.sym("y", 2)
    z = ca.SX.sym("z", 2, 2)
    return ca.Function("f", [x, y, z], [ca.mtimes(z, y) + x, ca.sin(y * x).T])


def _inputs(n, seed=0):
    rng = np.random.default_rng(seed)
    return rng.random((1, n)), rng.random((2, n)), rng.random((2, 2 * n))


@pytest.mark.parametrize("par", ["serial", "openmp", "thread", "unroll", "inline"])
def test_map_equals_loop(fun, par):
    n = 5
    X, Y, Z = _inputs(n)
    r0, r1 = fun.map(n, par)(X, Y, Z)
    for k in range(n):
        a, b = fun(X[:, k], Y[:, k], Z[:, 2 * k:2 * k + 2])
        assert np.array_equal(r0[:, k:k + 1], a)
        assert np.array_equal(r1[:, 2 * k:2 * k + 2], b)


def test_implicit_parallel_call(fun):
    n = 4
    X, Y, Z = _inputs(n, 1)
    r0, r1 = fun(X, Y, Z)                     # all three stacked: npar = 4
    m0, m1 = fun.map(n, "serial")(X, Y, Z)
    assert r0.shape == (2, n) and np.array_equal(r0, m0) and np.array_equal(r1, m1)
    # a non-stacked argument is shared by every evaluation
    r0, _ = fun(X, Y[:, :1], Z)
    for k in range(n):
        a, _ = fun(X[:, k], Y[:, 0], Z[:, 2 * k:2 * k + 2])
        assert np.array_equal(r0[:, k:k + 1], a)
    # a transposed vector and a scalar broadcast
    a, b = fun(0.5, [[1.0, 2.0]], np.eye(2))
    a2, b2 = fun(0.5, [1.0, 2.0], np.eye(2))
    assert np.array_equal(a, a2) and np.array_equal(b, b2)
    with pytest.raises(ValueError):
        fun(X[:, :1], np.ones((3, 1)), Z[:, :2])


def test_map_derivatives_are_mapped_derivatives(fun):
    n = 3
    X, Y, Z = _inputs(n, 2)
    F = fun.map(n, "serial")
    rng = np.random.default_rng(3)
    S0, S1 = rng.random((2, n)), rng.random((1, 2 * n))
    gx, gy, gz = F.reverse(1)(X, Y, Z, None, None, S0, S1)
    R = fun.reverse(1)
    for k in range(n):
        a, b, c = R(X[:, k], Y[:, k], Z[:, 2 * k:2 * k + 2], None, None, S0[:, k], S1[:, 2 * k:2 * k + 2])
        assert np.array_equal(gx[:, k:k + 1], a) and np.array_equal(gy[:, k:k + 1], b)
        assert np.array_equal(gz[:, 2 * k:2 * k + 2], c)
    T0, T1, T2 = rng.random((1, n)), rng.random((2, n)), rng.random((2, 2 * n))
    f0, f1 = F.forward(1)(X, Y, Z, None, None, T0, T1, T2)
    Fw = fun.forward(1)
    for k in range(n):
        a, b = Fw(X[:, k], Y[:, k], Z[:, 2 * k:2 * k + 2], None, None, T0[:, k], T1[:, k], T2[:, 2 * k:2 * k + 2])
        assert np.array_equal(f0[:, k:k + 1], a) and np.array_equal(f1[:, 2 * k:2 * k + 2], b)
