"""Reduced maps (MapSum): the host paths against the oracle, and the oracle's
MapSum restatement pinned by exact sums, rounding bounds and linearity.

The reference's tests have no stored vectors for MapSum; what fixes a sum is
arithmetic: sums of small integers are exact in any order, a sum of floats
computed term by term is within (terms - 1) * eps * sum|term| of the exact sum
(math.fsum), and the sensitivity of a summed output to a shared input is the
derivative of the sum (central differences).
"""
import math

import numpy as np
import pytest

import oracle_ref as O

EPS = 2.0 ** -53


@pytest.fixture(scope="module")
def ca():
    import casadi_amd
    return casadi_amd


def _fsum_rows(rows):
    return np.array([math.fsum(rows[:, j]) for j in range(rows.shape[1])])


def _step_inputs(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, (n, 4)), rng.uniform(-10, 10, (1,))


def test_oracle_mapsum_pinned(ca):
    from casadi_amd import workloads as W
    f = W.cartpole_step()
    n = 2000
    x, u = _step_inputs(n, 3)
    (s,) = O.mapsum_eval(f, n, [x, u], [False, True], [True])
    per = O.map_eval(f, n, [x, np.tile(u, (n, 1))])[0]
    exact = _fsum_rows(per)
    bound = (n - 1) * EPS * np.sum(np.abs(per), axis=0)
    assert np.all(np.abs(s - exact) <= bound)
    # a sum that drops one instance is caught by the same bound
    assert not np.all(np.abs(O._serial_sum(per[1:]) - exact) <= bound)
    # integer-valued outputs: exact in any order
    a = ca.SX.sym("a", 2)
    g = ca.Function("g", [a], [ca.vertcat(ca.floor(a[0] * 8), ca.floor(a[1] * 3) + 1)])
    av = np.random.default_rng(4).uniform(-50, 50, (n, 2))
    (si,) = O.mapsum_eval(g, n, [av], [False], [True])
    assert np.array_equal(si, np.floor(av * [8, 3]).sum(axis=0) + [0, n])


def test_oracle_mapsum_adjoint_is_the_derivative_of_the_sum(ca):
    from casadi_amd import workloads as W
    f = W.cartpole_step()
    n = 40
    x, u = _step_inputs(n, 5)
    seed = np.random.default_rng(6).standard_normal(4)
    (s,), (gx, gu) = O.mapsum_adj(f, n, [x, u], [seed], [False, True], [True])
    # the summed output's directional derivative along u, by central differences
    h = 1e-6
    (sp,) = O.mapsum_eval(f, n, [x, u + h], [False, True], [True])
    (sm,) = O.mapsum_eval(f, n, [x, u - h], [False, True], [True])
    fd = seed @ (sp - sm) / (2 * h)
    assert abs(gu[0] - fd) <= 1e-7 * max(1.0, abs(fd))
    # mapped input: the per-instance sensitivity with the shared seed
    _, (rx, _) = O.adj_map(f, n, [x, np.tile(u, (n, 1))], [np.tile(seed, (n, 1))])
    assert np.array_equal(gx, rx)
    assert np.array_equal(s, O.mapsum_eval(f, n, [x, u], [False, True], [True])[0])


@pytest.mark.parametrize("par", ["serial", "openmp"])
def test_host_reduced_maps_match_oracle(ca, par):
    from casadi_amd import workloads as W
    f = W.cartpole_step()
    n = 777
    x, u = _step_inputs(n, 7)
    (ref,) = O.mapsum_eval(f, n, [x, u], [False, True], [True])
    F = f.map(n, [False, True], [True])                   # MapSum, serial
    assert F.name() == f"mapsum_{n}_{f.name()}"
    assert np.array_equal(np.asarray(F(x.T, u)).reshape(-1), ref)
    G = f.map("G", par, n, ["u"], [0])                    # reduced map by names / indices
    assert G.name() == "G"
    y = np.asarray(G(x.T, u[0])).reshape(-1)
    assert np.array_equal(y, ref)
    # nothing reduced: the plain map
    H = f.map("H", par, n, [], [])
    assert np.array_equal(np.asarray(H(x.T, np.tile(u, n))), np.asarray(f.map(n, par)(x.T, np.tile(u, n))))


def test_reduced_map_arguments(ca):
    from casadi_amd import workloads as W
    f = W.cartpole_step()
    with pytest.raises(ValueError):
        f.map(10, [True], [True])                         # one flag per input
    with pytest.raises(ValueError):
        f.map("F", "gpu", 10, [], [])                     # unknown parallelization
    F = f.map("F", "hip", 10, [1], [0])
    assert F.reduce_in == [False, True] and F.reduce_out == [True]
    with pytest.raises(TypeError):
        f.map("H", "serial", 10, [1], [0]).to_torch()     # the differentiable op runs on the GPU
