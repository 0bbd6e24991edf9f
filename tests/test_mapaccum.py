"""mapaccum / fold on the host: the chain equals repeated calls
(casadi/core/function.cpp:661-754), and the rollout cost written as a
mapaccum of a cost-carrying step equals the unrolled rollout."""
import numpy as np
import pytest


def test_mapaccum_equals_repeated_calls():
    from casadi_amd import workloads as W
    f = W.cartpole_step()
    N = 12
    F = f.mapaccum("roll", N)
    assert F.sparsity_in(1).shape == (1, N) and F.sparsity_out(0).shape == (4, N)
    rng = np.random.default_rng(0)
    x0, U = rng.uniform(-1, 1, 4), rng.uniform(-5, 5, (1, N))
    X = F(x0, U)
    x = x0
    for k in range(N):
        x = f(x, U[0, k])
        assert np.array_equal(X[:, k:k + 1], x)
    assert np.array_equal(f.fold(N)(x0, U), x)
    with pytest.raises(ValueError):
        f.mapaccum("bad", 0)


def test_rollout_cost_as_mapaccum():
    from casadi_amd import workloads as W
    N = 15
    step = W.cartpole_cost_step()
    F = step.mapaccum(N)
    R = W.cartpole_rollout(N)
    rng = np.random.default_rng(1)
    x0, U = rng.uniform(-1, 1, 4), rng.uniform(-5, 5, (1, N))
    XC = F(np.append(x0, 0.0), U)
    J = XC[4, -1] + np.sum(XC[:4, -1] ** 2)
    Jr = float(np.asarray(R(x0, U.reshape(-1))).reshape(-1)[0])
    assert abs(J - Jr) <= 1e-13 * abs(Jr)
