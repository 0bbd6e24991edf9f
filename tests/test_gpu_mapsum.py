"""GPU parity of reduced maps (cxh_map_eval_sum, cxh_map_fwd_adj_sum).

Per-instance values are held to the oracle at the usual 1e-13 bar.  A summed
block is held to math.fsum of the GPU's own per-instance values within the
rounding bound of the kernel's fixed summation order (per-thread chains over
a chunk, 6 butterfly levels, 2 pairwise levels, then the same over the
chunks), and to the oracle's instance-by-instance MapSum within both bounds.
Sums of integers are exact in any order, so there the GPU equals the oracle bit
for bit.
"""
import math

import numpy as np
import pytest

import oracle_ref as O
from conftest import seed_of

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53
TOL = 1e-13


@pytest.fixture(scope="module")
def ca():
    import casadi_amd
    return casadi_amd


def _dev(*arrs):
    import torch
    return [torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda") for a in arrs]


def _tree_terms(n):
    """Longest chain of additions in the kernel's order for n instances."""
    nchunk = min(1024, max(1, -(-n // 2048)))
    chunk = max(1, -(-n // nchunk))
    return -(-chunk // 256) + 8 + -(-nchunk // 256) + 8


def _check_sum(s, per):
    """s (nnz,) against the exact sum of per (nnz, n) columns."""
    exact = np.array([math.fsum(row) for row in per])
    bound = _tree_terms(per.shape[1]) * EPS * np.sum(np.abs(per), axis=1)
    assert np.all(np.abs(s - exact) <= bound), (np.abs(s - exact), bound)


@pytest.mark.parametrize("engine", ["interp", "jit"])
@pytest.mark.parametrize("n", [0, 1, 63, 1000, 65536])
def test_eval_sum_shared_input_summed_output(cuda, engine, n):
    import torch
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_step()
    rng = np.random.default_rng(seed_of(30) + n)
    x, u = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (1,))
    print(f"[case] eval_sum cartpole engine={engine} n={n}", flush=True)
    t = HipTape(f, engine=engine)
    X, U = _dev(x, u)
    (s,) = t.eval_sum([X, U], n, [False, True], [True])
    assert s.shape == (4,)
    if n == 0:
        assert torch.count_nonzero(s) == 0
        return
    (per,) = t.eval([X, U.reshape(1, 1).expand(1, n).contiguous()], n)
    per = per.cpu().numpy()
    assert O.rel_err(per, O.map_eval(f, n, [x.T, np.tile(u, (n, 1))])[0].T) <= TOL
    s = s.cpu().numpy()
    _check_sum(s, per)
    (ref,) = O.mapsum_eval(f, n, [x.T, u], [False, True], [True])
    bound = (_tree_terms(n) + n) * EPS * np.sum(np.abs(per), axis=1) + TOL * np.sum(np.maximum(np.abs(per), 1), axis=1)
    assert np.all(np.abs(s - ref) <= bound)
    (s2,) = t.eval_sum([X, U], n, [False, True], [True])   # fixed order: same bits again
    assert np.array_equal(s2.cpu().numpy(), s)


def test_eval_sum_engines_agree_and_integer_sums_are_exact(ca, cuda):
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_step()
    n = 100003
    rng = np.random.default_rng(seed_of(31))
    x, u = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (1,))
    X, U = _dev(x, u)
    print(f"[case] eval_sum engines n={n}", flush=True)
    (si,) = HipTape(f, engine="interp").eval_sum([X, U], n, [False, True], [True])
    (sj,) = HipTape(f, engine="jit").eval_sum([X, U], n, [False, True], [True])
    assert np.array_equal(si.cpu().numpy(), sj.cpu().numpy())
    a = ca.SX.sym("a", 2)
    b = ca.SX.sym("b")
    g = ca.Function("g", [a, b], [ca.vertcat(ca.floor(a[0] * 8) * b, ca.floor(a[1] * 3) + 1), a[0] * b])
    av = rng.uniform(-50, 50, (2, n))
    for engine in ("interp", "jit"):
        A, B = _dev(av, np.array([2.0]))
        s_int, y = HipTape(g, engine=engine).eval_sum([A, B], n, [False, True], [True, False])
        (ref_int, ref_y) = O.mapsum_eval(g, n, [av.T, np.array([2.0])], [False, True], [True, False])
        assert np.array_equal(s_int.cpu().numpy(), ref_int)
        assert np.array_equal(y.cpu().numpy(), ref_y.T)


@pytest.mark.parametrize("engine", ["interp", "jit"])
def test_fwd_adj_sum_shared_parameters(cuda, engine):
    """Gradient of a summed output w.r.t. a shared input = sum over instances
    of the per-instance sensitivities; mapped sensitivities unchanged."""
    import torch
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_step()
    n = 50000
    rng = np.random.default_rng(seed_of(32))
    x, u = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (1,))
    seed = rng.standard_normal(4)
    X, U, S = _dev(x, u, seed)
    print(f"[case] fwd_adj_sum cartpole engine={engine} n={n}", flush=True)
    t = HipTape(f, engine=engine)
    (s,), (gx, gu) = t.fwd_adj_sum([X, U], [S], n, [False, True], [True])
    assert s.shape == (4,) and gu.shape == (1,) and gx.shape == (4, n)
    Ut = U.reshape(1, 1).expand(1, n).contiguous()
    St = S.reshape(4, 1).expand(4, n).contiguous()
    (per,), (ax, au) = t.fwd_adj([X, Ut], [St], n)
    assert torch.equal(gx, ax)
    _check_sum(s.cpu().numpy(), per.cpu().numpy())
    _check_sum(gu.cpu().numpy(), au.cpu().numpy())
    (ro,), (rx, ru) = O.mapsum_adj(f, n, [x.T, u], [seed], [False, True], [True])
    assert O.rel_err(gx.cpu().numpy(), rx.T) <= TOL
    aun = au.cpu().numpy()
    bound = (_tree_terms(n) + n) * EPS * np.sum(np.abs(aun)) + TOL * np.sum(np.maximum(np.abs(aun), 1))
    assert abs(gu.item() - ru[0]) <= bound


def test_fwd_adj_sum_rollout_checkpointed(cuda):
    """A 20-step rollout whose control sequence is shared by every instance
    (one control trajectory, many initial states): the gradient w.r.t. the
    shared controls is summed, checkpointed reverse sweep on both engines."""
    import torch
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_rollout(20)
    n = 4096
    rng = np.random.default_rng(seed_of(33))
    x0, us = rng.uniform(-1, 1, (4, n)), rng.uniform(-3, 3, (20,))
    X0, US, S = _dev(x0, us, np.ones(1))
    every = (f.n_instructions() + 3) // 4
    print(f"[case] fwd_adj_sum rollout20 n={n}", flush=True)
    res = {}
    for engine in ("interp", "jit"):
        t = HipTape(f, engine=engine)
        (J,), (gx, gu) = t.fwd_adj_sum([X0, US], [S], n, [False, True], [True], checkpoint_every=every)
        (Jp,), (ax, au) = t.fwd_adj([X0, US.reshape(20, 1).expand(20, n).contiguous()],
                                    [S.reshape(1, 1).expand(1, n).contiguous()], n, checkpoint_every=every)
        assert torch.equal(gx, ax)
        _check_sum(J.cpu().numpy(), Jp.cpu().numpy())
        _check_sum(gu.cpu().numpy(), au.cpu().numpy())
        res[engine] = (J.cpu().numpy(), gu.cpu().numpy())
    assert np.array_equal(res["interp"][0], res["jit"][0]) and np.array_equal(res["interp"][1], res["jit"][1])


def test_reduced_map_function_api(cuda):
    """Function.map(name, "hip", n, reduce_in, reduce_out) on host arrays and
    on device tensors, against the host MapSum."""
    from casadi_amd import workloads as W
    f = W.cartpole_step()
    n = 3000
    rng = np.random.default_rng(seed_of(34))
    x, u = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10)
    F = f.map("F", "hip", n, ["u"], [0])
    yh = np.asarray(F(x, u)).reshape(-1)
    (ref,) = O.mapsum_eval(f, n, [x.T, np.array([u])], [False, True], [True])
    assert np.all(np.abs(yh - ref) <= 1e-12 * np.maximum(np.abs(ref), 1) * np.sqrt(n))
    X, U = _dev(x, np.array([u]))
    (yd,) = F.call([X, U])
    assert np.array_equal(yd.cpu().numpy(), yh)


def test_reduced_map_autograd(cuda):
    """MapSumFunction.to_torch: gradients of a shared input come back summed
    over the instances.  Checked against the per-instance adjoints of the
    unreduced map on the replicated input (summed exactly), against the
    oracle's MapSum adjoint, and by torch's finite-difference gradcheck."""
    import torch
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_step()
    n = 2000
    rng = np.random.default_rng(seed_of(35))
    x, u = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (1,))
    w = rng.standard_normal(4)
    print(f"[case] mapsum autograd n={n}", flush=True)
    X, U, Wt = _dev(x, u, w)
    t = HipTape(f, engine="jit")
    Ut = U.reshape(1, 1).expand(1, n).contiguous()
    for red_out in ([], [0]):
        T = f.map("F", "hip", n, ["u"], red_out).to_torch()
        Xg, Ug = X.clone().requires_grad_(), U.clone().requires_grad_()
        y = T(Xg, Ug)
        loss = (y * Wt).sum() if red_out else (y ** 2).sum()
        loss.backward()
        if red_out:
            seed = Wt.reshape(4, 1).expand(4, n).contiguous()
        else:
            (yp,) = t.eval([X, Ut], n)
            assert torch.equal(y, yp)
            seed = 2 * yp
        _, (ax, au) = t.fwd_adj([X, Ut], [seed], n)
        assert O.rel_err(Xg.grad.cpu().numpy(), ax.cpu().numpy()) <= TOL
        _check_sum(Ug.grad.cpu().numpy(), au.cpu().numpy())
        if red_out:
            _, (rx, ru) = O.mapsum_adj(f, n, [x.T, u], [w], [False, True], [True])
            assert O.rel_err(Xg.grad.cpu().numpy(), rx.T) <= TOL
            aun = au.cpu().numpy()
            bound = (_tree_terms(n) + n) * EPS * np.sum(np.abs(aun)) + TOL * np.sum(np.maximum(np.abs(aun), 1))
            assert abs(Ug.grad.item() - ru[0]) <= bound
    m = 5
    Tm = f.map("G", "hip", m, ["u"], [0]).to_torch()
    xs = torch.tensor(x[:, :m], dtype=torch.float64, device="cuda", requires_grad=True)
    us = torch.tensor(u, dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b: Tm(a, b), (xs, us), eps=1e-6, atol=1e-6, rtol=1e-5)
