"""The checkpointed reverse sweep behind torch.autograd (cart-pole 100-step
rollout cost, batch 32768, checkpoints every 10 steps).

Checked against: the fused forward+adjoint launch (bit-exact), CasADi's
reverse(1) Function evaluated by the CPU oracle on 64 sampled instances, and
central finite differences (h = 1e-6) on the same 64 instances.
"""
import numpy as np
import pytest

import oracle_ref as O
from conftest import seed_of

pytestmark = pytest.mark.gpu
TOL = 1e-13
N_STEPS = 100


@pytest.fixture(scope="module")
def rollout():
    from casadi_amd import workloads as W
    return W.cartpole_rollout(N_STEPS)


def test_autograd_rollout_gradient(rollout, cuda):
    import torch
    f = rollout
    n = 32768
    every = (f.n_instructions() + 9) // 10          # 10 segments = every 10 steps
    rng = np.random.default_rng(seed_of(2))
    x0, us = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (N_STEPS, n))
    print(f"[case] autograd rollout n={n} every={every}", flush=True)
    F = f.to_torch(checkpoint_every=every)
    X0 = torch.tensor(x0, dtype=torch.float64, device="cuda", requires_grad=True)
    U = torch.tensor(us, dtype=torch.float64, device="cuda", requires_grad=True)
    J = F(X0, U)
    assert J.shape == (1, n)
    w = torch.tensor(rng.standard_normal((1, n)), dtype=torch.float64, device="cuda")
    (J * w).sum().backward()
    gx, gu = X0.grad.cpu().numpy(), U.grad.cpu().numpy()
    # the fused launch with the same seeds gives the same bits
    (Jf,), (fx, fu) = F.tape.fwd_adj([X0.detach(), U.detach()], [w], n, checkpoint_every=every)
    assert np.array_equal(J.detach().cpu().numpy(), Jf.cpu().numpy())
    assert np.array_equal(gx, fx.cpu().numpy()) and np.array_equal(gu, fu.cpu().numpy())
    # the interpreter's fused adjoint agrees bit for bit
    from casadi_amd.hip import HipTape
    (Ji,), (ix, iu) = HipTape(f).fwd_adj([X0.detach(), U.detach()], [w], n, checkpoint_every=every)
    assert np.array_equal(gx, ix.cpu().numpy()) and np.array_equal(gu, iu.cpu().numpy())
    # CasADi reverse(1) on 64 sampled instances (CPU oracle VM).  The rollout is
    # chaotic: a 1-ulp difference in one sin/cos (OCML vs glibc) grows along
    # the trajectory, so errors are per instance, scaled by its gradient norm.
    # Bars: the median meets 1e-13; the maximum stays inside the rounding
    # envelope -- how far a CPU evaluation moves when every transcendental
    # result is nudged by one ulp (env for the numeric sweep, env_r for the
    # symbolic reverse(1) tape).  GPU and CPU each round independently of the
    # exact value, so their distance is bounded by the sum of two envelopes:
    # 2 env against the same numeric sweep, env + env_r against reverse(1).
    idx = np.random.default_rng(5).choice(n, 64, replace=False)
    wv = w.cpu().numpy()[:, idx]
    R = f.reverse(1)
    ins = [x0[:, idx].T, us[:, idx].T]
    rx, ru = O.map_eval(R, 64, ins + [None, wv.T])
    gr = np.concatenate([rx, ru], axis=1)
    g = np.concatenate([gx[:, idx], gu[:, idx]]).T
    scale = np.maximum(np.max(np.abs(gr), axis=1, keepdims=True), 1.0)
    (Jo,), (ox, ou) = O.adj_map(f, 64, ins, [wv.T])
    go = np.concatenate([ox, ou], axis=1)
    with O.perturbed_transcendentals():
        (Jp,), (px, pu) = O.adj_map(f, 64, ins, [wv.T])
        qx, qu = O.map_eval(R, 64, ins + [None, wv.T])
    env = float(np.max(np.abs(np.concatenate([px, pu], axis=1) - go) / scale))
    env_r = float(np.max(np.abs(np.concatenate([qx, qu], axis=1) - gr) / scale))
    per_num = np.max(np.abs(g - go) / scale, axis=1)
    per = np.max(np.abs(g - gr) / scale, axis=1)
    print(f"[autograd] vs numeric sweep: median {np.median(per_num):.2e} max {per_num.max():.2e} (2 env {2 * env:.2e}); "
          f"vs reverse(1): median {np.median(per):.2e} max {per.max():.2e} (env {env + env_r:.2e})", flush=True)
    assert np.median(per_num) <= TOL and per_num.max() <= max(TOL, 2 * env)
    assert np.median(per) <= TOL and per.max() <= max(TOL, env + env_r)
    # central finite differences of the weighted cost
    h = 1e-6
    base = np.concatenate([x0[:, idx], us[:, idx]]).T
    m = base.shape[1]
    pert = np.repeat(base, 2 * m, axis=0)
    for j in range(m):
        pert[j::2 * m, j] += h
        pert[m + j::2 * m, j] -= h
    (vals,) = O.map_eval(f, pert.shape[0], [pert[:, :4], pert[:, 4:]])
    vals = vals.reshape(64, 2, m) * wv.T[:, :, None]
    fd = (vals[:, 0, :] - vals[:, 1, :]) / (2 * h)
    fd_err = np.max(np.abs(fd - g) / np.maximum(np.max(np.abs(g), axis=1, keepdims=True), 1.0), axis=1)
    print(f"[autograd] central FD: median {np.median(fd_err):.2e} max {fd_err.max():.2e}", flush=True)
    assert np.median(fd_err) <= 1e-6 and fd_err.max() <= 1e-2


def test_autograd_composes_with_torch(cuda):
    """Gradients flow through torch ops on both sides and only to the inputs
    that require them."""
    import torch
    from casadi_amd import workloads as W
    f = W.cartpole_step()
    n = 500
    rng = np.random.default_rng(seed_of(0))
    p = torch.tensor(rng.uniform(-1, 1, (4, n)), dtype=torch.float64, device="cuda", requires_grad=True)
    u = torch.tensor(rng.uniform(-10, 10, (1, n)), dtype=torch.float64, device="cuda")
    F = f.to_torch()
    y = F(2.0 * p, u)
    loss = (y ** 2).sum()
    loss.backward()
    assert u.grad is None
    # d/dp sum(y^2) = 2 * J_x(2p)^T (2 y)
    seed = (2 * y).detach()
    R = f.reverse(1)
    x = (2 * p).detach().cpu().numpy()
    rx, _ = O.map_eval(R, n, [x.T, u.cpu().numpy().T, None, seed.cpu().numpy().T])
    assert O.rel_err(p.grad.cpu().numpy(), 2 * rx.T) <= TOL
