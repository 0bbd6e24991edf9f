"""Batched mapaccum on the GPU (loop kernels around one specialised step):
forward chain and checkpointed adjoint against the CPU oracle on the unrolled
chain, and against the flat rollout tape."""
import numpy as np
import pytest

import oracle_ref as O
from conftest import seed_of

pytestmark = pytest.mark.gpu
TOL = 1e-13


def _dev(*arrs):
    import torch
    return [torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda") for a in arrs]


def test_mapaccum_forward_and_adjoint(cuda):
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipMapAccum
    step = W.cartpole_cost_step()
    N, n = 20, 4096
    F = step.mapaccum("roll", N)
    m = HipMapAccum.from_function(F)
    rng = np.random.default_rng(seed_of(2))
    x0 = np.vstack([rng.uniform(-1, 1, (4, n)), np.zeros((1, n))])
    U = rng.uniform(-3, 3, (N, n))
    seed = rng.standard_normal((5 * N, n))
    print(f"[case] mapaccum N={N} n={n}", flush=True)
    (y,) = m.eval(_dev(x0, U), n)
    (ref,) = O.map_eval(F, n, [x0.T, U.T])
    assert O.rel_err(y.cpu().numpy(), ref.T) <= TOL
    (_,), (r0, rU) = O.adj_map(F, n, [x0.T, U.T], [seed.T])
    grads = {}
    for K in (0, 5, 7):
        (o,), (g0, gU) = m.fwd_adj(_dev(x0, U), _dev(seed), n, checkpoint_steps=K)
        assert np.array_equal(o.cpu().numpy(), y.cpu().numpy())
        assert O.rel_err(g0.cpu().numpy(), r0.T) <= TOL and O.rel_err(gU.cpu().numpy(), rU.T) <= TOL
        grads[K] = (g0.cpu().numpy(), gU.cpu().numpy())
    # the checkpoint interval does not change a bit
    for K in (5, 7):
        assert all(np.array_equal(a, b) for a, b in zip(grads[K], grads[0]))


def test_mapaccum_rollout_matches_flat_tape(cuda):
    """The 100-step rollout cost gradient from the mapaccum (checkpoints every
    10 steps) equals the flat rollout tape's to rounding (per-instance scaled
    median at 1e-13; the trajectories are chaotic)."""
    import torch
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipMapAccum, HipTape
    N, n = 100, 8192
    m = HipMapAccum(W.cartpole_cost_step(), N)
    rng = np.random.default_rng(seed_of(2))
    x0, U = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (N, n))
    X0c = np.vstack([x0, np.zeros((1, n))])
    # seeds: the cost carried to the end has cotangent 1, the final state 2 x_N
    (y,) = m.eval(_dev(X0c, U), n)
    xN = y[5 * (N - 1):5 * (N - 1) + 4]
    S = torch.zeros((5 * N, n), dtype=torch.float64, device="cuda")
    S[5 * (N - 1):5 * (N - 1) + 4] = 2 * xN
    S[5 * (N - 1) + 4] = 1.0
    (_,), (g0, gU) = m.fwd_adj(_dev(X0c, U), [S], n, checkpoint_steps=10)
    f = W.cartpole_rollout(N)
    (J,), (fx, fu) = HipTape(f, engine="jit").fwd_adj(_dev(x0, U), _dev(np.ones((1, n))), n,
                                                       checkpoint_every=(f.n_instructions() + 9) // 10)
    Jm = (y[5 * (N - 1) + 4] + (xN ** 2).sum(0)).cpu().numpy()
    assert O.rel_err(Jm, J.cpu().numpy()[0]) <= 1e-12
    g = np.concatenate([g0.cpu().numpy()[:4], gU.cpu().numpy()]).T
    gf = np.concatenate([fx.cpu().numpy(), fu.cpu().numpy()]).T
    per = np.max(np.abs(g - gf) / np.maximum(np.max(np.abs(gf), axis=1, keepdims=True), 1.0), axis=1)
    print(f"[mapaccum] vs flat tape: median {np.median(per):.2e} max {per.max():.2e}", flush=True)
    assert np.median(per) <= TOL


def test_mapaccum_behind_autograd(cuda):
    """Function.mapaccum(...).to_torch: forward stores the state every 10
    steps, backward runs the loop adjoint from those checkpoints; equal to the
    fused launch bit for bit."""
    import torch
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipMapAccum
    N, n = 100, 4096
    F = W.cartpole_cost_step().mapaccum("roll", N)
    T = F.to_torch(checkpoint_every=10)
    rng = np.random.default_rng(seed_of(2))
    XC0 = torch.tensor(np.vstack([rng.uniform(-1, 1, (4, n)), np.zeros((1, n))]), dtype=torch.float64,
                       device="cuda", requires_grad=True)
    U = torch.tensor(rng.uniform(-10, 10, (N, n)), dtype=torch.float64, device="cuda", requires_grad=True)
    Y = T(XC0, U)
    last = Y[5 * (N - 1):]
    cost = last[4] + (last[:4] ** 2).sum(0)          # running cost + |x_N|^2
    cost.sum().backward()
    S = torch.zeros((5 * N, n), dtype=torch.float64, device="cuda")
    S[5 * (N - 1):5 * (N - 1) + 4] = 2 * last[:4].detach()
    S[5 * (N - 1) + 4] = 1.0
    (Yf,), (g0, gU) = HipMapAccum.from_function(F).fwd_adj([XC0.detach(), U.detach()], [S], n, checkpoint_steps=10)
    assert torch.equal(Y.detach(), Yf)
    assert torch.equal(XC0.grad, g0) and torch.equal(U.grad, gU)


def test_mapaccum_split_reverse_is_bit_identical(cuda, monkeypatch):
    """The reverse run per segment (each launch: one segment's reverse steps
    beside the recompute of the segment before it) against the single
    fused kernel (CXH_JIT_SPLIT=0): the same bits, fused and from stored
    checkpoints."""
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipMapAccum
    N, n, K = 30, 3000, 4
    m = HipMapAccum(W.cartpole_cost_step(), N)
    rng = np.random.default_rng(seed_of(2) + 7)
    X0 = np.vstack([rng.uniform(-1, 1, (4, n)), np.zeros((1, n))])
    U = rng.uniform(-3, 3, (N, n))
    S = rng.standard_normal((5 * N, n))
    got = {}
    for split in ("1", "0"):
        monkeypatch.setenv("CXH_JIT_SPLIT", split)
        print(f"[case] mapaccum split={split} N={N} K={K} n={n}", flush=True)
        (o,), (g0, gU) = m.fwd_adj(_dev(X0, U), _dev(S), n, checkpoint_steps=K)
        got[split] = [x.cpu().numpy() for x in (o, g0, gU)]
    assert all(np.array_equal(a, b) for a, b in zip(got["1"], got["0"]))


def test_mapaccum_fast_divisions_on_special_dividends(cuda, monkeypatch):
    """The split sweep's kernels divide the step's constants (1.1) through
    sxh_div_c and re-run a loop with IEEE division when a lane's dividend
    leaves the exact range; the fused kernel (CXH_JIT_SPLIT=0) divides as
    IEEE 754 does throughout.  Lanes whose controls or states make the
    dividends subnormal (5e-320), huge (1e300), infinite, NaN or zero, among
    ordinary lanes: the same bits, NaN where NaN, with and without the fast
    division (CXH_JIT_FAST_DIV=0)."""
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipMapAccum
    N, n, K = 12, 640, 4
    rng = np.random.default_rng(seed_of(2) + 17)
    X0 = np.vstack([rng.uniform(-1, 1, (4, n)), np.zeros((1, n))])
    U = rng.uniform(-3, 3, (N, n))
    for lane, u in ((64, 5e-320), (65, -5e-320), (130, 1e300), (200, np.inf), (260, np.nan)):
        X0[:4, lane] = 0.0
        U[:, lane] = u
    X0[:, 320:384] = 0.0            # a wavefront of zeros: exact without the re-run
    U[:, 320:384] = 0.0
    X0[1, 400] = np.nan
    U[5, 450] = 5e-320              # one step of one lane
    S = rng.standard_normal((5 * N, n))

    def same(a, b):
        return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))
    got = {}
    for name, split, fast in (("split", "1", "1"), ("fused", "0", "1"), ("split_ieee", "1", "0")):
        monkeypatch.setenv("CXH_JIT_SPLIT", split)
        monkeypatch.setenv("CXH_JIT_FAST_DIV", fast)
        m = HipMapAccum(W.cartpole_cost_step(), N)
        print(f"[case] mapaccum special dividends {name} N={N} K={K} n={n}", flush=True)
        (o,), (g0, gU) = m.fwd_adj(_dev(X0, U), _dev(S), n, checkpoint_steps=K)
        (e,) = m.eval(_dev(X0, U), n)
        got[name] = [x.cpu().numpy() for x in (o, g0, gU, e)]
    for other in ("fused", "split_ieee"):
        assert all(same(a, b) for a, b in zip(got["split"], got[other])), other
    assert np.isnan(got["split"][0]).any() and np.isinf(got["split"][0]).any()
