"""GPU parity of flat tapes rerolled into loops (jit_roll.hip).

A rerolled tape runs one loop body per iteration instead of straight-line
segments, with the same operations on the same operands and every adjoint
accumulated in the same order, so values and adjoints must equal the
straight-line kernels' (CXH_JIT_ROLL=0) and the interpreter's bit for bit.
The oracle check is the 1e-13 bar per nonzero where the workload is not
chaotic, and the rounding envelope where it is (the 100-step cart-pole
rollout with |u| <= 10; see test_gpu_tape.py).
"""
import numpy as np
import pytest

import oracle_ref as O
from conftest import seed_of
from test_jit_roll import division_inputs, division_rollout, parametric_rollout, second_order

pytestmark = pytest.mark.gpu
TOL = 1e-13


def _dev(*arrs):
    import torch
    return [torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda") for a in arrs]


def _np(ts):
    return [t.cpu().numpy() for t in ts]


def _sweeps(f, ins, seeds, every, roll, monkeypatch):
    """Fused forward + adjoint, the split form (forward storing checkpoints,
    reverse from them) and forward-only evaluation on the specialised engine."""
    from casadi_amd.hip import HipTape
    monkeypatch.setenv("CXH_JIT_ROLL", roll)
    n = ins[0].shape[1]
    t = HipTape(f, engine="jit")
    assert t.roll_plan(every)["enabled"] == int(roll)
    outs, g = t.fwd_adj(_dev(*ins), _dev(*seeds), checkpoint_every=every)
    o2, ck = t.fwd_ckpt(_dev(*ins), n, every)
    g2 = t.rev_ckpt(_dev(*ins), _dev(*seeds), n, every, ck)
    ev = t.eval(_dev(*ins))
    return _np(outs) + _np(g) + _np(o2) + _np(g2) + _np(ev)


def test_rollout_full_batch_rolled_flat_interpreter(cuda, monkeypatch):
    """Config 3 at its full size: the 100-step rollout cost + gradient over
    32768 instances, checkpoints every 10 steps.  Rolled, straight-line and
    interpreter results are the same bits; the oracle within the envelope."""
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_rollout(100)
    n = 32768
    every = (f.n_instructions() + 9) // 10
    rng = np.random.default_rng(seed_of(2))
    x0, us = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (100, n))
    seed = np.ones((1, n))
    print(f"[case] rolled rollout100 n={n} every={every}", flush=True)
    plan = HipTape(f, engine="jit").roll_plan(every)
    assert plan["ok"] and plan["iterations"] == 99 and plan["K"] == 10
    rolled = _sweeps(f, [x0, us], [seed], every, "1", monkeypatch)
    flat = _sweeps(f, [x0, us], [seed], every, "0", monkeypatch)
    assert all(np.array_equal(a, b) for a, b in zip(rolled, flat))
    ti = HipTape(f)
    (Ji,), gi = ti.fwd_adj(_dev(x0, us), _dev(seed), checkpoint_every=every)
    assert np.array_equal(rolled[0], Ji.cpu().numpy())
    assert np.array_equal(rolled[1], gi[0].cpu().numpy()) and np.array_equal(rolled[2], gi[1].cpu().numpy())
    # the oracle on a sample: median per nonzero at 1e-13, the maximum inside
    # the 1-ulp envelope of the chaotic trajectories
    m = 512
    (Jo,), (ox, ou) = O.adj_map(f, m, [x0[:, :m].T, us[:, :m].T], [seed[:, :m].T])
    with O.perturbed_transcendentals():
        _, (px, pu) = O.adj_map(f, m, [x0[:, :m].T, us[:, :m].T], [seed[:, :m].T])
    g = np.concatenate([rolled[1][:, :m], rolled[2][:, :m]]).T
    go = np.concatenate([ox, ou], axis=1)
    env = np.abs(np.concatenate([px, pu], axis=1) - go) / np.maximum(np.abs(go), 1.0)
    err = np.abs(g - go) / np.maximum(np.abs(go), 1.0)
    print(f"[rollout100] per nonzero: median {np.median(err):.2e} max {err.max():.2e}; envelope {env.max():.2e}",
          flush=True)
    assert np.median(err) <= TOL
    assert err.max() <= max(TOL, 2 * env.max())


def test_short_rollout_full_batch_per_nonzero():
    """A 5-step rollout over the full 32768 batch, every gradient nonzero
    held to 1e-13 of the oracle, unscaled: in five steps one ulp of a sin/cos
    cannot grow past the bar (moving every transcendental result of the
    oracle by one ulp moves the gradient by 2.7e-14 at most on this batch;
    at ten steps by 2.1e-13)."""
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    f = W.cartpole_rollout(5)
    n = 32768
    rng = np.random.default_rng(seed_of(2) + 10)
    x0, us = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (5, n))
    seed = rng.standard_normal((1, n))
    print(f"[case] rollout5 n={n} per nonzero", flush=True)
    t = HipTape(f, engine="jit")
    (J,), (gx, gu) = t.fwd_adj(_dev(x0, us), _dev(seed), checkpoint_every=(f.n_instructions() + 1) // 2)
    (Jo,), (ox, ou) = O.adj_map(f, n, [x0.T, us.T], [seed.T])
    assert O.rel_err(J.cpu().numpy(), Jo.T) <= TOL
    assert O.rel_err(gx.cpu().numpy(), ox.T) <= TOL
    assert O.rel_err(gu.cpu().numpy(), ou.T) <= TOL


@pytest.mark.parametrize("case", ["parametric", "order2"])
def test_rolled_structures_match_interpreter(cuda, monkeypatch, case):
    """A parameter passed through every iteration with an output per
    iteration, and a second-order recurrence rerolled two steps per body:
    rolled, straight-line and interpreter results are the same bits, the
    oracle within 1e-13."""
    from casadi_amd.hip import HipTape
    f, every = (parametric_rollout(24), 13 * 5) if case == "parametric" else (second_order(150), 80)
    n = 1000
    rng = np.random.default_rng(seed_of(2) + len(case))
    ins = [rng.uniform(-1, 1, (f.nnz_in(i), n)) for i in range(f.n_in())]
    seeds = [rng.standard_normal((f.nnz_out(j), n)) for j in range(f.n_out())]
    print(f"[case] rolled {case} n={n} every={every}", flush=True)
    assert HipTape(f, engine="jit").roll_plan(every)["ok"]
    rolled = _sweeps(f, ins, seeds, every, "1", monkeypatch)
    flat = _sweeps(f, ins, seeds, every, "0", monkeypatch)
    assert all(np.array_equal(a, b) for a, b in zip(rolled, flat))
    ti = HipTape(f)
    outs, g = ti.fwd_adj(_dev(*ins), _dev(*seeds), checkpoint_every=every)
    nout = f.n_out()
    for a, b in zip(rolled[:nout + f.n_in()], _np(outs) + _np(g)):
        assert np.array_equal(a, b)
    ro, rg = O.adj_map(f, n, [a.T for a in ins], [s.T for s in seeds])
    for a, b in zip(rolled[:nout], ro):
        assert O.rel_err(a, b.T) <= TOL
    for a, b in zip(rolled[nout:nout + f.n_in()], rg):
        assert O.rel_err(a, b.T) <= TOL


def test_fused_call_captured_in_a_graph(cuda):
    """The whole fused call of the rerolled rollout (checkpointing forward +
    the paired reverse launches) captured once in a HIP graph (torch's
    CUDAGraph: the library launches on the capturing stream, allocates
    nothing after its first call, records no memset) and replayed with new
    inputs copied in between, eager work and synchronisations between the
    replays: every replay gives the bits of a direct call."""
    import torch
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_rollout(20)
    n = 4096
    every = (f.n_instructions() + 3) // 4
    rng = np.random.default_rng(seed_of(2) + 77)
    t = HipTape(f, engine="jit")
    X0 = torch.zeros((4, n), dtype=torch.float64, device="cuda")
    U = torch.zeros((20, n), dtype=torch.float64, device="cuda")
    S = torch.ones((1, n), dtype=torch.float64, device="cuda")
    print(f"[case] graph capture of the fused rolled call n={n}", flush=True)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        t.fwd_adj([X0, U], [S], n, checkpoint_every=every)     # warm-up: compiles, sizes the buffers
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        (Jg,), (gxg, gug) = t.fwd_adj([X0, U], [S], n, checkpoint_every=every)
    for rep in range(3):
        X0.copy_(torch.tensor(rng.uniform(-1, 1, (4, n)), device="cuda"))
        U.copy_(torch.tensor(rng.uniform(-10, 10, (20, n)), device="cuda"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        (Je,), (gxe, gue) = t.fwd_adj([X0.clone(), U.clone()], [S], n, checkpoint_every=every)
        torch.cuda.synchronize()
        assert torch.equal(Jg, Je) and torch.equal(gxg, gxe) and torch.equal(gug, gue), rep


def test_fast_constant_divisions_special_dividends(cuda, monkeypatch):
    """The rerolled loop bodies divide by constants through sxh_div_c (the
    folded reciprocal and two FMA corrections, v_div_fixup for zeros,
    infinities and NaN) and re-run an iteration with IEEE division when a
    lane's dividend leaves the exact range.  Dividends that misround or
    overflow to NaN without that re-run, zeros, signed zeros, subnormals,
    infinities and NaN, from the start or at one step, in some wavefronts
    and not in others: the same bits as the straight-line kernels and the
    interpreter (IEEE division), NaN where NaN, for values, adjoints, the
    split form and evaluation."""
    from casadi_amd.hip import HipTape
    n_steps, n, every = 40, 1000, 9 * 4
    f = division_rollout(n_steps)
    rng = np.random.default_rng(seed_of(2) + 91)
    ins = division_inputs(n_steps, n, rng)
    seeds = [rng.standard_normal((f.nnz_out(0), n))]
    print(f"[case] fast constant divisions n={n} every={every}", flush=True)
    assert "sxh::sxh_div_c(" in HipTape(f, engine="jit").roll_source(every, 4)
    rolled = _sweeps(f, ins, seeds, every, "1", monkeypatch)
    flat = _sweeps(f, ins, seeds, every, "0", monkeypatch)

    def same(a, b):
        return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))
    assert all(same(a, b) for a, b in zip(rolled, flat))
    outs, g = HipTape(f).fwd_adj(_dev(*ins), _dev(*seeds), checkpoint_every=every)
    for a, b in zip(rolled[:1 + f.n_in()], _np(outs) + _np(g)):
        assert same(a, b)
    assert np.isnan(rolled[0]).any() and np.isinf(rolled[0]).any()


def test_lean_kernels_keep_the_bits(cuda, monkeypatch):
    """Batches above 128 instances per SIMD run the lean loop kernels (no
    one-iteration-ahead loads, IEEE division: fewer registers, more resident
    wavefronts); CXH_JIT_ROLL_FAT_MAX=0 forces them at any size.  Lean and
    full kernels give the same bits: the 20-step cart-pole rollout and the
    division rollout on special dividends, fused, split and evaluated."""
    from casadi_amd import workloads as W
    cases = [(W.cartpole_rollout(20), (W.cartpole_rollout(20).n_instructions() + 3) // 4, None),
             (division_rollout(40), 9 * 4, "special")]
    for f, every, kind in cases:
        n = 1000
        rng = np.random.default_rng(seed_of(2) + 55)
        if kind == "special":
            ins = division_inputs(40, n, rng)
        else:
            ins = [rng.uniform(-1, 1, (f.nnz_in(i), n)) for i in range(f.n_in())]
        seeds = [rng.standard_normal((f.nnz_out(j), n)) for j in range(f.n_out())]
        print(f"[case] lean kernels {f.name()} n={n}", flush=True)
        full = _sweeps(f, ins, seeds, every, "1", monkeypatch)
        monkeypatch.setenv("CXH_JIT_ROLL_FAT_MAX", "0")
        lean = _sweeps(f, ins, seeds, every, "1", monkeypatch)
        monkeypatch.delenv("CXH_JIT_ROLL_FAT_MAX")
        assert all(bool(np.all((a == b) | (np.isnan(a) & np.isnan(b)))) for a, b in zip(full, lean))
