"""GPU parity of the straight-line specialised kernels (engine 'jit').

The specialised kernels evaluate the same operations (op_fun_full /
op_der_full) in the same order as the interpreter, so forward values and
adjoints must agree bit for bit with the interpreter kernels, which are in
turn held to the CPU oracle in test_gpu_tape.py.  Every test also checks the
oracle directly at the 1e-13 bar.
"""
import numpy as np
import pytest

import oracle_ref as O
from conftest import seed_of

pytestmark = pytest.mark.gpu
TOL = 1e-13


@pytest.fixture(scope="module")
def ca():
    import casadi_amd
    return casadi_amd


def _dev(*arrs):
    import torch
    return [torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda") for a in arrs]


def _np(ts):
    return [t.cpu().numpy() for t in ts]


def _all_ops_function(ca):
    """One output per operation, chained through the inputs so that every
    partial derivative reaches an input."""
    a = ca.SX.sym("a")      # in (0.2, 0.8)
    b = ca.SX.sym("b")      # in (1.5, 3)
    c = ca.SX.sym("c", 2)   # in (-2, 2)
    c0, c1 = c[0], c[1]
    e = [a + c0, a - c1, a * c0, c1 / b, -c0, ca.exp(c0), ca.log(b), b ** c1, ca.constpow(b, 2.5),
         ca.sqrt(b), ca.sq(c0 + a), 2 * c1, ca.sin(c0 * b), ca.cos(c1 + a), ca.tan(a), ca.asin(a), ca.acos(a),
         ca.atan(c0), c0 < c1, c0 <= a, c0.eq(c1), c0.ne(a), ca.logic_not(c0 < 0), ca.logic_and(c0, c1),
         ca.logic_or(c0 < 0, c1), ca.floor(c0 * 3), ca.ceil(c1 * 3), ca.fmod(b * 5, a), ca.fabs(c0),
         ca.sign(c1), ca.copysign(b, c0), ca.if_else_zero(c0 < c1, b * c1), ca.erf(c0), ca.fmin(c0, c1),
         ca.fmax(a, c1), 1 / b, ca.sinh(c0), ca.cosh(c1), ca.tanh(c0 * 2), ca.asinh(c1), ca.acosh(b),
         ca.atanh(a), ca.atan2(c0, b), ca.erfinv(a - 0.5), ca.log1p(a), ca.expm1(c1), ca.hypot(c0, b)]
    return ca.Function("all_ops", [a, b, c], [ca.vertcat(*e)]), len(e)


def _all_ops_inputs(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.uniform(0.2, 0.8, (1, n)), rng.uniform(1.5, 3, (1, n)), rng.uniform(-2, 2, (2, n)))


def test_jit_all_operations_match_interpreter(ca, cuda):
    from casadi_amd.hip import HipTape
    f, m = _all_ops_function(ca)
    n = 3000
    a, b, c = _all_ops_inputs(n, seed_of(10))
    seed = np.random.default_rng(seed_of(11)).standard_normal((m, n))
    print(f"[case] jit all-ops n={n} ops={m}", flush=True)
    ti, tj = HipTape(f), HipTape(f, engine="jit")
    (yi,) = _np(ti.eval(_dev(a, b, c)))
    (yj,) = _np(tj.eval(_dev(a, b, c)))
    assert np.array_equal(yi, yj, equal_nan=True)
    ref = O.map_eval(f, n, [a.T, b.T, c.T])[0].T
    assert O.rel_err(yj, ref) <= TOL
    (oi,), gi = ti.fwd_adj(_dev(a, b, c), _dev(seed))
    (oj,), gj = tj.fwd_adj(_dev(a, b, c), _dev(seed))
    assert np.array_equal(oi.cpu().numpy(), oj.cpu().numpy(), equal_nan=True)
    for x, y in zip(_np(gi), _np(gj)):
        assert np.array_equal(x, y, equal_nan=True)
    R = f.reverse(1)
    rr = O.map_eval(R, n, [a.T, b.T, c.T, None, seed.T])
    for x, r in zip(_np(gj), rr):
        assert O.rel_err(x, r.T) <= TOL


@pytest.mark.parametrize("n", [1, 63, 1000, 65536])
def test_jit_cartpole_map(ca, cuda, n):
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_step()
    rng = np.random.default_rng(seed_of(0) + n)
    x, u = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (1, n))
    print(f"[case] jit cartpole map n={n}", flush=True)
    (yj,) = _np(HipTape(f, engine="jit").eval(_dev(x, u)))
    (yi,) = _np(HipTape(f).eval(_dev(x, u)))
    assert np.array_equal(yi, yj)
    assert O.rel_err(yj, O.map_eval(f, n, [x.T, u.T])[0].T) <= TOL


def test_jit_aos_layout_and_null_blocks(ca, cuda):
    """CasADi Map layout (instance-major) and NULL arg / res / seed blocks."""
    import torch
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape, CXH_AOS
    from casadi_amd import _lib as L
    f = W.cartpole_step()
    n = 777
    rng = np.random.default_rng(seed_of(12))
    x, u = rng.uniform(-1, 1, (n, 4)), rng.uniform(-10, 10, (n, 1))
    tj = HipTape(f, engine="jit")
    X, U = (torch.tensor(v, dtype=torch.float64, device="cuda") for v in (x, u))
    Y = torch.empty((n, 4), dtype=torch.float64, device="cuda")
    L.check(L.cxh_map_eval(tj._h, n, L.ptr_array([X.data_ptr(), U.data_ptr()]), L.ptr_array([Y.data_ptr()]),
                           CXH_AOS, n, torch.cuda.current_stream().cuda_stream))
    assert O.rel_err(Y.cpu().numpy(), O.map_eval(f, n, [x, u])[0]) <= TOL
    # u = NULL means zeros
    Y0 = torch.empty((n, 4), dtype=torch.float64, device="cuda")
    L.check(L.cxh_map_eval(tj._h, n, L.ptr_array([X.data_ptr(), None]), L.ptr_array([Y0.data_ptr()]),
                           CXH_AOS, n, torch.cuda.current_stream().cuda_stream))
    assert O.rel_err(Y0.cpu().numpy(), O.map_eval(f, n, [x, np.zeros((n, 1))])[0]) <= TOL
    # adjoint without primal outputs
    seed = np.random.default_rng(3).standard_normal((4, n))
    outs, (ax, au) = tj.fwd_adj(_dev(x.T, u.T), _dev(seed), res=False)
    R = f.reverse(1)
    rx, ru = O.map_eval(R, n, [x, u, None, seed.T])
    assert O.rel_err(ax.cpu().numpy(), rx.T) <= TOL and O.rel_err(au.cpu().numpy(), ru.T) <= TOL


@pytest.mark.parametrize("roll", ["1", "0"])
def test_jit_rollout_segments_match_interpreter(ca, cuda, monkeypatch, roll):
    """Checkpointed fused adjoint, split forward/reverse and forward-only
    evaluation of a 20-step rollout cut into 4 segments: as one loop body
    (roll = 1, the default) and as straight-line segment kernels."""
    monkeypatch.setenv("CXH_JIT_ROLL", roll)
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_rollout(20)
    n = 2048
    rng = np.random.default_rng(seed_of(2))
    x0, us = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (20, n))
    seed = np.ones((1, n))
    every = (f.n_instructions() + 3) // 4
    print(f"[case] jit rollout20 n={n} every={every} roll={roll}", flush=True)
    ti, tj = HipTape(f), HipTape(f, engine="jit")
    assert tj.roll_plan(every)["ok"] == 1 and tj.roll_plan(every)["enabled"] == int(roll)
    (Ji,), gi = ti.fwd_adj(_dev(x0, us), _dev(seed), checkpoint_every=every)
    (Jj,), gj = tj.fwd_adj(_dev(x0, us), _dev(seed), checkpoint_every=every)
    assert np.array_equal(Ji.cpu().numpy(), Jj.cpu().numpy())
    for x, y in zip(_np(gi), _np(gj)):
        assert np.array_equal(x, y)
    # one segment gives the same bits
    (J1,), g1 = tj.fwd_adj(_dev(x0, us), _dev(seed), checkpoint_every=0)
    for x, y in zip(_np(g1), _np(gj)):
        assert np.array_equal(x, y)
    # split through caller-owned (liveness-compacted) checkpoints
    outs, ck = tj.fwd_ckpt(_dev(x0, us), n, every)
    assert ck.numel() < ti.ckpt_size(n, every)
    g2 = tj.rev_ckpt(_dev(x0, us), _dev(seed), n, every, ck)
    assert np.array_equal(outs[0].cpu().numpy(), Jj.cpu().numpy())
    for x, y in zip(_np(g2), _np(gj)):
        assert np.array_equal(x, y)
    # oracle
    (Jo,), (ox, ou) = O.adj_map(f, n, [x0.T, us.T], [seed.T])
    g = np.concatenate(_np(gj)).T
    go = np.concatenate([ox, ou], axis=1)
    scale = np.maximum(np.max(np.abs(go), axis=1, keepdims=True), 1.0)
    assert np.median(np.max(np.abs(g - go) / scale, axis=1)) <= TOL
    # forward-only evaluation through the forward segment kernels
    (Je,) = _np(tj.eval(_dev(x0, us)))
    assert np.array_equal(Je, Jj.cpu().numpy())


def test_jit_jacobian_matches_interpreter(ca, cuda):
    """Specialised tangent sweeps (5 forward colours) for the dense 4x5
    Jacobian of the cart-pole step over 65536 instances."""
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipJacobian
    f = W.cartpole_step()
    n = 65536
    rng = np.random.default_rng(seed_of(0))
    x, u = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (1, n))
    print(f"[case] jit jacobian n={n}", flush=True)
    (yi,), ji = HipJacobian(f)(_dev(x, u))
    (yj,), jj = HipJacobian(f, engine="jit")(_dev(x, u))
    assert np.array_equal(yi.cpu().numpy(), yj.cpu().numpy())
    assert np.array_equal(ji.cpu().numpy(), jj.cpu().numpy())
    ref = np.concatenate(O.map_eval(f.jacobian(), n, [x.T, u.T, None]), axis=1).T
    assert O.rel_err(jj.cpu().numpy(), ref) <= TOL


@pytest.mark.parametrize("every", [900, 257])
def test_jit_split_reverse_ragged_segments(ca, cuda, monkeypatch, every):
    """The split reverse sweep (each launch: one segment's sub-block recompute
    + adjoint beside the boundary-state pass of the segment before it) with a
    short last segment of a single sub-block (every = 900: 900, 900, 900, 81
    instructions) or segments of one or two sub-blocks (every = 257): the
    bits of the unsplit kernels and of the interpreter, fused and from stored
    checkpoints, on a ragged batch."""
    monkeypatch.setenv("CXH_JIT_ROLL", "0")   # the straight-line segment kernels
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_rollout(20)
    n = 1000
    rng = np.random.default_rng(seed_of(2) + every)
    x0, us = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (20, n))
    seed = rng.standard_normal((1, n))
    print(f"[case] jit split rollout20 n={n} every={every}", flush=True)
    ti = HipTape(f)
    (Ji,), gi = ti.fwd_adj(_dev(x0, us), _dev(seed), checkpoint_every=every)
    got = {}
    for split in ("1", "0"):
        monkeypatch.setenv("CXH_JIT_SPLIT", split)
        tj = HipTape(f, engine="jit")
        (Jj,), gj = tj.fwd_adj(_dev(x0, us), _dev(seed), checkpoint_every=every)
        outs, ck = tj.fwd_ckpt(_dev(x0, us), n, every)
        g2 = tj.rev_ckpt(_dev(x0, us), _dev(seed), n, every, ck)
        got[split] = [Jj.cpu().numpy()] + _np(gj) + _np(g2)
    assert all(np.array_equal(a, b) for a, b in zip(got["1"], got["0"]))
    assert np.array_equal(got["1"][0], Ji.cpu().numpy())
    for x, y in zip(got["1"][1:3], _np(gi)):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("ipt", [2, 3, 8])
def test_jit_instances_per_thread(ca, cuda, monkeypatch, ipt):
    """CXH_JIT_IPT: each thread of the forward segment kernels takes `ipt`
    instances, one block width apart.  On a ragged batch (1000 = 15 blocks of
    64 + 40) the forward evaluation, the fused adjoint and the split
    checkpointed sweep give the bits of one instance per thread."""
    monkeypatch.setenv("CXH_JIT_ROLL", "0")   # the straight-line segment kernels
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_rollout(20)
    n = 1000
    every = (f.n_instructions() + 3) // 4
    rng = np.random.default_rng(seed_of(2) + ipt)
    x0, us = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (20, n))
    seed = rng.standard_normal((1, n))
    got = {}
    for k in ("1", str(ipt)):
        monkeypatch.setenv("CXH_JIT_IPT", k)
        print(f"[case] jit rollout20 n={n} every={every} ipt={k}", flush=True)
        tj = HipTape(f, engine="jit")
        (Je,) = _np(tj.eval(_dev(x0, us)))
        (Jj,), gj = tj.fwd_adj(_dev(x0, us), _dev(seed), checkpoint_every=every)
        outs, ck = tj.fwd_ckpt(_dev(x0, us), n, every)
        g2 = tj.rev_ckpt(_dev(x0, us), _dev(seed), n, every, ck)
        step = HipTape(W.cartpole_step(), engine="jit")
        (ys,) = _np(step.eval(_dev(x0, us[:1])))
        got[k] = [Je, Jj.cpu().numpy(), outs[0].cpu().numpy(), ys] + _np(gj) + _np(g2)
    assert all(np.array_equal(a, b) for a, b in zip(got["1"], got[str(ipt)]))
    (Jo,) = O.map_eval(f, n, [x0.T, us.T])
    assert O.rel_err(got[str(ipt)][0], Jo.T) <= TOL


@pytest.mark.parametrize("setting", [("CXH_JIT_SUBBLOCK", "128"), ("CXH_JIT_TRIG_CACHE", "0"),
                                     ("CXH_JIT_TRIG_CALLS", "0"), ("CXH_JIT_TRIG_CALLS", "-1")])
def test_jit_settings_keep_the_bits(ca, cuda, monkeypatch, setting):
    """The code-shape settings of the reverse kernels change neither the
    operations nor their order: sub-blocks of 128 instructions instead of 256,
    sin/cos recomputed instead of loaded from the first pass, and sin/cos as
    calls at every site or inline at every site (the programs compile with
    -ffp-contract=off) give the default's bits, fused and split."""
    monkeypatch.setenv("CXH_JIT_ROLL", "0")   # the straight-line segment kernels
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_rollout(20)
    n = 1000
    every = (f.n_instructions() + 1) // 2       # 2 segments of ~1380 instructions, ~40 sin/cos sites each
    rng = np.random.default_rng(seed_of(3))
    x0, us = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (20, n))
    seed = rng.standard_normal((1, n))
    got = []
    for env in (None, setting):
        if env:
            monkeypatch.setenv(*env)
        print(f"[case] jit rollout20 n={n} every={every} setting={env}", flush=True)
        tj = HipTape(f, engine="jit")
        (Jj,), gj = tj.fwd_adj(_dev(x0, us), _dev(seed), checkpoint_every=every)
        outs, ck = tj.fwd_ckpt(_dev(x0, us), n, every)
        g2 = tj.rev_ckpt(_dev(x0, us), _dev(seed), n, every, ck)
        got.append([Jj.cpu().numpy()] + _np(gj) + _np(g2))
    assert all(np.array_equal(a, b) for a, b in zip(*got))
    (Jo,), (ox, ou) = O.adj_map(f, n, [x0.T, us.T], [seed.T])
    g = np.concatenate(got[1][1:3]).T
    go = np.concatenate([ox, ou], axis=1)
    scale = np.maximum(np.max(np.abs(go), axis=1, keepdims=True), 1.0)
    assert np.median(np.max(np.abs(g - go) / scale, axis=1)) <= TOL


def test_jit_wide_live_set_reverse_function(ca, cuda):
    """A reverse Function (reverse(1) of a 40-step rollout: its forward half's
    values stay live through its reverse half, thousands at every segment
    boundary) on the specialised forward kernels: live-in values loaded where
    read, the rest copied through in loops.  Same bits as the interpreter; the
    oracle on a 10-step one at 1e-13."""
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    for steps, n in ((40, 2000), (10, 300)):
        f = W.cartpole_rollout(steps)
        R = f.reverse(1)
        rng = np.random.default_rng(seed_of(41) + steps)
        x0 = rng.uniform(-1, 1, (4, n))
        us = rng.uniform(-10, 10, (steps, n))
        sd = rng.standard_normal((1, n))
        print(f"[case] reverse Function of the {steps}-step rollout ({R.n_instructions()} instructions) n={n}",
              flush=True)
        args = _dev(x0, us) + [None] + _dev(sd)
        yi = [t.cpu().numpy() for t in HipTape(R).eval(args, n)]
        yj = [t.cpu().numpy() for t in HipTape(R, engine="jit").eval(args, n)]
        for a, b in zip(yi, yj):
            assert np.array_equal(a, b)
        if steps == 10:
            ref = O.map_eval(R, n, [x0.T, us.T, None, sd.T])
            for a, r in zip(yj, ref):
                assert O.rel_err(a, r.T) <= TOL


def test_jit_wide_live_set_fused_adjoint(ca, cuda):
    """The fused forward + adjoint sweep of that reverse Function (a
    second-order sweep).  Its reverse kernels would need scratch (every value
    of the forward half is live across them), so the specialised tape refuses
    them before launching anything and runs this adjoint on the interpreter:
    the same bits, and the tape says so."""
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_rollout(10)
    R = f.reverse(1)
    n = 300
    rng = np.random.default_rng(seed_of(42))
    args = _dev(rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (10, n))) + [None] + \
        _dev(rng.standard_normal((1, n)))
    seeds = _dev(rng.standard_normal((4, n)), rng.standard_normal((10, n)))
    every = 400
    print(f"[case] fused adjoint of the reverse Function ({R.n_instructions()} instructions) every={every}",
          flush=True)
    oi, gi = HipTape(R).fwd_adj(args, seeds, n, checkpoint_every=every)
    tj = HipTape(R, engine="jit")
    oj, gj = tj.fwd_adj(args, seeds, n, checkpoint_every=every)
    for a, b in zip(oi + gi, oj + gj):
        if a is None:
            assert b is None
            continue
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    fb = tj.jit_fallback()
    print(f"[step] fallback {fb}", flush=True)
    assert not fb["eval"]
    if fb["adjoint"]:
        assert "spills" in fb["why"]
    # the split sweep sizes and uses one layout, whichever engine runs it
    outs, ck = tj.fwd_ckpt(args, n, every)
    g2 = tj.rev_ckpt(args, seeds, n, every, ck)
    for a, b in zip(gi, g2):
        if a is not None:
            assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
