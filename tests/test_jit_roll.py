"""Flat tapes rerolled into loops: the structure found, and the generated
loop kernels executed on the host at the driver's exact launch geometry.

A tape unrolled from a recurrence repeats one body of instructions.  The
specialiser finds the prelude, the period and the values that cross an
iteration (jit_roll.hip) and emits one loop body instead of straight-line
segments.  These tests check

* the structure found for the cart-pole rollout, and its absence where a
  body would need more than the previous iteration;
* the kernels of every launch sequence the driver uses -- the forward chain
  (evaluation), the fused forward + adjoint (checkpointing forward, then one
  paired launch per checkpoint segment), and the split form behind
  torch.autograd (forward storing the checkpoints; reverse from them) --
  compiled with g++ against the checked memory macros of jit_launch_emu.h:
  every access inside the buffer its pointer came from (sized exactly as the
  driver sizes it), no address shared between lanes of one launch;
* their results against the oracle's evaluation and numeric reverse sweep of
  the same tape (same host libm: bit for bit).
"""
import numpy as np
import pytest

import oracle_ref as O
from test_jit_host_exec import CxjArgs, _table
from test_jit_launch_emu import Emu, _off, _ptr, _round64


def _roll(f):
    from casadi_amd.hip import HipTape
    return HipTape(f, engine="jit")


def parametric_rollout(n_steps):
    """x_{t+1} = x_t + 0.1 p sin(x_t) + u_t over two states with a shared
    parameter p read by every step (a value the loop passes through), every
    state an output (an OUTPUT in the body, rows advancing by 2)."""
    import casadi_amd as ca
    x0, p, us = ca.SX.sym("x0", 2), ca.SX.sym("p"), ca.SX.sym("u", n_steps)
    x, outs = x0, []
    for t in range(n_steps):
        x = ca.vertcat(x[0] + 0.1 * p * ca.sin(x[1]) + us[t], x[1] + 0.1 * p * ca.cos(x[0]) * x[0])
        outs.append(x)
    return ca.Function("prollout", [x0, p, us], [ca.vertcat(*outs), ca.sumsqr(x)])


def division_rollout(n_steps):
    """x_{t+1} = x_t / 1.1 + u_t, y_{t+1} = sin(y_t / -3) + x_t / 0.3: every
    iteration divides by constants, which the loop kernels do through
    sxh_div_c (exact re-run of an iteration with an out-of-range dividend)."""
    import casadi_amd as ca
    x0, us = ca.SX.sym("x0", 2), ca.SX.sym("u", n_steps)
    x, y, outs = x0[0], x0[1], []
    for t in range(n_steps):
        x, y = x / 1.1 + us[t], ca.sin(y / -3.0) + x / 0.3
        outs += [x, y]
    return ca.Function("drollout", [x0, us], [ca.vertcat(*outs)])


def division_inputs(n_steps, n, rng):
    """Inputs of division_rollout for n instances: uniform values, with lanes
    whose dividends are zero, signed zero, subnormal, out of the fast range
    (2^950, 2^-1000), infinite or NaN from the start, and lanes where one
    control u_t is such a value, so that only some iterations (and on the
    GPU some wavefronts) take the exact re-run."""
    x0 = rng.uniform(-1, 1, (2, n))
    u = rng.uniform(-1, 1, (n_steps, n))
    # with u = 0 the first three misround under the fast sequence alone
    # (subnormal quotients of 1.1 and 0.3; 1.7e308 / 0.3 overflows to NaN
    # there): only the re-run gives the IEEE bits
    special = [float.fromhex("0x0.3f20dab6def8dp-1022"), float.fromhex("0x0.452d93f3099fdp-1022"), 1.7e308,
               0.0, -0.0, 5e-324, -2.5e-310, 2.0 ** 950, -(2.0 ** -1000), np.inf, -np.inf, np.nan]
    lane = 3
    for i, v in enumerate(special):
        x0[:, lane] = v
        u[:, lane] = 0.0 if i < 3 or v != v or abs(v) == np.inf else v
        lane += 7
    for i, v in enumerate(special):
        u[(5 * i + 2) % n_steps, lane] = v
        lane += 3
    x0[:, 192:256] = 0.0          # a wavefront of zeros: exact without the re-run
    u[:, 192:256] = 0.0
    return [x0, u]


def second_order(n_steps):
    """x_{t+1} = sin(x_t) x_{t-1} + u_t: each step reads two steps back."""
    import casadi_amd as ca
    x0, x1, us = ca.SX.sym("x0"), ca.SX.sym("x1"), ca.SX.sym("u", n_steps)
    a, b = x0, x1
    for t in range(n_steps):
        a, b = b, ca.sin(b) * a + us[t]
    return ca.Function("order2", [x0, x1, us], [b])


def test_rollout_structure():
    from casadi_amd import workloads as W
    f = W.cartpole_rollout(100)
    every = (f.n_instructions() + 9) // 10
    P = _roll(f).roll_plan(every)
    # step 0 and the constants (at their first use) before the loop, one RK4
    # step + cost term per iteration, 4 states + the running cost crossing
    assert P["ok"] and P["prelude"] == 158 and P["period"] == 138 and P["iterations"] == 99
    assert P["state"] == 5 and P["K"] == 10 and P["segments"] == 10
    assert P["prelude"] + P["period"] * P["iterations"] + 1 == f.n_instructions()  # the epilogue: one OUTPUT


def test_no_loop_where_none_fits():
    from casadi_amd import workloads as W
    assert not _roll(W.cartpole_step()).roll_plan()["ok"]           # too short to repeat
    P = _roll(second_order(150)).roll_plan()
    # a body reading two iterations back is not taken as such; a period of
    # two steps reads only the iteration before and is
    assert not P["ok"] or P["period"] % 2 == 0


def _buffers(emu, f, n, rng, arrays):
    """Device buffers of the inputs (SoA [nz][n], ld = n) filled from arrays."""
    out = []
    for a in arrays:
        b = emu.buffer(a.size)
        b[:] = a.ravel()
        out.append(b)
    return out


def _launch_sequence(f, every, n, seed_val, mode, inputs=None):
    """The rerolled kernels launched as the driver launches them, on the
    emulator.  mode "eval": the forward chain; "fused": checkpointing forward
    then the pairs from segment S-1; "split": forward storing the checkpoints
    in a caller buffer of the size cxh_ckpt_size gives, then the pairs from
    segment S (the first recomputes the last segment only).  Returns (errors,
    outputs, sensitivities, inputs, seeds)."""
    tape = _roll(f)
    P = tape.roll_plan(every)
    assert P["ok"]
    K, S, ns = P["K"], P["segments"], max(P["state"], 1)
    rng = np.random.default_rng(seed_val)
    ins = [rng.uniform(-1, 1, (f.nnz_in(i), n)) for i in range(f.n_in())] if inputs is None else inputs
    seeds = [rng.standard_normal((f.nnz_out(j), n)) for j in range(f.n_out())]
    srcs = [tape.roll_source(every, 0), tape.roll_source(every, 2), tape.roll_source(every, 4)]
    lds = _round64(n)
    with __import__("tempfile").TemporaryDirectory() as d:
        emu = Emu(srcs, d)
        xb = _buffers(emu, f, n, rng, ins)
        yb = [emu.buffer(f.nnz_out(j) * n) for j in range(f.n_out())]
        sb = _buffers(emu, f, n, rng, seeds)
        gb = [emu.buffer(f.nnz_in(i) * n) for i in range(f.n_in())]
        A = CxjArgs()
        A.arg, A.n, A.lds = _table(xb, n), n, lds
        errors = 0
        if mode == "eval":
            A.res = _table(yb, n)
            errors += emu.launch(0, A, n)
            return errors, [y.reshape(-1, n).copy() for y in yb], None, ins, seeds
        # the driver's buffers, each at its exact size (one allocation there;
        # separate here, so that a spill from one into the next is caught)
        ck = emu.buffer((S + 1) * ns * lds)
        states = [emu.buffer(K * P["step_rows"] * lds) for _ in range(2)]
        adjb = [emu.buffer(ns * lds) for _ in range(2)]
        A.res = _table(yb, n)
        A.st_in = _ptr(ck)
        if mode == "fused":
            A.st_out = _ptr(states[(S - 1) % 2])
            first = S - 1
        else:
            A.st_out = None
            first = S
        errors += emu.launch(1, A, n)
        A.res = _table([], n)
        A.aseed, A.asens = _table(sb, n), _table(gb, n)
        B = CxjArgs.from_buffer_copy(A)
        for seg in range(first, -1, -1):
            A.st_out = _ptr(states[seg % 2])
            A.adj_in, A.adj_out = _ptr(adjb[(seg + 1) % 2]), _ptr(adjb[seg % 2])
            B.st_out = _ptr(states[(seg + 1) % 2])
            errors += emu.launch(2, A, n, B, gy=2, seg=seg)
        return (errors, [y.reshape(-1, n).copy() for y in yb], [g.reshape(-1, n).copy() for g in gb], ins, seeds)


def _cases():
    from casadi_amd import workloads as W
    return {
        # 19 iterations, K = 4: segments 4, 4, 4, 4, 3; the padded last wavefront
        "rollout20": (W.cartpole_rollout(20), 138 * 4 + 10, 1000),
        # K = 1: one iteration per segment, three tiles of 64 (the last ragged)
        "rollout6": (W.cartpole_rollout(6), 138, 130),
        # a parameter passed through every iteration, outputs in the body
        "parametric": (parametric_rollout(24), 13 * 5, 70),
        # a second-order recurrence: its loop body is two steps
        "order2": (second_order(150), 8 * 10, 100),
    }


@pytest.mark.parametrize("case", ["rollout20", "rollout6", "parametric", "order2"])
@pytest.mark.parametrize("mode", ["eval", "fused", "split"])
def test_rerolled_kernels_in_bounds_race_free_and_exact(case, mode):
    f, every, n = _cases()[case]
    errors, ys, gs, ins, seeds = _launch_sequence(f, every, n, 20261210 + n, mode)
    assert errors == 0
    yr = O.map_eval(f, n, [a.T for a in ins])
    for y, r in zip(ys, yr):
        assert np.array_equal(y, r.T)
    if mode == "eval":
        return
    _, gr = O.adj_map(f, n, [a.T for a in ins], [s.T for s in seeds])
    for g, r in zip(gs, gr):
        assert np.array_equal(g, r.T)


def test_trig_rows_of_the_loop_body(monkeypatch):
    """The cart-pole step takes sin and cos of the angle once per RK4 stage:
    four pairs, stored beside the 5 state values of every iteration.  A body
    that takes only sin of some operand gets none (its forward would not
    share the sincos), nor does any body with CXH_JIT_TRIG_CACHE=0."""
    from casadi_amd import workloads as W
    f = W.cartpole_rollout(6)
    assert _roll(f).roll_plan(138)["step_rows"] == 5 + 2 * 4
    assert _roll(parametric_rollout(24)).roll_plan()["step_rows"] == 3
    monkeypatch.setenv("CXH_JIT_TRIG_CACHE", "0")
    assert _roll(f).roll_plan(138)["step_rows"] == 5


@pytest.mark.parametrize("mode", ["fused", "split"])
def test_rerolled_kernels_without_trig_rows(monkeypatch, mode):
    monkeypatch.setenv("CXH_JIT_TRIG_CACHE", "0")
    f, every, n = _cases()["rollout6"]
    errors, ys, gs, ins, seeds = _launch_sequence(f, every, n, 20261210 + n, mode)
    assert errors == 0
    _, gr = O.adj_map(f, n, [a.T for a in ins], [s.T for s in seeds])
    for g, r in zip(gs, gr):
        assert np.array_equal(g, r.T)


def test_parametric_structure_passes_the_parameter_through():
    f = parametric_rollout(24)
    P = _roll(f).roll_plan()
    # two states carried, the parameter passed through every iteration
    assert P["ok"] and P["state"] == 3 and P["iterations"] >= 20


def _same(a, b):
    """Equal, or NaN where NaN (the sign of a zero adjoint is left open)."""
    return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


@pytest.mark.parametrize("mode", ["eval", "fused", "split"])
def test_fast_constant_divisions_exact_on_special_dividends(mode):
    """The loop body divides by 1.1, -3 and 0.3 through sxh_div_c; zeros,
    infinities and NaN come out of its fixup, subnormal and out-of-range
    dividends re-run the iteration with IEEE division.  Executed on the host
    (each lane its own vote), in bounds, and bit for bit the oracle's."""
    n_steps, n = 40, 256
    f = division_rollout(n_steps)
    src = _roll(f).roll_source(0, 4)
    assert "sxh::sxh_div_c(" in src and "cxj_step_rev_x" in src
    ins = division_inputs(n_steps, n, np.random.default_rng(5))
    errors, ys, gs, ins, seeds = _launch_sequence(f, 9 * 4, n, 11, mode, inputs=ins)
    assert errors == 0
    yr = O.map_eval(f, n, [a.T for a in ins])
    for y, r in zip(ys, yr):
        assert _same(y, r.T)
    if mode == "eval":
        return
    _, gr = O.adj_map(f, n, [a.T for a in ins], [s.T for s in seeds])
    for g, r in zip(gs, gr):
        assert _same(g, r.T)


def test_fast_division_switch(monkeypatch):
    """CXH_JIT_FAST_DIV=0: IEEE division in the loop bodies, only the `_x`
    iteration functions and no re-run."""
    src = _roll(division_rollout(40)).roll_source(0, 4)
    assert "sxh::sxh_div_c(" in src and "cxj_step_rev_f(" in src and "CXJ_ANY(wide)" in src
    monkeypatch.setenv("CXH_JIT_FAST_DIV", "0")
    src = _roll(division_rollout(40)).roll_source(0, 4)
    assert "sxh::sxh_div_c(" not in src and "_f(" not in src and "CXJ_ANY" not in src.split("cxj_pre_fwd")[-1]


def _roll_default_source(f, every):
    import os
    saved = {k: os.environ.pop(k, None) for k in ("CXH_JIT_ROLL_PREFETCH", "CXH_JIT_FAST_DIV")}
    try:
        return _roll(f).roll_source(every, 4)
    finally:
        os.environ.update({k: v for k, v in saved.items() if v is not None})


@pytest.mark.parametrize("case", ["rollout20", "division"])
def test_lean_variant_in_bounds_and_exact(monkeypatch, case):
    """The kernels large batches run (CXH_JIT_ROLL_PREFETCH=0 and
    CXH_JIT_FAST_DIV=0 give the same source): each iteration loading at its
    start, IEEE division, no re-run; in bounds, race-free and the oracle's
    bits on the host, split sweep."""
    monkeypatch.setenv("CXH_JIT_ROLL_PREFETCH", "0")
    monkeypatch.setenv("CXH_JIT_FAST_DIV", "0")
    if case == "division":
        f, every, n = division_rollout(40), 9 * 4, 256
        ins = division_inputs(40, n, np.random.default_rng(5))
    else:
        (f, every, n), ins = _cases()["rollout20"], None
    src = _roll(f).roll_source(every, 4)
    assert "sxh::sxh_div_c(" not in src and "for (;; i" not in src     # no fast pass, no prefetching loop
    assert "for (;; i" in _roll_default_source(f, every)
    errors, ys, gs, ins, seeds = _launch_sequence(f, every, n, 13, "split", inputs=ins)
    assert errors == 0
    _, gr = O.adj_map(f, n, [a.T for a in ins], [s.T for s in seeds])
    for g, r in zip(gs, gr):
        assert _same(g, r.T)
