"""Asynchronous compile of the specialised kernels (cxh_tape_jit_async): the
calls made while a worker thread builds them run on the interpreter, the
calls after on the specialised kernels, and every result has the
interpreter's bits (the engines agree bit for bit, tests/test_gpu_jit.py).
"""
import time

import numpy as np
import pytest

from conftest import seed_of

pytestmark = pytest.mark.gpu


def _wait(tape, what, limit=300.0):
    t0 = time.time()
    last = t0
    while tape.jit_pending():
        if time.time() - t0 > limit:
            raise AssertionError(f"{what}: still compiling after {limit} s")
        if time.time() - last > 10:
            print(f"[step] {what}: compiling, {time.time() - t0:.0f} s", flush=True)
            last = time.time()
        time.sleep(0.05)
    return time.time() - t0


@pytest.mark.parametrize("roll", ["1", "0"])
def test_async_compile_keeps_the_bits(cuda, monkeypatch, roll):
    import torch
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    monkeypatch.setenv("CXH_JIT_ROLL", roll)   # 0: straight-line segments, seconds to compile
    f = W.cartpole_rollout(30)
    n = 3000
    rng = np.random.default_rng(seed_of(30))
    args = [torch.tensor(rng.uniform(-1, 1, (4, n)), dtype=torch.float64, device="cuda"),
            torch.tensor(rng.uniform(-10, 10, (30, n)), dtype=torch.float64, device="cuda")]
    seed = [torch.ones((1, n), dtype=torch.float64, device="cuda")]
    every = (f.n_instructions() + 2) // 3
    print(f"[case] async compile roll={roll} n={n}", flush=True)
    ti = HipTape(f)
    (yi,) = ti.eval(args)
    (oi,), gi = ti.fwd_adj(args, seed, checkpoint_every=every)
    tj = HipTape(f, engine="jit")
    tj.jit_async(True)
    # evaluation: the first call starts the build and runs on the interpreter
    (y0,) = tj.eval(args)
    assert torch.equal(y0, yi)
    s = _wait(tj, "eval kernels")
    print(f"[step] eval kernels built in {s:.1f} s", flush=True)
    (y1,) = tj.eval(args)
    assert torch.equal(y1, yi) and not tj.jit_pending()
    # the fused adjoint needs other kernels: another build, the same way
    (o0,), g0 = tj.fwd_adj(args, seed, checkpoint_every=every)
    assert torch.equal(o0, oi) and all(torch.equal(a, b) for a, b in zip(g0, gi))
    _wait(tj, "adjoint kernels")
    (o1,), g1 = tj.fwd_adj(args, seed, checkpoint_every=every)
    assert torch.equal(o1, oi) and all(torch.equal(a, b) for a, b in zip(g1, gi))
    # the split sweep waits for a running build, then keeps one engine
    tj2 = HipTape(f, engine="jit")
    tj2.jit_async(True)
    tj2.eval(args)                               # a build in flight
    outs, ck = tj2.fwd_ckpt(args, n, every)
    g2 = tj2.rev_ckpt(args, seed, n, every, ck)
    assert all(torch.equal(a, b) for a, b in zip(g2, gi))
    del tj2                                     # destroy joins any worker


def test_async_map_option(cuda):
    import torch
    import casadi_amd as ca
    from casadi_amd import workloads as W
    f = W.cartpole_step()
    n = 1000
    rng = np.random.default_rng(seed_of(31))
    x = rng.uniform(-1, 1, (4, n))
    u = rng.uniform(-10, 10, (1, n))
    F = f.map(n, "hip", opts={"jit_async": True})
    Fi = f.map(n, "hip", opts={"engine": "interp"})
    ref = np.asarray(Fi(x, u))
    for _ in range(3):
        assert np.array_equal(np.asarray(F(x, u)), ref)
    _wait(F.tape, "map kernels")
    assert np.array_equal(np.asarray(F(x, u)), ref)
