"""Timings of the GPU path for each benchmark workload (one JSON line each).

    python -m casadi_amd.benchmarks [c1] [c1big] [c2] [c4] [c5]

c1     cart-pole RK4 step map over 65536 instances (interpreter, specialised)
       and its dense 4x5 Jacobian (fused tangent sweeps vs separate launches)
c1big  the same map over 8M instances (the HBM / FP64 roofline point)
c2     star-coloured Hessian of the 20-step quadrotor OCP Lagrangian, 16384
c4     heterogeneous batch: 40000 cart-pole + 12000 quadrotor + 25000 arm,
       forward + adjoint, one hipGraph launch (autotuned) vs plain launches
c5     batched KKT solve, n = 274, 8192 instances, ldl and qr
c3macc the 100-step rollout cost + gradient as a mapaccum of one step, 32768
c3interp the 100-step rollout cost + gradient on the interpreter's fused sweep, 32768
mapsum reduced map: cart-pole step over 1M states, shared control, summed
sparsity Jacobian patterns from dependency sweeps, host against GPU

Inputs are synthetic, from numpy.random.default_rng(20261202 + workload
index).  Times are HIP-event times on torch's current stream, device-resident
data, after one warm-up call.  Used to produce the profiles of the non-headline
workloads; the headline number is bench.py's.
"""
import json
import os
import sys
import time

import numpy as np


def _timed(fn, reps=10):
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _out(**kw):
    print(json.dumps(kw), flush=True)


def _dev(a):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def cartpole(n, jac=True):
    from . import workloads as W
    from .hip import HipTape, HipJacobian
    f = W.cartpole_step()
    rng = np.random.default_rng(20261202)
    X, U = _dev(rng.uniform(-1, 1, (4, n))), _dev(rng.uniform(-10, 10, (1, n)))
    for eng in ("interp", "jit"):
        tp = HipTape(f, engine=eng)
        ms = _timed(lambda: tp.eval([X, U], n), 50)
        _out(k="cartpole_map", engine=eng, n=n, ms=ms, GBs=72 * n / ms / 1e6)
    if not jac:
        return
    for eng in ("interp", "jit"):
        hj = HipJacobian(f, engine=eng)
        ms = _timed(lambda: hj([X, U], n), 20)
        _out(k="cartpole_jac_fused", engine=eng, n=n, ms=ms, GBs=(72 + 160) * n / ms / 1e6, ndir=hj.plan["ndir"])
    J = f.jacobian()
    for eng in ("interp", "jit"):
        tf, tJ = HipTape(f, engine=eng), HipTape(J, engine=eng)
        ms = _timed(lambda: (tf.eval([X, U], n), tJ.eval([X, U, None], n)), 20)
        _out(k="cartpole_jac_plain", engine=eng, n=n, ms=ms)


def hessian(n=16384):
    import torch
    from . import workloads as W
    from .function import Function
    from .hip import HipJacobian
    from .sx import gradient
    w, lam, Lag = W.quadrotor_ocp_lagrangian()
    G = Function("grad_lag", [w, lam], [gradient(Lag, w)])
    rng = np.random.default_rng(20261203)
    args = [_dev(rng.uniform(-1, 1, (w.nnz(), n))), _dev(rng.standard_normal((lam.nnz(), n)))]
    hj = HipJacobian(G, symmetric=True, first_block=True)
    ms = _timed(lambda: hj(args, n), 3)
    _out(k="hessian_fused", engine="interp", n=n, ms=ms, ndir=hj.plan["ndir"], nnz=hj.nnz_jac,
         instr=G.n_instructions())
    if os.environ.get("CXH_BENCH_HESSIAN_JIT", "1") == "1":
        hjj = HipJacobian(G, symmetric=True, first_block=True, engine="jit")
        t0 = time.time()
        hjj(args, n)
        torch.cuda.synchronize()
        ct = time.time() - t0
        ms = _timed(lambda: hjj(args, n), 5)
        (_, a), (_, b) = hj(args, n), hjj(args, n)
        _out(k="hessian_fused", engine="jit", n=n, ms=ms, compile_s=ct,
             max_abs_diff_vs_interp=float((a - b).abs().max()))


def hetero_batch():
    from . import workloads as W
    from .hip import HipBatch, HipTape
    fs = [W.cartpole_step(), W.quadrotor_step(), W.arm_dynamics()]
    counts = [40000, 12000, 25000]
    rng = np.random.default_rng(20261205)
    B = HipBatch(fs, counts)
    data = []
    for i, (f, n) in enumerate(zip(fs, counts)):
        A = [_dev(rng.uniform(-1, 1, (f.nnz_in(j), n))) for j in range(f.n_in())]
        S = [_dev(rng.standard_normal((f.nnz_out(j), n))) for j in range(f.n_out())]
        B.bind(i, A, S)
        data.append((A, S))
    cfg, tms = B.tune(reps=3)
    ms = _timed(lambda: B.launch(), 50)
    _out(k="batch_graph", cfg=cfg, tune_ms=tms, ms=ms, inst_per_s=sum(counts) / ms * 1e3)
    ms = _timed(lambda: B.launch(graph=False), 50)
    _out(k="batch_direct", ms=ms)
    tapes = [(HipTape(f), HipTape(f.reverse(1))) for f in fs]

    def plain():
        for (tf, tr), (A, S), n in zip(tapes, data, counts):
            tf.eval(A, n)
            tr.eval(A + [None] * len(S) + S, n)
    ms = _timed(plain, 20)
    _out(k="batch_plain_interp", ms=ms, inst_per_s=sum(counts) / ms * 1e3)
    jt = [HipTape(f, engine="jit") for f in fs]

    def fused_sep():
        for tj, (A, S), n in zip(jt, data, counts):
            tj.fwd_adj(A, S, n)
    ms = _timed(fused_sep, 20)
    _out(k="batch_fused_per_function", engine="jit", ms=ms)


def rollout_mapaccum(n=32768, N=100, K=10):
    """The 100-step rollout cost + gradient as a mapaccum of the cost-carrying
    step (one compiled step body, loops), checkpoints every K steps."""
    import torch
    from . import workloads as W
    from .hip import HipMapAccum
    step = W.cartpole_cost_step()
    t0 = time.time()
    m = HipMapAccum(step, N)
    rng = np.random.default_rng(20261204)
    X0 = _dev(np.vstack([rng.uniform(-1, 1, (4, n)), np.zeros((1, n))]))
    U = _dev(rng.uniform(-10, 10, (N, n)))
    S = torch.zeros((5 * N, n), dtype=torch.float64, device="cuda")
    S[5 * (N - 1) + 4] = 1.0      # d(cost carried to the end) -- the |x_N|^2 term is outside the chain
    m.fwd_adj([X0, U], [S], n, checkpoint_steps=K)
    torch.cuda.synchronize()
    ct = time.time() - t0
    ms = _timed(lambda: m.fwd_adj([X0, U], [S], n, checkpoint_steps=K), 20)
    _out(k="rollout_mapaccum_fwd_adj", n=n, N=N, K=K, ms=ms, inst_per_s=n / ms * 1e3, compile_s=ct)
    ms = _timed(lambda: m.eval([X0, U], n), 20)
    _out(k="rollout_mapaccum_eval", n=n, N=N, ms=ms)


def rollout_interp(n=32768, N=100, K=10):
    """The rollout cost + gradient on the tape interpreter's fused sweep
    (tape_adj_kernel), checkpoints every K steps: three calls."""
    from . import workloads as W
    from .hip import HipTape
    f = W.cartpole_rollout(N)
    rng = np.random.default_rng(20261204)
    X0, U = _dev(rng.uniform(-1, 1, (4, n))), _dev(rng.uniform(-10, 10, (N, n)))
    S = _dev(np.ones((1, n)))
    t = HipTape(f)
    every = (f.n_instructions() + K - 1) // K
    ms = _timed(lambda: t.fwd_adj([X0, U], [S], n, checkpoint_every=every), 2)
    _out(k="rollout_interp_fwd_adj", n=n, N=N, K=K, ms=ms, inst_per_s=n / ms * 1e3)


def kkt(n=8192):
    from . import workloads as W
    from .linsol import Linsol
    sp, nw, nl = W.ocp_kkt_sparsity()
    rng = np.random.default_rng(20261206)
    _, A = W.ocp_kkt_batch(rng, n)
    At, Bt = _dev(A), _dev(rng.standard_normal((sp.size1(), n)))
    for solver, engine in (("ldl", "jit"), ("ldl", "interp"), ("qr", "jit"), ("qr", "interp")):
        t0 = time.time()
        Ls = Linsol("kkt", solver, sp, {"engine": engine})
        Ls.solve_batch(At, Bt)
        ct = time.time() - t0
        ms = _timed(lambda: Ls.solve_batch(At, Bt), 5)
        _out(k="kkt_" + solver, engine=engine, n=n, ms=ms, nnz=sp.nnz(), inst_per_s=n / ms * 1e3,
             GBs=(sp.nnz() + 2 * sp.size1()) * 8 * n / ms / 1e6, first_call_s=ct)


def mapsum(n=1 << 20):
    """Reduced map: the cart-pole step over n initial states with the control
    shared by all of them (reduce_in), the next states summed (reduce_out);
    forward alone and forward + adjoint (the gradient w.r.t. the shared
    control summed over the instances).  Against the same results without the
    reduction support: the shared blocks replicated per instance and the sums
    taken by torch after a plain map."""
    import torch
    from . import workloads as W
    from .hip import HipTape
    f = W.cartpole_step()
    rng = np.random.default_rng(20261206)
    X = _dev(rng.uniform(-1, 1, (4, n)))
    U = _dev(rng.uniform(-10, 10, (1,)))
    S = _dev(rng.standard_normal(4))
    t = HipTape(f, engine="jit")
    ms = _timed(lambda: t.eval_sum([X, U], n, [False, True], [True]), 20)
    _out(k="mapsum_eval", n=n, ms=ms, GBs=32 * n / ms / 1e6)
    ms = _timed(lambda: t.fwd_adj_sum([X, U], [S], n, [False, True], [True]), 20)
    _out(k="mapsum_fwd_adj", n=n, ms=ms, GBs=64 * n / ms / 1e6)

    def unfused():
        Ut = U.reshape(1, 1).expand(1, n).contiguous()
        St = S.reshape(4, 1).expand(4, n).contiguous()
        (y,), (gx, gu) = t.fwd_adj([X, Ut], [St], n)
        return y.sum(dim=1), gx, gu.sum(dim=1)
    ms = _timed(unfused, 20)
    _out(k="mapsum_fwd_adj_unfused", n=n, ms=ms)


def sparsity():
    """Jacobian sparsity from dependency sweeps: the host's 64-column sweeps
    (SXFunction::get_jac_sparsity) against the GPU's, all words side by side
    (HipTape.jac_sparsity: seeds up, cxh_sp_forward / cxh_sp_reverse, words
    down, pattern assembled).  Patterns compared for equality."""
    import torch
    from . import SX, Function, sin
    from . import workloads as W
    from .hip import HipTape
    from .sx import gradient
    w, lam, Lag = W.quadrotor_ocp_lagrangian()
    x = SX.sym("x", 20000)
    cases = [("hessian_pattern_quadrotor_ocp", lambda: Function("grad_lag", [w, lam], [gradient(Lag, w)])),
             ("banded_20000", lambda: Function("band", [x], [sin(x[:-1]) * x[1:]]))]
    for name, make in cases:
        f = make()
        t0 = time.perf_counter()
        hs = f.jac_sparsity(0, 0)
        host_ms = (time.perf_counter() - t0) * 1e3
        t = HipTape(f)
        t.jac_sparsity(0, 0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gs = t.jac_sparsity(0, 0)
        torch.cuda.synchronize()
        gpu_ms = (time.perf_counter() - t0) * 1e3
        nw = -(-f.nnz_in(0) // 64)
        seeds = [torch.zeros((f.nnz_in(i), nw), dtype=torch.int64, device="cuda") for i in range(f.n_in())]
        kern_ms = _timed(lambda: t.sp_forward(seeds, nw), 5)
        _out(k="jac_sparsity", case=name, instr=f.n_instructions(), worksize=f.sz_w(), nnz_in=f.nnz_in(0),
             nnz_out=f.nnz_out(0), words=nw, nnz=hs.nnz(), equal=bool(gs == hs), host_ms=host_ms,
             gpu_ms=gpu_ms, sp_forward_ms=kern_ms)


def main(argv):
    which = argv or ["c1", "c2", "c4", "c5"]
    if "c1" in which:
        cartpole(65536)
    if "c1big" in which:
        cartpole(1 << 23, jac=False)
    if "c2" in which:
        hessian()
    if "c4" in which:
        hetero_batch()
    if "c5" in which:
        kkt()
    if "c3macc" in which:
        rollout_mapaccum()
    if "c3interp" in which:
        rollout_interp()
    if "mapsum" in which:
        mapsum()
    if "sparsity" in which:
        sparsity()


if __name__ == "__main__":
    main(sys.argv[1:])
