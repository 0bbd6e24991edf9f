"""Heterogeneous batch: 40000 cart-pole RK4 steps, 12000 quadrotor RK4 steps
and 25000 two-link arm forward-dynamics evaluations, forward + adjoint each,
in one persistent launch replayed from a hipGraph, launch shape from the
autotuner.

Each Function's outputs and adjoints must equal (bit for bit) its own fused
forward + adjoint launch, and the CPU oracle's evaluation of CasADi's
reverse(1) Function at the 1e-13 bar.
"""
import numpy as np
import pytest

import oracle_ref as O
from conftest import seed_of

pytestmark = pytest.mark.gpu
TOL = 1e-13
COUNTS = [40000, 12000, 25000]


def _dev(*arrs):
    import torch
    return [torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda") for a in arrs]


@pytest.fixture(scope="module")
def problem(cuda):
    from casadi_amd import workloads as W
    fs = [W.cartpole_step(), W.quadrotor_step(), W.arm_dynamics()]
    rng = np.random.default_rng(seed_of(3))
    data = []
    for f, n in zip(fs, COUNTS):
        args = [rng.uniform(-1, 1, (f.nnz_in(i), n)) for i in range(f.n_in())]
        seeds = [rng.standard_normal((f.nnz_out(o), n)) for o in range(f.n_out())]
        data.append((args, seeds))
    return fs, data


def test_batch_matches_single_launches_and_oracle(problem):
    import torch
    from casadi_amd.hip import HipBatch, HipTape
    fs, data = problem
    print(f"[case] batch counts={COUNTS}", flush=True)
    B = HipBatch(fs, COUNTS)
    bound = []
    for i, (args, seeds) in enumerate(data):
        A, S = _dev(*args), _dev(*seeds)
        bound.append((A, S) + B.bind(i, A, S))
    cfg, ms = B.tune(reps=2)
    print(f"[batch] autotuned {cfg} {ms:.4f} ms", flush=True)
    assert cfg["block"] in (64, 128, 256) and 0 <= cfg["order"] <= 2 and cfg["graph"] in (True, False)
    B.launch(graph=None)                 # the tuner's choice of graph replay or direct launch
    torch.cuda.synchronize()
    for f, n, (A, S, outs, adjs), (args, seeds) in zip(fs, COUNTS, bound, data):
        t = HipTape(f, engine="jit")
        o1, a1 = t.fwd_adj(A, S, n)
        for x, y in zip(outs + adjs, o1 + a1):
            assert np.array_equal(x.cpu().numpy(), y.cpu().numpy())
        ref = O.map_eval(f, n, [a.T for a in args])
        for x, r in zip(outs, ref):
            assert O.rel_err(x.cpu().numpy(), r.T) <= TOL
        R = f.reverse(1)
        rr = O.map_eval(R, n, [a.T for a in args] + [None] * f.n_out() + [s.T for s in seeds])
        for x, r in zip(adjs, rr):
            assert O.rel_err(x.cpu().numpy(), r.T) <= TOL, f.name()


def test_batch_graph_replays_track_inputs(problem):
    """Replays of the captured graph, with eager work and synchronisations in
    between and inputs changed in place, equal a direct launch every time."""
    import torch
    from casadi_amd.hip import HipBatch
    fs, data = problem
    B = HipBatch(fs, COUNTS)
    bound = []
    for i, (args, seeds) in enumerate(data):
        A, S = _dev(*args), _dev(*seeds)
        bound.append((A, S) + B.bind(i, A, S))
    B.config(block=128, grid_per_cu=2, order=1)
    for rep in range(3):
        for A, S, outs, adjs in bound:
            A[0].mul_(0.9)               # eager work on the inputs
        torch.cuda.synchronize()
        B.launch(graph=True)
        torch.cuda.synchronize()
        got = [[x.clone() for x in outs + adjs] for A, S, outs, adjs in bound]
        B.launch(graph=False)
        torch.cuda.synchronize()
        for g, (A, S, outs, adjs) in zip(got, bound):
            for x, y in zip(g, outs + adjs):
                assert torch.equal(x, y), rep
    # a new launch shape recaptures the graph
    B.config(block=64, grid_per_cu=0, order=2)
    B.launch()
    torch.cuda.synchronize()
    got = [[x.clone() for x in outs + adjs] for A, S, outs, adjs in bound]
    B.launch(graph=False)
    torch.cuda.synchronize()
    for g, (A, S, outs, adjs) in zip(got, bound):
        for x, y in zip(g, outs + adjs):
            assert torch.equal(x, y)
