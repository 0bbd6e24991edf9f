"""GPU parity of tapes with nested Function calls (cxh_tape_create_nested):
values against the oracle's call-by-call evaluation, adjoints against the
oracle's reverse sweep of the same computation written as one SX graph."""
import numpy as np
import pytest

import oracle_ref as O
from conftest import seed_of
from test_nested_tape import nested_program, _inputs

pytestmark = pytest.mark.gpu
TOL = 1e-13


@pytest.fixture(scope="module")
def ca():
    import casadi_amd
    return casadi_amd


def _dev(a):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("n", [1, 1000, 65536])
def test_nested_tape_eval_and_adjoint(ca, cuda, n):
    import torch
    from casadi_amd.hip import HipTape
    instr, ws, nin, nout, calls, F = nested_program(ca)
    args = _inputs(n, seed_of(50) + n)
    rng = np.random.default_rng(seed_of(51) + n)
    seeds = [rng.standard_normal((n, k)) for k in nout]
    small = min(n, 2000)
    ref = O.eval_calls(instr, ws, nin, nout, calls, [a[:small] for a in args])
    _, ref_adj = O.adj_map(F, small, [a[:small] for a in args], [s[:small] for s in seeds])
    got = {}
    for engine in ("interp", "jit"):
        print(f"[case] nested tape engine={engine} n={n}", flush=True)
        g = HipTape(calls[0][0])
        t = HipTape.nested(instr, ws, nin, nout, [(g, c[1], c[2]) for c in calls], engine=engine)
        X = [_dev(a.T) for a in args]
        S = [_dev(s.T) for s in seeds]
        outs = t.eval(X, n)
        (o2, adj) = t.fwd_adj(X, S, n)
        for a, b in zip(outs, o2):
            assert torch.equal(a, b)
        for a, r in zip(outs, ref):
            assert O.rel_err(a.cpu().numpy().T[:small], r) <= TOL
        for a, r in zip(adj, ref_adj):
            assert O.rel_err(a.cpu().numpy().T[:small], r) <= TOL
        got[engine] = [x.cpu().numpy() for x in outs + adj]
    for a, b in zip(got["interp"], got["jit"]):
        assert np.array_equal(a, b)
