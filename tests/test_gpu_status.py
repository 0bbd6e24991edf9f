"""Per-instance regularity status of batched evaluation on the GPU.

cxh_tape_regularity_check makes every evaluation scan what it wrote,
instance by instance: status[k] = 1 for a NaN, | 2 for an infinity; the call
returns the number of flagged instances (OmpMap::eval ORs a per-instance
flag, casadi/core/map.cpp:371-380; FunctionInternal's regularity_check
raises on a NaN / Inf output, casadi/core/function_internal.cpp:1016-1030).
The expected flags come from the oracle's evaluation of the same inputs.
"""
import numpy as np
import pytest

import oracle_ref as O
from conftest import seed_of

pytestmark = pytest.mark.gpu


def _dev(*arrs):
    import torch
    return [torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda") for a in arrs]


def _flags(*blocks):
    """Expected status per instance from (nnz, n) blocks."""
    f = np.zeros(blocks[0].shape[1], np.int32)
    for b in blocks:
        f |= np.any(np.isnan(b), axis=0).astype(np.int32)
        f |= 2 * np.any(np.isinf(b) & ~np.isnan(b), axis=0).astype(np.int32)
    return f


@pytest.mark.parametrize("engine", ["interp", "jit"])
def test_status_of_map_and_adjoint(cuda, engine):
    import casadi_amd as ca
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_step()
    n = 5000
    rng = np.random.default_rng(seed_of(0) + 40)
    x, u = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (1, n))
    x[1, 7] = np.nan                 # NaN through sin/cos
    u[0, 100] = np.inf               # inf force: inf - inf later
    x[2, 2000] = 1e308               # overflow to inf in the update
    x[3, 4999] = -np.inf
    seed = rng.standard_normal((4, n))
    print(f"[case] status engine={engine} n={n}", flush=True)
    t = HipTape(f, engine=engine)
    st = t.regularity_check(n)
    (yo,) = O.map_eval(f, n, [x.T, u.T])
    want = _flags(yo.T)
    assert want[[7, 100, 2000, 4999]].all() and want.sum() >= 4
    with pytest.raises(ca.CasadiHipError, match="instances wrote NaN or Inf"):
        t.eval(_dev(x, u))
    assert np.array_equal(st[:n].cpu().numpy(), want)
    # adjoint: outputs and sensitivities
    _, (ox, ou) = O.adj_map(f, n, [x.T, u.T], [seed.T])
    want_adj = _flags(yo.T, ox.T, ou.T)
    with pytest.raises(ca.CasadiHipError):
        t.fwd_adj(_dev(x, u), _dev(seed))
    assert np.array_equal(st[:n].cpu().numpy(), want_adj)
    # regular inputs: 0 flagged, no error, status cleared
    xr, ur = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (1, n))
    t.eval(_dev(xr, ur))
    assert not st[:n].cpu().numpy().any()
    # the return code itself through the C ABI: the count
    from casadi_amd import _lib as L
    from casadi_amd.hip import _ptrs
    y = t._alloc(t.nnz_out, n, _dev(x)[0])
    rc = L.cxh_map_eval(t._h, n, _ptrs(_dev(x, u)), _ptrs(y), 0, n, None)
    assert rc == int((want != 0).sum())
    # a batch beyond the status capacity fails
    assert L.cxh_map_eval(t._h, n + 1, _ptrs(_dev(np.zeros((4, n + 1)), np.zeros((1, n + 1)))),
                          _ptrs(t._alloc(t.nnz_out, n + 1, _dev(x)[0])), 0, n + 1, None) < 0
    t.regularity_check(on=False)
    t.eval(_dev(x, u))               # off: NaN outputs pass silently, as by default in CasADi


def test_status_of_rolled_split_sweep(cuda):
    """The rollout's split sweep (forward storing checkpoints, reverse from
    them) on the rerolled kernels: one instance starts from NaN, one from a
    control of 1e300 (overflow)."""
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    import casadi_amd as ca
    f = W.cartpole_rollout(20)
    n = 1000
    rng = np.random.default_rng(seed_of(2) + 41)
    x0, us = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (20, n))
    x0[0, 3] = np.nan
    us[5, 900] = 1e300
    seed = np.ones((1, n))
    every = (f.n_instructions() + 3) // 4
    print(f"[case] status rolled rollout20 n={n}", flush=True)
    t = HipTape(f, engine="jit")
    st = t.regularity_check(n)
    (Jo,), (ox, ou) = O.adj_map(f, n, [x0.T, us.T], [seed.T])
    with pytest.raises(ca.CasadiHipError):
        t.fwd_ckpt(_dev(x0, us), n, every)
    assert np.array_equal(st[:n].cpu().numpy(), _flags(Jo.T))
    t.regularity_check(on=False)     # checkpoints again, unchecked, for the reverse
    outs, ck = t.fwd_ckpt(_dev(x0, us), n, every)
    st = t.regularity_check(n)
    with pytest.raises(ca.CasadiHipError):
        t.rev_ckpt(_dev(x0, us), _dev(seed), n, every, ck)
    assert np.array_equal(st[:n].cpu().numpy(), _flags(ox.T, ou.T))
    assert st[:n].cpu().numpy()[[3, 900]].all()
