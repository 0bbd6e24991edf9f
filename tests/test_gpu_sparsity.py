"""GPU dependency propagation (cxh_sp_forward, cxh_sp_reverse) against the
oracle's sp_forward / sp_reverse and the host's jac_sparsity, bit for bit."""
import numpy as np
import pytest

import oracle_ref as O
from conftest import seed_of

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ca():
    import casadi_amd
    return casadi_amd


def _functions(ca):
    from casadi_amd import workloads as W
    from casadi_amd.sx import gradient
    w, lam, Lag = W.quadrotor_ocp_lagrangian()
    G = ca.Function("grad_lag", [w, lam], [gradient(Lag, w)])
    return {"cartpole": W.cartpole_step(), "quadrotor": W.quadrotor_step(), "arm": W.arm_dynamics(),
            "rollout20": W.cartpole_rollout(20), "grad_lag": G}


def _words(rng, rows, nw):
    return rng.integers(-2 ** 63, 2 ** 63 - 1, (rows, nw), dtype=np.int64, endpoint=True)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def test_jac_sparsity_matches_host(ca, cuda):
    from casadi_amd.hip import HipTape
    for name, f in _functions(ca).items():
        t = HipTape(f)
        for o in range(f.n_out()):
            for i in range(f.n_in()):
                print(f"[case] jac_sparsity {name} out={o} in={i}", flush=True)
                assert t.jac_sparsity(o, i) == f.jac_sparsity(o, i), (name, o, i)


@pytest.mark.parametrize("nw", [1, 63, 65, 1000])
def test_random_words_match_oracle(ca, cuda, nw):
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.quadrotor_step()
    t = HipTape(f)
    rng = np.random.default_rng(seed_of(40) + nw)
    seeds = [_words(rng, f.nnz_in(i), nw) for i in range(f.n_in())]
    seeds[1] = None                                            # a null input reads as zero
    print(f"[case] sp_forward quadrotor n_words={nw}", flush=True)
    out = t.sp_forward([None if s is None else _dev(s) for s in seeds], nw)
    ref = O.sp_forward(f, [None if s is None else s.view(np.uint64) for s in seeds])
    for o, r in zip(out, ref):
        assert np.array_equal(o.cpu().numpy().view(np.uint64), r)
    rs = [_words(rng, f.nnz_out(j), nw) for j in range(f.n_out())]
    print(f"[case] sp_reverse quadrotor n_words={nw}", flush=True)
    dev = [_dev(s) for s in rs]
    sens = t.sp_reverse(dev, nw)
    ref = O.sp_reverse(f, [s.view(np.uint64).copy() for s in rs])
    for a, r in zip(sens, ref):
        assert np.array_equal(a.cpu().numpy().view(np.uint64), r)
    assert all(np.array_equal(d.cpu().numpy(), s) for d, s in zip(dev, rs))   # the caller's seeds stay


def test_work_vector_in_device_memory(ca, cuda):
    """worksize 9001 > 8192 words: the work vector does not fit LDS for even
    one lane and lives in device memory."""
    from casadi_amd.hip import HipTape
    x = ca.SX.sym("x", 9000)
    f = ca.Function("wide", [x], [x * ca.sum1(x)])
    assert f.sz_w() > 8192
    t = HipTape(f)
    rng = np.random.default_rng(seed_of(41))
    for nw in (3, 70):
        s = _words(rng, 9000, nw)
        print(f"[case] sp sweeps wide n_words={nw}", flush=True)
        (out,) = t.sp_forward([_dev(s)], nw)
        (ref,) = O.sp_forward(f, [s.view(np.uint64)])
        assert np.array_equal(out.cpu().numpy().view(np.uint64), ref)
        (sens,) = t.sp_reverse([_dev(s)], nw)
        (ref,) = O.sp_reverse(f, [s.view(np.uint64).copy()])
        assert np.array_equal(sens.cpu().numpy().view(np.uint64), ref)


def test_no_words(ca, cuda):
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_step()
    t = HipTape(f)
    import torch
    out = t.sp_forward([torch.zeros((4, 0), dtype=torch.int64, device="cuda"), None], 0)
    assert out[0].shape == (4, 0)
