"""The no-allocation reduced map (cxh_map_eval_sum_noalloc) against the
allocating one (cxh_map_eval_sum), called through the C ABI.

The variant runs the same kernels in the same order with its scratch in the
library's static device array, so its results are the allocating call's
bits: two calls in a row, the second over a smaller batch (the array keeps
the first call's contents), on both engines.  Its refusals are checked too:
a batch whose summed rows do not fit the array, and a tape whose kernels the
allocating call has not loaded yet (nothing is launched in either case).
"""
import math

import numpy as np
import pytest

from conftest import seed_of

pytestmark = pytest.mark.gpu


def _call(fn, tape, args, outs, ri, ro, n, ld):
    from casadi_amd import _lib as L
    from casadi_amd.hip import _ptrs
    return fn(tape._h, n, _ptrs(args), _ptrs(outs), (L.I32 * len(ri))(*ri), (L.I32 * len(ro))(*ro), ld, None)


@pytest.mark.parametrize("engine", ["interp", "jit"])
def test_noalloc_reduced_map_matches_allocating(cuda, engine):
    import torch
    from casadi_amd import _lib as L
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_step()
    tape = HipTape(f, engine=engine)
    rng = np.random.default_rng(seed_of(0) + 7)
    ri, ro = [0, 1], [1]                       # the control shared, the next states summed
    for n in (300000, 70001):                  # the second call smaller, ragged
        print(f"[case] noalloc reduced map engine={engine} n={n}", flush=True)
        x = torch.tensor(rng.uniform(-1, 1, (4, n)), dtype=torch.float64, device="cuda")
        u = torch.tensor(rng.uniform(-10, 10, (1,)), dtype=torch.float64, device="cuda")
        got = []
        for fn in (L.cxh_map_eval_sum, L.cxh_map_eval_sum_noalloc):
            y = torch.full((4,), float("nan"), dtype=torch.float64, device="cuda")
            L.check(_call(fn, tape, [x, u], [y], ri, ro, n, n), fn.__name__)
            torch.cuda.synchronize()
            got.append(y.cpu().numpy())
        assert np.array_equal(got[0], got[1])
        # the sum itself: within the rounding bound of the instance values' fsum
        (yi,) = tape.eval([x, u.reshape(1, 1).expand(1, n).contiguous()], n)
        inst = yi.cpu().numpy()
        fs = np.array([math.fsum(r) for r in inst])
        bound = (n - 1) * np.finfo(float).eps * np.abs(inst).sum(axis=1)
        assert np.all(np.abs(got[1] - fs) <= bound)


def test_noalloc_refusals(cuda):
    import torch
    from casadi_amd import _lib as L
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    import casadi_amd as ca
    # a program no call in this process has compiled: the step scaled by a
    # constant of its own, so its kernel source is new
    xs, us = ca.SX.sym("x", 4), ca.SX.sym("u")
    scale = 1.0 + float(np.random.default_rng(seed_of(0) + 8).uniform(1e-3, 1e-2))
    f = ca.Function("scaled_step", [xs, us], [scale * W.cartpole_step()(xs, us)])
    x = torch.full((4, 1000), 0.25, dtype=torch.float64, device="cuda")
    u = torch.full((1,), 0.5, dtype=torch.float64, device="cuda")
    y = torch.zeros((4,), dtype=torch.float64, device="cuda")
    # its kernels are not loaded yet, and loading them would allocate
    fresh = HipTape(f, engine="jit")
    assert _call(L.cxh_map_eval_sum_noalloc, fresh, [x, u], [y], [0, 1], [1], 1000, 1000) < 0
    assert "no-allocation" in L.cxh_last_error().decode()
    # after one allocating call the same call goes through, with its bits
    y0 = torch.zeros((4,), dtype=torch.float64, device="cuda")
    L.check(_call(L.cxh_map_eval_sum, fresh, [x, u], [y0], [0, 1], [1], 1000, 1000), "cxh_map_eval_sum")
    L.check(_call(L.cxh_map_eval_sum_noalloc, fresh, [x, u], [y], [0, 1], [1], 1000, 1000), "noalloc")
    torch.cuda.synchronize()
    assert torch.equal(y, y0)
    # 4 summed rows over 4M instances need more than the 8 Mi doubles
    big = 4 * 1024 * 1024
    assert _call(L.cxh_map_eval_sum_noalloc, fresh, [x, u], [y], [0, 1], [1], big, big) < 0
    assert "static array holds" in L.cxh_last_error().decode()
