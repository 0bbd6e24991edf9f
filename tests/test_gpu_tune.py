"""Launch-shape autotuner of one map (cxh_tape_tune) and the settings it
searches: every setting gives the interpreter's / the specialised kernels'
same bits, ragged batches included, and the tuned setting stays on the tape.

Settings: threads per block; for the interpreter the lanes per instance
(instances per wavefront 64/32/16/8) and the work vector in LDS or HBM; for
the specialised kernels the instances per thread (1/2/4).
"""
import numpy as np
import pytest

import oracle_ref as O
from conftest import seed_of

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    import torch
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (4, n))
    u = rng.uniform(-10, 10, (1, n))
    return x, u, [torch.tensor(a, dtype=torch.float64, device="cuda") for a in (x, u)]


@pytest.mark.parametrize("engine", ["interp", "jit"])
def test_tune_keeps_the_bits(cuda, engine):
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_step()
    n = 20011                                    # not a multiple of any block x instances per thread
    x, u, dev = _inputs(n, seed_of(21))
    print(f"[case] tune engine={engine} n={n}", flush=True)
    tape = HipTape(f, engine=engine)
    (y0,) = [t.cpu().numpy() for t in tape.eval(dev)]
    cfg, ms = tape.tune(dev, reps=3)
    print(f"[step] tuned {cfg} {ms:.4f} ms", flush=True)
    assert ms > 0 and cfg["block"] in (64, 128, 256)
    (y1,) = [t.cpu().numpy() for t in tape.eval(dev)]
    assert np.array_equal(y0, y1)
    # every setting the tuner searches, set by hand, the same bits
    if engine == "jit":
        for ipt in (1, 2, 4):
            for block in (64, 256):
                tape.config(block=block)
                tape.ipt(ipt)
                (y,) = [t.cpu().numpy() for t in tape.eval(dev)]
                assert np.array_equal(y0, y), (ipt, block)
    else:
        for ipw in (8, 16, 32, 64):
            for lds in (True, False):
                tape.config(block=128, w_in_lds=lds, ipw=ipw)
                (y,) = [t.cpu().numpy() for t in tape.eval(dev)]
                assert np.array_equal(y0, y), (ipw, lds)
    # tuning a tape that compiles asynchronously times the specialised
    # kernels (built in the calling thread meanwhile), same bits after
    if engine == "jit":
        t2 = HipTape(f, engine="jit")
        t2.jit_async(True)
        cfg2, _ = t2.tune(dev, reps=2)
        assert not t2.jit_pending() and cfg2["block"] in (64, 128, 256)
        (y,) = [t.cpu().numpy() for t in t2.eval(dev)]
        assert np.array_equal(y0, y)
    # and the oracle's values on a sample
    m = 64
    (yo,) = O.map_eval(f, m, [x[:, :m].T, u[:, :m].T])
    assert O.rel_err(y0[:, :m], yo.T) <= 1e-13


def test_tune_settings_validated(cuda):
    from casadi_amd import _lib as L
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    tape = HipTape(W.cartpole_step(), engine="jit")
    assert L.cxh_tape_ipt(tape._h, 3) < 0
    assert "ipt must be" in L.cxh_last_error().decode()
    assert L.cxh_tape_tune(tape._h, 0, None, None, 0, 0, 1, None, None, None) < 0
