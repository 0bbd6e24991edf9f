"""Host side of the tape specialiser: segment plans and generated source.

No GPU is needed: tapes are created on the host, the generated HIP source is
compiled for gfx950 through hiprtc, and the structure of the checkpoint plan
is checked against the tape's own slot liveness computed here in numpy.
"""
import re

import numpy as np
import pytest

from casadi_amd.ops import OP


@pytest.fixture(scope="module")
def ca():
    import casadi_amd
    return casadi_amd


def _live_at(f, bounds):
    """Slots live at each boundary (written before, read at or after)."""
    a = f.instructions()
    live = set()
    out = {}
    b = sorted(bounds, reverse=True)
    if len(a) in b:
        out[len(a)] = set()
    for k in range(len(a) - 1, -1, -1):
        op, i0, i1, i2 = (int(a[k][x]) for x in ("op", "i0", "i1", "i2"))
        if op == OP.OUTPUT:
            live.add(i1)
        else:
            live.discard(i0)
            if op not in (OP.CONST, OP.INPUT):
                live.add(i1)
                if op in (OP.ADD, OP.SUB, OP.MUL, OP.DIV, OP.POW, OP.CONSTPOW, OP.LT, OP.LE, OP.EQ, OP.NE,
                          OP.AND, OP.OR, OP.COPYSIGN, OP.FMOD, OP.REMAINDER, OP.FMIN, OP.FMAX, OP.ATAN2,
                          OP.PRINTME, OP.LIFT, OP.HYPOT, OP.IF_ELSE_ZERO):
                    live.add(i2)
        if k in b:
            out[k] = set(live)
    return out


def test_segment_state_is_the_live_set(ca, monkeypatch):
    monkeypatch.setenv("CXH_JIT_ROLL", "0")   # the straight-line segments, not the loop the rollout rerolls into
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_rollout(5)
    t = HipTape(f, engine="jit")
    every = 200
    ni = f.n_instructions()
    nseg = (ni + every - 1) // every
    live = _live_at(f, [s * every for s in range(nseg + 1)] + [ni])
    for s in range(nseg):
        src, ns = t.jit_source(every, s, reverse=False)
        assert ns == nseg
        # a live-in value is loaded where the segment reads it, or copied to
        # its live-out row unread (one by one or in a loop over a run of rows)
        loads = {int(m) for m in re.findall(r"const double w(\d+) = (?:CXJ_ST_IN\(|cxj_q\d+\[)", src)}
        direct = re.findall(r"CXJ_ST_OUT\(\d+\) = CXJ_ST_IN\(\d+\);", src)
        runs = [int(m) for m in re.findall(r"for \(int i = 0; i < (\d+); \+\+i\) CXJ_ST_OUT\(\d+ \+ i\) = CXJ_ST_IN", src)]
        passed = len(direct) + sum(runs)
        stores = len(re.findall(r"CXJ_ST_OUT\(\d+\) =", src)) + sum(runs)
        lin = live[s * every] if s > 0 else set()
        assert loads <= lin and len(loads) + passed == len(lin)
        assert stores == (len(live[min((s + 1) * every, ni)]) if s + 1 < nseg else 0)
        rsrc, _ = t.jit_source(every, s, reverse=True)
        assert len(re.findall(r"CXJ_ADJ_OUT\(\d+\) =", rsrc)) == (len(live[s * every]) if s > 0 else 0)
    # checkpoint rows = sum of the interior live sets, far below nseg * worksize
    ck = t.ckpt_size(64, every)
    rows = sum(len(live[s * every]) for s in range(1, nseg))
    assert ck == rows * 64
    assert rows < (nseg - 1) * f.sz_w()


def test_generated_sources_compile_for_gfx950(ca):
    import sys
    import os
    sys.path.insert(0, os.path.dirname(__file__))
    from test_gpu_jit import _all_ops_function
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape, jit_compile_check
    f, _ = _all_ops_function(ca)
    t = HipTape(f, engine="jit")
    for rev in (False, True):
        src, ns = t.jit_source(0, 0, rev)
        assert ns == 1
        jit_compile_check(src)
    g = W.cartpole_rollout(3)
    t = HipTape(g, engine="jit")
    src, ns = t.jit_source(150, 1, True)
    assert ns == (g.n_instructions() + 149) // 150
    jit_compile_check(src)


def test_compile_errors_are_reported(ca):
    from casadi_amd.hip import jit_compile_check
    from casadi_amd import CasadiHipError
    with pytest.raises(CasadiHipError, match="hiprtc compile failed"):
        jit_compile_check("extern \"C\" __global__ void k() { undefined_symbol(); }")


def test_engine_selection_is_validated(ca):
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    t = HipTape(W.cartpole_step())
    with pytest.raises(ValueError):
        t.set_engine("cuda")
    t.set_engine("jit")
    assert t.engine == "jit"


def test_batch_kernel_source_compiles(ca):
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape, batch_source, jit_compile_check
    fs = [W.cartpole_step(), W.quadrotor_step(), W.arm_dynamics()]
    src = batch_source([HipTape(f) for f in fs])
    # one fused body per Function and a dispatch case for each
    for i in range(len(fs)):
        assert f"cxj_f{i}(A, k)" in src and f"void cxj_f{i}(const CxjArgs& A" in src
    jit_compile_check(src)


def test_long_programs_call_one_sincos(ca, monkeypatch):
    """Programs with more than 32 sin/cos sites call one out-of-line copy
    (CXH_JIT_TRIG_CALLS sets the limit; 0 always calls, negative never);
    shorter ones inline it.  Both forms compile."""
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape, jit_compile_check
    step = HipTape(W.cartpole_step(), engine="jit")
    src, _ = step.jit_source(0, 0, True)
    assert src.count("  CXJ_SINCOS(") == 4 and not src.startswith("#define CXJ_TRIG_CALLS")
    roll = W.cartpole_rollout(10)
    long_src, _ = HipTape(roll, engine="jit").jit_source(0, 0, True)
    assert long_src.count("  CXJ_SINCOS(") > 32 and long_src.startswith("#define CXJ_TRIG_CALLS\n")
    jit_compile_check(long_src)
    monkeypatch.setenv("CXH_JIT_TRIG_CALLS", "0")
    assert HipTape(W.cartpole_step(), engine="jit").jit_source(0, 0, True)[0].startswith("#define CXJ_TRIG_CALLS\n")
    monkeypatch.setenv("CXH_JIT_TRIG_CALLS", "-1")
    assert not HipTape(roll, engine="jit").jit_source(0, 0, True)[0].startswith("#define CXJ_TRIG_CALLS")
