"""Tapes with nested Function calls (OP_CALL): the inlining done by
cxh_tape_create_nested against the oracle's restatement of call_fwd, and the
oracle's restatement against the same computation written as one SX graph.

The reference's tests hold no tapes with call records.  What fixes the
semantics: SXFunction::call_fwd (casadi/core/sx_function.cpp:1596-1620) reads
every argument before the call and writes the results after it, so a result
slot may reuse an argument slot; a result slot of -1 is not stored.  The same
computation written as one SX graph (the callee's expressions substituted, as
SX does when a Function is called on symbols) is a second statement of it.
"""
import numpy as np
import pytest

import oracle_ref as O
from casadi_amd.function import INSTR_DTYPE
from casadi_amd.ops import OP


@pytest.fixture(scope="module")
def ca():
    import casadi_amd
    return casadi_amd


def _tape(recs):
    a = np.zeros(len(recs), INSTR_DTYPE)
    for k, (op, i0, i1, i2) in enumerate(recs):
        a[k] = (op, i0, i1, i2)
    return a


def nested_program(ca):
    """Outer tape over inputs X (4), U (1) calling the cart-pole step g twice:
        a = g(2 X, U)                       (results written over the argument slots)
        b = g([a0, a1, a2, 2 X3 * sin U], sin U)
        Y = b, Z = b3 + X0.
    Returns (instr, worksize, nnz_in, nnz_out, calls, the same as one SX Function)."""
    from casadi_amd import workloads as W
    g = W.cartpole_step()
    recs = [(OP.INPUT, i, 0, i) for i in range(4)] + [(OP.INPUT, 4, 1, 0)]
    recs += [(OP.TWICE, 5 + i, i, i) for i in range(4)]               # slots 5..8 = 2 X
    recs += [(OP.CALL, 5, 0, 0)]                                       # a -> slots 5, 6, 7 (a3 dropped)
    recs += [(OP.SIN, 9, 4, 4), (OP.MUL, 8, 8, 9)]                     # slot 8 = 2 X3 * sin U
    recs += [(OP.CALL, 10, 1, 0)]                                      # b -> slots 10..13
    recs += [(OP.OUTPUT, 0, 10 + i, i) for i in range(4)]
    recs += [(OP.ADD, 14, 13, 0), (OP.OUTPUT, 1, 14, 0)]
    calls = [(g, [5, 6, 7, 8, 4], [5, 6, 7, -1]), (g, [5, 6, 7, 8, 9], [10, 11, 12, 13])]
    x, u = ca.SX.sym("x", 4), ca.SX.sym("u")
    a = g.call([2 * x, u])[0]
    s = ca.sin(u)
    b = g.call([ca.vertcat(a[0], a[1], a[2], 2 * x[3] * s), s])[0]
    F = ca.Function("F", [x, u], [b, b[3] + x[0]])
    return _tape(recs), 15, [4, 1], [4, 1], calls, F


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-1, 1, (n, 4)), rng.uniform(-3, 3, (n, 1))]


def test_oracle_call_semantics_match_one_graph(ca):
    instr, ws, nin, nout, calls, F = nested_program(ca)
    args = _inputs(200, 1)
    got = O.eval_calls(instr, ws, nin, nout, calls, args)
    ref = O.map_eval(F, 200, args)
    for a, b in zip(got, ref):
        assert O.rel_err(a, b) <= 1e-14
    # reading the arguments after writing the results (a wrong call order) is caught:
    # call 0 writes slots 5..7, which are its own argument slots
    late = [(calls[0][0], [5, 6, 7, 8, 4], [5, 6, 7, -1])]
    seq = _tape([(OP.INPUT, i, 0, i) for i in range(4)] + [(OP.INPUT, 4, 1, 0)] +
                [(OP.TWICE, 5 + i, i, i) for i in range(4)] + [(OP.CALL, 5, 0, 0)] +
                [(OP.OUTPUT, 0, 5 + i, i) for i in range(3)] + [(OP.OUTPUT, 0, 8, 3), (OP.OUTPUT, 1, 4, 0)])
    (y, _) = O.eval_calls(seq, 9, [4, 1], [4, 1], late, args)
    step = O.map_eval(calls[0][0], 200, [2 * args[0], args[1]])[0]
    assert np.array_equal(y[:, :3], step[:, :3]) and np.array_equal(y[:, 3], 2 * args[0][:, 3])


def test_inlined_tape_equals_oracle_calls(ca):
    """cxh_tape_create_nested's flat tape, run on the oracle VM, gives the
    oracle's call-by-call result bit for bit."""
    from casadi_amd.hip import HipTape
    instr, ws, nin, nout, calls, _ = nested_program(ca)
    g = HipTape(calls[0][0])
    t = HipTape.nested(instr, ws, nin, nout, [(g, c[1], c[2]) for c in calls])
    flat, fws = t.instructions()
    assert not np.any(flat["op"] == OP.CALL) and fws > ws
    args = _inputs(300, 2)
    ref = O.eval_calls(instr, ws, nin, nout, calls, args)

    class Flat:  # the flat tape in the oracle VM's Function interface
        def instructions(self):
            return flat

        def sz_w(self):
            return fws

        def n_in(self):
            return 2

        def n_out(self):
            return 2

        def nnz_in(self, i):
            return nin[i]

        def nnz_out(self, i):
            return nout[i]
    got = O.map_eval(Flat(), 300, args)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    # a nested tape can be a callee in turn
    outer = _tape([(OP.INPUT, i, 0, i) for i in range(4)] + [(OP.INPUT, 4, 1, 0), (OP.CALL, 0, 0, 0)] +
                  [(OP.OUTPUT, 0, 5 + i, i) for i in range(4)] + [(OP.OUTPUT, 1, 9, 0)])
    t2 = HipTape.nested(outer, 10, nin, nout, [(t, [0, 1, 2, 3, 4], [5, 6, 7, 8, 9])])
    assert t2.instructions()[0].size > flat.size


def test_nested_tape_errors(ca):
    from casadi_amd import CasadiHipError
    from casadi_amd.hip import HipTape
    instr, ws, nin, nout, calls, _ = nested_program(ca)
    g = HipTape(calls[0][0])
    with pytest.raises(CasadiHipError, match="call index out of range"):
        HipTape.nested(instr, ws, nin, nout, [(g, calls[0][1], calls[0][2])])
    bad = [(g, [5, 6, 7, 8, 99], calls[0][2]), (g, calls[1][1], calls[1][2])]
    with pytest.raises(CasadiHipError, match="argument slot out of range"):
        HipTape.nested(instr, ws, nin, nout, bad)
    # without call records the tape is cxh_tape_create's
    from casadi_amd import workloads as W
    f = W.cartpole_step()
    t = HipTape.nested(f.instructions(), f.sz_w(), [4, 1], [4], [])
    flat, fws = t.instructions()
    assert np.array_equal(flat.view(np.uint8), f.instructions().view(np.uint8)) and fws == f.sz_w()
