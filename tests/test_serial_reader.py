"""Reader of CasADi's serialized SXFunction (cxh_tape_deserialize).

The reference holds no serialized Function to read (its compatibility test
reads serialize_3.5.5/, which is absent), so the streams are assembled by
hand in tests/serial_stream.py, field by field in the reference's order.
Pinned here: the hand-traced tape of a tiny Function comes out of a stream
whose every field was written by hand, in plain and in debug mode (type
letters and field names, which the reader checks); shared objects written
once and referenced after; a realistic graph (the cart-pole step, its SX
node graph walked from the host symbolic layer with shared subexpressions)
gives back the tape the host compiler emits; malformed streams are refused.
Beyond these, reading real CasADi output is parity unpinned.
"""
import numpy as np
import pytest

import casadi_amd as ca
import serial_stream as S
from casadi_amd.hip import HipTape

SCALAR = [1, 1, 0, 1, 0]            # compressed 1x1 dense pattern, canonical form


def tiny_stream(debug=False):
    """f(x, y) = [(x + y) sin(x) - y, x + y]: the Function of the hand-traced
    tape in test_symbolic.py."""
    s = S.Stream(debug)
    S.function_header(s, "f", [SCALAR, None], [None, None], ["x", "y"], ["o0", "o1"],
                      sp_refs_in=[None, 0], sp_refs_out=[0, 0])
    s.integer(1, "XFunction::serialization::version")
    # XFunction::in: two 1x1 SX matrices (casadi/core/matrix_impl.hpp:3049-3052)
    s._name("XFunction::in")
    s._deco("V")
    s.casadi_int(2)
    ids = {}
    for nm in ("x", "y"):
        s.sparsity(None, "Matrix::sparsity", ref=0)
        s._name("Matrix::nonzeros")
        s._deco("V")
        s.casadi_int(1)
        ids[nm] = s.shared_def("E", S.sx_symbol(s, nm))
    # the algorithm (hand traced: test_symbolic.py::test_tape_hand_traced)
    alg = [(45, 0, 0, 0), (45, 1, 1, 0), (1, 2, 0, 1), (13, 0, 0, 0), (3, 0, 2, 0), (2, 0, 0, 1),
           (46, 0, 0, 0), (46, 1, 2, 0)]
    S.sx_function_body(s, 3, alg, 0, 0)
    # XFunction::out: d = (a b) - y with a = x + y, b = sin x; then a again
    s._name("XFunction::out")
    s._deco("V")
    s.casadi_int(2)
    s.sparsity(None, "Matrix::sparsity", ref=0)
    s._name("Matrix::nonzeros")
    s._deco("V")
    s.casadi_int(1)
    a_ix = {}

    def a_def():
        S.sx_node(s, 1, [("ref", ids["x"]), ("ref", ids["y"])])()

    def c_def():
        s.casadi_int(3, "SXNode::op")
        a_ix["a"] = s.shared_def("E", a_def, "UnarySX::dep0")
        s.shared_def("E", S.sx_node(s, 13, [("ref", ids["x"])]), "UnarySX::dep1")

    s.shared_def("E", S.sx_node(s, 2, [("def", c_def), ("ref", ids["y"])]))
    s.sparsity(None, "Matrix::sparsity", ref=0)
    s._name("Matrix::nonzeros")
    s._deco("V")
    s.casadi_int(1)
    s.shared_ref("E", a_ix["a"])
    return s.text(), alg


def _records(tape):
    a, ws = tape.instructions()
    return [(int(e["op"]), int(e["i0"]), int(e["i1"]), int(e["i2"])) for e in a], ws


@pytest.mark.parametrize("debug", [False, True])
def test_tiny_function_read(debug):
    text, alg = tiny_stream(debug)
    assert set(text) <= set("abcdefghijklmnop")
    t = HipTape.deserialize(text)
    assert _records(t) == (alg, 3)
    assert t.nnz_in == [1, 1] and t.nnz_out == [1, 1]
    # the tape the host compiler emits for the same Function
    x, y = ca.SX.sym("x"), ca.SX.sym("y")
    a = x + y
    f = ca.Function("f", [x, y], [a * ca.sin(x) - y, a])
    assert _records(t) == _records(HipTape(f))


def test_debug_mode_checks_field_names():
    text, _ = tiny_stream(True)
    # rename one field in place: "ProtoFunction::name" -> "ProtoFunction::namE"
    enc = lambda b: "".join(chr(97 + (c & 15)) + chr(97 + (c >> 4)) for c in b)
    bad = text.replace(enc(b"ProtoFunction::name"), enc(b"ProtoFunction::namE"), 1)
    assert bad != text
    with pytest.raises(ca.CasadiHipError, match="field 'ProtoFunction::namE'"):
        HipTape.deserialize(bad)


def test_malformed_streams_refused():
    text, _ = tiny_stream(False)
    with pytest.raises(ca.CasadiHipError, match="sanity check"):
        HipTape.deserialize("b" + text[1:])
    with pytest.raises(ca.CasadiHipError, match="ends early"):
        HipTape.deserialize(text[:-40])
    with pytest.raises(ca.CasadiHipError, match="data after"):
        HipTape.deserialize(text + "aa")
    with pytest.raises(ca.CasadiHipError, match="outside"):
        HipTape.deserialize(text[:100] + "z" + text[101:])
    # another Function class
    s = S.Stream()
    s.boolean(False, "Function::null")
    s.string("MXFunction", "FunctionInternal::base_function")
    with pytest.raises(ca.CasadiHipError, match="only SXFunction"):
        HipTape.deserialize(s.text())
    # a null Function
    s = S.Stream()
    s.boolean(True, "Function::null")
    with pytest.raises(ca.CasadiHipError, match="null"):
        HipTape.deserialize(s.text())


def _graph_writer(s):
    """Writes SX nodes of the host symbolic layer as the serializer would:
    each node defined once (its dependencies first), referenced after."""
    from casadi_amd import _lib as L
    from casadi_amd.ops import OP
    index = {}

    def info(h):
        op, v, d0, d1 = L.I32(), L.D(), L.U64(), L.U64()
        L.check(L.cxs_node_info(h, L.C.byref(op), L.C.byref(v), L.C.byref(d0), L.C.byref(d1)))
        return op.value, v.value, d0.value, d1.value

    def put(h, name=""):
        if h in index:
            s.shared_ref("E", index[h], name)
            return
        op, v, d0, d1 = info(h)

        def body():
            if op == OP.CONST:
                S.sx_real(s, v)()
            elif op in (OP.PARAMETER, OP.INPUT) or (d0 == 0 and d1 == 0):
                S.sx_symbol(s, f"s{h % 1000}")()
            else:
                s.casadi_int(op, "SXNode::op")
                if d1 and op != OP.SQ and _binary(op):
                    put(d0, "UnarySX::dep0")
                    put(d1, "UnarySX::dep1")
                else:
                    put(d0, "UnarySX::dep")
        index[h] = s.shared_def("E", body, name)

    return put


def _binary(op):
    from casadi_amd.ops import BINARY
    return op in [int(b) for b in BINARY]


def test_cartpole_step_stream_gives_the_compiled_tape():
    """The cart-pole RK4 step: every node of its input and output graphs
    written (shared subexpressions by reference), the algorithm as the host
    compiler emits it, constants in CONST records.  The reader must skip the
    graphs exactly and hand over that algorithm."""
    from casadi_amd import workloads as W
    x, u = ca.SX.sym("x", 4), ca.SX.sym("u")
    xn = W.rk4(W.cartpole_ode, x, u, 0.05)
    f = ca.Function("cartpole", [x, u], [xn])
    a = f.instructions()
    alg = [(int(e["op"]), int(e["i0"]), int(e["i1"]), int(e["i2"])) for e in a]
    s = S.Stream()
    S.function_header(s, "cartpole", [[4, 1, 0, 4, 0, 1, 2, 3], SCALAR], [None], ["x", "u"], ["xf"],
                      sp_refs_out=[0])
    s.integer(1, "XFunction::serialization::version")
    put = _graph_writer(s)
    s._name("XFunction::in")
    s._deco("V")
    s.casadi_int(2)
    for ex, ref in ((x, 0), (u, 1)):
        s.sparsity(None, "Matrix::sparsity", ref=ref)
        s.vector([int(h) for h in ex.nz], put, "Matrix::nonzeros")
    S.sx_function_body(s, f.sz_w(), alg, 0, 0)
    s._name("XFunction::out")
    s._deco("V")
    s.casadi_int(1)
    s.sparsity(None, "Matrix::sparsity", ref=0)
    s.vector([int(h) for h in xn.nz], put, "Matrix::nonzeros")
    t = HipTape.deserialize(s.text())
    assert _records(t) == (alg, f.sz_w())
    assert t.nnz_in == [4, 1] and t.nnz_out == [4]
    assert s.n_shared > 100      # the graph was really written (and skipped by the reader)


def test_deep_expression_graph():
    """A chain of 200 000 nested sin nodes (an unrolled recurrence is this
    deep): the reader walks the graph without recursion, numbers the nodes
    after their dependencies, and hands over the algorithm."""
    from casadi_amd.ops import OP
    N = 200000
    s = S.Stream()
    S.function_header(s, "chain", [SCALAR], [None], ["x"], ["y"], sp_refs_out=[0])
    s.integer(1, "XFunction::serialization::version")
    s._name("XFunction::in")
    s._deco("V")
    s.casadi_int(1)
    s.sparsity(None, "Matrix::sparsity", ref=0)
    s._name("Matrix::nonzeros")
    s._deco("V")
    s.casadi_int(1)
    kx = s.shared_def("E", S.sx_symbol(s, "x"))
    alg = [(int(OP.INPUT), 0, 0, 0)] + [(int(OP.SIN), 0, 0, 0)] * N + [(int(OP.OUTPUT), 0, 0, 0)]
    S.sx_function_body(s, 1, alg, 0, 0)
    s._name("XFunction::out")
    s._deco("V")
    s.casadi_int(1)
    s.sparsity(None, "Matrix::sparsity", ref=0)
    s._name("Matrix::nonzeros")
    s._deco("V")
    s.casadi_int(1)
    for j in range(N):                  # sin(sin(...)): each node's dependency defined inside it
        s._name("" if j == 0 else "UnarySX::dep")
        s._deco("E")
        s.char("d", "Shared::flag")
        s.casadi_int(int(OP.SIN), "SXNode::op")
    s.shared_ref("E", kx, "UnarySX::dep")
    t = HipTape.deserialize(s.text())
    assert _records(t) == (alg, 1)
    # a reference to a node the chain has not finished defining is refused
    s2 = S.Stream()
    S.function_header(s2, "chain", [SCALAR], [None], ["x"], ["y"], sp_refs_out=[0])
    s2.integer(1, "XFunction::serialization::version")
    s2._name("XFunction::in")
    s2._deco("V")
    s2.casadi_int(1)
    s2.sparsity(None, "Matrix::sparsity", ref=0)
    s2._name("Matrix::nonzeros")
    s2._deco("V")
    s2.casadi_int(1)
    s2._deco("E")
    s2.char("d", "Shared::flag")
    s2.casadi_int(int(OP.SIN), "SXNode::op")
    s2.shared_ref("E", s2.n_shared, "UnarySX::dep")   # itself: not yet numbered
    with pytest.raises(Exception, match="not yet defined"):
        HipTape.deserialize(s2.text())
