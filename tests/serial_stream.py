"""A CasADi SerializingStream written field by field, for the reader's tests.

No serialized file of the reference is available here, so the tests assemble
the stream by hand, following the order in which the reference's classes
pack their fields.  Each method names the reference lines it follows.  Values
the reader skips (options, names) carry a fresh Function's defaults.
"""
import struct

CHECK = 123456789012345          # casadi/core/serializing_stream.cpp:34
PROTOCOL = 3                     # casadi/core/serializing_stream.cpp:33


class Stream:
    """SerializingStream (casadi/core/serializing_stream.cpp:56-160,
    serializing_stream.hpp:226-330): bytes as two letters 'a' + low nibble,
    'a' + high nibble; in debug mode a type letter before each value and the
    field's name before each named field."""

    def __init__(self, debug=False):
        self.out = []
        self.debug = False
        self.n_shared = 0
        self.casadi_int(CHECK)
        self.casadi_int(PROTOCOL)
        self.boolean(debug)
        self.debug = debug

    def text(self):
        return "".join(self.out)

    # -- primitives ------------------------------------------------------------
    def _bytes(self, b):
        for c in b:
            self.out.append(chr(ord("a") + (c & 15)) + chr(ord("a") + (c >> 4)))

    def _deco(self, e):
        if self.debug:
            self._bytes(e.encode())

    def _name(self, name):
        if self.debug and name:
            self.string(name)

    def char(self, c, name=""):                  # pack(char): no type letter
        self._name(name)
        self._bytes(c.encode() if isinstance(c, str) else bytes([c]))

    def integer(self, v, name=""):               # int: 'i', 4 bytes
        self._name(name)
        self._deco("i")
        self._bytes(struct.pack("<i", v))

    def casadi_int(self, v, name=""):            # casadi_int: 'J', 8 bytes
        self._name(name)
        self._deco("J")
        self._bytes(struct.pack("<q", v))

    def size_t(self, v, name=""):                # size_t: 'K', 8 bytes
        self._name(name)
        self._deco("K")
        self._bytes(struct.pack("<Q", v))

    def boolean(self, v, name=""):               # bool: 'b', 1 byte
        self._name(name)
        self._deco("b")
        self._bytes(bytes([1 if v else 0]))

    def real(self, v, name=""):                  # double: 'd', 8 bytes
        self._name(name)
        self._deco("d")
        self._bytes(struct.pack("<d", v))

    def string(self, s, name=""):                # string: 's', int length, bytes
        self._name(name)
        self._deco("s")
        self.integer(len(s))
        self._bytes(s.encode())

    def vector(self, items, put, name=""):       # vector: 'V', casadi_int length, elements
        self._name(name)
        self._deco("V")
        self.casadi_int(len(items))
        for x in items:
            put(x)

    def dict(self, d, name=""):                  # map: 'D', casadi_int size, pairs (only empty here)
        self._name(name)
        self._deco("D")
        self.casadi_int(len(d))
        assert not d

    # -- shared objects (serializing_stream.hpp:314-328) ---------------------------
    def shared_def(self, letter, body, name=""):
        """'d' and the definition; returns the object's index (assigned after
        its own nested definitions, as the serializer numbers them)."""
        self._name(name)
        self._deco(letter)
        self.char("d", "Shared::flag")
        body()
        k = self.n_shared
        self.n_shared += 1
        return k

    def shared_ref(self, letter, k, name=""):
        self._name(name)
        self._deco(letter)
        self.char("r", "Shared::flag")
        self.casadi_int(k, "Shared::reference")

    def sparsity(self, compressed, name="", ref=None):
        """Sparsity::serialize (casadi/core/sparsity.cpp:2079-2085)."""
        if ref is not None:
            self.shared_ref("S", ref, name)
            return ref
        return self.shared_def("S", lambda: self.vector(compressed, self.casadi_int, "SparsityInternal::compressed"),
                               name)

    def null_function(self, name="", ref=None):
        """A null Function member (pack(Function) -> shared -> Function::serialize,
        casadi/core/function.cpp:1482-1489)."""
        if ref is not None:
            self.shared_ref("F", ref, name)
            return ref
        return self.shared_def("F", lambda: self.boolean(True, "Function::null"), name)


# SX nodes (casadi/core/sx_node.cpp:238-241; symbolic_sx.hpp:73-75; constant_sx.hpp:171-174;
# unary_sx.hpp:123-125; binary_sx.hpp:135-138): "SXNode::op" as casadi_int, then the node's fields
def sx_symbol(s, name):
    return lambda: (s.casadi_int(47, "SXNode::op"), s.string(name, "SymbolicSX::name"))


def sx_real(s, value):
    return lambda: (s.casadi_int(44, "SXNode::op"), s.char("r", "ConstantSX::type"), s.real(value, "ConstantSX::value"))


def sx_node(s, op, deps):
    """deps: list of (kind, x) with kind 'def' (a body to define) or 'ref'."""
    def body():
        s.casadi_int(op, "SXNode::op")
        names = ["UnarySX::dep0", "UnarySX::dep1"] if len(deps) == 2 else ["UnarySX::dep"]
        for nm, (kind, x) in zip(names, deps):
            if kind == "ref":
                s.shared_ref("E", x, nm)
            else:
                s.shared_def("E", x, nm)
    return body


def function_header(s, name, sp_in, sp_out, names_in, names_out, sp_refs_in=None, sp_refs_out=None):
    """Function::serialize + serialize_type + ProtoFunction and FunctionInternal
    bodies (casadi/core/function.cpp:1482-1489; casadi/core/function_internal.cpp:
    4165-4172, 4185-4251), version 2 and 8, with a fresh Function's option values."""
    s.boolean(False, "Function::null")
    s.string("SXFunction", "FunctionInternal::base_function")
    s.integer(2, "ProtoFunction::serialization::version")
    s.string(name, "ProtoFunction::name")
    for f in ("verbose", "print_time", "record_time", "regularity_check", "error_on_fail"):
        s.boolean(f == "error_on_fail", "ProtoFunction::" + f)
    s.integer(8, "FunctionInternal::serialization::version")
    s.vector([True] * len(sp_in), s.boolean, "FunctionInternal::is_diff_in")
    s.vector([True] * len(sp_out), s.boolean, "FunctionInternal::is_diff_out")
    for lbl, sps, refs in (("sp_in", sp_in, sp_refs_in), ("sp_out", sp_out, sp_refs_out)):
        s._name("FunctionInternal::" + lbl)
        s._deco("V")
        s.casadi_int(len(sps))
        for j, sp in enumerate(sps):
            s.sparsity(sp, ref=None if refs is None else refs[j])
    s.vector(names_in, s.string, "FunctionInternal::name_in")
    s.vector(names_out, s.string, "FunctionInternal::name_out")
    s.boolean(False, "FunctionInternal::jit")
    s.boolean(True, "FunctionInternal::jit_cleanup")
    s.string("source", "FunctionInternal::jit_serialize")
    s.boolean(True, "FunctionInternal::jit_temp_suffix")
    s.string("jit_tmp", "FunctionInternal::jit_base_name")
    s.dict({}, "FunctionInternal::jit_options")
    s.string("clang", "FunctionInternal::compiler_plugin")
    s.boolean(False, "FunctionInternal::has_refcount")
    s.dict({}, "FunctionInternal::cache_init")
    null_fn = s.null_function("FunctionInternal::derivative_of")
    s.real(2.0, "FunctionInternal::jac_penalty")
    for f in ("forward", "reverse", "jacobian", "fd", "forward_op", "reverse_op", "jacobian_op", "fd_op"):
        s.boolean(f != "fd", "FunctionInternal::enable_" + f)
    s.real(-1.0, "FunctionInternal::ad_weight")
    s.real(-1.0, "FunctionInternal::ad_weight_sp")
    s.boolean(False, "FunctionInternal::always_inline")
    s.boolean(False, "FunctionInternal::never_inline")
    s.casadi_int(64, "FunctionInternal::max_num_dir")
    s.boolean(True, "FunctionInternal::inputs_check")
    s.real(1e-8, "FunctionInternal::fd_step")
    s.string("central", "FunctionInternal::fd_method")
    s.boolean(False, "FunctionInternal::print_in")
    s.boolean(False, "FunctionInternal::print_out")
    s.boolean(False, "FunctionInternal::print_canonical")
    s.casadi_int(10000, "FunctionInternal::max_io")
    s.boolean(False, "FunctionInternal::dump_in")
    s.boolean(False, "FunctionInternal::dump_out")
    s.string(".", "FunctionInternal::dump_dir")
    s.string("mtx", "FunctionInternal::dump_format")
    for f in ("forward_options", "reverse_options", "jacobian_options", "der_options"):
        s.dict({}, "FunctionInternal::" + f)
    s.null_function("FunctionInternal::custom_jacobian", ref=null_fn)   # the same null Function: a reference
    s.vector([], None, "FunctionInternal::registered_functions")
    for f in ("arg_per", "res_per", "iw_per", "w_per", "arg_tmp", "res_tmp", "iw_tmp", "w_tmp"):
        s.size_t(0, "FunctionInternal::sz_" + f)


def sx_function_body(s, worksize, algorithm, n_ops, n_consts):
    """SXFunction::serialize_body after XFunction::in
    (casadi/core/sx_function.cpp:1992-2031), version 3.  algorithm: list of
    (op, i0, i1, i2); free_vars empty, operations / constants left as empty
    vectors (the reader skips them), no calls."""
    s.integer(3, "SXFunction::serialization::version")
    s.size_t(len(algorithm), "SXFunction::n_instr")
    s.size_t(worksize, "SXFunction::worksize")
    s.vector([], None, "SXFunction::free_vars")
    s.vector([], None, "SXFunction::operations")
    s.vector([], None, "SXFunction::constants")
    s.vector([0.0, 0.0], s.real, "SXFunction::default_in")
    for f in ("call_sz_arg", "call_sz_res", "call_sz_iw", "call_sz_w", "call_sz_arg", "call_sz_res"):
        s.size_t(0, "SXFunction::" + f)
    s.size_t(0, "SXFunction::call_el_size")
    s.vector([False] * len(algorithm), s.boolean, "SXFunction::copy_elision")
    for op, i0, i1, i2 in algorithm:
        s.integer(op, "SXFunction::ScalarAtomic::op")
        s.integer(i0, "SXFunction::ScalarAtomic::i0")
        s.integer(i1, "SXFunction::ScalarAtomic::i1")
        s.integer(i2, "SXFunction::ScalarAtomic::i2")
    s.boolean(True, "SXFunction::live_variables")
    s.boolean(False, "SXFunction::print_instructions")
