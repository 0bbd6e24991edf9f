// Reader of CasADi's serialized SXFunction: the string Function::serialize()
// returns (and the stream Function::save writes after its file header).
//
// The format is CasADi's SerializingStream (casadi/core/serializing_stream.
// cpp:26-400, serializing_stream.hpp:226-330): after a header (the check
// number 123456789012345, protocol version 3, the debug flag), every field is
// written in the order the classes' serialize_body methods pack them, each
// byte as two letters 'a' + low nibble, 'a' + high nibble.  Scalars are
// little-endian (int 4 bytes, casadi_int / size_t / double 8, bool 1); a
// string is its int length and bytes; a vector its casadi_int length and
// elements; a map its casadi_int size and key / value pairs.  Shared objects
// (Sparsity, SX nodes, Functions, GenericTypes) are written once: a 'd' and
// the definition the first time, an 'r' and the index of the definition (one
// count over all shared objects) after that.  In debug mode every value is
// preceded by a type letter and every named field by its name.  Nothing is
// framed, so a reader has to walk every field in order.
//
// The reader walks an SXFunction body -- ProtoFunction
// (casadi/core/function_internal.cpp:4165-4183), FunctionInternal (:4189-4340),
// XFunction (casadi/core/x_function.hpp:306-332), SXFunction
// (casadi/core/sx_function.cpp:1926-2041) -- with the version branches of the
// reference's deserialising constructors, parses the SX node graphs of the
// inputs and outputs (casadi/core/sx_node.cpp:238-274, symbolic_sx.hpp,
// constant_sx.hpp, unary_sx.hpp, binary_sx.hpp), GenericType options
// (casadi/core/generic_type.cpp:773-818) and Sparsity (casadi/core/sparsity.
// cpp:2079-2095), and hands the algorithm (ScalarAtomic op / i0 / i1 / i2,
// the constant's double in i1 / i2), worksize and input / output nonzero
// counts to cxh_tape_create.  Nested Functions are parsed recursively.

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/casadi_hip.h"
#include "capi_common.hpp"
#include "sx_ops.h"

namespace sxh {
namespace {

constexpr int64_t kCheck = 123456789012345LL;
constexpr int64_t kProtocol = 3;

// What the reader keeps of an SXFunction.
struct SerialFunction {
  std::string name;
  std::vector<int64_t> nnz_in, nnz_out;
  std::vector<Instr> alg;
  int64_t worksize = 0;
  int64_t n_free = 0, n_calls = 0;
};

class Reader {
 public:
  Reader(const char* p, int64_t n) : p_(p), end_(p + n) {}

  void header() {
    const int64_t check = casadi_int("");
    if (check != kCheck) throw std::runtime_error("not a CasADi serialization (sanity check number differs)");
    const int64_t v = casadi_int("");
    if (v != kProtocol)
      throw std::runtime_error("serialization protocol " + std::to_string(v) + ", this reader reads " +
                               std::to_string(kProtocol));
    debug_ = boolean("");
  }

  // Function::serialize(SerializingStream&) (casadi/core/function.cpp:1482-1489).
  bool function(SerialFunction* out) {
    if (boolean("Function::null")) return false;
    const std::string base = str("FunctionInternal::base_function");
    if (base != "SXFunction")
      throw std::runtime_error("serialized Function of type " + base + ": only SXFunction is read");
    sx_function(out);
    return true;
  }

 private:
  const char* p_;
  const char* end_;
  bool debug_ = false;
  int64_t n_shared_ = 0;

  [[noreturn]] void fail(const std::string& what) const { throw std::runtime_error("serialized stream: " + what); }

  unsigned char byte() {
    if (end_ - p_ < 2) fail("ends early");
    const unsigned char lo = static_cast<unsigned char>(p_[0]) - 'a', hi = static_cast<unsigned char>(p_[1]) - 'a';
    if (lo > 15 || hi > 15) fail("character outside 'a'..'p'");
    p_ += 2;
    return static_cast<unsigned char>(lo | (hi << 4));
  }
  void bytes(void* dst, int n) {
    unsigned char* d = static_cast<unsigned char*>(dst);
    for (int j = 0; j < n; ++j) d[j] = byte();
  }
  void deco(char e) {
    if (!debug_) return;
    const char t = static_cast<char>(byte());
    if (t != e) fail(std::string("type letter '") + t + "' where '" + e + "' belongs");
  }
  // In debug mode a named field is preceded by its name.
  void field(const char* name) {
    if (!debug_ || !*name) return;
    const std::string got = raw_string();
    if (got != name) fail("field '" + got + "' where '" + name + "' belongs");
  }

  int32_t raw_int() {
    deco('i');
    int32_t v;
    bytes(&v, 4);
    return v;
  }
  std::string raw_string() {
    deco('s');
    const int32_t n = raw_int();
    if (n < 0 || n > (end_ - p_) / 2) fail("bad string length");
    std::string s(static_cast<size_t>(n), '\0');
    for (int32_t j = 0; j < n; ++j) s[j] = static_cast<char>(byte());
    return s;
  }

 public:
  int32_t integer(const char* name) {
    field(name);
    return raw_int();
  }
  int64_t casadi_int(const char* name) {
    field(name);
    deco('J');
    int64_t v;
    bytes(&v, 8);
    return v;
  }
  uint64_t size(const char* name) {
    field(name);
    deco('K');
    uint64_t v;
    bytes(&v, 8);
    return v;
  }
  bool boolean(const char* name) {
    field(name);
    deco('b');
    return byte() != 0;
  }
  double real(const char* name) {
    field(name);
    deco('d');
    double v;
    bytes(&v, 8);
    return v;
  }
  char character(const char* name) {
    field(name);
    return static_cast<char>(byte());
  }
  std::string str(const char* name) {
    field(name);
    return raw_string();
  }
  int64_t vector_length(const char* name) {
    field(name);
    deco('V');
    const int64_t n = casadi_int("");
    if (n < 0 || n > (end_ - p_)) fail("bad vector length");
    return n;
  }
  std::vector<int64_t> int64_vector(const char* name) {
    std::vector<int64_t> v(static_cast<size_t>(vector_length(name)));
    for (auto& x : v) x = casadi_int("");
    return v;
  }
  std::vector<int32_t> int_vector(const char* name) {
    std::vector<int32_t> v(static_cast<size_t>(vector_length(name)));
    for (auto& x : v) x = integer("");
    return v;
  }
  void bool_vector(const char* name) {
    const int64_t n = vector_length(name);
    for (int64_t j = 0; j < n; ++j) boolean("");
  }
  void double_vector(const char* name) {
    const int64_t n = vector_length(name);
    for (int64_t j = 0; j < n; ++j) real("");
  }
  std::vector<std::string> string_vector(const char* name) {
    std::vector<std::string> v(static_cast<size_t>(vector_length(name)));
    for (auto& s : v) s = str("");
    return v;
  }

  // A shared object: the definition (read by def) the first time, a
  // reference to an earlier one after that.  Returns the object's index.
  template <class Def>
  int64_t shared(char letter, Def def) {
    deco(letter);
    const char flag = character("Shared::flag");
    if (flag == 'd') {
      def();
      return n_shared_++;
    }
    if (flag == 'r') {
      const int64_t k = casadi_int("Shared::reference");
      if (k < 0 || k >= n_shared_) fail("reference to an object not yet defined");
      return k;
    }
    fail(std::string("shared-object flag '") + flag + "'");
  }

  // Sparsity::serialize: the compressed pattern [nrow, ncol, colind, row],
  // or [nrow, ncol, 1] for a dense one.  Returns its nonzero count (a
  // reference repeats an earlier pattern's).
  int64_t sparsity(const char* name) {
    field(name);
    int64_t n = -1;
    const int64_t idx = shared('S', [&] {
      const std::vector<int64_t> v = int64_vector("SparsityInternal::compressed");
      if (v.empty()) {  // the null pattern
        n = 0;
        return;
      }
      if (v.size() < 3) fail("compressed sparsity too short");
      const int64_t nrow = v[0], ncol = v[1];
      if (nrow < 0 || ncol < 0) fail("negative sparsity dimensions");
      if (ncol > 0 && nrow > INT64_MAX / ncol) fail("sparsity dimensions overflow");
      if (v[2] == 1 && ncol > 0 && v.size() == 3) {
        n = nrow * ncol;  // dense, short form
      } else {
        if (static_cast<int64_t>(v.size()) < 3 + ncol) fail("compressed sparsity too short");
        n = v[2 + ncol];
        if (static_cast<int64_t>(v.size()) != 3 + ncol + n && n != nrow * ncol) fail("bad compressed sparsity");
      }
    });
    if (n >= 0) {
      sp_nnz_.push_back({idx, n});
      return n;
    }
    for (const auto& e : sp_nnz_)
      if (e.first == idx) return e.second;
    fail("reference to an object that is not a sparsity pattern");
  }

  // SXElem: an SX node graph (SXNode::serialize, casadi/core/sx_node.cpp:238-274).
  // A node's definition holds its dependencies' (definitions or references)
  // and it is numbered after them.  Read with an explicit stack: the graph of
  // a long unrolled recurrence is thousands of nodes deep.
  void sx_elem(const char* name) {
    struct Open {
      int n_dep, next;  // dependencies of the node, the next one to read
    };
    std::vector<Open> open;
    const char* fname = name;
    for (;;) {
      field(fname);
      deco('E');
      const char flag = character("Shared::flag");
      bool complete = true;  // this element is read in full
      if (flag == 'r') {
        const int64_t k = casadi_int("Shared::reference");
        if (k < 0 || k >= n_shared_) fail("reference to an object not yet defined");
      } else if (flag == 'd') {
        const int64_t op = casadi_int("SXNode::op");
        if (op_ndeps_binary(op) || op_ndeps_unary(op)) {
          open.push_back({op_ndeps_binary(op) ? 2 : 1, 0});
          complete = false;
        } else if (op == OP_PARAMETER) {
          str("SymbolicSX::name");
        } else if (op == OP_CONST) {
          const char type = character("ConstantSX::type");
          if (type == 'r') real("ConstantSX::value");
          else if (type == 'i') integer("ConstantSX::value");
          else if (type == 0 || std::strchr("01nfFm", type) == nullptr) fail("constant type");
        } else {
          fail("SX node op " + std::to_string(op) + " (call / output nodes are not read)");
        }
        if (complete) ++n_shared_;
      } else {
        fail(std::string("shared-object flag '") + flag + "'");
      }
      // close the nodes whose last dependency this was
      while (complete && !open.empty() && open.back().next == open.back().n_dep) {
        open.pop_back();
        ++n_shared_;
      }
      if (open.empty()) return;
      Open& o = open.back();
      fname = o.n_dep == 1 ? "UnarySX::dep" : o.next == 0 ? "UnarySX::dep0" : "UnarySX::dep1";
      ++o.next;
    }
  }
  void sx_matrix() {  // Matrix<SXElem>::serialize (casadi/core/matrix_impl.hpp:3049-3052)
    sparsity("Matrix::sparsity");
    const int64_t n = vector_length("Matrix::nonzeros");
    for (int64_t j = 0; j < n; ++j) sx_elem("");
  }
  void sx_vector(const char* name) {
    const int64_t n = vector_length(name);
    for (int64_t j = 0; j < n; ++j) sx_matrix();
  }
  int64_t sx_elem_vector(const char* name) {
    const int64_t n = vector_length(name);
    for (int64_t j = 0; j < n; ++j) sx_elem("");
    return n;
  }

  // GenericType (casadi/core/generic_type.cpp:773-818, TypeID numbering of
  // casadi/core/generic_type.hpp:43-63).
  void generic(const char* name) {
    field(name);
    shared('G', [&] {
      const int32_t type = integer("GenericType::type");
      switch (type) {
        case 1: boolean("GenericType::d"); break;                  // OT_BOOL
        case 2: casadi_int("GenericType::d"); break;               // OT_INT
        case 3: real("GenericType::d"); break;                     // OT_DOUBLE
        case 4: str("GenericType::d"); break;                      // OT_STRING
        case 5: int64_vector("GenericType::d"); break;             // OT_INTVECTOR
        case 6: {                                                   // OT_INTVECTORVECTOR
          const int64_t n = vector_length("GenericType::d");
          for (int64_t j = 0; j < n; ++j) int64_vector("");
          break;
        }
        case 7: bool_vector("GenericType::d"); break;              // OT_BOOLVECTOR
        case 8: double_vector("GenericType::d"); break;            // OT_DOUBLEVECTOR
        case 9: {                                                   // OT_DOUBLEVECTORVECTOR
          const int64_t n = vector_length("GenericType::d");
          for (int64_t j = 0; j < n; ++j) double_vector("");
          break;
        }
        case 10: string_vector("GenericType::d"); break;           // OT_STRINGVECTOR
        case 11: dict("GenericType::d"); break;                    // OT_DICT
        case 12: nested_function("GenericType::d"); break;         // OT_FUNCTION
        case 13: {                                                  // OT_FUNCTIONVECTOR
          const int64_t n = vector_length("GenericType::d");
          for (int64_t j = 0; j < n; ++j) nested_function("");
          break;
        }
        case 16: {                                                  // OT_STRINGVECTORVECTOR
          const int64_t n = vector_length("GenericType::d");
          for (int64_t j = 0; j < n; ++j) string_vector("");
          break;
        }
        case 17: {                                                  // OT_DICTVECTOR
          const int64_t n = vector_length("GenericType::d");
          for (int64_t j = 0; j < n; ++j) dict("");
          break;
        }
        case 18: {                                                  // OT_VECTORVECTOR
          const int64_t n = vector_length("GenericType::d");
          for (int64_t j = 0; j < n; ++j) {
            const int64_t m = vector_length("");
            for (int64_t i = 0; i < m; ++i) generic("");
          }
          break;
        }
        case 19: {                                                  // OT_VECTOR
          const int64_t n = vector_length("GenericType::d");
          for (int64_t j = 0; j < n; ++j) generic("");
          break;
        }
        default: fail("GenericType of type " + std::to_string(type));
      }
    });
  }
  void dict(const char* name) {  // std::map<std::string, GenericType>
    field(name);
    deco('D');
    const int64_t n = casadi_int("");
    if (n < 0 || n > (end_ - p_)) fail("bad dictionary size");
    for (int64_t j = 0; j < n; ++j) {
      str("");
      generic("");
    }
  }
  void nested_function(const char* name) {  // a Function member: a shared object
    field(name);
    if (++fn_depth_ > kMaxFunctionDepth) fail("Functions nested more than 64 deep");
    shared('F', [&] {
      SerialFunction f;
      function(&f);
    });
    --fn_depth_;
  }

  // The SXFunction body with its bases, in the reference's order.
  void sx_function(SerialFunction* F) {
    // ProtoFunction (casadi/core/function_internal.cpp:4165-4183)
    const int32_t vp = integer("ProtoFunction::serialization::version");
    if (vp < 1 || vp > 2) fail("ProtoFunction version " + std::to_string(vp));
    F->name = str("ProtoFunction::name");
    boolean("ProtoFunction::verbose");
    boolean("ProtoFunction::print_time");
    boolean("ProtoFunction::record_time");
    if (vp >= 2) {
      boolean("ProtoFunction::regularity_check");
      boolean("ProtoFunction::error_on_fail");
    }
    // FunctionInternal (casadi/core/function_internal.cpp:4189-4340)
    const int32_t v = integer("FunctionInternal::serialization::version");
    if (v < 1 || v > 8) fail("FunctionInternal version " + std::to_string(v));
    bool_vector("FunctionInternal::is_diff_in");
    bool_vector("FunctionInternal::is_diff_out");
    for (int io = 0; io < 2; ++io) {
      const int64_t n = vector_length(io == 0 ? "FunctionInternal::sp_in" : "FunctionInternal::sp_out");
      std::vector<int64_t>& nnz = io == 0 ? F->nnz_in : F->nnz_out;
      for (int64_t j = 0; j < n; ++j) nnz.push_back(sparsity(""));
    }
    string_vector("FunctionInternal::name_in");
    string_vector("FunctionInternal::name_out");
    boolean("FunctionInternal::jit");
    boolean("FunctionInternal::jit_cleanup");
    const std::string jit_serialize = v >= 2 ? str("FunctionInternal::jit_serialize") : "source";
    if (jit_serialize == "link" || jit_serialize == "embed")
      fail("a Function serialized with a jit library (jit_serialize = " + jit_serialize + ")");
    boolean("FunctionInternal::jit_temp_suffix");
    str("FunctionInternal::jit_base_name");
    dict("FunctionInternal::jit_options");
    str("FunctionInternal::compiler_plugin");
    boolean("FunctionInternal::has_refcount");
    if (v >= 6) dict("FunctionInternal::cache_init");
    nested_function("FunctionInternal::derivative_of");
    real("FunctionInternal::jac_penalty");
    for (const char* e : {"FunctionInternal::enable_forward", "FunctionInternal::enable_reverse",
                          "FunctionInternal::enable_jacobian", "FunctionInternal::enable_fd",
                          "FunctionInternal::enable_forward_op", "FunctionInternal::enable_reverse_op",
                          "FunctionInternal::enable_jacobian_op", "FunctionInternal::enable_fd_op"})
      boolean(e);
    real("FunctionInternal::ad_weight");
    real("FunctionInternal::ad_weight_sp");
    boolean("FunctionInternal::always_inline");
    boolean("FunctionInternal::never_inline");
    casadi_int("FunctionInternal::max_num_dir");
    if (v < 3) boolean("FunctionInternal::regularity_check");
    boolean("FunctionInternal::inputs_check");
    real("FunctionInternal::fd_step");
    str("FunctionInternal::fd_method");
    boolean("FunctionInternal::print_in");
    boolean("FunctionInternal::print_out");
    if (v >= 7) boolean("FunctionInternal::print_canonical");
    if (v >= 4) casadi_int("FunctionInternal::max_io");
    boolean("FunctionInternal::dump_in");
    boolean("FunctionInternal::dump_out");
    str("FunctionInternal::dump_dir");
    str("FunctionInternal::dump_format");
    dict("FunctionInternal::forward_options");
    dict("FunctionInternal::reverse_options");
    if (v >= 5) {
      dict("FunctionInternal::jacobian_options");
      dict("FunctionInternal::der_options");
    }
    nested_function("FunctionInternal::custom_jacobian");
    if (v >= 8) {
      const int64_t n = vector_length("FunctionInternal::registered_functions");
      for (int64_t j = 0; j < n; ++j) nested_function("");
    }
    for (const char* e : {"FunctionInternal::sz_arg_per", "FunctionInternal::sz_res_per", "FunctionInternal::sz_iw_per",
                          "FunctionInternal::sz_w_per", "FunctionInternal::sz_arg_tmp", "FunctionInternal::sz_res_tmp",
                          "FunctionInternal::sz_iw_tmp", "FunctionInternal::sz_w_tmp"})
      size(e);
    // XFunction (casadi/core/x_function.hpp:306-332): the inputs
    const int32_t vx = integer("XFunction::serialization::version");
    if (vx != 1) fail("XFunction version " + std::to_string(vx));
    sx_vector("XFunction::in");
    // SXFunction (casadi/core/sx_function.cpp:1926-2041)
    const int32_t vs = integer("SXFunction::serialization::version");
    if (vs < 1 || vs > 3) fail("SXFunction version " + std::to_string(vs));
    const uint64_t n_instr = size("SXFunction::n_instr");
    F->worksize = static_cast<int64_t>(size("SXFunction::worksize"));
    F->n_free = sx_elem_vector("SXFunction::free_vars");
    sx_elem_vector("SXFunction::operations");
    sx_elem_vector("SXFunction::constants");
    double_vector("SXFunction::default_in");
    if (vs >= 2) {
      for (const char* e : {"SXFunction::call_sz_arg", "SXFunction::call_sz_res", "SXFunction::call_sz_iw",
                            "SXFunction::call_sz_w", "SXFunction::call_sz_arg", "SXFunction::call_sz_res"})
        size(e);
      F->n_calls = static_cast<int64_t>(size("SXFunction::call_el_size"));
      for (int64_t k = 0; k < F->n_calls; ++k) {
        nested_function("SXFunction::call_el_f");
        int_vector("SXFunction::call_el_dep");
        int_vector("SXFunction::call_el_res");
        int_vector("SXFunction::call_el_copy_elision_arg");
        int_vector("SXFunction::call_el_copy_elision_offset");
      }
      bool_vector("SXFunction::copy_elision");
    }
    if (n_instr > static_cast<uint64_t>(end_ - p_)) fail("bad instruction count");
    F->alg.resize(n_instr);
    for (Instr& e : F->alg) {
      e.op = integer("SXFunction::ScalarAtomic::op");
      e.i0 = integer("SXFunction::ScalarAtomic::i0");
      e.i1 = integer("SXFunction::ScalarAtomic::i1");
      e.i2 = integer("SXFunction::ScalarAtomic::i2");
    }
    boolean("SXFunction::live_variables");
    if (vs >= 3) boolean("SXFunction::print_instructions");
    sx_vector("XFunction::out");  // delayed_deserialize_members
  }

 private:
  // nonzeros of each Sparsity definition, by shared index
  std::vector<std::pair<int64_t, int64_t>> sp_nnz_;
  // Function members inside Function members (derivative_of, options):
  // each level recurses, so the depth is bounded
  static constexpr int kMaxFunctionDepth = 64;
  int fn_depth_ = 0;
  static bool op_ndeps_binary(int64_t op) { return op >= 0 && op < 1000 && op_is_binary(static_cast<int>(op)); }
  static bool op_ndeps_unary(int64_t op) {
    return op >= 0 && op < 1000 && !op_is_binary(static_cast<int>(op)) && op_ndeps(static_cast<int>(op)) == 1 &&
           op != OP_OUTPUT;
  }

 public:
  bool at_end() const { return p_ == end_; }
};

}  // namespace
}  // namespace sxh

using namespace sxh;

extern "C" {

int cxh_tape_deserialize(const char* data, int64_t len, cxh_tape* out) {
  CXH_TRY
  if (!data || len < 0 || !out) return cxh_fail("cxh_tape_deserialize: null argument");
  Reader r(data, len);
  r.header();
  SerialFunction f;
  if (!r.function(&f)) return cxh_fail("cxh_tape_deserialize: the serialized Function is null");
  if (!r.at_end()) return cxh_fail("cxh_tape_deserialize: data after the Function");
  if (f.n_free > 0) return cxh_fail("cxh_tape_deserialize: the Function has free variables");
  if (f.n_calls > 0) return cxh_fail("cxh_tape_deserialize: the Function calls other Functions (OP_CALL)");
  return cxh_tape_create(f.alg.data(), static_cast<int64_t>(f.alg.size()), f.worksize,
                         static_cast<int32_t>(f.nnz_in.size()), f.nnz_in.data(),
                         static_cast<int32_t>(f.nnz_out.size()), f.nnz_out.data(), out);
  CXH_CATCH
}

}  // extern "C"
