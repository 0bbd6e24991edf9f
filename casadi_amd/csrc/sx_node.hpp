// Scalar symbolic expression graph (host side).
//
// Restates CasADi's SXElem / SXNode layer (casadi/core/sx_elem.cpp,
// casadi/core/{unary,binary,constant,symbolic}_sx.hpp): hash-consed constants,
// symbols, unary and binary operation nodes, constant folding at creation and
// the on-the-fly simplification rules that decide what ends up on a tape
// (casadi/core/sx_elem.cpp:189-263, casadi/core/calculus.hpp:1812-2139).
//
// Nodes live in a process-wide arena and are never freed: a handle is a plain
// pointer, which keeps the C ABI trivial.  Like CasADi built without
// WITH_THREADSAFE_SYMBOLICS, graph construction is not thread-safe.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "sx_ops.h"

namespace sxh {

struct Node {
  int op;          // OP_CONST, OP_PARAMETER (symbol) or an arithmetic op
  double value;    // constant value (OP_CONST)
  Node* dep[2];    // operands (unary: dep[1] == dep[0])
  int temp;        // scratch used by graph algorithms; 0 when idle
  int64_t sym_id;  // symbol index (OP_PARAMETER), -1 otherwise
  int n_dep() const {
    if (op == OP_CONST || op == OP_PARAMETER) return 0;
    return op_is_binary(op) ? 2 : 1;
  }
  bool is_constant() const { return op == OP_CONST; }
  bool is_symbolic() const { return op == OP_PARAMETER; }
};

using SXE = Node*;

// --- creation --------------------------------------------------------------
SXE sx_const(double v);
SXE sx_sym(const std::string& name);
const std::string& sx_name(SXE x);
SXE sx_unary(int op, SXE x);
SXE sx_binary(int op, SXE x, SXE y);
SXE sx_neg(SXE x);  // SXElem::operator-() (casadi/core/sx_elem.cpp:150-161)
SXE sx_if_else(SXE c, SXE a, SXE b);

// --- predicates (SXElem helpers) ------------------------------------------
bool sx_is_zero(SXE x);
bool sx_is_one(SXE x);
bool sx_is_minus_one(SXE x);
bool sx_is_half(SXE x);
bool sx_is_inf(SXE x);
bool sx_is_minus_inf(SXE x);
bool sx_is_integer(SXE x);
bool sx_is_op(SXE x, int op);
bool sx_is_equal(SXE x, SXE y, int depth = 0);
bool sx_is_nonnegative(SXE x);
bool sx_is_doubled(SXE x);

// Default depth of the structural equality test used by the simplifier
// (SXNode::eq_depth_, casadi/core/sx_node.cpp:232).
extern int eq_depth;
// GlobalOptions::simplification_on_the_fly (casadi/core/global_options.cpp)
extern bool simplification_on_the_fly;

// Number of nodes allocated so far (diagnostics).
int64_t sx_node_count();

// Symbolic partial derivatives d[0], d[1] of f = op(x, y), built with the same
// formulas as casadi_math<SXElem>::der (casadi/core/calculus.hpp:598-1018).
void sx_der(int op, SXE x, SXE y, SXE f, SXE* d);
// f = op(x, y) evaluated on symbolic operands the way casadi_math<SXElem>::fun
// does it (e.g. TWICE goes through 2*x, NEG through operator-).
SXE sx_fun(int op, SXE x, SXE y);

}  // namespace sxh
