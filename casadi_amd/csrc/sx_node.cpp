// Scalar symbolic nodes, constant folding and on-the-fly simplification.
// Reference: casadi/core/sx_elem.cpp:71-263, casadi/core/constant_sx.hpp,
// casadi/core/binary_sx.hpp:55-67, casadi/core/unary_sx.hpp:55-66,
// casadi/core/calculus.hpp:1812-2139.
#include "sx_node.hpp"

#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <stdexcept>
#include <unordered_map>

namespace sxh {

int eq_depth = 1;
bool simplification_on_the_fly = true;

namespace {

std::deque<Node>& arena() {
  static std::deque<Node> a;
  return a;
}
std::vector<std::string>& sym_names() {
  static std::vector<std::string> v;
  return v;
}

Node* new_node(int op, double value, Node* d0, Node* d1) {
  arena().push_back(Node{op, value, {d0, d1}, 0, -1});
  return &arena().back();
}

// Constants are unique per value, as CasADi's cached constant nodes
// (RealtypeSX / IntegerSX caches plus the 0, 1, 2, -1, nan, inf singletons).
Node* cached_const(double v) {
  static std::unordered_map<uint64_t, Node*> cache;
  static Node* nan_node = nullptr;
  if (std::isnan(v)) {
    if (!nan_node) nan_node = new_node(OP_CONST, std::numeric_limits<double>::quiet_NaN(),
                                       nullptr, nullptr);
    return nan_node;
  }
  if (v == 0) v = 0.0;  // -0.0 maps onto the zero singleton
  uint64_t bits;
  std::memcpy(&bits, &v, sizeof bits);
  auto it = cache.find(bits);
  if (it != cache.end()) return it->second;
  Node* n = new_node(OP_CONST, v, nullptr, nullptr);
  cache.emplace(bits, n);
  return n;
}

inline double cval(SXE x) { return x->value; }
inline bool is_const(SXE x) { return x->op == OP_CONST; }
inline SXE dep(SXE x, int i = 0) { return x->dep[i]; }

// Operator shorthands going through the simplifying constructors, so that the
// rewrites below compose exactly like the C++ operator overloads on SXElem.
inline SXE C(double v) { return sx_const(v); }
inline SXE add(SXE a, SXE b) { return sx_binary(OP_ADD, a, b); }
inline SXE sub(SXE a, SXE b) { return sx_binary(OP_SUB, a, b); }
inline SXE mul(SXE a, SXE b) { return sx_binary(OP_MUL, a, b); }
inline SXE dvd(SXE a, SXE b) { return sx_binary(OP_DIV, a, b); }
inline SXE neg(SXE a) { return sx_neg(a); }
inline SXE un(int op, SXE a) { return sx_unary(op, a); }
inline SXE inv(SXE a) { return sx_unary(OP_INV, a); }

// BinarySX::create / UnarySX::create: fold constants, else allocate.
SXE create_binary(int op, SXE x, SXE y) {
  if (is_const(x) && is_const(y)) return C(op_fun(op, cval(x), cval(y)));
  return new_node(op, 0, x, y);
}
SXE create_unary(int op, SXE x) {
  if (is_const(x)) return C(op_fun(op, cval(x), cval(x)));
  return new_node(op, 0, x, x);
}

// common_simp_unary (casadi/core/calculus.hpp:1812-1864)
SXE simp_unary(int op, SXE x, bool& hit) {
  hit = true;
  switch (op) {
    case OP_TWICE:
      if (sx_is_op(x, OP_MUL) && is_const(dep(x, 0)) && sx_is_half(dep(x, 0))) return dep(x, 1);
      if (sx_is_op(x, OP_MUL) && is_const(dep(x, 1)) && sx_is_half(dep(x, 1))) return dep(x, 0);
      break;
    case OP_SQ:
      if (sx_is_op(x, OP_SQRT)) return dep(x);
      if (sx_is_op(x, OP_NEG)) return un(OP_SQ, dep(x));
      if (sx_is_op(x, OP_FABS)) return un(OP_SQ, dep(x));
      break;
    case OP_FABS:
      if (sx_is_nonnegative(x)) return x;
      if (sx_is_op(x, OP_NEG)) return un(OP_FABS, dep(x));
      break;
    case OP_LOG:
      if (sx_is_op(x, OP_EXP)) return dep(x);
      break;
    case OP_INV:
      if (sx_is_op(x, OP_INV)) return dep(x);
      break;
    case OP_SQRT:
      if (sx_is_op(x, OP_SQ)) return un(OP_FABS, dep(x));
      break;
    case OP_COS:
      if (sx_is_op(x, OP_NEG)) return un(OP_COS, dep(x));
      if (sx_is_op(x, OP_FABS)) return un(OP_COS, dep(x));
      break;
    case OP_NEG:
      if (sx_is_op(x, OP_NEG)) return dep(x);
      break;
  }
  hit = false;
  return nullptr;
}

// common_simp_binary (casadi/core/calculus.hpp:1866-2139), unique_x/unique_y
// false as in every SXElem operator call site.
SXE simp_binary(int op, SXE x, SXE y, bool& hit) {
  const int depth = eq_depth;
  hit = true;
  switch (op) {
    case OP_ADD:
      if (sx_is_zero(x)) return y;
      if (sx_is_zero(y)) return x;
      if (sx_is_op(y, OP_NEG)) return sub(x, neg(y));
      if (sx_is_op(x, OP_NEG)) return sub(y, dep(x));
      if (sx_is_equal(x, y, depth)) return mul(C(2), x);
      if (sx_is_op(x, OP_MUL) && sx_is_op(y, OP_MUL) && is_const(dep(x, 0)) &&
          cval(dep(x, 0)) == 0.5 && is_const(dep(y, 0)) && cval(dep(y, 0)) == 0.5 &&
          sx_is_equal(dep(y, 1), dep(x, 1), depth))
        return dep(x, 1);
      if (sx_is_op(x, OP_DIV) && sx_is_op(y, OP_DIV) && is_const(dep(x, 1)) &&
          cval(dep(x, 1)) == 2 && is_const(dep(y, 1)) && cval(dep(y, 1)) == 2 &&
          sx_is_equal(dep(y, 0), dep(x, 0), depth))
        return dep(x, 0);
      if (sx_is_op(x, OP_MUL) && sx_is_op(y, OP_MUL) && is_const(dep(x, 0)) &&
          is_const(dep(y, 0)) && cval(dep(x, 0)) + cval(dep(y, 0)) == 1 &&
          sx_is_equal(dep(y, 1), dep(x, 1), depth))
        return dep(x, 1);
      if (sx_is_op(x, OP_SUB) && sx_is_equal(dep(x, 1), y, depth)) return dep(x, 0);
      if (sx_is_op(y, OP_SUB) && sx_is_equal(x, dep(y, 1), depth)) return dep(y, 0);
      if (sx_is_op(x, OP_SQ) && sx_is_op(y, OP_SQ) &&
          ((sx_is_op(dep(x), OP_SIN) && sx_is_op(dep(y), OP_COS)) ||
           (sx_is_op(dep(x), OP_COS) && sx_is_op(dep(y), OP_SIN))) &&
          sx_is_equal(dep(dep(x)), dep(dep(y)), depth))
        return C(1);
      if (sx_is_doubled(x) && sx_is_equal(dep(x, 0), y, depth)) return mul(C(3), dep(x, 0));
      if (sx_is_doubled(y) && sx_is_equal(dep(y, 0), x, depth)) return mul(C(3), dep(y, 0));
      if (sx_is_op(x, OP_MUL) && sx_is_op(y, OP_MUL) && sx_is_equal(dep(x, 0), dep(y, 0), depth) &&
          sx_is_one(add(dep(x, 1), dep(y, 1))))
        return dep(x, 0);
      break;
    case OP_SUB:
      if (sx_is_zero(y)) return x;
      if (sx_is_zero(x)) return neg(y);
      if (sx_is_equal(x, y, depth)) return C(0);
      if (sx_is_op(y, OP_NEG)) return add(x, dep(y));
      if (sx_is_op(x, OP_ADD) && sx_is_equal(dep(x, 1), y, depth)) return dep(x, 0);
      if (sx_is_op(x, OP_ADD) && sx_is_equal(dep(x, 0), y, depth)) return dep(x, 1);
      if (sx_is_op(y, OP_ADD) && sx_is_equal(x, dep(y, 1), depth)) return neg(dep(y, 0));
      if (sx_is_op(y, OP_ADD) && sx_is_equal(x, dep(y, 0), depth)) return neg(dep(y, 1));
      if (sx_is_op(x, OP_NEG)) return neg(add(dep(x), y));
      if (sx_is_op(x, OP_MUL) && sx_is_op(y, OP_MUL) && is_const(dep(x, 0)) &&
          is_const(dep(y, 0)) && cval(dep(x, 0)) - cval(dep(y, 0)) == 1 &&
          sx_is_equal(dep(y, 1), dep(x, 1), depth))
        return dep(x, 1);
      if (sx_is_doubled(x) && sx_is_equal(dep(x, 0), y, depth)) return y;
      if (sx_is_doubled(y) && sx_is_equal(dep(y, 0), x, depth)) return neg(x);
      break;
    case OP_MUL:
      if (sx_is_equal(y, x, depth)) return un(OP_SQ, x);
      if (!is_const(x) && is_const(y)) return mul(y, x);
      if (sx_is_zero(x) || sx_is_zero(y)) return C(0);
      if (sx_is_one(x)) return y;
      if (sx_is_one(y)) return x;
      if (sx_is_minus_one(y)) return neg(x);
      if (sx_is_minus_one(x)) return neg(y);
      if (sx_is_op(y, OP_INV)) return dvd(x, inv(y));
      if (sx_is_op(x, OP_INV)) return dvd(y, inv(x));
      if (is_const(x) && sx_is_op(y, OP_MUL) && is_const(dep(y, 0)) &&
          cval(x) * cval(dep(y, 0)) == 1)
        return dep(y, 1);
      if (is_const(x) && sx_is_doubled(y) && cval(x) * 2 == 1) return dep(y, 0);
      if (is_const(x) && sx_is_op(y, OP_DIV) && is_const(dep(y, 1)) &&
          cval(x) == cval(dep(y, 1)))
        return dep(y, 0);
      if (is_const(x) && cval(x) == 2) return sx_unary(OP_TWICE, y);
      if (is_const(y) && cval(y) == 2) return sx_unary(OP_TWICE, x);
      if (sx_is_op(x, OP_DIV) && sx_is_equal(dep(x, 1), y, depth)) return dep(x, 0);
      if (sx_is_op(y, OP_DIV) && sx_is_equal(dep(y, 1), x, depth)) return dep(y, 0);
      if (sx_is_op(x, OP_NEG)) return neg(mul(dep(x), y));
      if (sx_is_op(y, OP_NEG)) return neg(mul(x, dep(y)));
      // (the four unique_x / unique_y rewrites are disabled: unique flags are false)
      if (is_const(x) && sx_is_op(y, OP_MUL) && is_const(dep(y, 0)))
        return mul(mul(x, dep(y, 0)), dep(y, 1));
      if (is_const(y) && sx_is_op(x, OP_MUL) && is_const(dep(x, 0)))
        return mul(mul(y, dep(x, 0)), dep(x, 1));
      break;
    case OP_DIV:
      if (sx_is_zero(y)) return C(std::numeric_limits<double>::infinity());
      if (sx_is_zero(x)) return C(0);
      if (sx_is_one(y)) return x;
      if (sx_is_minus_one(y)) return neg(x);
      if (sx_is_half(y)) return mul(C(2), x);
      if (sx_is_equal(x, y, depth)) return C(1);
      if (sx_is_doubled(x) && is_const(y) && cval(y) == 2) return dep(x, 0);
      if (sx_is_doubled(x) && sx_is_equal(y, dep(x, 0), depth)) return C(2);
      if (sx_is_op(x, OP_MUL) && sx_is_equal(y, dep(x, 0), depth)) return dep(x, 1);
      if (sx_is_op(x, OP_MUL) && sx_is_equal(y, dep(x, 1), depth)) return dep(x, 0);
      if (sx_is_one(x)) return inv(y);
      if (sx_is_op(y, OP_INV)) return mul(x, inv(y));
      if (sx_is_doubled(x) && sx_is_doubled(y)) return dvd(dep(x, 0), dep(y, 0));
      if (is_const(y) && sx_is_op(x, OP_DIV) && is_const(dep(x, 1)) &&
          cval(y) * cval(dep(x, 1)) == 1)
        return dep(x, 0);
      if (sx_is_op(y, OP_MUL) && sx_is_equal(dep(y, 1), x, depth)) return dvd(C(1), dep(y, 0));
      if (sx_is_doubled(y) && sx_is_equal(dep(y, 0), x, depth)) return C(0.5);
      if (sx_is_op(x, OP_NEG) && sx_is_equal(dep(x, 0), y, depth)) return C(-1);
      if (sx_is_op(y, OP_NEG) && sx_is_equal(dep(y, 0), x, depth)) return C(-1);
      if (sx_is_op(y, OP_NEG) && sx_is_op(x, OP_NEG) && sx_is_equal(dep(x, 0), dep(y, 0), depth))
        return C(1);
      if (sx_is_op(x, OP_DIV) && sx_is_equal(y, dep(x, 0), depth)) return inv(dep(x, 1));
      if (sx_is_op(x, OP_NEG)) return neg(dvd(dep(x), y));
      if (sx_is_op(y, OP_NEG)) return neg(dvd(x, dep(y)));
      break;
    case OP_POW:
      if (is_const(y)) {
        if (sx_is_integer(y)) {
          int64_t nn = static_cast<int64_t>(cval(y));
          if (nn == 0) return C(1);
          if (nn > 100 || nn < -100) return sx_binary(OP_CONSTPOW, x, C(double(nn)));
          if (nn < 0) return dvd(C(1), sx_binary(OP_POW, x, C(double(-nn))));
          if (nn % 2 == 1) return mul(x, sx_binary(OP_POW, x, C(double(nn - 1))));
          SXE rt = sx_binary(OP_POW, x, C(double(nn / 2)));
          return mul(rt, rt);
        } else if (cval(y) == 0.5) {
          return sx_unary(OP_SQRT, x);
        } else {
          return sx_binary(OP_CONSTPOW, x, y);
        }
      }
      break;
    case OP_LE:
      if (sx_is_nonnegative(sub(y, x))) return C(1);
      break;
    case OP_FMIN:
      if (sx_is_inf(x)) return y;
      if (sx_is_inf(y)) return x;
      if (sx_is_minus_inf(x) || sx_is_minus_inf(y))
        return C(-std::numeric_limits<double>::infinity());
      if (sx_is_equal(x, y, depth)) return x;
      break;
    case OP_FMAX:
      if (sx_is_minus_inf(x)) return y;
      if (sx_is_minus_inf(y)) return x;
      if (sx_is_inf(x) || sx_is_inf(y)) return C(std::numeric_limits<double>::infinity());
      if (sx_is_equal(x, y, depth)) return x;
      break;
    case OP_LT:
      if (sx_is_nonnegative(sub(x, y))) return C(0);
      break;
    case OP_EQ:
      if (sx_is_equal(x, y, depth)) return C(1);
      break;
    case OP_NE:
      if (sx_is_equal(x, y, depth)) return C(0);
      break;
    case OP_IF_ELSE_ZERO:
      if (sx_is_zero(y)) return y;
      if (is_const(x)) return cval(x) != 0 ? y : C(0);
      break;
  }
  hit = false;
  return nullptr;
}

}  // namespace

int64_t sx_node_count() { return static_cast<int64_t>(arena().size()); }

// SXElem(double) (casadi/core/sx_elem.cpp:71-93)
SXE sx_const(double v) { return cached_const(v); }

SXE sx_sym(const std::string& name) {
  Node* n = new_node(OP_PARAMETER, 0, nullptr, nullptr);
  n->sym_id = static_cast<int64_t>(sym_names().size());
  sym_names().push_back(name);
  return n;
}

const std::string& sx_name(SXE x) {
  if (!x->is_symbolic()) throw std::runtime_error("sx_name: not a symbolic primitive");
  return sym_names().at(static_cast<size_t>(x->sym_id));
}

bool sx_is_zero(SXE x) { return is_const(x) && cval(x) == 0; }
bool sx_is_one(SXE x) { return is_const(x) && cval(x) == 1; }
bool sx_is_minus_one(SXE x) { return is_const(x) && cval(x) == -1; }
bool sx_is_half(SXE x) { return is_const(x) && cval(x) == 0.5; }
bool sx_is_inf(SXE x) { return is_const(x) && cval(x) == std::numeric_limits<double>::infinity(); }
bool sx_is_minus_inf(SXE x) {
  return is_const(x) && cval(x) == -std::numeric_limits<double>::infinity();
}
bool sx_is_integer(SXE x) {
  if (!is_const(x)) return false;
  double v = cval(x);
  if (!std::isfinite(v)) return false;
  return v == static_cast<double>(static_cast<int64_t>(v));
}
bool sx_is_op(SXE x, int op) { return !is_const(x) && !x->is_symbolic() && x->op == op; }

// SXElem::is_equal + the per-node is_equal overrides
// (casadi/core/sx_elem.cpp:355-364, binary_sx.hpp:90-100, unary_sx.hpp:85-88,
// constant_sx.hpp:84-87).
bool sx_is_equal(SXE x, SXE y, int depth) {
  if (x == y) return true;
  if (depth <= 0) return false;
  if (is_const(x) || is_const(y)) {
    return is_const(x) && is_const(y) && cval(x) == cval(y);
  }
  if (x->is_symbolic() || y->is_symbolic()) return false;
  if (x->op != y->op) return false;
  if (op_is_binary(x->op)) {
    if (sx_is_equal(x->dep[0], y->dep[0], depth - 1) && sx_is_equal(x->dep[1], y->dep[1], depth - 1))
      return true;
    bool comm = x->op == OP_ADD || x->op == OP_MUL || x->op == OP_EQ || x->op == OP_NE ||
                x->op == OP_AND || x->op == OP_OR || x->op == OP_HYPOT;
    return comm && sx_is_equal(x->dep[1], y->dep[0], depth - 1) &&
           sx_is_equal(x->dep[0], y->dep[1], depth - 1);
  }
  return sx_is_equal(x->dep[0], y->dep[0], depth - 1);
}

// SXElem::is_nonnegative (casadi/core/sx_elem.cpp:366-374) with
// NonnegativeChecker (casadi/core/calculus.hpp)
bool sx_is_nonnegative(SXE x) {
  if (is_const(x)) return cval(x) >= 0;
  if (x->is_symbolic()) return false;
  if (op_is_unary(x->op)) {
    switch (x->op) {
      case OP_SQRT: case OP_SQ: case OP_EXP: case OP_LT: case OP_LE: case OP_EQ:
      case OP_NE: case OP_NOT: case OP_AND: case OP_OR: case OP_HYPOT: case OP_FABS:
        return true;
      default:
        return false;
    }
  }
  return false;
}

// SXElem::is_doubled (casadi/core/sx_elem.cpp:168-171)
bool sx_is_doubled(SXE x) {
  return (sx_is_op(x, OP_ADD) && sx_is_equal(x->dep[0], x->dep[1], eq_depth)) ||
         sx_is_op(x, OP_TWICE);
}

// SXElem::operator-() (casadi/core/sx_elem.cpp:150-161)
SXE sx_neg(SXE x) {
  if (sx_is_op(x, OP_NEG)) return x->dep[0];
  if (sx_is_zero(x)) return C(0);
  if (sx_is_minus_one(x)) return C(1);
  if (sx_is_one(x)) return C(-1);
  return sx_unary(OP_NEG, x);
}

// SXElem::binary (casadi/core/sx_elem.cpp:189-230)
SXE sx_binary(int op, SXE x, SXE y) {
  if (sx_is_op(y, OP_IF_ELSE_ZERO)) {
    if (op == OP_MUL) {
      return sx_binary(OP_IF_ELSE_ZERO, y->dep[0], mul(x, y->dep[1]));
    } else if (op == OP_ADD && sx_is_op(x, OP_IF_ELSE_ZERO) && sx_is_equal(x->dep[0], y->dep[0])) {
      return sx_binary(OP_IF_ELSE_ZERO, y->dep[0], add(x->dep[1], y->dep[1]));
    }
  } else if (sx_is_op(x, OP_IF_ELSE_ZERO) && op == OP_MUL) {
    return sx_binary(OP_IF_ELSE_ZERO, x->dep[0], mul(x->dep[1], y));
  }
  if (simplification_on_the_fly) {
    bool hit;
    SXE ret = simp_binary(op, x, y, hit);
    if (hit) return ret;
    if (op == OP_MUL) {
      if (is_const(x) && sx_is_op(y, OP_TWICE)) return mul(mul(x, C(2)), y->dep[0]);
      if (is_const(y) && sx_is_op(x, OP_TWICE)) return mul(mul(C(2), y), x->dep[0]);
    }
  }
  return create_binary(op, x, y);
}

// SXElem::unary (casadi/core/sx_elem.cpp:237-263)
SXE sx_unary(int op, SXE x) {
  if (simplification_on_the_fly) {
    switch (op) {
      case OP_NOT:
        if (sx_is_op(x, OP_NOT)) return x->dep[0];
        break;
      case OP_TWICE:
        if (sx_is_op(x, OP_MUL) && is_const(x->dep[0])) return mul(mul(C(2), x->dep[0]), x->dep[1]);
        if (sx_is_op(x, OP_MUL) && is_const(x->dep[1])) return mul(mul(C(2), x->dep[1]), x->dep[0]);
        if (sx_is_op(x, OP_TWICE)) return mul(C(4), x->dep[0]);
        break;
    }
    bool hit;
    SXE ret = simp_unary(op, x, hit);
    if (hit) return ret;
  }
  return create_unary(op, x);
}

// if_else(x, y, z) = if_else_zero(x, y) + if_else_zero(!x, z)
// (casadi/core/sx_elem.hpp:288-290)
SXE sx_if_else(SXE c, SXE a, SXE b) {
  return add(sx_binary(OP_IF_ELSE_ZERO, c, a),
             sx_binary(OP_IF_ELSE_ZERO, sx_unary(OP_NOT, c), b));
}

SXE sx_fun(int op, SXE x, SXE y) {
  switch (op) {
    case OP_ASSIGN: return x;
    case OP_NEG: return neg(x);
    case OP_TWICE: return mul(C(2), x);
    case OP_INV: return dvd(C(1), x);
    default:
      if (op_is_binary(op)) return sx_binary(op, x, y);
      return sx_unary(op, x);
  }
}

void sx_der(int op, SXE x, SXE y, SXE f, SXE* d) {
  const double pi = 3.14159265358979323846;
  d[1] = C(0);
  switch (op) {
    case OP_ASSIGN: d[0] = C(1); break;
    case OP_ADD: d[0] = C(1); d[1] = C(1); break;
    case OP_SUB: d[0] = C(1); d[1] = C(-1); break;
    case OP_MUL: d[0] = y; d[1] = x; break;
    case OP_DIV: d[0] = dvd(C(1), y); d[1] = dvd(neg(f), y); break;
    case OP_NEG: d[0] = C(-1); break;
    case OP_EXP: d[0] = f; break;
    case OP_LOG: d[0] = dvd(C(1), x); break;
    case OP_POW:
      d[0] = mul(y, sx_binary(OP_POW, x, sub(y, C(1))));
      d[1] = mul(un(OP_LOG, x), f);
      break;
    case OP_CONSTPOW: d[0] = mul(y, sx_binary(OP_POW, x, sub(y, C(1)))); break;
    case OP_SQRT: d[0] = dvd(C(1), add(f, f)); break;
    case OP_SQ: d[0] = add(x, x); break;
    case OP_TWICE: d[0] = C(2); break;
    case OP_SIN: d[0] = un(OP_COS, x); break;
    case OP_COS: d[0] = neg(un(OP_SIN, x)); break;
    case OP_TAN: d[0] = dvd(C(1), un(OP_SQ, un(OP_COS, x))); break;
    case OP_ASIN: d[0] = dvd(C(1), un(OP_SQRT, sub(C(1), mul(x, x)))); break;
    case OP_ACOS: d[0] = dvd(C(-1), un(OP_SQRT, sub(C(1), mul(x, x)))); break;
    case OP_ATAN: d[0] = dvd(C(1), add(C(1), mul(x, x))); break;
    case OP_LT: case OP_LE: case OP_EQ: case OP_NE: case OP_AND: case OP_OR:
      d[0] = C(0); d[1] = C(0); break;
    case OP_NOT: case OP_FLOOR: case OP_CEIL: case OP_SIGN: d[0] = C(0); break;
    case OP_FMOD: case OP_REMAINDER: d[0] = C(1); d[1] = dvd(sub(f, x), y); break;
    case OP_FABS: d[0] = un(OP_SIGN, x); break;
    case OP_COPYSIGN: d[0] = sx_binary(OP_COPYSIGN, C(1), y); break;
    case OP_IF_ELSE_ZERO: d[0] = C(0); d[1] = x; break;
    case OP_ERF: d[0] = mul(C(2 / std::sqrt(pi)), un(OP_EXP, mul(neg(x), x))); break;
    case OP_FMIN: {
      SXE a = sx_binary(OP_LE, x, y), b = sx_binary(OP_LE, y, x), c = add(a, b);
      d[0] = dvd(a, c); d[1] = dvd(b, c);
    } break;
    case OP_FMAX: {
      SXE a = sx_binary(OP_LE, y, x), b = sx_binary(OP_LE, x, y), c = add(a, b);
      d[0] = dvd(a, c); d[1] = dvd(b, c);
    } break;
    case OP_INV: d[0] = mul(neg(f), f); break;
    case OP_SINH: d[0] = un(OP_COSH, x); break;
    case OP_COSH: d[0] = un(OP_SINH, x); break;
    case OP_TANH: d[0] = sub(C(1), mul(f, f)); break;
    case OP_ASINH: d[0] = dvd(C(1), un(OP_SQRT, add(C(1), mul(x, x)))); break;
    case OP_ACOSH:
      d[0] = dvd(dvd(C(1), un(OP_SQRT, sub(x, C(1)))), un(OP_SQRT, add(x, C(1))));
      break;
    case OP_ATANH: d[0] = dvd(C(1), sub(C(1), mul(x, x))); break;
    case OP_ATAN2: {
      SXE t = add(mul(x, x), mul(y, y));
      d[0] = dvd(y, t); d[1] = dvd(neg(x), t);
    } break;
    case OP_ERFINV: d[0] = mul(C(std::sqrt(pi) / 2), un(OP_EXP, mul(f, f))); break;
    case OP_PRINTME: case OP_LIFT: d[0] = C(1); break;
    case OP_LOG1P: d[0] = dvd(C(1), add(C(1), x)); break;
    case OP_EXPM1: d[0] = un(OP_EXP, x); break;
    case OP_HYPOT: d[0] = dvd(x, f); d[1] = dvd(y, f); break;
    default: throw std::runtime_error("sx_der: unsupported operation " + std::to_string(op));
  }
}

}  // namespace sxh
