// SX Function: tape compiler, VM, symbolic evaluation and derivatives.
#include "sx_function.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <numeric>
#include <stack>
#include <stdexcept>

namespace sxh {

bool SXMat::is_zero() const {
  for (SXE e : nz)
    if (!sx_is_zero(e)) return false;
  return true;
}

namespace {

// XFunction::sort_depth_first (casadi/core/x_function.hpp:412-438)
void sort_depth_first(std::stack<Node*>& s, std::vector<Node*>& nodes) {
  while (!s.empty()) {
    Node* t = s.top();
    if (t && t->temp >= 0) {
      int next_dep = t->temp++;
      if (next_dep < t->n_dep()) {
        s.push(t->dep[next_dep]);
      } else {
        nodes.push_back(t);
        t->temp = -1;
        s.pop();
      }
    } else {
      s.pop();
    }
  }
}

// linear indices of the nonzeros (Sparsity::find)
ivec find_nz(const Sparsity& sp) {
  ivec f(sp.nnz());
  for (int64_t c = 0; c < sp.ncol; ++c)
    for (int64_t k = sp.colind[c]; k < sp.colind[c + 1]; ++k) f[k] = sp.row[k] + c * sp.nrow;
  return f;
}

// veccat: vertical concatenation of vec(x_i), keeping structural zeros
SXMat veccat(const std::vector<SXMat>& v) {
  int64_t n = 0;
  ivec rows;
  std::vector<SXE> nz;
  for (const auto& m : v) {
    ivec f = find_nz(m.sp);
    for (size_t k = 0; k < f.size(); ++k) {
      rows.push_back(n + f[k]);
      nz.push_back(m.nz[k]);
    }
    n += m.sp.numel();
  }
  Sparsity sp(n, 1);
  sp.row = rows;
  sp.colind[1] = static_cast<int64_t>(rows.size());
  return SXMat(sp, nz);
}

SXMat horzcat(const std::vector<SXMat>& v) {
  if (v.size() == 1) return v[0];
  int64_t nr = v.empty() ? 0 : v[0].sp.nrow;
  Sparsity sp(nr, 0);
  std::vector<SXE> nz;
  for (const auto& m : v) {
    if (m.sp.nrow != nr) throw std::runtime_error("horzcat: row mismatch");
    for (int64_t c = 0; c < m.sp.ncol; ++c) {
      for (int64_t k = m.sp.colind[c]; k < m.sp.colind[c + 1]; ++k) sp.row.push_back(m.sp.row[k]);
      sp.colind.push_back(static_cast<int64_t>(sp.row.size()));
    }
    sp.ncol += m.sp.ncol;
    nz.insert(nz.end(), m.nz.begin(), m.nz.end());
  }
  return SXMat(sp, nz);
}

SXMat sym_mat(const std::string& name, const Sparsity& sp) {
  std::vector<SXE> nz(sp.nnz());
  for (int64_t k = 0; k < sp.nnz(); ++k)
    nz[k] = sx_sym(sp.nnz() == 1 ? name : name + "_" + std::to_string(k));
  return SXMat(sp, nz);
}

}  // namespace

SXFunction::SXFunction(std::string name, std::vector<SXMat> in, std::vector<SXMat> out,
                       std::vector<std::string> name_in, std::vector<std::string> name_out,
                       FunctionOptions opts)
    : name_(std::move(name)), in_(std::move(in)), out_(std::move(out)),
      name_in_(std::move(name_in)), name_out_(std::move(name_out)), opts_(std::move(opts)) {
  if (name_in_.empty())
    for (size_t i = 0; i < in_.size(); ++i) name_in_.push_back("i" + std::to_string(i));
  if (name_out_.empty())
    for (size_t i = 0; i < out_.size(); ++i) name_out_.push_back("o" + std::to_string(i));
  if (name_in_.size() != in_.size() || name_out_.size() != out_.size())
    throw std::runtime_error(name_ + ": mismatching number of input/output names");
  if (opts_.is_diff_in.empty()) opts_.is_diff_in.assign(in_.size(), true);
  if (opts_.is_diff_out.empty()) opts_.is_diff_out.assign(out_.size(), true);
  for (const auto& m : in_)
    if (static_cast<int64_t>(m.nz.size()) != m.sp.nnz())
      throw std::runtime_error(name_ + ": input nonzero count mismatch");
  for (const auto& m : out_)
    if (static_cast<int64_t>(m.nz.size()) != m.sp.nnz())
      throw std::runtime_error(name_ + ": output nonzero count mismatch");
  // XFunction::init: inputs must be purely symbolic and unique
  // (casadi/core/x_function.hpp:336-409)
  std::vector<Node*> seen;
  for (const auto& m : in_) {
    for (SXE e : m.nz) {
      if (!e->is_symbolic())
        throw std::runtime_error(name_ + ": Xfunction input arguments must be purely symbolic.");
      seen.push_back(e);
    }
  }
  std::sort(seen.begin(), seen.end());
  if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
    throw std::runtime_error(name_ + ": The input expressions are not independent.");
  init();
}

int64_t SXFunction::nnz_in() const {
  int64_t n = 0;
  for (const auto& m : in_) n += m.nnz();
  return n;
}
int64_t SXFunction::nnz_out() const {
  int64_t n = 0;
  for (const auto& m : out_) n += m.nnz();
  return n;
}

// SXFunction::init (casadi/core/sx_function.cpp:476-841)
void SXFunction::init() {
  std::stack<Node*> s;
  std::vector<Node*> nodes;
  for (const auto& o : out_) {
    for (SXE e : o.nz) {
      s.push(e);
      sort_depth_first(s, nodes);
      nodes.push_back(nullptr);  // output marker
    }
  }
  for (size_t i = 0; i < nodes.size(); ++i)
    if (nodes[i]) nodes[i]->temp = static_cast<int>(i);

  operations_.clear();
  for (Node* t : nodes)
    if (t && !t->is_constant() && !t->is_symbolic()) operations_.push_back(t);

  std::vector<std::pair<int, Node*>> symb_loc;
  int curr_oind = 0, curr_nz = 0;
  for (curr_oind = 0; curr_oind < static_cast<int>(out_.size()); ++curr_oind)
    if (out_[curr_oind].nnz() != 0) break;

  std::vector<int64_t> refcount(nodes.size(), 0);
  algorithm_.clear();
  algorithm_.reserve(nodes.size());
  for (Node* n : nodes) {
    Instr ae;
    ae.op = n == nullptr ? static_cast<int>(OP_OUTPUT) : n->op;
    ae.i0 = 0;
    ae.i1 = 0;
    ae.i2 = 0;
    int ndeps = op_ndeps(ae.op);
    switch (ae.op) {
      case OP_CONST:
        ae.d = n->value;
        ae.i0 = n->temp;
        break;
      case OP_PARAMETER:
        symb_loc.emplace_back(static_cast<int>(algorithm_.size()), n);
        ae.i0 = n->temp;
        ae.d = 0;
        break;
      case OP_OUTPUT:
        ae.i0 = curr_oind;
        ae.i1 = out_[curr_oind].nz[curr_nz]->temp;
        ae.i2 = curr_nz;
        curr_nz++;
        if (curr_nz >= out_[curr_oind].nnz()) {
          curr_nz = 0;
          curr_oind++;
          for (; curr_oind < static_cast<int>(out_.size()); ++curr_oind)
            if (out_[curr_oind].nnz() != 0) break;
        }
        break;
      default:
        ae.i0 = n->temp;
        ae.i1 = n->dep[0]->temp;
        ae.i2 = n->dep[1]->temp;
    }
    if (ae.op == OP_OUTPUT) ndeps = 1;
    int* dep = &ae.i1;
    for (int c = 0; c < ndeps; ++c) refcount.at(dep[c])++;
    algorithm_.push_back(ae);
  }

  // Live-variable slot assignment with a LIFO stack of free slots
  std::vector<int> place(nodes.size());
  std::stack<int> unused;
  int worksize = 0;
  for (auto& a : algorithm_) {
    int ndeps = a.op == OP_OUTPUT ? 1 : op_ndeps(a.op);
    int* dep = &a.i1;
    for (int c = ndeps - 1; c >= 0; --c) {
      int ch = dep[c];
      if (--refcount.at(ch) == 0) unused.push(place[ch]);
    }
    if (a.op != OP_OUTPUT) {
      if (opts_.live_variables && !unused.empty()) {
        a.i0 = place[a.i0] = unused.top();
        unused.pop();
      } else {
        a.i0 = place[a.i0] = worksize++;
      }
    }
    for (int c = 0; c < ndeps; ++c) dep[c] = place[dep[c]];
    if (ndeps == 1 && a.op != OP_OUTPUT) a.i2 = a.i1;
  }
  worksize_ = worksize;

  for (Node* n : nodes)
    if (n) n->temp = 0;
  for (auto& sl : symb_loc) sl.second->temp = sl.first + 1;
  for (int ind = 0; ind < static_cast<int>(in_.size()); ++ind) {
    int nz = 0;
    for (SXE e : in_[ind].nz) {
      int i = e->temp - 1;
      if (i >= 0) {
        algorithm_[i].op = OP_INPUT;
        algorithm_[i].i1 = ind;
        algorithm_[i].i2 = nz;
        e->temp = 0;
      }
      ++nz;
    }
  }
  free_vars_.clear();
  for (auto& sl : symb_loc) {
    if (sl.second->temp != 0) {
      algorithm_[sl.first].i1 = static_cast<int>(free_vars_.size());
      algorithm_[sl.first].i2 = 0;
      free_vars_.push_back(sl.second);
      sl.second->temp = 0;
    }
  }
  if (!opts_.allow_free && !free_vars_.empty()) {
    std::string names;
    for (SXE e : free_vars_) names += (names.empty() ? "" : ", ") + sx_name(e);
    throw std::runtime_error(name_ + "::init: Initialization failed since variables [" + names +
                             "] are free. These symbols occur in the output expressions but you "
                             "forgot to declare these as inputs. Set option 'allow_free' to allow "
                             "free variables.");
  }
}

void SXFunction::eval(const double* const* arg, double* const* res, double* w) const {
  if (!free_vars_.empty())
    throw std::runtime_error("Cannot evaluate \"" + name_ + "\" since variables are free.");
  for (const Instr& e : algorithm_) {
    switch (e.op) {
      case OP_CONST: w[e.i0] = e.d; break;
      case OP_INPUT: w[e.i0] = arg[e.i1] == nullptr ? 0 : arg[e.i1][e.i2]; break;
      case OP_OUTPUT:
        if (res[e.i0] != nullptr) res[e.i0][e.i2] = w[e.i1];
        break;
      default: w[e.i0] = op_fun(e.op, w[e.i1], w[e.i2]);
    }
  }
}

std::vector<SXMat> SXFunction::eval_sx(const std::vector<SXMat>& arg) const {
  if (arg.size() != in_.size()) throw std::runtime_error(name_ + ": wrong number of arguments");
  for (size_t i = 0; i < arg.size(); ++i)
    if (arg[i].sp != in_[i].sp && arg[i].nnz() != 0)
      throw std::runtime_error(name_ + ": argument sparsity mismatch for input " + std::to_string(i));
  std::vector<SXMat> res;
  for (const auto& o : out_) res.push_back(SXMat::zeros(o.sp));
  std::vector<SXE> w(worksize_, sx_const(0));
  for (const Instr& e : algorithm_) {
    switch (e.op) {
      case OP_CONST: w[e.i0] = sx_const(e.d); break;
      case OP_INPUT: w[e.i0] = arg[e.i1].nnz() == 0 ? sx_const(0) : arg[e.i1].nz[e.i2]; break;
      case OP_PARAMETER: w[e.i0] = free_vars_.at(e.i1); break;
      case OP_OUTPUT: res[e.i0].nz[e.i2] = w[e.i1]; break;
      default: w[e.i0] = sx_fun(e.op, w[e.i1], w[e.i2]);
    }
  }
  return res;
}

// SXFunction::ad_forward (casadi/core/sx_function.cpp:1194-1366)
void SXFunction::ad_forward(const std::vector<std::vector<SXMat>>& fseed,
                            std::vector<std::vector<SXMat>>& fsens) const {
  const size_t nfwd = fseed.size();
  fsens.assign(nfwd, {});
  if (nfwd == 0) return;
  for (const auto& r : fseed) {
    if (r.size() != in_.size()) throw std::runtime_error("ad_forward: wrong number of seeds");
    for (size_t i = 0; i < in_.size(); ++i)
      if (r[i].sp != in_[i].sp) throw std::runtime_error("ad_forward: seed sparsity mismatch");
  }
  for (size_t d = 0; d < nfwd; ++d)
    for (const auto& o : out_) fsens[d].push_back(SXMat::zeros(o.sp));
  // partial derivatives of every operation node
  std::vector<std::array<SXE, 2>> pd(operations_.size());
  for (size_t k = 0; k < operations_.size(); ++k) {
    SXE f = operations_[k];
    sx_der(f->op, f->dep[0], f->dep[1], f, pd[k].data());
  }
  std::vector<SXE> w(worksize_);
  for (size_t dir = 0; dir < nfwd; ++dir) {
    size_t it = 0;
    for (const Instr& a : algorithm_) {
      switch (a.op) {
        case OP_INPUT: w[a.i0] = fseed[dir][a.i1].nz[a.i2]; break;
        case OP_OUTPUT: fsens[dir][a.i0].nz[a.i2] = w[a.i1]; break;
        case OP_CONST: case OP_PARAMETER: w[a.i0] = sx_const(0); break;
        case OP_IF_ELSE_ZERO:
          w[a.i0] = sx_binary(OP_IF_ELSE_ZERO, pd[it++][1], w[a.i2]);
          break;
        default:
          if (op_is_binary(a.op)) {
            w[a.i0] = sx_binary(OP_ADD, sx_binary(OP_MUL, pd[it][0], w[a.i1]),
                                sx_binary(OP_MUL, pd[it][1], w[a.i2]));
          } else {
            w[a.i0] = sx_binary(OP_MUL, pd[it][0], w[a.i1]);
          }
          it++;
      }
    }
  }
}

// SXFunction::ad_reverse (casadi/core/sx_function.cpp:1368-1569)
void SXFunction::ad_reverse(const std::vector<std::vector<SXMat>>& aseed,
                            std::vector<std::vector<SXMat>>& asens) const {
  const size_t nadj = aseed.size();
  asens.assign(nadj, {});
  if (nadj == 0) return;
  for (const auto& r : aseed) {
    if (r.size() != out_.size()) throw std::runtime_error("ad_reverse: wrong number of seeds");
    for (size_t i = 0; i < out_.size(); ++i)
      if (r[i].sp != out_[i].sp) throw std::runtime_error("ad_reverse: seed sparsity mismatch");
  }
  for (size_t d = 0; d < nadj; ++d)
    for (const auto& m : in_) asens[d].push_back(SXMat::zeros(m.sp));
  std::vector<std::array<SXE, 2>> pd(operations_.size());
  for (size_t k = 0; k < operations_.size(); ++k) {
    SXE f = operations_[k];
    sx_der(f->op, f->dep[0], f->dep[1], f, pd[k].data());
  }
  std::vector<SXE> w(worksize_, sx_const(0));
  for (size_t dir = 0; dir < nadj; ++dir) {
    size_t it = operations_.size();
    for (auto a = algorithm_.rbegin(); a != algorithm_.rend(); ++a) {
      SXE seed;
      switch (a->op) {
        case OP_INPUT:
          asens[dir][a->i1].nz[a->i2] = w[a->i0];
          w[a->i0] = sx_const(0);
          break;
        case OP_OUTPUT:
          w[a->i1] = sx_binary(OP_ADD, w[a->i1], aseed[dir][a->i0].nz[a->i2]);
          break;
        case OP_CONST: case OP_PARAMETER: w[a->i0] = sx_const(0); break;
        case OP_IF_ELSE_ZERO:
          seed = w[a->i0];
          w[a->i0] = sx_const(0);
          --it;
          w[a->i2] = sx_binary(OP_ADD, w[a->i2], sx_binary(OP_IF_ELSE_ZERO, pd[it][1], seed));
          break;
        default:
          --it;
          seed = w[a->i0];
          w[a->i0] = sx_const(0);
          if (op_is_binary(a->op)) {
            w[a->i1] = sx_binary(OP_ADD, w[a->i1], sx_binary(OP_MUL, pd[it][0], seed));
            w[a->i2] = sx_binary(OP_ADD, w[a->i2], sx_binary(OP_MUL, pd[it][1], seed));
          } else {
            w[a->i1] = sx_binary(OP_ADD, w[a->i1], sx_binary(OP_MUL, pd[it][0], seed));
          }
      }
    }
  }
  // structurally zero sensitivities lose their pattern (casadi #4345)
  for (size_t d = 0; d < nadj; ++d)
    for (size_t i = 0; i < in_.size(); ++i)
      if (asens[d][i].is_zero())
        asens[d][i] = SXMat::empty(in_[i].sp.nrow, in_[i].sp.ncol);
}

// --- dependency (sparsity) propagation -----------------------------------

// SXFunction::sp_forward (casadi/core/sx_function.cpp:1657-1681) on 64-bit
// blocks: arg[i][nz] -> res[o][nz]
void SXFunction::sp_forward_blocks(const std::vector<std::vector<uint64_t>>& arg,
                                   std::vector<std::vector<uint64_t>>& res) const {
  std::vector<uint64_t> w(worksize_, 0);
  for (const Instr& e : algorithm_) {
    switch (e.op) {
      case OP_CONST: case OP_PARAMETER: w[e.i0] = 0; break;
      case OP_INPUT:
        w[e.i0] = (!arg[e.i1].empty() && opts_.is_diff_in[e.i1]) ? arg[e.i1][e.i2] : 0;
        break;
      case OP_OUTPUT:
        if (!res[e.i0].empty()) res[e.i0][e.i2] = opts_.is_diff_out[e.i0] ? w[e.i1] : 0;
        break;
      default: w[e.i0] = w[e.i1] | w[e.i2];
    }
  }
}

// SXFunction::sp_reverse (casadi/core/sx_function.cpp:1733-1782)
void SXFunction::sp_reverse_blocks(std::vector<std::vector<uint64_t>>& arg,
                                   std::vector<std::vector<uint64_t>>& res) const {
  std::vector<uint64_t> w(worksize_, 0);
  for (auto it = algorithm_.rbegin(); it != algorithm_.rend(); ++it) {
    uint64_t seed;
    switch (it->op) {
      case OP_CONST: case OP_PARAMETER: w[it->i0] = 0; break;
      case OP_INPUT:
        if (!arg[it->i1].empty() && opts_.is_diff_in[it->i1]) arg[it->i1][it->i2] |= w[it->i0];
        w[it->i0] = 0;
        break;
      case OP_OUTPUT:
        if (!res[it->i0].empty() && opts_.is_diff_out[it->i0]) {
          w[it->i1] |= res[it->i0][it->i2];
          res[it->i0][it->i2] = 0;
        }
        break;
      default:
        seed = w[it->i0];
        w[it->i0] = 0;
        w[it->i1] |= seed;
        w[it->i2] |= seed;
    }
  }
}

// FunctionInternal::get_jac_sparsity (casadi/core/function_internal.cpp:1928-1976),
// compact (nnz_out x nnz_in) pattern of output oind w.r.t. input iind.  The
// hierarchical variants used by CasADi for large blocks compute the same
// structural pattern, so plain 64-wide sweeps are used here.
Sparsity SXFunction::get_jac_sparsity(size_t oind, size_t iind, bool symmetric) const {
  const int64_t nz_in = nnz_in(iind), nz_out = nnz_out(oind);
  const int64_t nsweep_fwd = (nz_in + 63) / 64, nsweep_adj = (nz_out + 63) / 64;
  ivec rr, cc;
  if (nsweep_fwd <= nsweep_adj) {
    for (int64_t s = 0; s < nsweep_fwd; ++s) {
      std::vector<std::vector<uint64_t>> arg(in_.size()), res(out_.size());
      for (size_t i = 0; i < in_.size(); ++i) arg[i].assign(nnz_in(i), 0);
      for (size_t i = 0; i < out_.size(); ++i) res[i].assign(nnz_out(i), 0);
      for (int64_t b = 0; b < 64 && s * 64 + b < nz_in; ++b) arg[iind][s * 64 + b] = uint64_t(1) << b;
      sp_forward_blocks(arg, res);
      for (int64_t r = 0; r < nz_out; ++r)
        for (int64_t b = 0; b < 64; ++b)
          if (res[oind][r] >> b & 1) {
            rr.push_back(r);
            cc.push_back(s * 64 + b);
          }
    }
  } else {
    for (int64_t s = 0; s < nsweep_adj; ++s) {
      std::vector<std::vector<uint64_t>> arg(in_.size()), res(out_.size());
      for (size_t i = 0; i < in_.size(); ++i) arg[i].assign(nnz_in(i), 0);
      for (size_t i = 0; i < out_.size(); ++i) res[i].assign(nnz_out(i), 0);
      for (int64_t b = 0; b < 64 && s * 64 + b < nz_out; ++b) res[oind][s * 64 + b] = uint64_t(1) << b;
      sp_reverse_blocks(arg, res);
      for (int64_t c = 0; c < nz_in; ++c)
        for (int64_t b = 0; b < 64; ++b)
          if (arg[iind][c] >> b & 1) {
            rr.push_back(s * 64 + b);
            cc.push_back(c);
          }
    }
  }
  (void)symmetric;
  return Sparsity::triplet(nz_out, nz_in, rr, cc);
}

Sparsity SXFunction::jac_sparsity(size_t oind, size_t iind, bool compact, bool symmetric) const {
  auto key = std::make_pair(oind, iind);
  Sparsity jsp;
  auto it = jsp_cache_[1].find(key);
  if (it != jsp_cache_[1].end()) {
    jsp = it->second;
  } else {
    if (!opts_.is_diff_out.at(oind) || !opts_.is_diff_in.at(iind)) {
      jsp = Sparsity(nnz_out(oind), nnz_in(iind));
    } else {
      if (symmetric && !out_[oind].sp.is_dense())
        throw std::runtime_error(
            "Symmetry exploitation in Jacobian assumes dense expression. A potential workaround "
            "is to apply densify().");
      jsp = get_jac_sparsity(oind, iind, symmetric);
    }
    jsp_cache_[1][key] = jsp;
  }
  // from_compact: enlarge rows/columns to the full dimensions
  // (FunctionInternal::from_compact, casadi/core/function_internal.cpp:1985-2001)
  ivec fo = find_nz(out_[oind].sp), fi = find_nz(in_[iind].sp);
  Sparsity full = jsp.enlarge(out_[oind].sp.numel(), in_[iind].sp.numel(), fo, fi);
  // A symmetric Jacobian keeps only the entries present in both J and J^T, in
  // the full space (the pattern operator* is an intersection;
  // casadi/core/function_internal.cpp:2051-2062).  A dependency that an
  // operation such as floor or sign drops in one direction only is removed.
  if (symmetric && !full.is_symmetric()) {
    full = full.intersect(full.T());
    if (compact) jsp = full.sub(fo, fi);  // to_compact (function_internal.cpp:1978-1983)
  }
  return compact ? jsp : full;
}

void SXFunction::get_partition(size_t iind, size_t oind, Sparsity& D1, Sparsity& D2, bool compact,
                               bool symmetric, bool allow_forward, bool allow_reverse) const {
  Sparsity AT = jac_sparsity(oind, iind, compact, symmetric);
  Sparsity A = symmetric ? AT : AT.T();
  D1 = Sparsity::null();
  D2 = Sparsity::null();
  if (symmetric) {
    D1 = A.star_coloring();
    return;
  }
  const double w = opts_.ad_weight;
  if (w == 1) allow_forward = false;
  if (w == 0) allow_reverse = false;
  double best = std::numeric_limits<double>::infinity();
  bool test_fwd_first = allow_forward && w * static_cast<double>(A.nrow) <= (1 - w) * static_cast<double>(A.ncol);
  int mode_fwd = test_fwd_first ? 0 : 1;
  for (int mode = 0; mode < 2; ++mode) {
    bool fwd = mode == mode_fwd;
    if (!allow_forward && fwd) continue;
    if (!allow_reverse && !fwd) continue;
    if (fwd) {
      bool d = best >= w * static_cast<double>(A.nrow);
      int64_t maxc = d ? A.nrow : static_cast<int64_t>(std::floor(best / w));
      Sparsity c = AT.uni_coloring(A, maxc);
      if (!c.is_null()) {
        D1 = c;
        D2 = Sparsity::null();
        best = w * static_cast<double>(D1.ncol);
      }
    } else {
      bool d = best >= (1 - w) * static_cast<double>(A.ncol);
      int64_t maxc = d ? A.ncol : static_cast<int64_t>(std::floor(best / (1 - w)));
      Sparsity c = A.uni_coloring(AT, maxc);
      if (!c.is_null()) {
        D2 = c;
        D1 = Sparsity::null();
        best = (1 - w) * static_cast<double>(D2.ncol);
      }
    }
  }
}

// Compressed Jacobian bookkeeping shared by XFunction::jac and the numeric
// seed-sweep kernels (casadi/core/x_function.hpp:567-826).
SXFunction::JacPlan SXFunction::jac_plan(bool symmetric) const {
  if (in_.size() < 1 || out_.size() != 1)
    throw std::runtime_error("jac: Function must have one differentiable input and one output");
  JacPlan plan;
  plan.jsp = jac_sparsity(0, 0, true, symmetric);  // nnz_out x nnz_in
  Sparsity D1, D2;
  if (plan.jsp.nnz() == 0) return plan;
  get_partition(0, 0, D1, D2, true, symmetric, true, true);
  const Sparsity& J = plan.jsp;
  const int64_t nnzJ = J.nnz();
  plan.src_dir.assign(nnzJ, -1);
  plan.src_nz.assign(nnzJ, -1);
  Sparsity JT = J.T();  // nnz_in x nnz_out: column r lists the inputs of output r
  if (!D1.is_null()) {
    plan.forward = true;
    plan.ndir = D1.ncol;
    for (int64_t d = 0; d < D1.ncol; ++d) {
      ivec seed(D1.row.begin() + D1.colind[d], D1.row.begin() + D1.colind[d + 1]);
      plan.seeds.push_back(seed);
      ivec cnt;
      if (symmetric) {
        cnt.assign(J.nrow, 0);
        for (int64_t c : seed)
          for (int64_t k = J.colind[c]; k < J.colind[c + 1]; ++k) cnt[J.row[k]]++;
      }
      for (int64_t c : seed) {
        for (int64_t k = J.colind[c]; k < J.colind[c + 1]; ++k) {
          int64_t r = J.row[k];
          if (symmetric) {
            if (cnt[r] == 1) {
              plan.src_dir[k] = d;
              plan.src_nz[k] = r;
              int64_t kt = J.get_nz(c, r);
              plan.src_dir[kt] = d;
              plan.src_nz[kt] = r;
            }
          } else {
            plan.src_dir[k] = d;
            plan.src_nz[k] = r;
          }
        }
      }
    }
  } else {
    plan.forward = false;
    plan.ndir = D2.ncol;
    for (int64_t d = 0; d < D2.ncol; ++d) {
      ivec seed(D2.row.begin() + D2.colind[d], D2.row.begin() + D2.colind[d + 1]);
      plan.seeds.push_back(seed);
      for (int64_t r : seed) {
        for (int64_t kt = JT.colind[r]; kt < JT.colind[r + 1]; ++kt) {
          int64_t c = JT.row[kt];
          int64_t k = J.get_nz(r, c);
          plan.src_dir[k] = d;
          plan.src_nz[k] = c;
        }
      }
    }
  }
  return plan;
}

// Plan over all inputs / outputs flattened (veccat), as Function::jacobian
// builds it (casadi/core/x_function.hpp:939-960).
SXFunction::JacPlan SXFunction::jac_plan_all(bool symmetric) const {
  if (in_.size() == 1 && out_.size() == 1) return jac_plan(symmetric);
  FunctionOptions o;
  o.allow_free = true;
  SXFunction tmp("flattened_" + name_, {veccat(in_)}, {veccat(out_)}, {}, {}, o);
  return tmp.jac_plan(symmetric);
}

// XFunction::jac (casadi/core/x_function.hpp:442-840)
SXMat SXFunction::jac(bool symmetric, bool allow_forward, bool allow_reverse) const {
  (void)allow_forward;
  (void)allow_reverse;
  if (nnz_in() == 0 || nnz_out() == 0)
    return SXMat::empty(out_.at(0).sp.numel(), in_.at(0).sp.numel());
  for (size_t i = 1; i < in_.size(); ++i)
    if (opts_.is_diff_in[i]) throw std::runtime_error("jac: Not implemented (several inputs)");
  JacPlan plan = jac_plan(symmetric);
  const Sparsity& J = plan.jsp;
  std::vector<SXE> jnz(J.nnz(), sx_const(0));
  if (J.nnz() > 0) {
    std::vector<std::vector<SXMat>> sens;
    if (plan.forward) {
      std::vector<std::vector<SXMat>> fseed(plan.ndir);
      for (int64_t d = 0; d < plan.ndir; ++d) {
        for (size_t i = 0; i < in_.size(); ++i) fseed[d].push_back(SXMat::zeros(in_[i].sp));
        for (int64_t c : plan.seeds[d]) fseed[d][0].nz[c] = sx_const(1);
      }
      ad_forward(fseed, sens);
      for (int64_t k = 0; k < J.nnz(); ++k)
        if (plan.src_dir[k] >= 0) jnz[k] = sens[plan.src_dir[k]][0].nz[plan.src_nz[k]];
    } else {
      std::vector<std::vector<SXMat>> aseed(plan.ndir);
      for (int64_t d = 0; d < plan.ndir; ++d) {
        aseed[d].push_back(SXMat::zeros(out_[0].sp));
        for (int64_t r : plan.seeds[d]) aseed[d][0].nz[r] = sx_const(1);
      }
      ad_reverse(aseed, sens);
      for (int64_t k = 0; k < J.nnz(); ++k) {
        if (plan.src_dir[k] < 0) continue;
        const SXMat& as = sens[plan.src_dir[k]][0];
        if (as.nnz() == 0) continue;
        jnz[k] = as.nz[plan.src_nz[k]];
      }
    }
  }
  // enlarge compact -> full dimensions (rows: out numel, cols: in numel)
  ivec fo = find_nz(out_[0].sp), fi = find_nz(in_[0].sp);
  ivec col = J.get_col(), rr(J.nnz()), cc(J.nnz()), map;
  for (int64_t k = 0; k < J.nnz(); ++k) {
    rr[k] = fo[J.row[k]];
    cc[k] = fi[col[k]];
  }
  Sparsity Jfull = Sparsity::triplet(out_[0].sp.numel(), in_[0].sp.numel(), rr, cc, &map);
  std::vector<SXE> nz(Jfull.nnz());
  for (int64_t k = 0; k < J.nnz(); ++k) nz[map[k]] = jnz[k];
  return SXMat(Jfull, nz);
}

// XFunction::get_forward (casadi/core/x_function.hpp:844-888)
FunctionPtr SXFunction::forward(int64_t nfwd) const {
  std::vector<std::vector<SXMat>> fseed(nfwd), fsens;
  for (int64_t d = 0; d < nfwd; ++d)
    for (size_t i = 0; i < in_.size(); ++i)
      fseed[d].push_back(sym_mat("f" + std::to_string(d) + "_" + name_in_[i],
                                 opts_.is_diff_in[i] ? in_[i].sp : Sparsity(in_[i].sp.nrow, in_[i].sp.ncol)));
  ad_forward(fseed, fsens);
  std::vector<SXMat> ret_in(in_);
  std::vector<std::string> inames(name_in_), onames;
  for (size_t i = 0; i < out_.size(); ++i) {
    ret_in.push_back(SXMat::empty(out_[i].sp.nrow, out_[i].sp.ncol));
    inames.push_back("out_" + name_out_[i]);
  }
  for (size_t i = 0; i < in_.size(); ++i) {
    std::vector<SXMat> v;
    for (int64_t d = 0; d < nfwd; ++d) v.push_back(fseed[d][i]);
    ret_in.push_back(horzcat(v));
    inames.push_back("fwd_" + name_in_[i]);
  }
  std::vector<SXMat> ret_out;
  for (size_t i = 0; i < out_.size(); ++i) {
    std::vector<SXMat> v;
    for (int64_t d = 0; d < nfwd; ++d) v.push_back(fsens[d][i]);
    ret_out.push_back(horzcat(v));
    onames.push_back("fwd_" + name_out_[i]);
  }
  FunctionOptions o;
  return std::make_shared<SXFunction>("fwd" + std::to_string(nfwd) + "_" + name_, ret_in, ret_out,
                                      inames, onames, o);
}

// XFunction::get_reverse (casadi/core/x_function.hpp:892-935)
FunctionPtr SXFunction::reverse(int64_t nadj) const {
  std::vector<std::vector<SXMat>> aseed(nadj), asens;
  for (int64_t d = 0; d < nadj; ++d)
    for (size_t i = 0; i < out_.size(); ++i) {
      std::string nm = "a" + (nadj > 1 ? std::to_string(d) + "_" : std::string()) + std::to_string(i);
      aseed[d].push_back(sym_mat(nm, opts_.is_diff_out[i] ? out_[i].sp
                                                          : Sparsity(out_[i].sp.nrow, out_[i].sp.ncol)));
    }
  ad_reverse(aseed, asens);
  std::vector<SXMat> ret_in(in_);
  std::vector<std::string> inames(name_in_), onames;
  for (size_t i = 0; i < out_.size(); ++i) {
    ret_in.push_back(SXMat::empty(out_[i].sp.nrow, out_[i].sp.ncol));
    inames.push_back("out_" + name_out_[i]);
  }
  for (size_t i = 0; i < out_.size(); ++i) {
    std::vector<SXMat> v;
    for (int64_t d = 0; d < nadj; ++d) v.push_back(aseed[d][i]);
    ret_in.push_back(horzcat(v));
    inames.push_back("adj_" + name_out_[i]);
  }
  std::vector<SXMat> ret_out;
  for (size_t i = 0; i < in_.size(); ++i) {
    std::vector<SXMat> v;
    bool all_empty = true;
    for (int64_t d = 0; d < nadj; ++d) {
      v.push_back(asens[d][i]);
      if (asens[d][i].nnz() != 0) all_empty = false;
    }
    if (all_empty) {
      ret_out.push_back(SXMat::empty(in_[i].sp.nrow, in_[i].sp.ncol * nadj));
    } else {
      // ensure_stacked: every direction on the input's pattern
      std::vector<SXMat> w;
      for (int64_t d = 0; d < nadj; ++d)
        w.push_back(asens[d][i].nnz() == 0 ? SXMat::zeros(in_[i].sp) : asens[d][i]);
      ret_out.push_back(horzcat(w));
    }
    onames.push_back("adj_" + name_in_[i]);
  }
  FunctionOptions o;
  return std::make_shared<SXFunction>("adj" + std::to_string(nadj) + "_" + name_, ret_in, ret_out,
                                      inames, onames, o);
}

// XFunction::get_jacobian (casadi/core/x_function.hpp:939-1019)
FunctionPtr SXFunction::jacobian() const {
  std::vector<SXMat> diff_in, diff_out, non_diff_in;
  for (size_t i = 0; i < in_.size(); ++i) (opts_.is_diff_in[i] ? diff_in : non_diff_in).push_back(in_[i]);
  for (size_t i = 0; i < out_.size(); ++i)
    if (opts_.is_diff_out[i]) diff_out.push_back(out_[i]);
  FunctionOptions tmp_opts;
  tmp_opts.allow_free = true;
  tmp_opts.is_diff_in = {true};
  std::vector<SXMat> tmp_args = {veccat(diff_in)};
  if (!non_diff_in.empty()) {
    tmp_args.push_back(veccat(non_diff_in));
    tmp_opts.is_diff_in.push_back(false);
  }
  SXFunction tmp("flattened_" + name_, tmp_args, {veccat(diff_out)}, {}, {}, tmp_opts);
  SXMat J = tmp.jac();
  // block split by numel offsets
  ivec r_off = {0}, c_off = {0};
  for (auto& e : diff_out) r_off.push_back(r_off.back() + e.sp.numel());
  for (auto& e : diff_in) c_off.push_back(c_off.back() + e.sp.numel());
  auto block = [&](size_t bi, size_t bj) {
    ivec rr, cc;
    std::vector<SXE> nz;
    ivec col = J.sp.get_col();
    for (int64_t k = 0; k < J.sp.nnz(); ++k) {
      int64_t r = J.sp.row[k], c = col[k];
      if (r >= r_off[bi] && r < r_off[bi + 1] && c >= c_off[bj] && c < c_off[bj + 1]) {
        rr.push_back(r - r_off[bi]);
        cc.push_back(c - c_off[bj]);
        nz.push_back(J.nz[k]);
      }
    }
    ivec map;
    Sparsity sp = Sparsity::triplet(r_off[bi + 1] - r_off[bi], c_off[bj + 1] - c_off[bj], rr, cc, &map);
    std::vector<SXE> snz(sp.nnz());
    for (size_t k = 0; k < map.size(); ++k) snz[map[k]] = nz[k];
    return SXMat(sp, snz);
  };
  std::vector<SXMat> ret_out;
  std::vector<std::string> onames;
  size_t diff_i = 0;
  for (size_t i = 0; i < out_.size(); ++i) {
    size_t diff_j = 0;
    for (size_t j = 0; j < in_.size(); ++j) {
      if (opts_.is_diff_out[i] && opts_.is_diff_in[j]) {
        ret_out.push_back(block(diff_i, diff_j));
        diff_j++;
      } else {
        ret_out.push_back(SXMat::empty(out_[i].sp.numel(), in_[j].sp.numel()));
      }
      onames.push_back("jac_" + name_out_[i] + "_" + name_in_[j]);
    }
    if (opts_.is_diff_out[i]) diff_i++;
  }
  std::vector<SXMat> ret_in(in_);
  std::vector<std::string> inames(name_in_);
  for (size_t i = 0; i < out_.size(); ++i) {
    ret_in.push_back(SXMat::empty(out_[i].sp.nrow, out_[i].sp.ncol));
    inames.push_back("out_" + name_out_[i]);
  }
  FunctionOptions o;
  o.allow_free = true;
  return std::make_shared<SXFunction>("jac_" + name_, ret_in, ret_out, inames, onames, o);
}

SXMat sx_jacobian(const SXMat& f, const SXMat& x, bool symmetric) {
  FunctionOptions o;
  o.allow_free = true;
  SXFunction h("jac_helper", {x}, {f}, {}, {}, o);
  return h.jac(symmetric);
}

// gradient = jtimes(ex, arg, 1, tr=true) projected on arg's pattern
SXMat sx_gradient(const SXMat& f, const SXMat& x) {
  if (f.sp.numel() != 1) throw std::runtime_error("'gradient' only defined for scalar outputs");
  FunctionOptions o;
  o.allow_free = true;
  SXFunction h("reverse_temp", {x}, {f}, {}, {}, o);
  std::vector<std::vector<SXMat>> aseed(1), asens;
  aseed[0].push_back(SXMat(f.sp, std::vector<SXE>(f.sp.nnz(), sx_const(1))));
  h.ad_reverse(aseed, asens);
  SXMat g = asens[0][0];
  if (g.nnz() == 0) return SXMat::zeros(x.sp);
  return g;
}

SXMat sx_hessian(const SXMat& f, const SXMat& x, SXMat* g_out) {
  SXMat g = sx_gradient(f, x);
  if (g_out) *g_out = g;
  return sx_jacobian(g, x, true);
}

}  // namespace sxh
