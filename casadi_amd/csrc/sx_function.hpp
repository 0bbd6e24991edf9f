// SX Function: tape compiler, numeric VM and derivative construction (host).
//
// Restates casadi::SXFunction (casadi/core/sx_function.cpp) and the parts of
// XFunction / FunctionInternal that build derivative Functions
// (casadi/core/x_function.hpp:412-1019, casadi/core/function_internal.cpp:1928-2440).
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "sparsity.hpp"
#include "sx_node.hpp"

namespace sxh {

// A sparse matrix of scalar expressions: a pattern plus one node per nonzero.
struct SXMat {
  Sparsity sp;
  std::vector<SXE> nz;
  SXMat() = default;
  SXMat(Sparsity s, std::vector<SXE> n) : sp(std::move(s)), nz(std::move(n)) {}
  static SXMat zeros(const Sparsity& s) { return SXMat(s, std::vector<SXE>(s.nnz(), sx_const(0))); }
  static SXMat empty(int64_t nr, int64_t nc) { return SXMat(Sparsity(nr, nc), {}); }
  int64_t nnz() const { return sp.nnz(); }
  bool is_zero() const;
};

struct FunctionOptions {
  bool live_variables = true;
  bool allow_free = false;
  std::vector<bool> is_diff_in, is_diff_out;
  double ad_weight = 0.33;     // FunctionInternal default (function_internal.cpp:79)
  int64_t max_num_dir = 64;    // GlobalOptions::max_num_dir
};

class SXFunction {
 public:
  SXFunction(std::string name, std::vector<SXMat> in, std::vector<SXMat> out,
             std::vector<std::string> name_in = {}, std::vector<std::string> name_out = {},
             FunctionOptions opts = {});

  const std::string& name() const { return name_; }
  size_t n_in() const { return in_.size(); }
  size_t n_out() const { return out_.size(); }
  int64_t nnz_in(size_t i) const { return in_.at(i).nnz(); }
  int64_t nnz_out(size_t i) const { return out_.at(i).nnz(); }
  int64_t nnz_in() const;
  int64_t nnz_out() const;
  const Sparsity& sparsity_in(size_t i) const { return in_.at(i).sp; }
  const Sparsity& sparsity_out(size_t i) const { return out_.at(i).sp; }
  const std::vector<std::string>& name_in() const { return name_in_; }
  const std::vector<std::string>& name_out() const { return name_out_; }
  const std::vector<SXMat>& sx_in() const { return in_; }
  const std::vector<SXMat>& sx_out() const { return out_; }
  const std::vector<Instr>& algorithm() const { return algorithm_; }
  int64_t worksize() const { return worksize_; }
  const std::vector<SXE>& free_vars() const { return free_vars_; }
  const std::vector<bool>& is_diff_in() const { return opts_.is_diff_in; }
  const std::vector<bool>& is_diff_out() const { return opts_.is_diff_out; }

  // SXFunction::eval (casadi/core/sx_function.cpp:72-127); w has worksize()
  // doubles.  A null arg[i] reads as zeros, a null res[i] is skipped.
  void eval(const double* const* arg, double* const* res, double* w) const;

  // Symbolic evaluation (SXFunction::eval_sx) -- calling f on expressions.
  std::vector<SXMat> eval_sx(const std::vector<SXMat>& arg) const;

  // Symbolic forward / reverse sweeps (casadi/core/sx_function.cpp:1194-1569).
  void ad_forward(const std::vector<std::vector<SXMat>>& fseed,
                  std::vector<std::vector<SXMat>>& fsens) const;
  void ad_reverse(const std::vector<std::vector<SXMat>>& aseed,
                  std::vector<std::vector<SXMat>>& asens) const;

  // Structural Jacobian pattern of output oind w.r.t. input iind
  // (FunctionInternal::jac_sparsity, casadi/core/function_internal.cpp:2003-2066).
  Sparsity jac_sparsity(size_t oind, size_t iind, bool compact, bool symmetric) const;
  // FunctionInternal::get_partition (casadi/core/function_internal.cpp:2068-2165)
  void get_partition(size_t iind, size_t oind, Sparsity& D1, Sparsity& D2, bool compact,
                     bool symmetric, bool allow_forward, bool allow_reverse) const;

  // XFunction::jac (casadi/core/x_function.hpp:442-840): Jacobian of the
  // single output w.r.t. the single differentiable input, as an SX matrix.
  SXMat jac(bool symmetric = false, bool allow_forward = true, bool allow_reverse = true) const;

  // Derivative Functions with CasADi's signatures and names.
  std::shared_ptr<SXFunction> forward(int64_t nfwd) const;   // fwd<n>_<name>
  std::shared_ptr<SXFunction> reverse(int64_t nadj) const;   // adj<n>_<name>
  std::shared_ptr<SXFunction> jacobian() const;              // jac_<name>

  // Compressed-Jacobian plan used by the numeric seed-sweep kernels: the
  // colouring, the seeds per direction and where every compressed
  // sensitivity lands in the Jacobian nonzeros.
  struct JacPlan {
    Sparsity jsp;            // Jacobian pattern (nnz_out x nnz_in, compact)
    bool forward = true;     // forward (D1) or adjoint (D2) sweeps
    int64_t ndir = 0;
    // seed_of[d] = input nonzeros (forward) or output nonzeros (adjoint)
    std::vector<ivec> seeds;
    // for every Jacobian nonzero: (direction, sensitivity nonzero)
    ivec src_dir, src_nz;
  };
  JacPlan jac_plan(bool symmetric) const;
  // the same over all inputs and outputs concatenated (flat nonzero indices)
  JacPlan jac_plan_all(bool symmetric) const;

  int64_t n_instructions() const { return static_cast<int64_t>(algorithm_.size()); }

 private:
  void init();
  void sp_forward_blocks(const std::vector<std::vector<uint64_t>>& arg,
                         std::vector<std::vector<uint64_t>>& res) const;
  void sp_reverse_blocks(std::vector<std::vector<uint64_t>>& arg,
                         std::vector<std::vector<uint64_t>>& res) const;
  Sparsity get_jac_sparsity(size_t oind, size_t iind, bool symmetric) const;

  std::string name_;
  std::vector<SXMat> in_, out_;
  std::vector<std::string> name_in_, name_out_;
  FunctionOptions opts_;
  std::vector<Instr> algorithm_;
  int64_t worksize_ = 0;
  std::vector<SXE> operations_;
  std::vector<SXE> free_vars_;
  mutable std::map<std::pair<size_t, size_t>, Sparsity> jsp_cache_[2];
};

using FunctionPtr = std::shared_ptr<SXFunction>;

// SX-level helpers (casadi/core/sx_instantiator.cpp:1155-1177,
// casadi/core/generic_matrix.hpp:1695-1703)
SXMat sx_jacobian(const SXMat& f, const SXMat& x, bool symmetric = false);
SXMat sx_gradient(const SXMat& f, const SXMat& x);
SXMat sx_hessian(const SXMat& f, const SXMat& x, SXMat* g = nullptr);

}  // namespace sxh
