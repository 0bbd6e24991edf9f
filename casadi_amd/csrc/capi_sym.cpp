// C ABI of the host symbolic layer (cxs_*).  No exception crosses it: every
// call catches and records the message for cxh_last_error().
#include <omp.h>

#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/casadi_hip.h"
#include "capi_common.hpp"
#include "sx_function.hpp"

using namespace sxh;

struct cxs_fun_s {
  FunctionPtr f;
};
struct cxs_buf_s {
  std::vector<int64_t> i;  // int64 payload, or node pointers as uint64
};

namespace {

inline SXE N(cxs_sx h) { return reinterpret_cast<SXE>(static_cast<uintptr_t>(h)); }
inline cxs_sx H(SXE n) { return static_cast<cxs_sx>(reinterpret_cast<uintptr_t>(n)); }

cxs_buf make_buf(const ivec& v) { return new cxs_buf_s{v}; }
cxs_buf make_nodes(const std::vector<SXE>& v) {
  auto* b = new cxs_buf_s;
  b->i.resize(v.size());
  for (size_t k = 0; k < v.size(); ++k) b->i[k] = static_cast<int64_t>(H(v[k]));
  return b;
}

SXMat make_mat(const int64_t* sp, const cxs_sx* nz) {
  Sparsity s = Sparsity::from_compressed(sp);
  std::vector<SXE> n(s.nnz());
  for (int64_t k = 0; k < s.nnz(); ++k) n[k] = N(nz[k]);
  return SXMat(s, n);
}

}  // namespace

extern "C" {

cxs_sx cxs_sym(const char* name) { return H(sx_sym(name ? name : "")); }
cxs_sx cxs_const(double v) { return H(sx_const(v)); }
cxs_sx cxs_unary(int32_t op, cxs_sx x) { return H(sx_fun(op, N(x), N(x))); }
cxs_sx cxs_binary(int32_t op, cxs_sx x, cxs_sx y) { return H(sx_fun(op, N(x), N(y))); }
cxs_sx cxs_neg(cxs_sx x) { return H(sx_neg(N(x))); }
cxs_sx cxs_if_else(cxs_sx c, cxs_sx a, cxs_sx b) { return H(sx_if_else(N(c), N(a), N(b))); }

int cxs_unary_n(int32_t op, int64_t n, const cxs_sx* x, cxs_sx* out) {
  CXH_TRY
  for (int64_t k = 0; k < n; ++k) out[k] = H(sx_fun(op, N(x[k]), N(x[k])));
  return 0;
  CXH_CATCH
}

int cxs_binary_n(int32_t op, int64_t n, const cxs_sx* x, int64_t incx, const cxs_sx* y, int64_t incy,
                 cxs_sx* out) {
  CXH_TRY
  for (int64_t k = 0; k < n; ++k) out[k] = H(sx_fun(op, N(x[k * incx]), N(y[k * incy])));
  return 0;
  CXH_CATCH
}

int cxs_node_info(cxs_sx x, int32_t* op, double* value, cxs_sx* dep0, cxs_sx* dep1) {
  SXE n = N(x);
  if (!n) return cxh_fail("null node");
  if (op) *op = n->op;
  if (value) *value = n->value;
  int nd = n->n_dep();
  if (dep0) *dep0 = nd > 0 ? H(n->dep[0]) : 0;
  if (dep1) *dep1 = nd > 0 ? H(n->dep[1]) : 0;
  return 0;
}

const char* cxs_sym_name(cxs_sx x) {
  SXE n = N(x);
  if (!n || !n->is_symbolic()) return "";
  return sx_name(n).c_str();
}

int cxs_is_equal(cxs_sx x, cxs_sx y, int32_t depth) { return sx_is_equal(N(x), N(y), depth) ? 1 : 0; }
int64_t cxs_node_count(void) { return sx_node_count(); }

int cxs_function_new(const char* name, int32_t n_in, const int64_t* const* sp_in,
                     const cxs_sx* const* nz_in, const char* const* names_in, int32_t n_out,
                     const int64_t* const* sp_out, const cxs_sx* const* nz_out,
                     const char* const* names_out, int32_t allow_free, cxs_fun* out) {
  CXH_TRY
  std::vector<SXMat> in, o;
  std::vector<std::string> ni, no;
  for (int32_t i = 0; i < n_in; ++i) {
    in.push_back(make_mat(sp_in[i], nz_in[i]));
    if (names_in) ni.push_back(names_in[i]);
  }
  for (int32_t i = 0; i < n_out; ++i) {
    o.push_back(make_mat(sp_out[i], nz_out[i]));
    if (names_out) no.push_back(names_out[i]);
  }
  FunctionOptions opts;
  opts.allow_free = allow_free != 0;
  auto f = std::make_shared<SXFunction>(name ? name : "f", in, o, ni, no, opts);
  *out = new cxs_fun_s{f};
  return 0;
  CXH_CATCH
}

int cxs_function_free(cxs_fun f) {
  delete f;
  return 0;
}

int cxs_function_info(cxs_fun f, int32_t* n_in, int32_t* n_out, int64_t* n_instr, int64_t* worksize,
                      int64_t* n_free) {
  CXH_TRY
  if (n_in) *n_in = static_cast<int32_t>(f->f->n_in());
  if (n_out) *n_out = static_cast<int32_t>(f->f->n_out());
  if (n_instr) *n_instr = f->f->n_instructions();
  if (worksize) *worksize = f->f->worksize();
  if (n_free) *n_free = static_cast<int64_t>(f->f->free_vars().size());
  return 0;
  CXH_CATCH
}

const char* cxs_function_name(cxs_fun f) { return f->f->name().c_str(); }
const char* cxs_function_name_in(cxs_fun f, int32_t i) { return f->f->name_in().at(i).c_str(); }
const char* cxs_function_name_out(cxs_fun f, int32_t i) { return f->f->name_out().at(i).c_str(); }

int cxs_function_sparsity_in(cxs_fun f, int32_t i, cxs_buf* out) {
  CXH_TRY
  *out = make_buf(f->f->sparsity_in(i).compressed());
  return 0;
  CXH_CATCH
}
int cxs_function_sparsity_out(cxs_fun f, int32_t i, cxs_buf* out) {
  CXH_TRY
  *out = make_buf(f->f->sparsity_out(i).compressed());
  return 0;
  CXH_CATCH
}
int cxs_function_nz_out(cxs_fun f, int32_t i, cxs_buf* out) {
  CXH_TRY
  *out = make_nodes(f->f->sx_out().at(i).nz);
  return 0;
  CXH_CATCH
}
int cxs_function_nz_in(cxs_fun f, int32_t i, cxs_buf* out) {
  CXH_TRY
  *out = make_nodes(f->f->sx_in().at(i).nz);
  return 0;
  CXH_CATCH
}

int cxs_function_tape(cxs_fun f, void* instr_out) {
  CXH_TRY
  const auto& a = f->f->algorithm();
  std::memcpy(instr_out, a.data(), a.size() * sizeof(Instr));
  return 0;
  CXH_CATCH
}

int cxs_function_eval(cxs_fun f, const double* const* arg, double* const* res) {
  CXH_TRY
  std::vector<double> w(f->f->worksize() + 1);
  f->f->eval(arg, res, w.data());
  return 0;
  CXH_CATCH
}

// Map serial (n_threads == 1) / OpenMP (n_threads > 1) over AoS host data:
// instance k of input i at arg[i] + k*nnz_in(i) (casadi/core/map.cpp:141-157,
// 340-386).
int cxs_function_map_eval(cxs_fun f, int64_t n, const double* const* arg, double* const* res,
                          int32_t n_threads) {
  CXH_TRY
  const SXFunction& fn = *f->f;
  const size_t ni = fn.n_in(), no = fn.n_out();
  const int64_t ws = fn.worksize() + 1;
  if (n_threads <= 1) {
    std::vector<const double*> a(ni);
    std::vector<double*> r(no);
    std::vector<double> w(ws);
    for (int64_t k = 0; k < n; ++k) {
      for (size_t i = 0; i < ni; ++i) a[i] = arg[i] ? arg[i] + k * fn.nnz_in(i) : nullptr;
      for (size_t i = 0; i < no; ++i) r[i] = res[i] ? res[i] + k * fn.nnz_out(i) : nullptr;
      fn.eval(a.data(), r.data(), w.data());
    }
  } else {
#pragma omp parallel num_threads(n_threads)
    {
      std::vector<const double*> a(ni);
      std::vector<double*> r(no);
      std::vector<double> w(ws);
#pragma omp for schedule(static)
      for (int64_t k = 0; k < n; ++k) {
        for (size_t i = 0; i < ni; ++i) a[i] = arg[i] ? arg[i] + k * fn.nnz_in(i) : nullptr;
        for (size_t i = 0; i < no; ++i) r[i] = res[i] ? res[i] + k * fn.nnz_out(i) : nullptr;
        fn.eval(a.data(), r.data(), w.data());
      }
    }
  }
  return 0;
  CXH_CATCH
}

int cxs_function_call(cxs_fun f, const cxs_sx* const* arg_nz, cxs_sx* const* res_nz) {
  CXH_TRY
  const SXFunction& fn = *f->f;
  std::vector<SXMat> arg;
  for (size_t i = 0; i < fn.n_in(); ++i) {
    const Sparsity& sp = fn.sparsity_in(i);
    std::vector<SXE> nz(sp.nnz());
    for (int64_t k = 0; k < sp.nnz(); ++k) nz[k] = N(arg_nz[i][k]);
    arg.emplace_back(sp, nz);
  }
  auto res = fn.eval_sx(arg);
  for (size_t i = 0; i < res.size(); ++i)
    for (int64_t k = 0; k < res[i].nnz(); ++k) res_nz[i][k] = H(res[i].nz[k]);
  return 0;
  CXH_CATCH
}

int cxs_function_jacobian(cxs_fun f, cxs_fun* out) {
  CXH_TRY
  *out = new cxs_fun_s{f->f->jacobian()};
  return 0;
  CXH_CATCH
}
int cxs_function_forward(cxs_fun f, int64_t nfwd, cxs_fun* out) {
  CXH_TRY
  *out = new cxs_fun_s{f->f->forward(nfwd)};
  return 0;
  CXH_CATCH
}
int cxs_function_reverse(cxs_fun f, int64_t nadj, cxs_fun* out) {
  CXH_TRY
  *out = new cxs_fun_s{f->f->reverse(nadj)};
  return 0;
  CXH_CATCH
}

int cxs_function_jac_plan(cxs_fun f, int32_t symmetric, int32_t first_block, int32_t* forward, cxs_buf* jsp,
                          cxs_buf* seed_ptr, cxs_buf* seed_idx, cxs_buf* src_dir, cxs_buf* src_nz) {
  CXH_TRY
  auto plan = first_block ? f->f->jac_plan(symmetric != 0) : f->f->jac_plan_all(symmetric != 0);
  *forward = plan.forward ? 1 : 0;
  *jsp = make_buf(plan.jsp.compressed());
  ivec ptr{0}, idx;
  for (const auto& s : plan.seeds) {
    idx.insert(idx.end(), s.begin(), s.end());
    ptr.push_back(static_cast<int64_t>(idx.size()));
  }
  *seed_ptr = make_buf(ptr);
  *seed_idx = make_buf(idx);
  *src_dir = make_buf(plan.src_dir);
  *src_nz = make_buf(plan.src_nz);
  return 0;
  CXH_CATCH
}

int cxs_function_jac_sparsity(cxs_fun f, int32_t oind, int32_t iind, int32_t compact, int32_t symmetric,
                              cxs_buf* out) {
  CXH_TRY
  *out = make_buf(f->f->jac_sparsity(oind, iind, compact != 0, symmetric != 0).compressed());
  return 0;
  CXH_CATCH
}

int cxs_jacobian(const int64_t* sp_f, const cxs_sx* nz_f, const int64_t* sp_x, const cxs_sx* nz_x,
                 int32_t symmetric, cxs_buf* sp_out, cxs_buf* nz_out) {
  CXH_TRY
  SXMat J = sx_jacobian(make_mat(sp_f, nz_f), make_mat(sp_x, nz_x), symmetric != 0);
  *sp_out = make_buf(J.sp.compressed());
  *nz_out = make_nodes(J.nz);
  return 0;
  CXH_CATCH
}

int cxs_gradient(const int64_t* sp_f, const cxs_sx* nz_f, const int64_t* sp_x, const cxs_sx* nz_x,
                 cxs_buf* sp_out, cxs_buf* nz_out) {
  CXH_TRY
  SXMat g = sx_gradient(make_mat(sp_f, nz_f), make_mat(sp_x, nz_x));
  *sp_out = make_buf(g.sp.compressed());
  *nz_out = make_nodes(g.nz);
  return 0;
  CXH_CATCH
}

int cxs_hessian(const int64_t* sp_f, const cxs_sx* nz_f, const int64_t* sp_x, const cxs_sx* nz_x,
                cxs_buf* sp_out, cxs_buf* nz_out) {
  CXH_TRY
  SXMat H_ = sx_hessian(make_mat(sp_f, nz_f), make_mat(sp_x, nz_x));
  *sp_out = make_buf(H_.sp.compressed());
  *nz_out = make_nodes(H_.nz);
  return 0;
  CXH_CATCH
}

int cxs_sp_uni_coloring(const int64_t* sp, int64_t cutoff, cxs_buf* out) {
  CXH_TRY
  Sparsity A = Sparsity::from_compressed(sp);
  // colour the columns of sp (rows of the transpose drive the conflicts)
  Sparsity D = A.uni_coloring(A.T(), cutoff <= 0 ? INT64_MAX : cutoff);
  *out = make_buf(D.is_null() ? ivec{} : D.compressed());
  return 0;
  CXH_CATCH
}
int cxs_sp_star_coloring(const int64_t* sp, int64_t ordering, cxs_buf* out) {
  CXH_TRY
  *out = make_buf(Sparsity::from_compressed(sp).star_coloring(ordering).compressed());
  return 0;
  CXH_CATCH
}
int cxs_sp_largest_first(const int64_t* sp, cxs_buf* out) {
  CXH_TRY
  *out = make_buf(Sparsity::from_compressed(sp).largest_first());
  return 0;
  CXH_CATCH
}
int cxs_sp_amd(const int64_t* sp, cxs_buf* out) {
  CXH_TRY
  *out = make_buf(Sparsity::from_compressed(sp).amd());
  return 0;
  CXH_CATCH
}
int cxs_sp_etree(const int64_t* sp, int32_t ata, cxs_buf* out) {
  CXH_TRY
  *out = make_buf(Sparsity::from_compressed(sp).etree(ata != 0));
  return 0;
  CXH_CATCH
}
int cxs_sp_ldl(const int64_t* sp, int32_t amd, cxs_buf* sp_lt, cxs_buf* p) {
  CXH_TRY
  ivec perm;
  Sparsity Lt = Sparsity::from_compressed(sp).ldl(perm, amd != 0);
  *sp_lt = make_buf(Lt.compressed());
  *p = make_buf(perm);
  return 0;
  CXH_CATCH
}
int cxs_sp_qr(const int64_t* sp, int32_t amd, cxs_buf* sp_v, cxs_buf* sp_r, cxs_buf* prinv,
              cxs_buf* pc) {
  CXH_TRY
  Sparsity V, R;
  ivec pri, pcol;
  Sparsity::from_compressed(sp).qr_sparse(V, R, pri, pcol, amd != 0);
  *sp_v = make_buf(V.compressed());
  *sp_r = make_buf(R.compressed());
  *prinv = make_buf(pri);
  *pc = make_buf(pcol);
  return 0;
  CXH_CATCH
}
int cxs_sp_mtimes(const int64_t* sp_a, const int64_t* sp_b, cxs_buf* out) {
  CXH_TRY
  *out = make_buf(Sparsity::from_compressed(sp_a).mtimes(Sparsity::from_compressed(sp_b)).compressed());
  return 0;
  CXH_CATCH
}

int64_t cxs_buf_size(cxs_buf b) { return b ? static_cast<int64_t>(b->i.size()) : 0; }
int cxs_buf_copy(cxs_buf b, void* dst) {
  if (b && !b->i.empty()) std::memcpy(dst, b->i.data(), b->i.size() * sizeof(int64_t));
  return 0;
}
int cxs_buf_free(cxs_buf b) {
  delete b;
  return 0;
}

}  // extern "C"
