// Batched sparse LDL^T and Householder QR factor + solve on gfx950.
//
// Every instance shares CasADi's symbolic factorisation (pattern of L^T and
// the AMD permutation from Sparsity::ldl, or V / R / row and column
// permutations from Sparsity::qr_sparse); only values differ.  One lane
// factors one instance, so all structure reads (column pointers, row indices,
// permutation) are wave-uniform scalar loads, and every value access of a
// wave touches 64 consecutive doubles of a [nonzero][instance] array.
// The dense accumulator w of casadi_ldl lives in LDS ([row][lane]) when it
// fits, otherwise in HBM scratch.
//
// Numerics restated (same operation order):
//   casadi_ldl, casadi_ldl_trs, casadi_ldl_solve  casadi/core/runtime/casadi_ldl.hpp:24-109
//   casadi_house, casadi_qr, casadi_qr_mv, casadi_qr_trs, casadi_qr_solve,
//   casadi_qr_singular                             casadi/core/runtime/casadi_qr.hpp:24-230
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/casadi_hip.h"
#include "capi_common.hpp"
#include "tape_internal.hpp"

using namespace sxh;

namespace {

// [index][lane] view over LDS or HBM
struct Col {
  double* base;
  int64_t stride;
  __device__ double& operator[](int64_t i) const { return base[i * stride]; }
};

// v - sum over k in [b, e) of x[k] * w[row[k]], subtracted in increasing k
// (the reference's order).  Four terms' loads are issued before their
// subtractions so that the memory latencies of a column overlap; the
// arithmetic is unchanged.
__device__ __forceinline__ double sub_dot(double v, Col x, Col w, const int* __restrict__ row, int b, int e) {
  int k = b;
  for (; k + 4 <= e; k += 4) {
    const int r0 = row[k], r1 = row[k + 1], r2 = row[k + 2], r3 = row[k + 3];
    const double x0 = x[k], x1 = x[k + 1], x2 = x[k + 2], x3 = x[k + 3];
    const double w0 = w[r0], w1 = w[r1], w2 = w[r2], w3 = w[r3];
    v -= x0 * w0;
    v -= x1 * w1;
    v -= x2 * w2;
    v -= x3 * w3;
  }
  for (; k < e; ++k) v -= x[k] * w[row[k]];
  return v;
}

// s + sum over k in [b, e) of x[k] * w[row[k]], added in increasing k.
__device__ __forceinline__ double add_dot(double s, Col x, Col w, const int* __restrict__ row, int b, int e) {
  int k = b;
  for (; k + 4 <= e; k += 4) {
    const int r0 = row[k], r1 = row[k + 1], r2 = row[k + 2], r3 = row[k + 3];
    const double x0 = x[k], x1 = x[k + 1], x2 = x[k + 2], x3 = x[k + 3];
    const double w0 = w[r0], w1 = w[r1], w2 = w[r2], w3 = w[r3];
    s += x0 * w0;
    s += x1 * w1;
    s += x2 * w2;
    s += x3 * w3;
  }
  for (; k < e; ++k) s += x[k] * w[row[k]];
  return s;
}

// w[row[k]] -= a * x[k] for k in [b, e); the rows of one column are
// distinct, so the updates are independent and their loads are batched.
__device__ __forceinline__ void sub_axpy(double a, Col x, Col w, const int* __restrict__ row, int b, int e) {
  int k = b;
  for (; k + 4 <= e; k += 4) {
    const int r0 = row[k], r1 = row[k + 1], r2 = row[k + 2], r3 = row[k + 3];
    const double x0 = x[k], x1 = x[k + 1], x2 = x[k + 2], x3 = x[k + 3];
    const double w0 = w[r0], w1 = w[r1], w2 = w[r2], w3 = w[r3];
    w[r0] = w0 - a * x0;
    w[r1] = w1 - a * x1;
    w[r2] = w2 - a * x2;
    w[r3] = w3 - a * x3;
  }
  for (; k < e; ++k) w[row[k]] -= a * x[k];
}

struct LdlArgs {
  int n;
  const int* a_colind;
  const int* a_row;
  const int* lt_colind;
  const int* lt_row;
  const int* p;
  const double* a;      // nnz_a values, element k of instance j at a[k*a_s + j*a_t]
  int64_t a_s, a_t;
  double* x;            // n*nrhs values, element i at x[i*x_s + j*x_t]
  int64_t x_s, x_t;
  int nrhs;
  double* lt;           // [nnz_lt][lanes]
  double* d;            // [n][lanes]
  double* wg;           // [n][lanes] when w is not in LDS
  int64_t lanes;
  int* status;
  int64_t n_inst;
};

// MODE: LS_FUSED factor + solve, LS_FACTOR the factor alone (L^T and D stay
// in A.lt / A.d), LS_SOLVE the solve alone from the L^T and D stored there.
template <bool WLDS, int MODE>
__global__ void __launch_bounds__(64) ldl_kernel(LdlArgs A) {
  extern __shared__ double lds[];
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool active = j < A.n_inst;
  const int64_t jj = active ? j : A.n_inst - 1;
  Col lt{A.lt + j, A.lanes}, d{A.d + j, A.lanes};
  Col w = WLDS ? Col{lds + threadIdx.x, static_cast<int64_t>(blockDim.x)} : Col{A.wg + j, A.lanes};
  const int n = A.n;
  if (MODE != LS_SOLVE) {
  const double* a = A.a + jj * A.a_t;
  for (int r = 0; r < n; ++r) w[r] = 0;
  // scatter the permuted A into L^T and D
  for (int c = 0; c < n; ++c) {
    const int c1 = A.p[c];
    for (int k = A.a_colind[c1]; k < A.a_colind[c1 + 1]; ++k) w[A.a_row[k]] = a[k * A.a_s];
    for (int k = A.lt_colind[c]; k < A.lt_colind[c + 1]; ++k) lt[k] = w[A.p[A.lt_row[k]]];
    d[c] = w[A.p[c]];
    for (int k = A.a_colind[c1]; k < A.a_colind[c1 + 1]; ++k) w[A.a_row[k]] = 0;
  }
  // up-looking factorisation, column by column
  for (int c = 0; c < n; ++c) {
    double dc = d[c];
    for (int k = A.lt_colind[c]; k < A.lt_colind[c + 1]; ++k) {
      const int r = A.lt_row[k];
      const double v0 = lt[k];
      double v = sub_dot(v0, lt, w, A.lt_row, A.lt_colind[r], A.lt_colind[r + 1]);
      w[r] = v;
      v /= d[r];
      lt[k] = v;
      dc -= w[r] * v;
    }
    d[c] = dc;
    for (int k = A.lt_colind[c]; k < A.lt_colind[c + 1]; ++k) w[A.lt_row[k]] = 0;
  }
  int zero_pivots = 0;
  // zero pivots, or a breakdown that already turned D non-finite
  for (int c = 0; c < n; ++c) zero_pivots += !(d[c] != 0 && isfinite(d[c]));
  if (active && A.status) A.status[j] = zero_pivots;
  }
  if (MODE == LS_FACTOR) return;
  // solve: x = P' L'^-1 D^-1 L^-1 P b for every right-hand side
  for (int q = 0; q < A.nrhs; ++q) {
    double* x = A.x + jj * A.x_t + static_cast<int64_t>(q) * n * A.x_s;
    for (int i = 0; i < n; ++i) w[i] = x[A.p[i] * A.x_s];
    for (int c = 0; c < n; ++c) w[c] = sub_dot(w[c], lt, w, A.lt_row, A.lt_colind[c], A.lt_colind[c + 1]);
    for (int i = 0; i < n; ++i) w[i] = w[i] / d[i];
    for (int c = n - 1; c >= 0; --c) {
      const double wc = w[c];
      sub_axpy(wc, lt, w, A.lt_row, A.lt_colind[c], A.lt_colind[c + 1]);
    }
    if (active)
      for (int i = 0; i < n; ++i) x[A.p[i] * A.x_s] = w[i];
  }
}

struct QrArgs {
  int ncol, nrow_ext;
  const int* a_colind;
  const int* a_row;
  const int* v_colind;
  const int* v_row;
  const int* r_colind;
  const int* r_row;
  const int* prinv;
  const int* pc;
  const double* a;
  int64_t a_s, a_t;
  double* x;
  int64_t x_s, x_t;
  int nrhs, tr;
  double eps;
  double* v;     // [nnz_v][lanes]
  double* r;     // [nnz_r][lanes]
  double* beta;  // [ncol][lanes]
  double* wg;    // [max(nrow_ext, ncol)][lanes]
  int64_t lanes;
  int* status;
  int64_t n_inst;
};

__device__ double qr_house(Col v, int64_t off, double* beta, int nv) {
  const double v0 = v[off];
  double sigma = 0;
  for (int i = 1; i < nv; ++i) sigma += v[off + i] * v[off + i];
  const double s = sqrt(v0 * v0 + sigma);
  const double sigma_is_zero = sigma == 0, v0_nonpos = v0 <= 0;
  const double v0n = sigma_is_zero != 0 ? 1 : (v0_nonpos != 0 ? v0 - s : -sigma / (v0 + s));
  v[off] = v0n;
  *beta = sigma_is_zero != 0 ? 2 * v0_nonpos : -1 / (s * v0n);
  return s;
}

// MODE as ldl_kernel: V, R and beta stay in A.v / A.r / A.beta.
template <bool WLDS, int MODE>
__global__ void __launch_bounds__(64) qr_kernel(QrArgs A) {
  extern __shared__ double lds[];
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool active = j < A.n_inst;
  const int64_t jj = active ? j : A.n_inst - 1;
  Col v{A.v + j, A.lanes}, r{A.r + j, A.lanes}, beta{A.beta + j, A.lanes};
  Col x = WLDS ? Col{lds + threadIdx.x, static_cast<int64_t>(blockDim.x)} : Col{A.wg + j, A.lanes};
  const int ncol = A.ncol, nrow = A.nrow_ext;
  // the factor leaves x all zero (every entry it fills is cleared again):
  // the state the solve starts from
  for (int i = 0; i < nrow; ++i) x[i] = 0;
  if (MODE != LS_SOLVE) {
  const double* a = A.a + jj * A.a_t;
  int nz_r = 0;
  for (int c = 0; c < ncol; ++c) {
    const int pcc = A.pc[c];
    for (int k = A.a_colind[pcc]; k < A.a_colind[pcc + 1]; ++k) x[A.prinv[A.a_row[k]]] = a[k * A.a_s];
    int rr;
    for (int k = A.r_colind[c]; k < A.r_colind[c + 1] && (rr = A.r_row[k]) < c; ++k) {
      double alpha = add_dot(0.0, v, x, A.v_row, A.v_colind[rr], A.v_colind[rr + 1]);
      alpha *= beta[rr];
      sub_axpy(alpha, v, x, A.v_row, A.v_colind[rr], A.v_colind[rr + 1]);
      r[nz_r++] = x[rr];
      x[rr] = 0;
    }
    for (int k = A.v_colind[c]; k < A.v_colind[c + 1]; ++k) {
      v[k] = x[A.v_row[k]];
      x[A.v_row[k]] = 0;
    }
    double b;
    r[nz_r++] = qr_house(v, A.v_colind[c], &b, A.v_colind[c + 1] - A.v_colind[c]);
    beta[c] = b;
  }
  // casadi_qr_singular: number of |R_cc| < eps
  int nullity = 0;
  for (int c = 0; c < ncol; ++c) nullity += !(fabs(r[A.r_colind[c + 1] - 1]) >= A.eps);
  if (active && A.status) A.status[j] = nullity;
  }
  if (MODE == LS_FACTOR) return;
  // casadi_qr_solve
  Col w = x;
  for (int q = 0; q < A.nrhs; ++q) {
    double* xs = A.x + jj * A.x_t + static_cast<int64_t>(q) * ncol * A.x_s;
    if (A.tr) {
      for (int c = 0; c < ncol; ++c) w[c] = xs[A.pc[c] * A.x_s];
      for (int c = 0; c < ncol; ++c) {
        for (int k = A.r_colind[c]; k < A.r_colind[c + 1]; ++k) {
          const int rr = A.r_row[k];
          if (rr == c) w[c] /= r[k];
          else w[c] -= r[k] * w[rr];
        }
      }
      for (int c1 = 0; c1 < ncol; ++c1) {
        const int c = ncol - 1 - c1;
        double alpha = add_dot(0.0, v, w, A.v_row, A.v_colind[c], A.v_colind[c + 1]);
        alpha *= beta[c];
        sub_axpy(alpha, v, w, A.v_row, A.v_colind[c], A.v_colind[c + 1]);
      }
      if (active)
        for (int c = 0; c < ncol; ++c) xs[c * A.x_s] = w[A.prinv[c]];
    } else {
      for (int c = 0; c < nrow; ++c) w[c] = 0;
      for (int c = 0; c < ncol; ++c) w[A.prinv[c]] = xs[c * A.x_s];
      for (int c = 0; c < ncol; ++c) {
        double alpha = add_dot(0.0, v, w, A.v_row, A.v_colind[c], A.v_colind[c + 1]);
        alpha *= beta[c];
        sub_axpy(alpha, v, w, A.v_row, A.v_colind[c], A.v_colind[c + 1]);
      }
      for (int c = ncol - 1; c >= 0; --c) {
        for (int k = A.r_colind[c + 1] - 1; k >= A.r_colind[c]; --k) {
          const int rr = A.r_row[k];
          if (rr == c) w[rr] /= r[k];
          else w[rr] -= r[k] * w[c];
        }
      }
      if (active)
        for (int c = 0; c < ncol; ++c) xs[A.pc[c] * A.x_s] = w[c];
    }
  }
}

// Per instance: the number of D entries that are negative (what = 0, the
// inertia LinsolLdl::neig reports) or nonzero (what = 1, LinsolLdl::rank),
// d laid out [row][ld].
__global__ void ldl_count_kernel(const double* __restrict__ d, int n, int64_t ld, int64_t n_inst, int what,
                                 int32_t* __restrict__ out) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n_inst) return;
  int c = 0;
  for (int i = 0; i < n; ++i) {
    const double v = d[i * ld + k];
    c += what == 0 ? (v < 0) : (v != 0);
  }
  out[k] = c;
}

// Per instance: casadi_det (casadi/core/runtime/casadi_det.hpp:26-45) of the
// QR factors, times the sign of the row and column permutations
// (LinsolQr::det, casadi/solvers/linsol_qr.cpp).  V, R, beta laid out
// [nonzero][ld]; the products run in the reference's order.
__global__ void qr_det_kernel(const double* __restrict__ v, const double* __restrict__ r,
                              const double* __restrict__ beta, const int* __restrict__ v_colind,
                              const int* __restrict__ r_colind, int ncol, int64_t ld, int64_t n_inst, double sign,
                              double* __restrict__ det) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n_inst) return;
  double dt = 1;
  for (int c = 0; c < ncol; ++c) dt *= r[static_cast<int64_t>(r_colind[c + 1] - 1) * ld + k];
  for (int c = 0; c < ncol; ++c) {
    double vtv = 0;
    for (int q = v_colind[c]; q < v_colind[c + 1]; ++q) {
      const double x = v[static_cast<int64_t>(q) * ld + k];
      vtv += x * x;
    }
    dt *= 1 - beta[static_cast<int64_t>(c) * ld + k] * vtv;
  }
  det[k] = sign * dt;
}

// flag |= 1 when any a[i] != b[i] (numeric comparison, as casadi_cache_check:
// a NaN never matches)
__global__ void differ_kernel(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                              int* __restrict__ flag) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  bool differ = false;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    differ |= a[i] != b[i];
  if (__any(differ) && threadIdx.x % 64 == 0) atomicOr(flag, 1);
}

// CasADi compact sparsity -> (n, colind, row) as int32
struct SpI {
  int nrow = 0, ncol = 0;
  std::vector<int> colind, row;
};
SpI parse_sp(const int64_t* sp) {
  SpI s;
  if (sp[0] < 0 || sp[1] < 0 || sp[0] > (1 << 28) || sp[1] > (1 << 28)) throw std::runtime_error("bad sparsity");
  s.nrow = static_cast<int>(sp[0]);
  s.ncol = static_cast<int>(sp[1]);
  const int64_t* ci = sp + 2;
  if (ci[0] != 0) throw std::runtime_error("sparsity colind must start at 0");
  for (int c = 0; c < s.ncol; ++c)
    if (ci[c + 1] < ci[c]) throw std::runtime_error("sparsity colind not monotone");
  const int64_t nnz = ci[s.ncol];
  if (nnz > (int64_t(1) << 30)) throw std::runtime_error("sparsity too large");
  s.colind.assign(ci, ci + s.ncol + 1);
  s.row.resize(nnz);
  for (int64_t k = 0; k < nnz; ++k) {
    int64_t r = sp[3 + s.ncol + k];
    if (r < 0 || r >= s.nrow) throw std::runtime_error("sparsity row index out of range");
    s.row[k] = static_cast<int>(r);
  }
  return s;
}

// (-1)^(n - number of cycles): the determinant of a permutation matrix
double permutation_sign(const std::vector<int>& p) {
  std::vector<char> seen(p.size(), 0);
  int64_t cycles = 0;
  for (size_t i = 0; i < p.size(); ++i) {
    if (seen[i]) continue;
    ++cycles;
    for (size_t j = i; !seen[j]; j = static_cast<size_t>(p[j])) seen[j] = 1;
  }
  return (static_cast<int64_t>(p.size()) - cycles) % 2 ? -1.0 : 1.0;
}

std::vector<int> to_perm(const int64_t* p, int n, int range) {
  std::vector<int> v(n);
  for (int i = 0; i < n; ++i) {
    if (p[i] < 0 || p[i] >= range) throw std::runtime_error("permutation index out of range");
    v[i] = static_cast<int>(p[i]);
  }
  return v;
}

}  // namespace

// The factors of the last nfact stay in the handle's device memory, laid out
// [nonzero][ld] with ld = n_inst rounded up to 64 (both engines write and read
// this layout), until the next nfact or factor + solve.
struct cxh_ldl_s {
  SpI a, lt;
  std::vector<int> p;
  DeviceBuffer d_a_colind, d_a_row, d_lt_colind, d_lt_row, d_p;
  DeviceBuffer lt_vals, d_vals, w;
  bool uploaded = false;                               // pattern arrays on the device
  int engine = CXH_ENGINE_JIT;                          // specialised program or the generic kernel
  std::map<int, std::shared_ptr<sxh::LinsolJit>> jit;  // specialised factor + solve per nrhs
  std::shared_ptr<sxh::LinsolJit> jit_fact;            // specialised factor alone
  std::shared_ptr<sxh::LinsolJit> jit_solve;  // specialised solve alone, one right-hand side
  int64_t fact_n = -1;                                 // instances factored by the last nfact (-1: none)
};

struct cxh_qr_s {
  SpI a, v, r;
  std::vector<int> prinv, pc;
  double eps = 1e-12;
  double pivot_sign = 1;  // sign of the row and column permutations
  DeviceBuffer d_a_colind, d_a_row, d_v_colind, d_v_row, d_r_colind, d_r_row, d_prinv, d_pc;
  DeviceBuffer v_vals, r_vals, beta, w;
  bool uploaded = false;                               // pattern arrays on the device
  int engine = CXH_ENGINE_JIT;                         // specialised program or the generic kernel
  std::map<int, std::shared_ptr<sxh::LinsolJit>> jit;  // specialised factor + solve per (nrhs, tr)
  std::shared_ptr<sxh::LinsolJit> jit_fact;            // specialised factor alone
  std::shared_ptr<sxh::LinsolJit> jit_solve[2];  // specialised solve alone, one right-hand side, per tr
  int64_t fact_n = -1;                                 // instances factored by the last nfact (-1: none)
  // factorisations remembered (LinsolQr's "cache" option): each holds a copy
  // of the batch of A values it was computed from, and its V, R, beta
  struct CacheEntry {
    DeviceBuffer a, v, r, beta;
    int64_t n_inst = 0;
    int32_t layout = 0;
  };
  std::vector<CacheEntry> cache;
  std::vector<int> cache_loc;  // most recently used first; -1: never filled
  DeviceBuffer cmp_flag;
};

static sxh::LinsolJit& qr_program(cxh_qr_s* q, int nrhs, bool tr) {
  auto& J = q->jit[2 * nrhs + (tr ? 1 : 0)];
  if (!J)
    J = sxh::jit_qr_create(q->a.ncol, q->v.nrow, q->a.colind, q->a.row, q->v.colind, q->v.row, q->r.colind,
                           q->r.row, q->prinv, q->pc, q->eps, nrhs, tr);
  return *J;
}

static void upload_ints(DeviceBuffer& b, const std::vector<int>& v) {
  b.upload(v.data(), v.size() * sizeof(int));
}

extern "C" {

int cxh_ldl_create(const int64_t* sp_a, const int64_t* sp_lt, const int64_t* p, cxh_ldl* out) {
  CXH_TRY
  if (!sp_a || !sp_lt || !p || !out) return cxh_fail("cxh_ldl_create: null argument");
  auto l = std::make_unique<cxh_ldl_s>();
  l->a = parse_sp(sp_a);
  l->lt = parse_sp(sp_lt);
  const int n = l->lt.ncol;
  if (l->a.nrow != n || l->a.ncol != n || l->lt.nrow != n)
    return cxh_fail("cxh_ldl_create: A and L^T must be square of the same size");
  for (int c = 0; c < n; ++c)
    for (int k = l->lt.colind[c]; k < l->lt.colind[c + 1]; ++k)
      if (l->lt.row[k] >= c) return cxh_fail("cxh_ldl_create: L^T must be strictly upper triangular");
  l->p = to_perm(p, n, n);
  *out = l.release();
  return 0;
  CXH_CATCH
}

int cxh_ldl_destroy(cxh_ldl l) {
  delete l;
  return 0;
}

// The generic kernel in one of its modes (LS_FUSED, LS_FACTOR, LS_SOLVE).
static void ldl_generic(cxh_ldl_s* l, int mode, int64_t n_inst, const double* a, double* x, int32_t nrhs,
                        int32_t* status, int32_t layout, hipStream_t s) {
  if (!l->uploaded) {  // the generic kernel reads the pattern from device memory
    upload_ints(l->d_a_colind, l->a.colind);
    upload_ints(l->d_a_row, l->a.row);
    upload_ints(l->d_lt_colind, l->lt.colind);
    upload_ints(l->d_lt_row, l->lt.row);
    upload_ints(l->d_p, l->p);
    l->uploaded = true;
  }
  const int n = l->lt.ncol;
  const int64_t nnz_a = l->a.colind.back(), nnz_lt = l->lt.colind.back();
  const int bs = 64;
  const int64_t grid = (n_inst + bs - 1) / bs;
  LdlArgs A{};
  A.n = n;
  A.a_colind = l->d_a_colind.as<int>();
  A.a_row = l->d_a_row.as<int>();
  A.lt_colind = l->d_lt_colind.as<int>();
  A.lt_row = l->d_lt_row.as<int>();
  A.p = l->d_p.as<int>();
  A.a = a;
  A.x = x;
  A.nrhs = nrhs;
  if (layout == CXH_SOA) {
    A.a_s = n_inst; A.a_t = 1; A.x_s = n_inst; A.x_t = 1;
  } else {
    A.a_s = 1; A.a_t = nnz_a; A.x_s = 1; A.x_t = static_cast<int64_t>(n) * nrhs;
  }
  A.lanes = grid * bs;  // = n_inst rounded up to 64
  if (mode != LS_SOLVE) {
    l->lt_vals.reserve(std::max<size_t>(nnz_lt, 1) * A.lanes * sizeof(double));
    l->d_vals.reserve(std::max<size_t>(n, 1) * A.lanes * sizeof(double));
  }
  A.lt = l->lt_vals.as<double>();
  A.d = l->d_vals.as<double>();
  A.status = status;
  A.n_inst = n_inst;
  const size_t lds = static_cast<size_t>(std::max(n, 1)) * bs * sizeof(double);
  const bool in_lds = lds <= static_cast<size_t>(device_lds_bytes());
  if (!in_lds) {
    l->w.reserve(std::max<size_t>(n, 1) * A.lanes * sizeof(double));
    A.wg = l->w.as<double>();
  }
  auto go = [&](auto kernel) {
    if (in_lds) allow_lds(reinterpret_cast<const void*>(kernel), lds);
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(bs), in_lds ? lds : 0, s, A);
  };
  if (in_lds) {
    if (mode == LS_FUSED) go(ldl_kernel<true, LS_FUSED>);
    else if (mode == LS_FACTOR) go(ldl_kernel<true, LS_FACTOR>);
    else go(ldl_kernel<true, LS_SOLVE>);
  } else {
    if (mode == LS_FUSED) go(ldl_kernel<false, LS_FUSED>);
    else if (mode == LS_FACTOR) go(ldl_kernel<false, LS_FACTOR>);
    else go(ldl_kernel<false, LS_SOLVE>);
  }
  hip_check(hipGetLastError(), "ldl_kernel launch");
}

static int ldl_check(cxh_ldl_s* l, int64_t n_inst, int32_t nrhs, int32_t layout) {
  if (!l) return cxh_fail("null ldl handle");
  if (n_inst < 0 || nrhs < 0) return cxh_fail("negative size");
  if (layout != CXH_SOA && layout != CXH_AOS) return cxh_fail("bad layout");
  return 0;
}

int cxh_ldl_factor_solve(cxh_ldl l, int64_t n_inst, const double* a, double* x, int32_t nrhs, int32_t* status,
                         int32_t layout, void* stream) {
  CXH_TRY
  if (int rc = ldl_check(l, n_inst, nrhs, layout)) return rc;
  if (n_inst == 0) return 0;
  if (!a || (!x && nrhs > 0)) return cxh_fail("null data pointer");
  const int n = l->lt.ncol;
  const int64_t nnz_a = l->a.colind.back();
  l->fact_n = -1;  // the fused call keeps no factors
  if (l->engine == CXH_ENGINE_JIT) {
    auto& J = l->jit[nrhs];
    if (!J) J = jit_ldl_create(n, l->a.colind, l->a.row, l->lt.colind, l->lt.row, l->p, nrhs);
    const bool soa = layout == CXH_SOA;
    return jit_linsol_run(*J, n_inst, a, soa ? n_inst : 1, soa ? 1 : nnz_a, x, soa ? n_inst : 1,
                       soa ? 1 : static_cast<int64_t>(n) * nrhs, status, static_cast<hipStream_t>(stream));
  }
  ldl_generic(l, LS_FUSED, n_inst, a, x, nrhs, status, layout, static_cast<hipStream_t>(stream));
  return 0;
  CXH_CATCH
}

int cxh_ldl_nfact(cxh_ldl l, int64_t n_inst, const double* a, int32_t* status, int32_t layout, void* stream) {
  CXH_TRY
  if (int rc = ldl_check(l, n_inst, 0, layout)) return rc;
  l->fact_n = -1;
  if (n_inst == 0) {
    l->fact_n = 0;
    return 0;
  }
  if (!a) return cxh_fail("null data pointer");
  const int n = l->lt.ncol;
  const int64_t nnz_a = l->a.colind.back(), nnz_lt = l->lt.colind.back();
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (l->engine == CXH_ENGINE_JIT) {
    if (!l->jit_fact) l->jit_fact = jit_ldl_create(n, l->a.colind, l->a.row, l->lt.colind, l->lt.row, l->p, 0,
                                                   LS_FACTOR);
    const int64_t ld = round_up(n_inst, 64);
    l->lt_vals.reserve(std::max<size_t>(nnz_lt, 1) * ld * sizeof(double));
    l->d_vals.reserve(std::max<size_t>(n, 1) * ld * sizeof(double));
    double* fac[2] = {l->lt_vals.as<double>(), l->d_vals.as<double>()};
    const bool soa = layout == CXH_SOA;
    if (int rc = jit_linsol_factor(*l->jit_fact, n_inst, a, soa ? n_inst : 1, soa ? 1 : nnz_a, fac, 2, ld, status, s))
      return rc;
  } else {
    ldl_generic(l, LS_FACTOR, n_inst, a, nullptr, 0, status, layout, s);
  }
  l->fact_n = n_inst;
  return 0;
  CXH_CATCH
}

int cxh_ldl_solve(cxh_ldl l, int64_t n_inst, double* x, int32_t nrhs, int32_t tr, int32_t layout, void* stream) {
  CXH_TRY
  (void)tr;  // A is symmetric: the transposed solve is the same (LinsolLdl::solve ignores tr)
  if (int rc = ldl_check(l, n_inst, nrhs, layout)) return rc;
  if (l->fact_n < 0) return cxh_fail("cxh_ldl_solve: the linear system has not been factorized (call cxh_ldl_nfact)");
  if (n_inst != l->fact_n)
    return cxh_fail("cxh_ldl_solve: " + std::to_string(n_inst) + " instances, the last nfact factored " +
                    std::to_string(l->fact_n));
  if (n_inst == 0 || nrhs == 0) return 0;
  if (!x) return cxh_fail("null data pointer");
  const int n = l->lt.ncol;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (l->engine == CXH_ENGINE_JIT) {
    auto& J = l->jit_solve;
    if (!J) J = jit_ldl_create(n, l->a.colind, l->a.row, l->lt.colind, l->lt.row, l->p, 1, LS_SOLVE);
    const int64_t ld = round_up(n_inst, 64);
    const double* fac[2] = {l->lt_vals.as<double>(), l->d_vals.as<double>()};
    const bool soa = layout == CXH_SOA;
    // the right-hand sides one after another: each is solved on its own, in
    // the order the fused program solves it, so the bits are the same
    for (int q = 0; q < nrhs; ++q)
      if (int rc = jit_linsol_solve(*J, n_inst, fac, 2, ld, x + static_cast<int64_t>(q) * n * (soa ? n_inst : 1),
                                    soa ? n_inst : 1, soa ? 1 : static_cast<int64_t>(n) * nrhs, s))
        return rc;
    return 0;
  }
  ldl_generic(l, LS_SOLVE, n_inst, nullptr, x, nrhs, nullptr, layout, s);
  return 0;
  CXH_CATCH
}

static int ldl_count(cxh_ldl l, int32_t* out, int what, void* stream) {
  CXH_TRY
  if (!l) return cxh_fail("null ldl handle");
  if (l->fact_n < 0) return cxh_fail("the linear system has not been factorized (call cxh_ldl_nfact)");
  if (l->fact_n == 0) return 0;
  if (!out) return cxh_fail("null output array");
  const int64_t ld = round_up(l->fact_n, 64);
  hipLaunchKernelGGL(ldl_count_kernel, dim3(static_cast<unsigned>((l->fact_n + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), l->d_vals.as<double>(), l->lt.ncol, ld, l->fact_n, what, out);
  hip_check(hipGetLastError(), "ldl_count_kernel launch");
  return 0;
  CXH_CATCH
}

int cxh_ldl_neig(cxh_ldl l, int32_t* neig, void* stream) { return ldl_count(l, neig, 0, stream); }
int cxh_ldl_rank(cxh_ldl l, int32_t* rank, void* stream) { return ldl_count(l, rank, 1, stream); }

int cxh_ldl_engine(cxh_ldl l, int32_t engine) {
  CXH_TRY
  if (!l) return cxh_fail("null ldl handle");
  if (engine != CXH_ENGINE_INTERP && engine != CXH_ENGINE_JIT) return cxh_fail("cxh_ldl_engine: unknown engine");
  l->engine = engine;
  return 0;
  CXH_CATCH
}

int cxh_ldl_jit_source(cxh_ldl l, int32_t nrhs, char* buf, int64_t cap, int64_t* len) {
  CXH_TRY
  if (!l) return cxh_fail("null ldl handle");
  if (nrhs < 0) return cxh_fail("negative nrhs");
  auto& J = l->jit[nrhs];
  if (!J) J = jit_ldl_create(l->lt.ncol, l->a.colind, l->a.row, l->lt.colind, l->lt.row, l->p, nrhs);
  const std::string src = jit_linsol_source(*J, 0);
  if (len) *len = static_cast<int64_t>(src.size());
  if (buf && cap > 0) {
    const size_t m = std::min<size_t>(src.size(), static_cast<size_t>(cap - 1));
    std::memcpy(buf, src.data(), m);
    buf[m] = '\0';
  }
  return 0;
  CXH_CATCH
}

int cxh_qr_engine(cxh_qr q, int32_t engine) {
  CXH_TRY
  if (!q) return cxh_fail("null qr handle");
  if (engine != CXH_ENGINE_INTERP && engine != CXH_ENGINE_JIT) return cxh_fail("cxh_qr_engine: unknown engine");
  q->engine = engine;
  return 0;
  CXH_CATCH
}

int cxh_qr_jit_source(cxh_qr q, int32_t nrhs, int32_t tr, int32_t segment, char* buf, int64_t cap, int64_t* len,
                      int32_t* n_segments) {
  CXH_TRY
  if (!q) return cxh_fail("null qr handle");
  if (nrhs < 0) return cxh_fail("negative nrhs");
  sxh::LinsolJit& J = qr_program(q, nrhs, tr != 0);
  if (n_segments) *n_segments = jit_linsol_segments(J);
  const std::string src = jit_linsol_source(J, segment);
  if (len) *len = static_cast<int64_t>(src.size());
  if (buf && cap > 0) {
    const size_t m = std::min<size_t>(src.size(), static_cast<size_t>(cap - 1));
    std::memcpy(buf, src.data(), m);
    buf[m] = '\0';
  }
  return 0;
  CXH_CATCH
}

int cxh_qr_create(const int64_t* sp_a, const int64_t* sp_v, const int64_t* sp_r, const int64_t* prinv,
                  const int64_t* pc, double eps, cxh_qr* out) {
  CXH_TRY
  if (!sp_a || !sp_v || !sp_r || !prinv || !pc || !out) return cxh_fail("cxh_qr_create: null argument");
  auto q = std::make_unique<cxh_qr_s>();
  q->a = parse_sp(sp_a);
  q->v = parse_sp(sp_v);
  q->r = parse_sp(sp_r);
  const int ncol = q->a.ncol, nrow_ext = q->v.nrow;
  if (q->v.ncol != ncol || q->r.ncol != ncol || nrow_ext < q->a.nrow || q->r.nrow != nrow_ext)
    return cxh_fail("cxh_qr_create: inconsistent V / R / A dimensions");
  if (q->a.nrow != ncol) return cxh_fail("cxh_qr_create: the batched solve needs a square A");
  for (int c = 0; c < ncol; ++c) {
    if (q->r.colind[c + 1] <= q->r.colind[c] || q->r.row[q->r.colind[c + 1] - 1] != c)
      return cxh_fail("cxh_qr_create: R columns must end with their diagonal entry");
    for (int k = q->v.colind[c]; k < q->v.colind[c + 1]; ++k)
      if (q->v.row[k] >= nrow_ext) return cxh_fail("cxh_qr_create: V row out of range");
  }
  q->prinv = to_perm(prinv, nrow_ext, nrow_ext);
  q->pc = to_perm(pc, ncol, ncol);
  q->eps = eps;
  q->pivot_sign = permutation_sign(q->prinv) * permutation_sign(q->pc);
  *out = q.release();
  return 0;
  CXH_CATCH
}

int cxh_qr_destroy(cxh_qr q) {
  delete q;
  return 0;
}

// The generic kernel in one of its modes (LS_FUSED, LS_FACTOR, LS_SOLVE).
static void qr_generic(cxh_qr_s* q, int mode, int64_t n_inst, const double* a, double* x, int32_t nrhs, int32_t tr,
                       int32_t* status, int32_t layout, hipStream_t s) {
  if (!q->uploaded) {  // the generic kernel reads the pattern from device memory
    upload_ints(q->d_a_colind, q->a.colind);
    upload_ints(q->d_a_row, q->a.row);
    upload_ints(q->d_v_colind, q->v.colind);
    upload_ints(q->d_v_row, q->v.row);
    upload_ints(q->d_r_colind, q->r.colind);
    upload_ints(q->d_r_row, q->r.row);
    upload_ints(q->d_prinv, q->prinv);
    upload_ints(q->d_pc, q->pc);
    q->uploaded = true;
  }
  const int ncol = q->a.ncol, nrow_ext = q->v.nrow;
  const int64_t nnz_a = q->a.colind.back();
  const int bs = 64;
  const int64_t grid = (n_inst + bs - 1) / bs;
  QrArgs A{};
  A.ncol = ncol;
  A.nrow_ext = nrow_ext;
  A.a_colind = q->d_a_colind.as<int>();
  A.a_row = q->d_a_row.as<int>();
  A.v_colind = q->d_v_colind.as<int>();
  A.v_row = q->d_v_row.as<int>();
  A.r_colind = q->d_r_colind.as<int>();
  A.r_row = q->d_r_row.as<int>();
  A.prinv = q->d_prinv.as<int>();
  A.pc = q->d_pc.as<int>();
  A.a = a;
  A.x = x;
  A.nrhs = nrhs;
  A.tr = tr != 0;
  A.eps = q->eps;
  if (layout == CXH_SOA) {
    A.a_s = n_inst; A.a_t = 1; A.x_s = n_inst; A.x_t = 1;
  } else {
    A.a_s = 1; A.a_t = nnz_a; A.x_s = 1; A.x_t = static_cast<int64_t>(ncol) * nrhs;
  }
  A.lanes = grid * bs;  // = n_inst rounded up to 64
  if (mode != LS_SOLVE) {
    q->v_vals.reserve(std::max<size_t>(q->v.colind.back(), 1) * A.lanes * sizeof(double));
    q->r_vals.reserve(std::max<size_t>(q->r.colind.back(), 1) * A.lanes * sizeof(double));
    q->beta.reserve(std::max<size_t>(ncol, 1) * A.lanes * sizeof(double));
  }
  A.v = q->v_vals.as<double>();
  A.r = q->r_vals.as<double>();
  A.beta = q->beta.as<double>();
  A.status = status;
  A.n_inst = n_inst;
  const int wlen = std::max(nrow_ext, ncol);
  const size_t lds = static_cast<size_t>(std::max(wlen, 1)) * bs * sizeof(double);
  const bool in_lds = lds <= static_cast<size_t>(device_lds_bytes());
  if (!in_lds) {
    q->w.reserve(std::max<size_t>(wlen, 1) * A.lanes * sizeof(double));
    A.wg = q->w.as<double>();
  }
  auto go = [&](auto kernel) {
    if (in_lds) allow_lds(reinterpret_cast<const void*>(kernel), lds);
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(bs), in_lds ? lds : 0, s, A);
  };
  if (in_lds) {
    if (mode == LS_FUSED) go(qr_kernel<true, LS_FUSED>);
    else if (mode == LS_FACTOR) go(qr_kernel<true, LS_FACTOR>);
    else go(qr_kernel<true, LS_SOLVE>);
  } else {
    if (mode == LS_FUSED) go(qr_kernel<false, LS_FUSED>);
    else if (mode == LS_FACTOR) go(qr_kernel<false, LS_FACTOR>);
    else go(qr_kernel<false, LS_SOLVE>);
  }
  hip_check(hipGetLastError(), "qr_kernel launch");
}

static int qr_check(cxh_qr_s* q, int64_t n_inst, int32_t nrhs, int32_t layout) {
  if (!q) return cxh_fail("null qr handle");
  if (n_inst < 0 || nrhs < 0) return cxh_fail("negative size");
  if (layout != CXH_SOA && layout != CXH_AOS) return cxh_fail("bad layout");
  return 0;
}

int cxh_qr_factor_solve(cxh_qr q, int64_t n_inst, const double* a, double* x, int32_t nrhs, int32_t tr,
                        int32_t* status, int32_t layout, void* stream) {
  CXH_TRY
  if (int rc = qr_check(q, n_inst, nrhs, layout)) return rc;
  if (n_inst == 0) return 0;
  if (!a || (!x && nrhs > 0)) return cxh_fail("null data pointer");
  const int ncol = q->a.ncol;
  const int64_t nnz_a = q->a.colind.back();
  q->fact_n = -1;  // the fused call keeps no factors
  if (q->engine == CXH_ENGINE_JIT) {
    const bool soa = layout == CXH_SOA;
    return jit_linsol_run(qr_program(q, nrhs, tr != 0), n_inst, a, soa ? n_inst : 1, soa ? 1 : nnz_a, x,
                          soa ? n_inst : 1, soa ? 1 : static_cast<int64_t>(ncol) * nrhs, status,
                          static_cast<hipStream_t>(stream));
  }
  qr_generic(q, LS_FUSED, n_inst, a, x, nrhs, tr, status, layout, static_cast<hipStream_t>(stream));
  return 0;
  CXH_CATCH
}

// casadi_cache_check (casadi/core/runtime/casadi_cache.hpp) over whole
// batches: the entries are walked most recently used first; a never-filled
// entry ends the walk (a miss that stores there), an entry whose batch of A
// values equals a (same count, same layout, every value ==) is a hit and moves
// to the front.  A miss with every entry filled stores in the last entry
// walked, and the order is left as it is, as the reference does.
static cxh_qr_s::CacheEntry* qr_cache_check(cxh_qr_s* q, int64_t n_inst, const double* a, int32_t layout,
                                            hipStream_t s, bool* hit) {
  *hit = false;
  const int sz = static_cast<int>(q->cache.size());
  if (sz == 0) return nullptr;
  const int64_t na = q->a.colind.back() * n_inst;
  q->cmp_flag.reserve(sizeof(int));
  int* flag = q->cmp_flag.as<int>();
  cxh_qr_s::CacheEntry* e = nullptr;
  for (int i = 0; i < sz; ++i) {
    if (q->cache_loc[i] < 0) {
      q->cache_loc[i] = i;
      return &q->cache[i];
    }
    e = &q->cache[q->cache_loc[i]];
    bool match = e->n_inst == n_inst && e->layout == layout;
    if (match && na > 0) {
      int h = 0;
      hip_check(hipMemcpyAsync(flag, &h, sizeof h, hipMemcpyHostToDevice, s), "cache flag");
      const unsigned grid = static_cast<unsigned>(std::min<int64_t>((na + 255) / 256, 4096));
      hipLaunchKernelGGL(differ_kernel, dim3(grid), dim3(256), 0, s, a, e->a.as<double>(), na, flag);
      hip_check(hipGetLastError(), "differ_kernel launch");
      hip_check(hipMemcpyAsync(&h, flag, sizeof h, hipMemcpyDeviceToHost, s), "cache flag");
      hip_check(hipStreamSynchronize(s), "cache check");
      match = h == 0;
    }
    if (match) {
      const int c = q->cache_loc[i];
      for (int k = i; k > 0; --k) q->cache_loc[k] = q->cache_loc[k - 1];
      q->cache_loc[0] = c;
      *hit = true;
      return e;
    }
  }
  return e;
}

int cxh_qr_nfact(cxh_qr q, int64_t n_inst, const double* a, int32_t* status, int32_t layout, void* stream) {
  CXH_TRY
  if (int rc = qr_check(q, n_inst, 0, layout)) return rc;
  q->fact_n = -1;
  if (n_inst == 0) {
    q->fact_n = 0;
    return 0;
  }
  if (!a) return cxh_fail("null data pointer");
  const int ncol = q->a.ncol;
  const int64_t nnz_a = q->a.colind.back(), nnz_v = q->v.colind.back(), nnz_r = q->r.colind.back();
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t ld = round_up(n_inst, 64);
  const size_t bv = static_cast<size_t>(std::max<int64_t>(nnz_v, 1)) * ld * sizeof(double),
               br = static_cast<size_t>(std::max<int64_t>(nnz_r, 1)) * ld * sizeof(double),
               bb = static_cast<size_t>(std::max(ncol, 1)) * ld * sizeof(double);
  bool hit = false;
  cxh_qr_s::CacheEntry* e = qr_cache_check(q, n_inst, a, layout, s, &hit);
  if (hit) {  // the remembered factors; status is the nullity they were stored with (zero)
    q->v_vals.reserve(bv);
    q->r_vals.reserve(br);
    q->beta.reserve(bb);
    hip_check(hipMemcpyAsync(q->v_vals.as<double>(), e->v.as<double>(), bv, hipMemcpyDeviceToDevice, s), "cache");
    hip_check(hipMemcpyAsync(q->r_vals.as<double>(), e->r.as<double>(), br, hipMemcpyDeviceToDevice, s), "cache");
    hip_check(hipMemcpyAsync(q->beta.as<double>(), e->beta.as<double>(), bb, hipMemcpyDeviceToDevice, s), "cache");
    if (status) hip_check(hipMemsetAsync(status, 0, static_cast<size_t>(n_inst) * sizeof(int32_t), s), "status");
    q->fact_n = n_inst;
    return 0;
  }
  DeviceBuffer nullity;  // the cache stores only factorisations of nonsingular batches
  int32_t* st = status;
  if (e && !st) {
    nullity.reserve(static_cast<size_t>(n_inst) * sizeof(int32_t));
    st = nullity.as<int32_t>();
  }
  if (q->engine == CXH_ENGINE_JIT) {
    if (!q->jit_fact)
      q->jit_fact = jit_qr_create(q->a.ncol, q->v.nrow, q->a.colind, q->a.row, q->v.colind, q->v.row, q->r.colind,
                                  q->r.row, q->prinv, q->pc, q->eps, 0, false, LS_FACTOR);
    q->v_vals.reserve(bv);
    q->r_vals.reserve(br);
    q->beta.reserve(bb);
    double* fac[3] = {q->v_vals.as<double>(), q->r_vals.as<double>(), q->beta.as<double>()};
    const bool soa = layout == CXH_SOA;
    if (int rc = jit_linsol_factor(*q->jit_fact, n_inst, a, soa ? n_inst : 1, soa ? 1 : nnz_a, fac, 3, ld, st, s))
      return rc;
  } else {
    qr_generic(q, LS_FACTOR, n_inst, a, nullptr, 0, 0, st, layout, s);
  }
  q->fact_n = n_inst;
  if (e) {  // remember this factorisation unless an instance is singular (LinsolQr::nfact returns 1 first)
    std::vector<int32_t> h(static_cast<size_t>(n_inst));
    hip_check(hipMemcpyAsync(h.data(), st, h.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s), "status");
    hip_check(hipStreamSynchronize(s), "status");
    if (std::all_of(h.begin(), h.end(), [](int32_t v) { return v == 0; })) {
      e->a.reserve(static_cast<size_t>(std::max<int64_t>(nnz_a * n_inst, 1)) * sizeof(double));
      e->v.reserve(bv);
      e->r.reserve(br);
      e->beta.reserve(bb);
      hip_check(hipMemcpyAsync(e->a.as<double>(), a, static_cast<size_t>(nnz_a * n_inst) * sizeof(double),
                               hipMemcpyDeviceToDevice, s), "cache");
      hip_check(hipMemcpyAsync(e->v.as<double>(), q->v_vals.as<double>(), bv, hipMemcpyDeviceToDevice, s), "cache");
      hip_check(hipMemcpyAsync(e->r.as<double>(), q->r_vals.as<double>(), br, hipMemcpyDeviceToDevice, s), "cache");
      hip_check(hipMemcpyAsync(e->beta.as<double>(), q->beta.as<double>(), bb, hipMemcpyDeviceToDevice, s), "cache");
      e->n_inst = n_inst;
      e->layout = layout;
      hip_check(hipStreamSynchronize(s), "cache");
    } else {
      e->n_inst = -1;  // holds nothing that can match
    }
  }
  return 0;
  CXH_CATCH
}

int cxh_qr_solve(cxh_qr q, int64_t n_inst, double* x, int32_t nrhs, int32_t tr, int32_t layout, void* stream) {
  CXH_TRY
  if (int rc = qr_check(q, n_inst, nrhs, layout)) return rc;
  if (q->fact_n < 0) return cxh_fail("cxh_qr_solve: the linear system has not been factorized (call cxh_qr_nfact)");
  if (n_inst != q->fact_n)
    return cxh_fail("cxh_qr_solve: " + std::to_string(n_inst) + " instances, the last nfact factored " +
                    std::to_string(q->fact_n));
  if (n_inst == 0 || nrhs == 0) return 0;
  if (!x) return cxh_fail("null data pointer");
  const int ncol = q->a.ncol;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (q->engine == CXH_ENGINE_JIT) {
    auto& J = q->jit_solve[tr ? 1 : 0];
    if (!J)
      J = jit_qr_create(q->a.ncol, q->v.nrow, q->a.colind, q->a.row, q->v.colind, q->v.row, q->r.colind, q->r.row,
                        q->prinv, q->pc, q->eps, 1, tr != 0, LS_SOLVE);
    const int64_t ld = round_up(n_inst, 64);
    const double* fac[3] = {q->v_vals.as<double>(), q->r_vals.as<double>(), q->beta.as<double>()};
    const bool soa = layout == CXH_SOA;
    for (int k = 0; k < nrhs; ++k)  // one right-hand side at a time, as the fused program orders them
      if (int rc = jit_linsol_solve(*J, n_inst, fac, 3, ld, x + static_cast<int64_t>(k) * ncol * (soa ? n_inst : 1),
                                    soa ? n_inst : 1, soa ? 1 : static_cast<int64_t>(ncol) * nrhs, s))
        return rc;
    return 0;
  }
  qr_generic(q, LS_SOLVE, n_inst, nullptr, x, nrhs, tr, nullptr, layout, s);
  return 0;
  CXH_CATCH
}

int cxh_qr_det(cxh_qr q, double* det, void* stream) {
  CXH_TRY
  if (!q) return cxh_fail("null qr handle");
  if (q->fact_n < 0) return cxh_fail("cxh_qr_det: the linear system has not been factorized (call cxh_qr_nfact)");
  if (q->fact_n == 0) return 0;
  if (!det) return cxh_fail("null output array");
  if (!q->uploaded) {
    upload_ints(q->d_a_colind, q->a.colind);
    upload_ints(q->d_a_row, q->a.row);
    upload_ints(q->d_v_colind, q->v.colind);
    upload_ints(q->d_v_row, q->v.row);
    upload_ints(q->d_r_colind, q->r.colind);
    upload_ints(q->d_r_row, q->r.row);
    upload_ints(q->d_prinv, q->prinv);
    upload_ints(q->d_pc, q->pc);
    q->uploaded = true;
  }
  const int64_t ld = round_up(q->fact_n, 64);
  hipLaunchKernelGGL(qr_det_kernel, dim3(static_cast<unsigned>((q->fact_n + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), q->v_vals.as<double>(), q->r_vals.as<double>(),
                     q->beta.as<double>(), q->d_v_colind.as<int>(), q->d_r_colind.as<int>(), q->a.ncol, ld,
                     q->fact_n, q->pivot_sign, det);
  hip_check(hipGetLastError(), "qr_det_kernel launch");
  return 0;
  CXH_CATCH
}

int cxh_qr_cache(cxh_qr q, int32_t n_cache) {
  CXH_TRY
  if (!q) return cxh_fail("null qr handle");
  if (n_cache < 0) return cxh_fail("cxh_qr_cache: negative size");
  q->cache.clear();
  q->cache.resize(static_cast<size_t>(n_cache));
  q->cache_loc.assign(static_cast<size_t>(n_cache), -1);
  return 0;
  CXH_CATCH
}

}  // extern "C"
