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

template <bool WLDS>
__global__ void __launch_bounds__(64) ldl_kernel(LdlArgs A) {
  extern __shared__ double lds[];
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool active = j < A.n_inst;
  const int64_t jj = active ? j : A.n_inst - 1;
  Col lt{A.lt + j, A.lanes}, d{A.d + j, A.lanes};
  Col w = WLDS ? Col{lds + threadIdx.x, static_cast<int64_t>(blockDim.x)} : Col{A.wg + j, A.lanes};
  const int n = A.n;
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

template <bool WLDS>
__global__ void __launch_bounds__(64) qr_kernel(QrArgs A) {
  extern __shared__ double lds[];
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool active = j < A.n_inst;
  const int64_t jj = active ? j : A.n_inst - 1;
  Col v{A.v + j, A.lanes}, r{A.r + j, A.lanes}, beta{A.beta + j, A.lanes};
  Col x = WLDS ? Col{lds + threadIdx.x, static_cast<int64_t>(blockDim.x)} : Col{A.wg + j, A.lanes};
  const double* a = A.a + jj * A.a_t;
  const int ncol = A.ncol, nrow = A.nrow_ext;
  for (int i = 0; i < nrow; ++i) x[i] = 0;
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

std::vector<int> to_perm(const int64_t* p, int n, int range) {
  std::vector<int> v(n);
  for (int i = 0; i < n; ++i) {
    if (p[i] < 0 || p[i] >= range) throw std::runtime_error("permutation index out of range");
    v[i] = static_cast<int>(p[i]);
  }
  return v;
}

}  // namespace

struct cxh_ldl_s {
  SpI a, lt;
  std::vector<int> p;
  DeviceBuffer d_a_colind, d_a_row, d_lt_colind, d_lt_row, d_p;
  DeviceBuffer lt_vals, d_vals, w;
  bool uploaded = false;                               // pattern arrays on the device
  int engine = CXH_ENGINE_JIT;                          // specialised program or the generic kernel
  std::map<int, std::shared_ptr<sxh::LinsolJit>> jit;  // specialised factor + solve per nrhs
};

struct cxh_qr_s {
  SpI a, v, r;
  std::vector<int> prinv, pc;
  double eps = 1e-12;
  DeviceBuffer d_a_colind, d_a_row, d_v_colind, d_v_row, d_r_colind, d_r_row, d_prinv, d_pc;
  DeviceBuffer v_vals, r_vals, beta, w;
  bool uploaded = false;                               // pattern arrays on the device
  int engine = CXH_ENGINE_JIT;                         // specialised program or the generic kernel
  std::map<int, std::shared_ptr<sxh::LinsolJit>> jit;  // specialised factor + solve per (nrhs, tr)
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

int cxh_ldl_factor_solve(cxh_ldl l, int64_t n_inst, const double* a, double* x, int32_t nrhs, int32_t* status,
                         int32_t layout, void* stream) {
  CXH_TRY
  if (!l) return cxh_fail("null ldl handle");
  if (n_inst < 0 || nrhs < 0) return cxh_fail("negative size");
  if (layout != CXH_SOA && layout != CXH_AOS) return cxh_fail("bad layout");
  if (n_inst == 0) return 0;
  if (!a || (!x && nrhs > 0)) return cxh_fail("null data pointer");
  const int n = l->lt.ncol;
  const int64_t nnz_a = l->a.colind.back(), nnz_lt = l->lt.colind.back();
  if (l->engine == CXH_ENGINE_JIT) {
    auto& J = l->jit[nrhs];
    if (!J) J = jit_ldl_create(n, l->a.colind, l->a.row, l->lt.colind, l->lt.row, l->p, nrhs);
    const bool soa = layout == CXH_SOA;
    return jit_linsol_run(*J, n_inst, a, soa ? n_inst : 1, soa ? 1 : nnz_a, x, soa ? n_inst : 1,
                       soa ? 1 : static_cast<int64_t>(n) * nrhs, status, static_cast<hipStream_t>(stream));
  }
  if (!l->uploaded) {  // the generic kernel reads the pattern from device memory
    upload_ints(l->d_a_colind, l->a.colind);
    upload_ints(l->d_a_row, l->a.row);
    upload_ints(l->d_lt_colind, l->lt.colind);
    upload_ints(l->d_lt_row, l->lt.row);
    upload_ints(l->d_p, l->p);
    l->uploaded = true;
  }
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
  A.lanes = grid * bs;
  l->lt_vals.reserve(std::max<size_t>(nnz_lt, 1) * A.lanes * sizeof(double));
  l->d_vals.reserve(std::max<size_t>(n, 1) * A.lanes * sizeof(double));
  A.lt = l->lt_vals.as<double>();
  A.d = l->d_vals.as<double>();
  A.status = status;
  A.n_inst = n_inst;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t lds = static_cast<size_t>(std::max(n, 1)) * bs * sizeof(double);
  if (lds <= static_cast<size_t>(device_lds_bytes())) {
    allow_lds(reinterpret_cast<const void*>(ldl_kernel<true>), lds);
    hipLaunchKernelGGL(ldl_kernel<true>, dim3(grid), dim3(bs), lds, s, A);
  } else {
    l->w.reserve(std::max<size_t>(n, 1) * A.lanes * sizeof(double));
    A.wg = l->w.as<double>();
    hipLaunchKernelGGL(ldl_kernel<false>, dim3(grid), dim3(bs), 0, s, A);
  }
  hip_check(hipGetLastError(), "ldl_kernel launch");
  return 0;
  CXH_CATCH
}

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
  *out = q.release();
  return 0;
  CXH_CATCH
}

int cxh_qr_destroy(cxh_qr q) {
  delete q;
  return 0;
}

int cxh_qr_factor_solve(cxh_qr q, int64_t n_inst, const double* a, double* x, int32_t nrhs, int32_t tr,
                        int32_t* status, int32_t layout, void* stream) {
  CXH_TRY
  if (!q) return cxh_fail("null qr handle");
  if (n_inst < 0 || nrhs < 0) return cxh_fail("negative size");
  if (layout != CXH_SOA && layout != CXH_AOS) return cxh_fail("bad layout");
  if (n_inst == 0) return 0;
  if (!a || (!x && nrhs > 0)) return cxh_fail("null data pointer");
  const int ncol = q->a.ncol, nrow_ext = q->v.nrow;
  const int64_t nnz_a = q->a.colind.back();
  if (q->engine == CXH_ENGINE_JIT) {
    const bool soa = layout == CXH_SOA;
    return jit_linsol_run(qr_program(q, nrhs, tr != 0), n_inst, a, soa ? n_inst : 1, soa ? 1 : nnz_a, x,
                          soa ? n_inst : 1, soa ? 1 : static_cast<int64_t>(ncol) * nrhs, status,
                          static_cast<hipStream_t>(stream));
  }
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
  A.lanes = grid * bs;
  q->v_vals.reserve(std::max<size_t>(q->v.colind.back(), 1) * A.lanes * sizeof(double));
  q->r_vals.reserve(std::max<size_t>(q->r.colind.back(), 1) * A.lanes * sizeof(double));
  q->beta.reserve(std::max<size_t>(ncol, 1) * A.lanes * sizeof(double));
  A.v = q->v_vals.as<double>();
  A.r = q->r_vals.as<double>();
  A.beta = q->beta.as<double>();
  A.status = status;
  A.n_inst = n_inst;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int wlen = std::max(nrow_ext, ncol);
  const size_t lds = static_cast<size_t>(std::max(wlen, 1)) * bs * sizeof(double);
  if (lds <= static_cast<size_t>(device_lds_bytes())) {
    allow_lds(reinterpret_cast<const void*>(qr_kernel<true>), lds);
    hipLaunchKernelGGL(qr_kernel<true>, dim3(grid), dim3(bs), lds, s, A);
  } else {
    q->w.reserve(std::max<size_t>(wlen, 1) * A.lanes * sizeof(double));
    A.wg = q->w.as<double>();
    hipLaunchKernelGGL(qr_kernel<false>, dim3(grid), dim3(bs), 0, s, A);
  }
  hip_check(hipGetLastError(), "qr_kernel launch");
  return 0;
  CXH_CATCH
}

}  // extern "C"
