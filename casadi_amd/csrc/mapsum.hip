// Reduced maps on the GPU: Function::map(name, parallelization, n, reduce_in,
// reduce_out) (casadi/core/function.cpp:796-817, a parallel map wrapped with
// repmat on the reduced inputs and repsum on the reduced outputs) and the
// serial MapSum::eval_gen it stands for (casadi/core/mapsum.cpp:155-186).
//
// A reduced input is one block shared by every instance: the kernels read it
// with instance stride 0, nothing is replicated.  A reduced output is the sum
// over the n instances of the instance outputs.  The instance outputs go to a
// transient SoA buffer [nz][ld] and two kernels sum its rows in a fixed order
// that depends on n only (per-thread strided sums, a butterfly over the 64
// lanes of a wavefront, the wavefront partials pairwise), so a result does not
// vary from run to run.  MapSum adds instance by instance; the two orders
// differ by rounding only (both are within (terms - 1) * eps * sum|term| of
// the exact sum).
//
// The fused forward + adjoint of a reduced map is the reverse of MapSum
// (MapSum::get_reverse, casadi/core/mapsum.cpp:354-415): the seed of a reduced
// output is the same for every instance (stride 0 again) and the sensitivity
// of a reduced input is the sum over instances of the instance sensitivities.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/casadi_hip.h"
#include "capi_common.hpp"
#include "tape_internal.hpp"

// Scratch of the no-allocation reduced map (cxh_map_eval_sum_noalloc): the
// summed instance outputs ([row][round_up(n, 64)]) and the first pass's
// partial sums.  64 MiB, held by every GPU the library runs on until the
// process ends.  Host code never uses the array's name as an address: the
// device address comes from hipGetSymbolAddress, per device.
constexpr int64_t kScratchDoubles = int64_t(8) << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace sxh {
namespace {

constexpr int kRedBlock = 256;        // threads per reduction block (4 wavefronts)

// Device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array at its own address).
double* scratch_ptr() {
  constexpr int kMaxDev = 64;
  static std::mutex mu;
  static double* ptr[kMaxDev] = {};
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDev) throw std::runtime_error("cxh_map_eval_sum_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw std::runtime_error("hipGetSymbolAddress(scratch) gave a null address");
    ptr[dev] = static_cast<double*>(p);
  }
  return ptr[dev];
}
constexpr int64_t kRedChunk = 2048;   // instances per first-pass block

// Sum over the block in a fixed order; the result is valid in thread 0.
__device__ double block_sum(double v, double* lds) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);  // a + b == b + a: all lanes agree
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  return (lds[0] + lds[1]) + (lds[2] + lds[3]);
}

// First pass: block (c, r) sums instances [c*chunk, min(n, (c+1)*chunk)) of
// row r of y ([row][ld]) into part[r*nchunk + c].
__global__ void __launch_bounds__(kRedBlock)
    sum_rows_pass1(const double* __restrict__ y, int64_t ld, int64_t n, int64_t chunk, double* __restrict__ part,
                   int64_t nchunk) {
  __shared__ double lds[kRedBlock / 64];
  const int64_t r = blockIdx.y, c = blockIdx.x;
  const int64_t k0 = c * chunk, k1 = k0 + chunk < n ? k0 + chunk : n;
  const double* __restrict__ p = y + r * ld;
  double v = 0.0;
  for (int64_t k = k0 + threadIdx.x; k < k1; k += kRedBlock) v += p[k];
  const double s = block_sum(v, lds);
  if (threadIdx.x == 0) part[r * nchunk + c] = s;
}

// Second pass: block r sums part[r*nchunk .. r*nchunk + nchunk) into out[r].
__global__ void __launch_bounds__(kRedBlock)
    sum_rows_pass2(const double* __restrict__ part, int64_t nchunk, double* __restrict__ out) {
  __shared__ double lds[kRedBlock / 64];
  const int64_t r = blockIdx.x;
  double v = 0.0;
  for (int64_t k = threadIdx.x; k < nchunk; k += kRedBlock) v += part[r * nchunk + k];
  const double s = block_sum(v, lds);
  if (threadIdx.x == 0) out[r] = s;
}

struct Reduced {
  double* out;    // caller's block (nnz doubles)
  int64_t row0;   // first row in the transient buffer
  int64_t rows;   // nnz of the block
};

// First-pass blocks per summed row: one per kRedChunk instances, at most 1024.
int64_t n_chunks(int64_t n) { return std::min<int64_t>(1024, std::max<int64_t>(1, (n + kRedChunk - 1) / kRedChunk)); }

// Sums rows [0, rows) of y ([row][ld], n instances) into the caller's blocks,
// the first pass's partial sums in part (rows * n_chunks(n) doubles).
void sum_rows(const double* y, int64_t ld, int64_t n, int64_t rows, const std::vector<Reduced>& blocks, double* part,
              hipStream_t s) {
  if (rows == 0) return;
  const int64_t nchunk = n_chunks(n);
  const int64_t chunk = std::max<int64_t>(1, (n + nchunk - 1) / nchunk);
  hipLaunchKernelGGL(sum_rows_pass1, dim3(static_cast<unsigned>(nchunk), static_cast<unsigned>(rows)),
                     dim3(kRedBlock), 0, s, y, ld, n, chunk, part, nchunk);
  hip_check(hipGetLastError(), "sum_rows_pass1 launch");
  for (const Reduced& b : blocks) {
    if (b.rows == 0) continue;
    hipLaunchKernelGGL(sum_rows_pass2, dim3(static_cast<unsigned>(b.rows)), dim3(kRedBlock), 0, s,
                       part + b.row0 * nchunk, nchunk, b.out);
    hip_check(hipGetLastError(), "sum_rows_pass2 launch");
  }
}

// Points the reduced blocks of tab at rows of the transient buffer
// (instance stride 1, nonzero stride ld); returns their rows.
int64_t redirect(IOTable& tab, double* const* user, const std::vector<int64_t>& nnz, const int32_t* flags,
                 double* base, int64_t ld, int64_t row0, std::vector<Reduced>& blocks) {
  int64_t r = row0;
  for (size_t j = 0; j < nnz.size(); ++j) {
    if (!flags || !flags[j] || !user || !user[j] || nnz[j] == 0) continue;
    tab.a[j].p = base ? base + r * ld : nullptr;
    tab.a[j].s_nz = ld;
    tab.a[j].s_inst = 1;
    blocks.push_back({user[j], r, nnz[j]});
    r += nnz[j];
  }
  return r - row0;
}

int64_t count_rows(double* const* user, const std::vector<int64_t>& nnz, const int32_t* flags) {
  int64_t r = 0;
  for (size_t j = 0; j < nnz.size(); ++j)
    if (flags && flags[j] && user && user[j]) r += nnz[j];
  return r;
}

// The first pass has one grid row per summed nonzero.
void check_rows(int64_t rows) {
  if (rows > 65535) throw std::runtime_error("reduced map: more than 65535 summed nonzeros");
}

void broadcast(IOTable& tab, const std::vector<int64_t>& nnz, const int32_t* flags) {
  for (size_t j = 0; j < nnz.size(); ++j)
    if (flags && flags[j]) {
      tab.a[j].s_nz = 1;
      tab.a[j].s_inst = 0;
    }
}

// Reduced map evaluation.  With no_alloc the summed instance outputs and the
// partial sums live in the library's static device array instead of the
// tape's own buffers, and nothing is allocated or compiled (NoAlloc).
int eval_sum(cxh_tape t, int64_t n, const double* const* arg, double* const* res, const int32_t* reduce_in,
             const int32_t* reduce_out, int64_t ld, hipStream_t s, bool no_alloc) {
  if (!t) return cxh_fail("null tape");
  if (n < 0) return cxh_fail("negative n");
  if (ld < n) return cxh_fail("ld < n");
  IOTable A = make_io(arg, t->nnz_in, CXH_SOA, ld);
  IOTable R = make_io(res, t->nnz_out, CXH_SOA, ld);
  broadcast(A, t->nnz_in, reduce_in);
  const int64_t ldr = round_up(std::max<int64_t>(n, 1), 64);
  const int64_t rows = count_rows(res, t->nnz_out, reduce_out);
  check_rows(rows);
  double* red = nullptr;
  double* part = nullptr;
  if (no_alloc) {
    const int64_t need = rows * (ldr + n_chunks(n));
    if (need > kScratchDoubles)
      return cxh_fail("cxh_map_eval_sum_noalloc: " + std::to_string(rows) + " summed nonzeros over " +
                      std::to_string(n) + " instances need " + std::to_string(need) +
                      " doubles of scratch; the static array holds " + std::to_string(kScratchDoubles));
    double* base = scratch_ptr();
    red = base;
    part = base + rows * ldr;
  } else if (rows > 0) {
    t->red.reserve(static_cast<size_t>(rows * ldr) * sizeof(double));
    t->partial.reserve(static_cast<size_t>(rows * n_chunks(n)) * sizeof(double));
    red = t->red.as<double>();
    part = t->partial.as<double>();
  }
  std::vector<Reduced> blocks;
  redirect(R, res, t->nnz_out, reduce_out, red, ldr, 0, blocks);
  std::unique_ptr<NoAlloc> guard(no_alloc ? new NoAlloc : nullptr);
  const int rc = map_eval_io(t, n, A, R, s);
  if (rc) return rc;
  sum_rows(red, ldr, n, rows, blocks, part, s);
  return 0;
}

}  // namespace
}  // namespace sxh

using namespace sxh;

extern "C" {

int cxh_map_eval_sum(cxh_tape t, int64_t n, const double* const* arg, double* const* res, const int32_t* reduce_in,
                     const int32_t* reduce_out, int64_t ld, void* stream) {
  CXH_TRY
  return eval_sum(t, n, arg, res, reduce_in, reduce_out, ld, static_cast<hipStream_t>(stream), false);
  CXH_CATCH
}

int cxh_map_eval_sum_noalloc(cxh_tape t, int64_t n, const double* const* arg, double* const* res,
                             const int32_t* reduce_in, const int32_t* reduce_out, int64_t ld, void* stream) {
  CXH_TRY
  return eval_sum(t, n, arg, res, reduce_in, reduce_out, ld, static_cast<hipStream_t>(stream), true);
  CXH_CATCH
}

int cxh_map_fwd_adj_sum(cxh_tape t, int64_t n, const double* const* arg, const double* const* adj_out,
                        double* const* res, double* const* adj_in, const int32_t* reduce_in,
                        const int32_t* reduce_out, int64_t ld, int64_t checkpoint_every, void* stream) {
  CXH_TRY
  if (!t) return cxh_fail("null tape");
  if (n < 0) return cxh_fail("negative n");
  if (ld < n) return cxh_fail("ld < n");
  hipStream_t s = static_cast<hipStream_t>(stream);
  IOTable A = make_io(arg, t->nnz_in, CXH_SOA, ld);
  IOTable S = make_io(adj_out, t->nnz_out, CXH_SOA, ld);
  IOTable R = make_io(res, t->nnz_out, CXH_SOA, ld);
  IOTable G = make_io(adj_in, t->nnz_in, CXH_SOA, ld);
  broadcast(A, t->nnz_in, reduce_in);
  broadcast(S, t->nnz_out, reduce_out);
  const int64_t ldr = round_up(std::max<int64_t>(n, 1), 64);
  const int64_t rows = count_rows(res, t->nnz_out, reduce_out) + count_rows(adj_in, t->nnz_in, reduce_in);
  check_rows(rows);
  if (rows > 0) t->red.reserve(static_cast<size_t>(rows * ldr) * sizeof(double));
  std::vector<Reduced> blocks;
  const int64_t r0 = redirect(R, res, t->nnz_out, reduce_out, t->red.as<double>(), ldr, 0, blocks);
  redirect(G, adj_in, t->nnz_in, reduce_in, t->red.as<double>(), ldr, r0, blocks);
  const int rc = adj_io(t, 0, n, A, S, R, G, checkpoint_every, nullptr, s);
  if (rc) return rc;
  if (rows > 0) t->partial.reserve(static_cast<size_t>(rows * n_chunks(n)) * sizeof(double));
  sum_rows(t->red.as<double>(), ldr, n, rows, blocks, t->partial.as<double>(), s);
  return 0;
  CXH_CATCH
}

}  // extern "C"
