// Dependency (sparsity) propagation over a tape on the GPU: SXFunction::
// sp_forward and sp_reverse (casadi/core/sx_function.cpp:1657-1681 and
// 1733-1782) evaluated for many 64-bit seed words at once.
//
// A word is 64 seed directions (CasADi's bvec_t); lane k of the grid runs the
// whole tape for word k, so n_words words run side by side.  The rules are the
// reference's: a constant carries no dependency, an input slot takes the seed
// word of its nonzero, every unary or binary operation ORs its operands (a
// unary operation's second operand is its first), an output copies its slot;
// the reverse sweep ORs a slot's word into its operands and clears it, an
// output seeds its slot and is cleared, an input ORs its slot into the input
// word.  Inputs and outputs are SoA [nz][ld] words; a null input reads as zero,
// a null output is skipped.  Every input and output of a tape is
// differentiable (the boundary's tapes carry no is_diff flags).
//
// The work vector is one word per slot and lane: in LDS ([slot][lane]) when
// worksize * lanes * 8 bytes fits 63 KB (64 KB less the staged tape chunk)
// with at least one lane per block, else in device memory ([slot][ld_w],
// coalesced across lanes).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>

#include "../../include/casadi_hip.h"
#include "capi_common.hpp"
#include "tape_internal.hpp"

namespace sxh {
namespace {

constexpr int kChunk = 64;                           // tape instructions staged per step
constexpr int64_t kLdsBytes = 64 * 1024 - kChunk * 16;  // work vector budget beside the staged chunk

struct SpTable {
  uint64_t* p[CXH_MAX_IO];
};

SpTable make_sp(const uint64_t* const* p, size_t n) {
  if (n > CXH_MAX_IO) throw std::runtime_error("too many input/output blocks");
  SpTable t{};
  for (size_t i = 0; i < n; ++i) t.p[i] = p ? const_cast<uint64_t*>(p[i]) : nullptr;
  return t;
}

// One lane per seed word.  kLds: work vector in LDS, else at wg ([slot][ldw]).
// The block stages the tape in LDS kChunk instructions at a time (one
// coalesced load and two barriers per chunk) and reads each instruction from
// there one step ahead, so no instruction waits on device memory unless it
// reads an input word.  Every lane of the block runs the sweep (the barriers
// need them all); lanes past n_words touch only their own work column.
template <bool kLds, bool kRev>
__global__ void __launch_bounds__(64)
    sp_sweep_kernel(const Instr* __restrict__ tape, int64_t n_instr, SpTable arg, SpTable res, int64_t ld,
                    int64_t n_words, uint64_t* __restrict__ wg, int64_t ldw, int64_t ws) {
  extern __shared__ uint64_t sp_lds[];
  __shared__ int4 chunk[kChunk];
  const int nb = blockDim.x;
  const int64_t k = static_cast<int64_t>(blockIdx.x) * nb + threadIdx.x;
  const bool active = k < n_words;
  uint64_t* w = kLds ? sp_lds + threadIdx.x : wg + k;
  const int64_t sw = kLds ? nb : ldw;  // stride between slots
  const int4* __restrict__ code = reinterpret_cast<const int4*>(tape);
#define SPW(s) w[static_cast<int64_t>(s) * sw]
  if (kRev)
    for (int64_t s = 0; s < ws; ++s) SPW(s) = 0;
  const int64_t n_chunk = (n_instr + kChunk - 1) / kChunk;
  for (int64_t c = 0; c < n_chunk; ++c) {
    const int64_t t0 = (kRev ? n_chunk - 1 - c : c) * kChunk;
    const int cnt = static_cast<int>(n_instr - t0 < kChunk ? n_instr - t0 : kChunk);
    __syncthreads();  // the previous chunk has been read by every lane
    for (int j = threadIdx.x; j < cnt; j += nb) chunk[j] = code[t0 + j];
    __syncthreads();
    int4 next = chunk[kRev ? cnt - 1 : 0];
    for (int jj = 0; jj < cnt; ++jj) {
      const int4 e = next;
      if (jj + 1 < cnt) next = chunk[kRev ? cnt - 2 - jj : jj + 1];
      const int op = __builtin_amdgcn_readfirstlane(e.x);
      const int i0 = __builtin_amdgcn_readfirstlane(e.y);
      const int i1 = __builtin_amdgcn_readfirstlane(e.z);
      const int i2 = __builtin_amdgcn_readfirstlane(e.w);
      if (!kRev) {
        switch (op) {
          case OP_CONST:
          case OP_PARAMETER:
            SPW(i0) = 0;
            break;
          case OP_INPUT: {
            const uint64_t* a = arg.p[i1];
            SPW(i0) = (a && active) ? a[static_cast<int64_t>(i2) * ld + k] : 0;
            break;
          }
          case OP_OUTPUT: {
            uint64_t* r = res.p[i0];
            if (r && active) r[static_cast<int64_t>(i2) * ld + k] = SPW(i1);
            break;
          }
          default:
            SPW(i0) = SPW(i1) | SPW(i2);
        }
      } else {
        switch (op) {
          case OP_CONST:
          case OP_PARAMETER:
            SPW(i0) = 0;
            break;
          case OP_INPUT: {
            uint64_t* a = arg.p[i1];
            if (a && active) a[static_cast<int64_t>(i2) * ld + k] |= SPW(i0);
            SPW(i0) = 0;
            break;
          }
          case OP_OUTPUT: {
            uint64_t* r = res.p[i0];
            if (r && active) {
              uint64_t& x = r[static_cast<int64_t>(i2) * ld + k];
              SPW(i1) |= x;
              x = 0;
            }
            break;
          }
          default: {
            const uint64_t seed = SPW(i0);
            SPW(i0) = 0;
            SPW(i1) |= seed;
            SPW(i2) |= seed;
          }
        }
      }
    }
  }
#undef SPW
}

// Lanes per block with the work vector in LDS (0: it does not fit).
int lds_lanes(int64_t ws) {
  const int64_t lanes = kLdsBytes / (std::max<int64_t>(ws, 1) * 8);
  return static_cast<int>(std::min<int64_t>(64, lanes));
}

int sp_sweep(cxh_tape_s* t, bool rev, int64_t n_words, const uint64_t* const* arg, const uint64_t* const* res,
             int64_t ld, hipStream_t s) {
  if (n_words < 0) return cxh_fail("negative n_words");
  if (ld < n_words) return cxh_fail("ld < n_words");
  if (n_words == 0 || t->alg.empty()) return 0;
  const SpTable A = make_sp(arg, t->nnz_in.size());
  const SpTable R = make_sp(res, t->nnz_out.size());
  const Instr* tape = device_tape(t);
  const int64_t ni = static_cast<int64_t>(t->alg.size()), ws = t->ws;
  const int lanes = lds_lanes(ws);
  if (lanes > 0) {
    const unsigned grid = static_cast<unsigned>((n_words + lanes - 1) / lanes);
    const size_t lds = static_cast<size_t>(ws) * lanes * 8;
    if (rev)
      hipLaunchKernelGGL((sp_sweep_kernel<true, true>), dim3(grid), dim3(lanes), lds, s, tape, ni, A, R, ld, n_words,
                         nullptr, 0, ws);
    else
      hipLaunchKernelGGL((sp_sweep_kernel<true, false>), dim3(grid), dim3(lanes), lds, s, tape, ni, A, R, ld,
                         n_words, nullptr, 0, ws);
  } else {
    const int64_t ldw = round_up(n_words, 64);
    t->spw.reserve(static_cast<size_t>(ws * ldw) * sizeof(uint64_t));
    const unsigned grid = static_cast<unsigned>((n_words + 63) / 64);
    if (rev)
      hipLaunchKernelGGL((sp_sweep_kernel<false, true>), dim3(grid), dim3(64), 0, s, tape, ni, A, R, ld, n_words,
                         t->spw.as<uint64_t>(), ldw, ws);
    else
      hipLaunchKernelGGL((sp_sweep_kernel<false, false>), dim3(grid), dim3(64), 0, s, tape, ni, A, R, ld, n_words,
                         t->spw.as<uint64_t>(), ldw, ws);
  }
  hip_check(hipGetLastError(), "sp_sweep_kernel launch");
  return 0;
}

}  // namespace
}  // namespace sxh

using namespace sxh;

extern "C" {

int cxh_sp_forward(cxh_tape t, int64_t n_words, const uint64_t* const* arg, uint64_t* const* res, int64_t ld,
                   void* stream) {
  CXH_TRY
  if (!t) return cxh_fail("null tape");
  return sp_sweep(t, false, n_words, arg, res, ld, static_cast<hipStream_t>(stream));
  CXH_CATCH
}

int cxh_sp_reverse(cxh_tape t, int64_t n_words, uint64_t* const* arg, uint64_t* const* res, int64_t ld,
                   void* stream) {
  CXH_TRY
  if (!t) return cxh_fail("null tape");
  return sp_sweep(t, true, n_words, arg, res, ld, static_cast<hipStream_t>(stream));
  CXH_CATCH
}

}  // extern "C"
