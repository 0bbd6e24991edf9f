// Code generation and launch helpers shared by the specialiser's units
// (jit.hip: straight-line segments, tangent sweeps, mapaccum, batch;
// jit_roll.hip: flat tapes rerolled into loops).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cmath>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <utility>

#include "sx_ops.h"
#include "tape_internal.hpp"

namespace sxh {

// Kernel argument block; the generated source declares the same layout.
struct CxjArgs {
  IOTable arg, res, aseed, asens;
  const double* st_in;
  double* st_out;
  const double* adj_in;
  double* adj_out;
  long long n, lds;
  double* spill;
};
static_assert(sizeof(IOArg) == 24, "IOArg layout");
static_assert(sizeof(CxjArgs) == 4 * CXH_MAX_IO * 24 + 7 * 8, "CxjArgs layout");

// sx_ops.h as source text, and the declarations every generated program
// starts with (argument block, memory macros, register barriers, sincos).
extern const char kSxOpsSource[];
extern const char kPrologue[];

// Exact hexadecimal literal of a double.
std::string lit(double d);
// The finished program (sin/cos pairs switched to out-of-line calls above a
// number of sites).
std::string finish_source(std::string o);
// Whether reverse chains run as paired launches (CXH_JIT_SPLIT=0: not).
bool split_reverse();
// Whether reverse sweeps load the sin / cos pairs their forward pass stored
// (CXH_JIT_TRIG_CACHE, default on) instead of computing them again.
bool trig_cache_enabled();
// Operand codes whose sin and cos are both taken in the tape.
std::unordered_set<int64_t> step_trig_pairs(const cxh_tape_s* t);
// `const double v<k> = sin/cos(...)` through one sincos per operand.
bool emit_trig(std::string& o, const Instr& e, int64_t k, int64_t x, const std::string& xname, bool all_trig,
               const std::unordered_set<int64_t>& pairs,
               std::unordered_map<int64_t, std::pair<std::string, std::string>>& trig);

// Divisions by a tape constant in the loop kernels: whether they go
// through sxh_div_c (CXH_JIT_FAST_DIV, default on; the same bits either way),
// and the divisors it takes (finite, 2^-64 <= |c| <= 2^64).
bool fast_div_enabled();
inline bool fast_divisor(double c) { return std::isfinite(c) && std::fabs(c) >= 0x1p-64 && std::fabs(c) <= 0x1p+64; }

// hiprtc compile (cached by source) and load; the checked form refuses a
// kernel that needs scratch, naming the remedy.
hipFunction_t jit_function(const std::string& src, const char* name);
int scratch_bytes(hipFunction_t f);
hipFunction_t jit_function_checked(const std::string& src, const char* name, const char* remedy);
// One lane per instance, ceil(n / block) blocks; the pair form adds
// blockIdx.y = 1 with argument block B.
void launch(hipFunction_t f, int64_t n, int block, CxjArgs& A, hipStream_t s);
void launch_pair(hipFunction_t f, int64_t n, int block, CxjArgs& A, CxjArgs& B, hipStream_t s);

}  // namespace sxh
