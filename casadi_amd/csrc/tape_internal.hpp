// Handle types behind the cxh_* C ABI (host side, shared by the .hip units).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "sx_ops.h"

namespace sxh {

void hip_check(hipError_t e, const char* what);
int device_lds_bytes();
void allow_lds(const void* kernel, size_t bytes);
int64_t round_up(int64_t a, int64_t b);

// While a NoAlloc guard lives on the calling thread, every device allocation
// and kernel compile the library would make throws instead: the
// no-allocation entry points run the same code as their allocating
// counterparts under it.
struct NoAlloc {
  NoAlloc();
  ~NoAlloc();
  NoAlloc(const NoAlloc&) = delete;
  NoAlloc& operator=(const NoAlloc&) = delete;
};
bool no_alloc_active();

// Owning device allocation that only grows.
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : ptr(o.ptr), bytes(o.bytes) {
    o.ptr = nullptr;
    o.bytes = 0;
  }
  ~DeviceBuffer();
  void reserve(size_t nbytes);
  void upload(const void* src, size_t nbytes);
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

// One input or output block: nonzero j of instance k at p[j*s_nz + k*s_inst].
struct IOArg {
  const double* p;
  int64_t s_nz;
  int64_t s_inst;
};

// At most this many input / output blocks per tape; the tables travel by
// value in the kernel arguments.
#define CXH_MAX_IO 16
struct IOTable {
  IOArg a[CXH_MAX_IO];
};
IOTable make_io(const double* const* p, const std::vector<int64_t>& nnz, int32_t layout, int64_t ld);

void validate_tape(const std::vector<Instr>& a, int64_t ws, const std::vector<int64_t>& nnz_in,
                   const std::vector<int64_t>& nnz_out);

struct JitTape;
struct RollTape;
struct JitAsync;

}  // namespace sxh

struct cxh_tape_s {
  std::vector<sxh::Instr> alg;
  int64_t ws = 1;
  std::vector<int64_t> nnz_in, nnz_out;
  std::vector<int2> unused_in;
  sxh::DeviceBuffer d_tape, d_unused;  // uploaded at first interpreter launch
  sxh::DeviceBuffer d_live;            // adjoint_live bits, uploaded at the first interpreter adjoint
  bool uploaded = false;
  sxh::DeviceBuffer scratch, trail, ckpt;
  sxh::DeviceBuffer red, partial;  // summed blocks of a reduced map (mapsum.hip)
  sxh::DeviceBuffer spw;           // work words of a sparsity sweep too large for LDS (sparsity_kernels.hip)
  // Regularity check (cxh_tape_regularity_check): the caller's per-instance
  // status array and its capacity, and the flagged-instance counter.
  int32_t* status = nullptr;
  int64_t status_cap = 0;
  sxh::DeviceBuffer status_count;
  int block = 64;
  int w_lds = 1;
  int ipw = 64;  // instances per wavefront (8, 16, 32, 64)
  int ipt = 0;   // instances per thread of the specialised forward kernels (0: CXH_JIT_IPT, else 1)
  int engine = 0;  // CXH_ENGINE_INTERP or CXH_ENGINE_JIT
  std::shared_ptr<sxh::JitTape> jit;  // specialised kernels (jit.hip)
  std::shared_ptr<sxh::RollTape> roll;  // the tape rerolled into a loop, when it repeats (jit_roll.hip)
  std::shared_ptr<sxh::JitAsync> async;  // compiles on a worker thread (cxh_tape_jit_async), else null
  // Specialised kernels that would need scratch are refused; the calls they
  // would serve take the interpreter instead (same bits): evaluation, and
  // the adjoint at these checkpoint intervals (cxh_tape_jit_fallback).
  bool eval_interp = false;
  std::vector<int64_t> adj_interp;
  std::string fallback_why;
};

namespace sxh {
struct GridShape {
  int bs, ipw, cols;
  int64_t grid, slots;
};
GridShape grid_shape(const cxh_tape_s* t, int64_t n);
const Instr* device_tape(cxh_tape_s* t);

// Evaluation and fused forward + adjoint over explicit I/O tables (any
// per-block strides: s_inst = 0 broadcasts a block to every instance).
int map_eval_io(cxh_tape_s* t, int64_t n, const IOTable& arg, const IOTable& res, hipStream_t s);

// Asynchronous compile of the specialised kernels (jit.hip).  jit_ready
// says whether a call may take the specialised path for (every, fwd, rev)
// now: with cxh_tape_jit_async on, the first call that needs kernels not yet
// built starts a worker thread compiling them on a copy of the program and
// is told no (the caller runs the interpreter, which gives the same bits);
// later calls take them once the worker is done.  every = -1 names
// forward-only evaluation.  jit_settle waits for a running worker and takes
// over what it built: every other path that touches the specialised state
// calls it first.
bool jit_ready(cxh_tape_s* t, int64_t every, bool fwd, bool rev);
void jit_settle(cxh_tape_s* t);
// Turns asynchronous compile off (after settling) and says whether it was on;
// jit_async_set(t, on) puts it back.
bool jit_async_pause(cxh_tape_s* t);
void jit_async_set(cxh_tape_s* t, bool on);

// Per instruction (bit k of word k / 32), whether its adjoint is
// structurally nonzero in the reverse sweep CasADi builds symbolically
// (SXFunction::ad_reverse, casadi/core/sx_function.cpp:1454-1560, every output
// seeded): an OUTPUT marks its operand's slot, an operation whose adjoint is
// marked marks its operands through every partial that is not the constant
// zero (op_zero_partial), and the mark of a slot ends where its value is
// defined.  The reverse sweeps make no contribution from unmarked values.
std::vector<uint32_t> adjoint_live(const std::vector<Instr>& alg, int64_t ws);
int adj_io(cxh_tape_s* t, int mode, int64_t n, const IOTable& arg, const IOTable& aseed, const IOTable& res,
           const IOTable& asens, int64_t every, double* ckpt_ext, hipStream_t s);

// Straight-line specialised kernels (jit.hip).
int jit_map_eval(cxh_tape_s* t, int64_t n, const IOTable& arg, const IOTable& res, hipStream_t s);
int jit_adj(cxh_tape_s* t, int mode, int64_t n, const IOTable& arg, const IOTable& aseed, const IOTable& res,
            const IOTable& asens, int64_t every, double* ckpt_ext, hipStream_t s);
int64_t jit_ckpt_rows(cxh_tape_s* t, int64_t every);
void jit_prepare(cxh_tape_s* t, int64_t every, bool fwd, bool rev);

// The same over a tape whose middle repeats, as one loop body (jit_roll.hip).
bool roll_active(cxh_tape_s* t);
int64_t roll_ckpt_rows(cxh_tape_s* t, int64_t every);
void roll_prepare(cxh_tape_s* t, int64_t every, bool fwd, bool rev);
void roll_prepare_eval(cxh_tape_s* t);
// Takes over the loop kernels another tape of the same program built.
void roll_adopt(cxh_tape_s* t, cxh_tape_s& from);
int roll_map_eval(cxh_tape_s* t, int64_t n, const IOTable& arg, const IOTable& res, hipStream_t s);
int roll_adj(cxh_tape_s* t, int mode, int64_t n, const IOTable& arg, const IOTable& aseed, const IOTable& res,
             const IOTable& asens, int64_t every, double* ckpt_ext, hipStream_t s);
}  // namespace sxh

namespace sxh {
struct JitJac;
}

struct cxh_jac_s {
  cxh_tape_s* tape = nullptr;
  int ndir = 0;
  int64_t nnz_jac = 0;
  std::vector<int> in_off, out_off;
  std::vector<double> seedval;  // [flat_in_nz][ndir]
  std::vector<int2> scatter;    // [flat_out_nz][ndir] -> up to two Jacobian nonzeros
  sxh::DeviceBuffer d_seed, d_scatter, d_in_off, d_out_off;
  std::shared_ptr<sxh::JitJac> jit;  // specialised tangent kernels (jit.hip)
  int64_t seg_len = 0;               // instructions per kernel of the tangent sweep (0: default)
  int reg_budget = 0;                // starting register budget of its kernels (0: default)
  int prefetch = -1;                 // reload lookahead in instructions (-1: default)
  const char* kernel = "cxj_tan";    // name of its kernels (profiles tell the programs apart)
};

namespace sxh {
int jit_map_jac(cxh_jac_s* j, int64_t n, const IOTable& arg, const IOTable& res, double* jac, int64_t jac_s_nz,
                int64_t jac_s_inst, hipStream_t s);

// Batched LDL / QR factor + solve specialised to one sparsity pattern (jit.hip).
// part: the fused factor + solve program (inputs A, B; outputs X, count), the
// factor alone (input A; outputs the factor blocks -- L^T, D for LDL; V, R,
// beta for QR -- then the count) or the solve alone (inputs the factor
// blocks, then B; output X).
enum LinsolPart { LS_FUSED = 0, LS_FACTOR = 1, LS_SOLVE = 2 };
struct LinsolJit;
std::shared_ptr<LinsolJit> jit_ldl_create(int n, const std::vector<int>& a_colind, const std::vector<int>& a_row,
                                          const std::vector<int>& lt_colind, const std::vector<int>& lt_row,
                                          const std::vector<int>& p, int nrhs, int part = LS_FUSED);
std::shared_ptr<LinsolJit> jit_qr_create(int ncol, int nrow, const std::vector<int>& a_colind,
                                         const std::vector<int>& a_row, const std::vector<int>& v_colind,
                                         const std::vector<int>& v_row, const std::vector<int>& r_colind,
                                         const std::vector<int>& r_row, const std::vector<int>& prinv,
                                         const std::vector<int>& pc, double eps, int nrhs, bool tr,
                                         int part = LS_FUSED);
int jit_linsol_run(LinsolJit& L, int64_t n_inst, const double* a, int64_t a_s, int64_t a_t, double* x,
                   int64_t x_s, int64_t x_t, int32_t* status, hipStream_t s);
// The factor program: fac[b] receives factor block b ([rows][ld]); status the count.
int jit_linsol_factor(LinsolJit& L, int64_t n_inst, const double* a, int64_t a_s, int64_t a_t, double* const* fac,
                      int n_fac, int64_t ld, int32_t* status, hipStream_t s);
// The solve program: factor blocks fac[b] ([rows][ld]), x overwritten in place.
int jit_linsol_solve(LinsolJit& L, int64_t n_inst, const double* const* fac, int n_fac, int64_t ld, double* x,
                     int64_t x_s, int64_t x_t, hipStream_t s);
std::string jit_linsol_source(LinsolJit& L, int segment);
int jit_linsol_segments(LinsolJit& L);
}
