// gfx950 kernels that evaluate CasADi SX tapes over batches of instances,
// and the C ABI entry points that launch them.
//
// One lane evaluates one instance.  Every lane of a wavefront walks the same
// instruction stream, so instruction fetch, the op dispatch and all slot
// indices are wave-uniform (scalar loads and branches); only the values differ
// per lane.  The work vector w[worksize] of every lane lives in LDS laid out
// [slot][lane] (a wave reads 64 consecutive doubles per slot: conflict-free),
// or, when it does not fit, in HBM scratch laid out [slot][global lane]
// (coalesced 512 B per wave access).  Inputs and outputs are per-nonzero
// structure-of-arrays [nz][instance] or CasADi's instance-major Map layout.
//
// Semantics restated:
//   forward  SXFunction::eval          casadi/core/sx_function.cpp:72-127
//   adjoint  SXFunction::ad_reverse    casadi/core/sx_function.cpp:1368-1569
//   tangent  SXFunction::ad_forward    casadi/core/sx_function.cpp:1194-1366
// evaluated numerically instead of symbolically.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/casadi_hip.h"
#include "capi_common.hpp"
#include "sx_ops.h"
#include "tape_internal.hpp"

using namespace sxh;

namespace sxh {

thread_local std::string g_last_error;
int cxh_fail(const std::string& msg) {
  g_last_error = msg;
  return -1;
}
void cxh_clear_error() { g_last_error.clear(); }

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

DeviceBuffer::~DeviceBuffer() {
  if (ptr) (void)hipFree(ptr);
}
namespace {
thread_local int g_no_alloc = 0;
}
NoAlloc::NoAlloc() { ++g_no_alloc; }
NoAlloc::~NoAlloc() { --g_no_alloc; }
bool no_alloc_active() { return g_no_alloc > 0; }

void DeviceBuffer::reserve(size_t nbytes) {
  if (nbytes <= bytes) return;
  if (g_no_alloc)
    throw std::runtime_error("no-allocation call: a device buffer of the tape would have to grow to " +
                             std::to_string(nbytes) + " bytes; make the allocating call once first with a batch "
                             "at least as large");
  if (ptr) hip_check(hipFree(ptr), "hipFree");
  ptr = nullptr;
  bytes = 0;
  hip_check(hipMalloc(&ptr, nbytes), "hipMalloc");
  bytes = nbytes;
}
void DeviceBuffer::upload(const void* src, size_t nbytes) {
  reserve(std::max<size_t>(nbytes, 16));
  if (nbytes) hip_check(hipMemcpy(ptr, src, nbytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

// Largest dynamic LDS allocation a block may opt into (160 KB on gfx950).
int device_lds_bytes() {
  static int lds = [] {
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess || v <= 0)
      if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
        v = 64 * 1024;
    return v;
  }();
  return lds;
}

// Raise a kernel's dynamic-LDS limit before launching it with `bytes`.
void allow_lds(const void* kernel, size_t bytes) {
  if (bytes > 64 * 1024)
    hip_check(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)),
              "hipFuncSetAttribute");
}

}  // namespace sxh

extern "C" const char* cxh_last_error(void) { return sxh::g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// Work-vector views: [slot][lane] in LDS or [slot][global lane] in HBM.
// ---------------------------------------------------------------------------
struct WLds {
  double* base;  // LDS, already offset by the lane
  int stride;
  __device__ double& operator[](int s) const { return base[s * stride]; }
};
struct WGlb {
  double* base;  // HBM scratch, already offset by the global lane
  int64_t stride;
  __device__ double& operator[](int s) const { return base[s * stride]; }
};

// Per-instance address of nonzero j: p + j*s_nz + k*s_inst
__device__ __forceinline__ const double* io_addr(const IOArg& a, int j, int64_t k) {
  return a.p + j * a.s_nz + k * a.s_inst;
}

__device__ __forceinline__ double instr_const(const Instr& e) { return e.d; }

// ---------------------------------------------------------------------------
// Instruction streaming.  The tape is read 64 instructions at a time: lane j
// of the wavefront loads instruction (chunk + j) into four VGPRs with one
// coalesced 16-byte load (1 KB per wave), and instruction j is broadcast to
// SGPRs with v_readlane.  The next chunk is in flight while the current one
// executes, so instruction fetch is off the dependent-latency path and the op
// dispatch is a scalar branch.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return static_cast<int>(threadIdx.x & 63); }

// Instances per wavefront.  With few instances the interpreter is bound by
// the latency of its dependent LDS round trips, not by VALU throughput, so a
// wave may carry only `ipw` (8, 16, 32 or 64) instances: lanes >= ipw mirror
// lane (lane % ipw) -- same instance, same LDS column, identical writes -- and
// store nothing.  Fewer instances per wave means proportionally more waves to
// hide each other's latency.
struct LaneMap {
  int64_t slot;  // instance slot of this lane (index into per-slot scratch)
  int64_t kk;    // instance read by this lane (clamped into range)
  int col;       // LDS column of this lane within the block
  int stride;    // LDS columns per block
  bool active;   // this lane owns its instance and writes results
};

__device__ __forceinline__ LaneMap map_lane(int ipw, int64_t n) {
  const int lane = lane_id(), wave = static_cast<int>(threadIdx.x >> 6);
  const int waves = static_cast<int>(blockDim.x >> 6);
  const int sub = lane & (ipw - 1);
  LaneMap m;
  m.slot = (static_cast<int64_t>(blockIdx.x) * waves + wave) * ipw + sub;
  m.kk = m.slot < n ? m.slot : n - 1;
  m.col = wave * ipw + sub;
  m.stride = waves * ipw;
  m.active = lane < ipw && m.slot < n;
  return m;
}

__device__ __forceinline__ int4 fetch_chunk(const Instr* __restrict__ tape, int base, int end) {
  const int t = base + lane_id();
  return t < end ? reinterpret_cast<const int4*>(tape)[t] : make_int4(0, 0, 0, 0);
}

__device__ __forceinline__ Instr bcast(const int4& v, int j) {
  Instr e;
  e.op = __builtin_amdgcn_readlane(v.x, j);
  e.i0 = __builtin_amdgcn_readlane(v.y, j);
  e.i1 = __builtin_amdgcn_readlane(v.z, j);
  e.i2 = __builtin_amdgcn_readlane(v.w, j);
  return e;
}

// f(instr, t) for t = t0 .. t1-1
template <class F>
__device__ __forceinline__ void stream_fwd(const Instr* __restrict__ tape, int t0, int t1, F&& f) {
  int4 next = fetch_chunk(tape, t0, t1);
  for (int c = t0; c < t1; c += 64) {
    const int4 cur = next;
    if (c + 64 < t1) next = fetch_chunk(tape, c + 64, t1);
    const int cnt = min(64, t1 - c);
    for (int j = 0; j < cnt; ++j) f(bcast(cur, j), c + j);
  }
}

// f(instr, t, j) for t = t1-1 down to t0, chunk-aligned from t0; pre(c0, c1)
// runs before the instructions [c0, c1) of a chunk are visited.
template <class P, class F>
__device__ __forceinline__ void stream_rev(const Instr* __restrict__ tape, int t0, int t1, P&& pre, F&& f) {
  if (t1 <= t0) return;
  int c = t0 + ((t1 - t0 - 1) / 64) * 64;  // start of the last chunk
  int4 next = fetch_chunk(tape, c, t1);
  for (; c >= t0; c -= 64) {
    const int4 cur = next;
    if (c - 64 >= t0) next = fetch_chunk(tape, c - 64, c);
    const int cnt = min(64, t1 - c);
    pre(c, c + cnt);
    for (int j = cnt - 1; j >= 0; --j) f(bcast(cur, j), c + j, j);
  }
}

// Execute instruction e of the primal sweep on work vector W.
template <class W>
__device__ __forceinline__ void primal_step(const Instr& e, W& w, const IOTable& arg, const IOTable& res,
                                            int64_t k, bool active) {
  switch (e.op) {
    case OP_CONST: w[e.i0] = instr_const(e); break;
    case OP_INPUT: {
      const IOArg a = arg.a[e.i1];
      w[e.i0] = a.p ? *io_addr(a, e.i2, k) : 0.0;
    } break;
    case OP_OUTPUT: {
      const IOArg r = res.a[e.i0];
      if (r.p && active) *const_cast<double*>(io_addr(r, e.i2, k)) = w[e.i1];
    } break;
    default: {
      const double x = w[e.i1], y = w[e.i2];
      w[e.i0] = op_fun(e.op, x, y);
    }
  }
}

// ---------------------------------------------------------------------------
// Forward evaluation
// ---------------------------------------------------------------------------
template <bool LDS>
__global__ void __launch_bounds__(256)
tape_fwd_kernel(const Instr* __restrict__ tape, int n_instr, int ws, const IOTable arg, const IOTable res,
                int64_t n, double* __restrict__ wglob, int64_t wstride, int ipw) {
  extern __shared__ double lds[];
  const LaneMap m = map_lane(ipw, n);
  if constexpr (LDS) {
    WLds w{lds + m.col, m.stride};
    stream_fwd(tape, 0, n_instr, [&](const Instr& e, int) { primal_step(e, w, arg, res, m.kk, m.active); });
  } else {
    WGlb w{wglob + m.slot, wstride};
    stream_fwd(tape, 0, n_instr, [&](const Instr& e, int) { primal_step(e, w, arg, res, m.kk, m.active); });
  }
}

// ---------------------------------------------------------------------------
// Fused forward + adjoint with checkpointed segments
//
// Forward: run the tape; at the start of every segment but the first, save
// the work vector (checkpoint).  Reverse: for each segment, last first,
// restore its checkpoint and re-run it forward recording, for every
// instruction, the value its destination slot held before (the trail); then
// walk the segment backwards: read f = w[i0], restore w[i0] from the trail,
// read the operands x = w[i1], y = w[i2], form the partials and push the
// adjoint seed of slot i0 into slots i1 / i2.  The trail of the last segment
// is recorded during the forward pass already.
// ---------------------------------------------------------------------------
struct AdjArgs {
  const Instr* tape;
  int n_instr, ws, seg_len, n_seg;
  IOTable arg;
  IOTable res;           // primal outputs (forward pass)
  IOTable aseed;         // adjoint seeds, one per output
  IOTable asens;         // adjoint sensitivities, one per input
  const int2* unused_in; // (input, nz) pairs read by no INPUT instruction
  int n_unused;
  int64_t n;
  double* trail;         // [seg_len][lanes]
  double* ckpt;          // [n_seg][ws][lanes]
  double* wglob;         // [2*ws][lanes] when the work vectors live in HBM
  int64_t lanes;         // stride of the HBM buffers (instance slots of the grid)
  int mode;              // 0 fwd+adj, 1 fwd + all checkpoints, 2 adj from checkpoints
  int ipw;               // instances per wavefront
  const uint32_t* live;  // bit t: instruction t's adjoint is structurally nonzero (adjoint_live)
};

// lane: index into the per-lane scratch (trail); inst: instance for I/O
// (equal to lane for active lanes, clamped to a valid instance otherwise).
template <class W>
__device__ __forceinline__ void fwd_segment(const AdjArgs& A, W& w, int t0, int t1, int64_t lane,
                                            int64_t inst, bool active, bool record, bool write_res) {
  stream_fwd(A.tape, t0, t1, [&](const Instr& e, int t) {
    if (record && e.op != OP_OUTPUT) A.trail[static_cast<int64_t>(t - t0) * A.lanes + lane] = w[e.i0];
    if (e.op == OP_OUTPUT) {
      if (write_res) {
        const IOArg r = A.res.a[e.i0];
        if (r.p && active) *const_cast<double*>(io_addr(r, e.i2, inst)) = w[e.i1];
      }
    } else {
      primal_step(e, w, A.arg, A.res, inst, active);
    }
  });
}

// tr: this lane's column of a 64-row LDS staging buffer for the trail values
// of the chunk being reversed (loaded with 64 independent coalesced loads
// before the chunk starts).
template <class W>
__device__ __forceinline__ void rev_segment(const AdjArgs& A, W& w, W& aw, WLds tr, int t0, int t1,
                                            int64_t lane, int64_t k, bool active) {
  stream_rev(
      A.tape, t0, t1,
      [&](int c0, int c1) {
        const double* src = A.trail + static_cast<int64_t>(c0 - t0) * A.lanes + lane;
        for (int j = 0; j < c1 - c0; ++j) tr[j] = src[static_cast<int64_t>(j) * A.lanes];
      },
      [&](const Instr& e, int t, int j) {
        if (e.op == OP_OUTPUT) {
          const IOArg s = A.aseed.a[e.i0];
          if (s.p) aw[e.i1] += *io_addr(s, e.i2, k);
          return;
        }
        const double told = tr[j];
        if (e.op == OP_INPUT) {
          const IOArg s = A.asens.a[e.i1];
          if (s.p && active) *const_cast<double*>(io_addr(s, e.i2, k)) = aw[e.i0];
          aw[e.i0] = 0.0;
          w[e.i0] = told;
          return;
        }
        if (e.op == OP_CONST) {
          aw[e.i0] = 0.0;
          w[e.i0] = told;
          return;
        }
        const double f = w[e.i0];
        w[e.i0] = told;
        const double x = w[e.i1], y = w[e.i2];
        const double seed = aw[e.i0];
        aw[e.i0] = 0.0;
        if (!((A.live[t >> 5] >> (t & 31)) & 1u)) return;  // no output reaches this value
        double d[2];
        op_der(e.op, x, y, f, d);
        if (e.op == OP_IF_ELSE_ZERO) {
          aw[e.i2] += sxh_if_else_zero(d[1], seed);
        } else {  // a constant-zero partial makes no contribution
          if (!op_zero_partial(e.op, 0)) aw[e.i1] += d[0] * seed;
          if (op_is_binary(e.op) && !op_zero_partial(e.op, 1)) aw[e.i2] += d[1] * seed;
        }
      });
}

template <class W>
__device__ void adj_body(const AdjArgs& A, W& w, W& aw, WLds tr, int64_t k, int64_t kk, bool active) {
  const int ws = A.ws;
  for (int s = 0; s < ws; ++s) {
    w[s] = 0.0;
    aw[s] = 0.0;
  }
  const int last = A.n_seg - 1;
  if (A.mode != 2) {
    // forward pass
    for (int sg = 0; sg < A.n_seg; ++sg) {
      const int t0 = sg * A.seg_len, t1 = min(A.n_instr, t0 + A.seg_len);
      if (sg > 0 || A.mode == 1) {
        double* c = A.ckpt + static_cast<int64_t>(sg) * ws * A.lanes + k;
        for (int s = 0; s < ws; ++s) c[static_cast<int64_t>(s) * A.lanes] = w[s];
      }
      fwd_segment(A, w, t0, t1, k, kk, active, A.mode == 0 && sg == last, true);
    }
    if (A.mode == 1) return;
  }
  // reverse pass
  for (int sg = last; sg >= 0; --sg) {
    const int t0 = sg * A.seg_len, t1 = min(A.n_instr, t0 + A.seg_len);
    if (A.mode == 2 || sg != last) {
      const double* c = A.ckpt + static_cast<int64_t>(sg) * ws * A.lanes + k;
      if (sg > 0 || A.mode == 2)
        for (int s = 0; s < ws; ++s) w[s] = c[static_cast<int64_t>(s) * A.lanes];
      fwd_segment(A, w, t0, t1, k, kk, active, true, false);
    }
    rev_segment(A, w, aw, tr, t0, t1, k, kk, active);
  }
  if (active) {
    for (int u = 0; u < A.n_unused; ++u) {
      const int2 p = A.unused_in[u];
      const IOArg s = A.asens.a[p.x];
      if (s.p) *const_cast<double*>(io_addr(s, p.y, k)) = 0.0;
    }
  }
}

template <bool LDS>
__global__ void __launch_bounds__(256) tape_adj_kernel(AdjArgs A) {
  extern __shared__ double lds[];
  const LaneMap m = map_lane(A.ipw, A.n);
  // LDS: [trail chunk: 64 rows][w: ws rows][aw: ws rows], one column per slot
  WLds tr{lds + m.col, m.stride};
  if constexpr (LDS) {
    WLds w{lds + 64 * m.stride + m.col, m.stride};
    WLds aw{lds + (64 + static_cast<int64_t>(A.ws)) * m.stride + m.col, m.stride};
    adj_body(A, w, aw, tr, m.slot, m.kk, m.active);
  } else {
    WGlb w{A.wglob + m.slot, A.lanes};
    WGlb aw{A.wglob + static_cast<int64_t>(A.ws) * A.lanes + m.slot, A.lanes};
    adj_body(A, w, aw, tr, m.slot, m.kk, m.active);
  }
}

// ---------------------------------------------------------------------------
// Fused primal + multi-direction tangent sweep + Jacobian scatter
// Layout of the tangents: direction d of slot s at tw[s*ndir + d].
// ---------------------------------------------------------------------------
struct JacArgs {
  const Instr* tape;
  int n_instr, ws, ndir;
  IOTable arg;
  IOTable res;
  const int* in_off;      // flat offset of input i's nonzeros
  const int* out_off;     // flat offset of output i's nonzeros
  const double* seedval;  // [flat_in_nz][ndir]
  const int2* scatter;    // [flat_out_nz][ndir] -> up to two Jacobian nonzeros
  double* jac;            // Jacobian nonzeros
  int64_t jac_s_nz, jac_s_inst;
  int64_t n;
  double* wglob;          // [(1+ndir)*ws][lanes] when not in LDS
  int64_t lanes;
  int ipw;
};

template <class W>
__device__ void jac_body(const JacArgs& A, W& w, W& tw, int64_t k, int64_t kk, bool active) {
  const int nd = A.ndir;
  stream_fwd(A.tape, 0, A.n_instr, [&](const Instr& e, int) {
    switch (e.op) {
      case OP_CONST:
        w[e.i0] = instr_const(e);
        for (int d = 0; d < nd; ++d) tw[e.i0 * nd + d] = 0.0;
        break;
      case OP_INPUT: {
        const IOArg a = A.arg.a[e.i1];
        w[e.i0] = a.p ? *io_addr(a, e.i2, kk) : 0.0;
        const double* sv = A.seedval + static_cast<int64_t>(A.in_off[e.i1] + e.i2) * nd;
        for (int d = 0; d < nd; ++d) tw[e.i0 * nd + d] = sv[d];
      } break;
      case OP_OUTPUT: {
        const IOArg r = A.res.a[e.i0];
        if (r.p && active) *const_cast<double*>(io_addr(r, e.i2, k)) = w[e.i1];
        const int2* sc = A.scatter + static_cast<int64_t>(A.out_off[e.i0] + e.i2) * nd;
        if (active) {
          for (int d = 0; d < nd; ++d) {
            const int2 tgt = sc[d];
            const double v = tw[e.i1 * nd + d];
            if (tgt.x >= 0) A.jac[tgt.x * A.jac_s_nz + k * A.jac_s_inst] = v;
            if (tgt.y >= 0) A.jac[tgt.y * A.jac_s_nz + k * A.jac_s_inst] = v;
          }
        }
      } break;
      default: {
        const double x = w[e.i1], y = w[e.i2];
        const double f = op_fun(e.op, x, y);
        double dd[2];
        op_der(e.op, x, y, f, dd);
        if (e.op == OP_IF_ELSE_ZERO) {
          for (int d = 0; d < nd; ++d) tw[e.i0 * nd + d] = sxh_if_else_zero(dd[1], tw[e.i2 * nd + d]);
        } else if (op_is_binary(e.op)) {
          for (int d = 0; d < nd; ++d)
            tw[e.i0 * nd + d] = dd[0] * tw[e.i1 * nd + d] + dd[1] * tw[e.i2 * nd + d];
        } else {
          for (int d = 0; d < nd; ++d) tw[e.i0 * nd + d] = dd[0] * tw[e.i1 * nd + d];
        }
        w[e.i0] = f;
      }
    }
  });
}

template <bool LDS>
__global__ void __launch_bounds__(256) tape_jac_kernel(JacArgs A) {
  extern __shared__ double lds[];
  const LaneMap m = map_lane(A.ipw, A.n);
  if constexpr (LDS) {
    WLds w{lds + m.col, m.stride};
    WLds tw{lds + static_cast<int64_t>(A.ws) * m.stride + m.col, m.stride};
    jac_body(A, w, tw, m.slot, m.kk, m.active);
  } else {
    WGlb w{A.wglob + m.slot, A.lanes};
    WGlb tw{A.wglob + static_cast<int64_t>(A.ws) * A.lanes + m.slot, A.lanes};
    jac_body(A, w, tw, m.slot, m.kk, m.active);
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace sxh {

// Validate a tape so that no kernel can index outside its buffers.
void validate_tape(const std::vector<Instr>& a, int64_t ws, const std::vector<int64_t>& nnz_in,
                   const std::vector<int64_t>& nnz_out) {
  auto slot_ok = [&](int s) { return s >= 0 && s < ws; };
  for (size_t t = 0; t < a.size(); ++t) {
    const Instr& e = a[t];
    const std::string where = "tape instruction " + std::to_string(t) + ": ";
    switch (e.op) {
      case OP_CONST:
        if (!slot_ok(e.i0)) throw std::runtime_error(where + "slot out of range");
        break;
      case OP_INPUT:
        if (!slot_ok(e.i0) || e.i1 < 0 || e.i1 >= static_cast<int>(nnz_in.size()) || e.i2 < 0 ||
            e.i2 >= nnz_in[e.i1])
          throw std::runtime_error(where + "input index out of range");
        break;
      case OP_OUTPUT:
        if (!slot_ok(e.i1) || e.i0 < 0 || e.i0 >= static_cast<int>(nnz_out.size()) || e.i2 < 0 ||
            e.i2 >= nnz_out[e.i0])
          throw std::runtime_error(where + "output index out of range");
        break;
      case OP_PARAMETER:
        throw std::runtime_error(where + "free parameter: the Function has free variables");
      case OP_CALL:
        throw std::runtime_error(where + "OP_CALL is not supported on the GPU tape");
      default:
        if (!op_is_arith(e.op)) throw std::runtime_error(where + "unsupported operation " + std::to_string(e.op));
        if (!slot_ok(e.i0) || !slot_ok(e.i1) || !slot_ok(e.i2))
          throw std::runtime_error(where + "slot out of range");
    }
  }
}

// A validated tape: its input nonzeros read by no INPUT instruction are noted
// (their adjoint is zero).
std::unique_ptr<cxh_tape_s> make_tape(std::vector<Instr> alg, int64_t worksize, int32_t n_in, const int64_t* nnz_in,
                                      int32_t n_out, const int64_t* nnz_out) {
  if (n_in > CXH_MAX_IO || n_out > CXH_MAX_IO)
    throw std::runtime_error("cxh_tape_create: at most " + std::to_string(CXH_MAX_IO) +
                             " input and output blocks");
  auto t = std::make_unique<cxh_tape_s>();
  t->alg = std::move(alg);
  t->ws = std::max<int64_t>(worksize, 1);
  t->nnz_in.assign(nnz_in, nnz_in + n_in);
  t->nnz_out.assign(nnz_out, nnz_out + n_out);
  validate_tape(t->alg, t->ws, t->nnz_in, t->nnz_out);
  std::vector<std::vector<char>> used(n_in);
  for (int i = 0; i < n_in; ++i) used[i].assign(t->nnz_in[i], 0);
  for (const Instr& e : t->alg)
    if (e.op == OP_INPUT) used[e.i1][e.i2] = 1;
  for (int i = 0; i < n_in; ++i)
    for (int64_t j = 0; j < t->nnz_in[i]; ++j)
      if (!used[i][j]) t->unused_in.push_back({i, static_cast<int>(j)});
  return t;
}

// Kernel-argument table of the I/O blocks (passed by value, no staging copy).
IOTable make_io(const double* const* p, const std::vector<int64_t>& nnz, int32_t layout, int64_t ld) {
  IOTable v{};
  if (nnz.size() > CXH_MAX_IO) throw std::runtime_error("too many input/output blocks");
  for (size_t i = 0; i < nnz.size(); ++i) {
    v.a[i].p = p ? p[i] : nullptr;
    if (layout == CXH_SOA) {
      v.a[i].s_nz = ld;
      v.a[i].s_inst = 1;
    } else {
      v.a[i].s_nz = 1;
      v.a[i].s_inst = nnz[i];
    }
  }
  return v;
}

int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// The instruction stream (and the list of unread input nonzeros) go to the
// device at the first interpreter launch, so that creating a tape, generating
// specialised source from it and the host VM need no GPU.
const Instr* device_tape(cxh_tape_s* t) {
  if (!t->uploaded) {
    t->d_tape.upload(t->alg.data(), t->alg.size() * sizeof(Instr));
    t->d_unused.upload(t->unused_in.data(), t->unused_in.size() * sizeof(int2));
    t->uploaded = true;
  }
  return t->d_tape.as<Instr>();
}

// Launch geometry for n instances: `cols` instance slots per block (waves
// per block times instances per wave), `slots` in the whole grid.
GridShape grid_shape(const cxh_tape_s* t, int64_t n) {
  GridShape g;
  g.bs = t->block;
  g.ipw = t->ipw;
  g.cols = (g.bs / 64) * g.ipw;
  g.grid = (n + g.cols - 1) / g.cols;
  g.slots = g.grid * g.cols;
  return g;
}

}  // namespace sxh

namespace sxh {

// Blocks scanned by the regularity check: what a call wrote (outputs,
// sensitivities), with their nonzero counts.
struct StatusIO {
  IOTable t[2];
  int64_t nnz[2][CXH_MAX_IO];
  int nb[2];
};

__global__ void status_clear(int* count) { *count = 0; }

// status[k]: 1 when instance k wrote a NaN, | 2 when it wrote an infinity.
__global__ void __launch_bounds__(256) status_kernel(StatusIO io, int64_t n, int32_t* status, int* count) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int flag = 0;
  for (int b = 0; b < 2; ++b)
    for (int i = 0; i < io.nb[b]; ++i) {
      const IOArg& a = io.t[b].a[i];
      if (!a.p) continue;
      for (int64_t j = 0; j < io.nnz[b][i]; ++j) {
        const double v = a.p[j * a.s_nz + k * a.s_inst];
        if (isnan(v)) flag |= 1;
        else if (isinf(v)) flag |= 2;
      }
    }
  status[k] = flag;
  if (flag) atomicAdd(count, 1);
}

void check_status_cap(const cxh_tape_s* t, int64_t n) {
  if (t->status && n > t->status_cap)
    throw std::runtime_error("regularity check: the status array holds " + std::to_string(t->status_cap) +
                             " instances, the call has " + std::to_string(n));
}

// With the regularity check on: scan the blocks, write the per-instance
// status, synchronize and return the number of flagged instances.
int regularity(cxh_tape_s* t, int64_t n, const IOTable& a, const std::vector<int64_t>& nnz_a, const IOTable& b,
               const std::vector<int64_t>& nnz_b, hipStream_t s) {
  if (!t->status || n == 0) return 0;
  StatusIO io{};
  io.t[0] = a;
  io.t[1] = b;
  io.nb[0] = static_cast<int>(nnz_a.size());
  io.nb[1] = static_cast<int>(nnz_b.size());
  for (size_t i = 0; i < nnz_a.size(); ++i) io.nnz[0][i] = nnz_a[i];
  for (size_t i = 0; i < nnz_b.size(); ++i) io.nnz[1][i] = nnz_b[i];
  t->status_count.reserve(sizeof(int));
  int* count = t->status_count.as<int>();
  hipLaunchKernelGGL(status_clear, dim3(1), dim3(1), 0, s, count);
  hip_check(hipGetLastError(), "status_clear launch");
  hipLaunchKernelGGL(status_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, io, n, t->status,
                     count);
  hip_check(hipGetLastError(), "status_kernel launch");
  int h = 0;
  hip_check(hipMemcpyAsync(&h, count, sizeof(int), hipMemcpyDeviceToHost, s), "hipMemcpyAsync status count");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  return h;
}

}  // namespace sxh

extern "C" {

int cxh_tape_create(const void* instr, int64_t n_instr, int64_t worksize, int32_t n_in,
                    const int64_t* nnz_in, int32_t n_out, const int64_t* nnz_out, cxh_tape* out) {
  CXH_TRY
  if (!out) return cxh_fail("cxh_tape_create: null output handle");
  if (n_instr < 0 || worksize < 0 || n_in < 0 || n_out < 0) return cxh_fail("cxh_tape_create: negative size");
  if (n_instr > (1 << 30) || worksize > (1 << 24)) return cxh_fail("cxh_tape_create: tape too large");
  std::vector<Instr> alg(n_instr);
  if (n_instr) std::memcpy(alg.data(), instr, n_instr * sizeof(Instr));
  *out = make_tape(std::move(alg), worksize, n_in, nnz_in, n_out, nnz_out).release();
  return 0;
  CXH_CATCH
}

int cxh_tape_create_nested(const void* instr, int64_t n_instr, int64_t worksize, int32_t n_in,
                           const int64_t* nnz_in, int32_t n_out, const int64_t* nnz_out, int32_t n_calls,
                           const cxh_tape* callees, const int32_t* const* call_dep,
                           const int32_t* const* call_res, cxh_tape* out) {
  CXH_TRY
  if (!out) return cxh_fail("cxh_tape_create_nested: null output handle");
  if (n_instr < 0 || worksize < 0 || n_in < 0 || n_out < 0 || n_calls < 0)
    return cxh_fail("cxh_tape_create_nested: negative size");
  if (n_instr > (1 << 30) || worksize > (1 << 24)) return cxh_fail("cxh_tape_create_nested: tape too large");
  if (n_calls > 0 && (!callees || !call_dep || !call_res))
    return cxh_fail("cxh_tape_create_nested: null call table");
  const Instr* a = static_cast<const Instr*>(instr);
  std::vector<Instr> flat;
  flat.reserve(static_cast<size_t>(n_instr));
  int64_t ws = std::max<int64_t>(worksize, 1);
  auto slot_ok = [&](int64_t v) { return v >= 0 && v < worksize; };
  auto assign = [&](int64_t dst, int64_t src) {
    Instr e{};
    e.op = OP_ASSIGN;
    e.i0 = static_cast<int>(dst);
    e.i1 = e.i2 = static_cast<int>(src);  // a unary operation's second operand is its first
    flat.push_back(e);
  };
  for (int64_t t = 0; t < n_instr; ++t) {
    const Instr& e = a[t];
    if (e.op != OP_CALL) {
      flat.push_back(e);
      continue;
    }
    const std::string where = "cxh_tape_create_nested: instruction " + std::to_string(t) + ": ";
    if (e.i1 < 0 || e.i1 >= n_calls) throw std::runtime_error(where + "call index out of range");
    const cxh_tape_s* c = callees[e.i1];
    if (!c) throw std::runtime_error(where + "null callee");
    const int32_t* dep = call_dep[e.i1];
    const int32_t* res = call_res[e.i1];
    std::vector<int64_t> in_off(c->nnz_in.size() + 1, 0), out_off(c->nnz_out.size() + 1, 0);
    for (size_t i = 0; i < c->nnz_in.size(); ++i) in_off[i + 1] = in_off[i] + c->nnz_in[i];
    for (size_t i = 0; i < c->nnz_out.size(); ++i) out_off[i + 1] = out_off[i] + c->nnz_out[i];
    const int64_t nd = in_off.back(), nr = out_off.back();
    if ((nd > 0 && !dep) || (nr > 0 && !res)) throw std::runtime_error(where + "null dep or res slots");
    // Slots after the caller's work vector, shared by every call (one call at a
    // time, as call_fwd): the arguments, the callee's work vector, the results.
    const int64_t in0 = worksize, cw = in0 + nd, out0 = cw + c->ws;
    ws = std::max(ws, out0 + nr);
    if (ws > (1 << 24)) throw std::runtime_error(where + "work vector too large");
    // call_fwd gathers every argument before the call ...
    for (int64_t i = 0; i < nd; ++i) {
      if (!slot_ok(dep[i])) throw std::runtime_error(where + "argument slot out of range");
      assign(in0 + i, dep[i]);
    }
    // ... the callee runs on its own work vector; an output nonzero its tape
    // does not write reads as zero ...
    std::vector<char> written(static_cast<size_t>(nr), 0);
    for (const Instr& ce : c->alg)
      if (ce.op == OP_OUTPUT) written[out_off[ce.i0] + ce.i2] = 1;
    for (int64_t i = 0; i < nr; ++i)
      if (!written[i]) {
        Instr z{};
        z.op = OP_CONST;
        z.i0 = static_cast<int>(out0 + i);
        z.d = 0.0;
        flat.push_back(z);
      }
    for (const Instr& ce : c->alg) {
      Instr f = ce;
      switch (ce.op) {
        case OP_CONST:
          f.i0 = static_cast<int>(cw + ce.i0);
          flat.push_back(f);
          break;
        case OP_INPUT:
          assign(cw + ce.i0, in0 + in_off[ce.i1] + ce.i2);
          break;
        case OP_OUTPUT:
          assign(out0 + out_off[ce.i0] + ce.i2, cw + ce.i1);
          break;
        default:
          f.i0 = static_cast<int>(cw + ce.i0);
          f.i1 = static_cast<int>(cw + ce.i1);
          f.i2 = static_cast<int>(cw + ce.i2);
          flat.push_back(f);
      }
    }
    // ... and its results are stored after it returns, where they are needed.
    for (int64_t i = 0; i < nr; ++i) {
      if (res[i] < 0) continue;
      if (!slot_ok(res[i])) throw std::runtime_error(where + "result slot out of range");
      assign(res[i], out0 + i);
    }
  }
  if (static_cast<int64_t>(flat.size()) > (1 << 30)) throw std::runtime_error("cxh_tape_create_nested: tape too large");
  *out = make_tape(std::move(flat), ws, n_in, nnz_in, n_out, nnz_out).release();
  return 0;
  CXH_CATCH
}

int cxh_tape_instructions(cxh_tape t, void* instr, int64_t cap, int64_t* n_instr, int64_t* worksize) {
  CXH_TRY
  if (!t) return cxh_fail("null tape");
  const int64_t n = static_cast<int64_t>(t->alg.size());
  if (n_instr) *n_instr = n;
  if (worksize) *worksize = t->ws;
  if (instr && cap > 0) std::memcpy(instr, t->alg.data(), static_cast<size_t>(std::min(cap, n)) * sizeof(Instr));
  return 0;
  CXH_CATCH
}

int cxh_tape_destroy(cxh_tape t) {
  if (t) jit_settle(t);  // a worker still compiling for the tape is waited for
  delete t;
  return 0;
}

int cxh_tape_config(cxh_tape t, int32_t block, int32_t w_in_lds, int32_t ipw) {
  if (!t) return cxh_fail("null tape");
  if (block != -1) {
    if (block != 64 && block != 128 && block != 256) return cxh_fail("block must be 64, 128 or 256");
    t->block = block;
  }
  if (w_in_lds != -1) t->w_lds = w_in_lds;
  if (ipw != -1) {
    if (ipw != 8 && ipw != 16 && ipw != 32 && ipw != 64) return cxh_fail("ipw must be 8, 16, 32 or 64");
    t->ipw = ipw;
  }
  return 0;
}

int cxh_tape_jit_fallback(cxh_tape t, char* why, int64_t cap) {
  if (!t) return cxh_fail("null tape");
  if (why && cap > 0) {
    const size_t m = std::min<size_t>(t->fallback_why.size(), static_cast<size_t>(cap - 1));
    std::memcpy(why, t->fallback_why.data(), m);
    why[m] = '\0';
  }
  return (t->eval_interp ? 1 : 0) | (t->adj_interp.empty() ? 0 : 2);
}

int cxh_tape_ipt(cxh_tape t, int32_t ipt) {
  if (!t) return cxh_fail("null tape");
  if (ipt != 0 && ipt != 1 && ipt != 2 && ipt != 4 && ipt != 8) return cxh_fail("ipt must be 0, 1, 2, 4 or 8");
  t->ipt = ipt;
  return 0;
}

int cxh_tape_io(cxh_tape t, int32_t* n_in, int32_t* n_out, int64_t* nnz_in, int64_t* nnz_out) {
  if (!t) return cxh_fail("null tape");
  if (n_in) *n_in = static_cast<int32_t>(t->nnz_in.size());
  if (n_out) *n_out = static_cast<int32_t>(t->nnz_out.size());
  for (size_t i = 0; nnz_in && i < t->nnz_in.size(); ++i) nnz_in[i] = t->nnz_in[i];
  for (size_t i = 0; nnz_out && i < t->nnz_out.size(); ++i) nnz_out[i] = t->nnz_out[i];
  return 0;
}

int cxh_tape_regularity_check(cxh_tape t, int32_t on, int32_t* status, int64_t cap) {
  if (!t) return cxh_fail("null tape");
  if (on && (!status || cap <= 0)) return cxh_fail("cxh_tape_regularity_check: a status array is needed");
  t->status = on ? status : nullptr;
  t->status_cap = on ? cap : 0;
  return 0;
}

int cxh_map_eval(cxh_tape t, int64_t n, const double* const* arg, double* const* res, int32_t layout,
                 int64_t ld, void* stream) {
  CXH_TRY
  if (!t) return cxh_fail("null tape");
  if (n < 0) return cxh_fail("negative n");
  if (layout != CXH_SOA && layout != CXH_AOS) return cxh_fail("bad layout");
  if (layout == CXH_SOA && ld < n) return cxh_fail("ld < n");
  check_status_cap(t, n);
  const IOTable R = make_io(res, t->nnz_out, layout, ld);
  const int rc = map_eval_io(t, n, make_io(arg, t->nnz_in, layout, ld), R, static_cast<hipStream_t>(stream));
  if (rc) return rc;
  return regularity(t, n, R, t->nnz_out, IOTable{}, {}, static_cast<hipStream_t>(stream));
  CXH_CATCH
}

int cxh_tape_tune(cxh_tape t, int64_t n, const double* const* arg, double* const* res, int32_t layout, int64_t ld,
                  int32_t reps, void* stream, int32_t* best, double* best_ms) {
  CXH_TRY
  if (!t) return cxh_fail("null tape");
  if (n <= 0) return cxh_fail("cxh_tape_tune: n must be positive");
  if (layout != CXH_SOA && layout != CXH_AOS) return cxh_fail("bad layout");
  if (layout == CXH_SOA && ld < n) return cxh_fail("ld < n");
  reps = std::max(reps, 1);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const IOTable A = make_io(arg, t->nnz_in, layout, ld), R = make_io(res, t->nnz_out, layout, ld);
  // the settings each engine reads: block size, and lanes per instance
  // (instances per wavefront) and the work vector's place for the
  // interpreter, instances per thread for the specialised kernels
  struct Cfg {
    int block, w_lds, ipw, ipt;
  };
  std::vector<Cfg> cfgs;
  for (int blk : {64, 128, 256}) {
    if (t->engine == CXH_ENGINE_JIT) {
      for (int ipt : {1, 2, 4}) cfgs.push_back({blk, t->w_lds, t->ipw, ipt});
    } else {
      for (int w : {1, 0})
        for (int ipw : {64, 32, 16, 8}) cfgs.push_back({blk, w, ipw, t->ipt});
    }
  }
  struct Events {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    Events() {
      hip_check(hipEventCreate(&e0), "hipEventCreate");
      hip_check(hipEventCreate(&e1), "hipEventCreate");
    }
    ~Events() {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    }
  } ev;
  // the tape's settings come back unless a fastest one is chosen
  struct Restore {
    cxh_tape_s* t;
    Cfg c;
    bool async;  // compiles stay in this thread while tuning: the timings are the specialised kernels'
    bool on = true;
    ~Restore() {
      jit_async_set(t, async);
      if (!on) return;
      t->block = c.block;
      t->w_lds = c.w_lds;
      t->ipw = c.ipw;
      t->ipt = c.ipt;
    }
  } restore{t, {t->block, t->w_lds, t->ipw, t->ipt}, jit_async_pause(t)};
  Cfg pick = restore.c;
  float best_t = -1.0f;
  std::string last_err;
  for (const Cfg& c : cfgs) {
    t->block = c.block;
    t->w_lds = c.w_lds;
    t->ipw = c.ipw;
    t->ipt = c.ipt;
    try {
      if (int rc = map_eval_io(t, n, A, R, s)) return rc;  // warm-up (generates the kernels)
      hip_check(hipEventRecord(ev.e0, s), "hipEventRecord");
      for (int r = 0; r < reps; ++r)
        if (int rc = map_eval_io(t, n, A, R, s)) return rc;
      hip_check(hipEventRecord(ev.e1, s), "hipEventRecord");
      hip_check(hipEventSynchronize(ev.e1), "hipEventSynchronize");
    } catch (const std::runtime_error& e) {  // a setting whose kernel needs scratch: not a candidate
      if (std::string(e.what()).find("spills") == std::string::npos) throw;
      last_err = e.what();
      continue;
    }
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, ev.e0, ev.e1), "hipEventElapsedTime");
    ms /= static_cast<float>(reps);
    if (best_t < 0 || ms < best_t) {
      best_t = ms;
      pick = c;
    }
  }
  if (best_t < 0) return cxh_fail("cxh_tape_tune: no setting could be launched: " + last_err);
  restore.on = false;
  t->block = pick.block;
  t->w_lds = pick.w_lds;
  t->ipw = pick.ipw;
  t->ipt = pick.ipt;
  if (best) {
    best[0] = pick.block;
    best[1] = pick.w_lds;
    best[2] = pick.ipw;
    best[3] = pick.ipt;
  }
  if (best_ms) *best_ms = best_t;
  return 0;
  CXH_CATCH
}

}  // extern "C"

namespace sxh {

// A refusal of a specialised kernel that would need scratch (the message of
// jit_function_checked and of the segment-length search).
static bool spill_refusal(const std::exception& e) {
  return std::string(e.what()).find("spills") != std::string::npos;
}

// Whether the specialised adjoint kernels at this checkpoint interval can be
// built (they are, in the calling thread); if not, the tape's adjoint at this
// interval runs on the interpreter from now on.  Decided before any launch,
// so that a split sweep keeps one engine's checkpoint layout.
static bool jit_adj_usable(cxh_tape_s* t, int64_t every) {
  if (std::find(t->adj_interp.begin(), t->adj_interp.end(), every) != t->adj_interp.end()) return false;
  try {
    jit_prepare(t, every, true, true);
    return true;
  } catch (const std::runtime_error& e) {
    if (!spill_refusal(e)) throw;
    t->adj_interp.push_back(every);
    t->fallback_why = e.what();
    return false;
  }
}

int map_eval_io(cxh_tape t, int64_t n, const IOTable& d_arg, const IOTable& d_res, hipStream_t s) {
  if (n == 0 || t->alg.empty()) return 0;
  if (t->engine == CXH_ENGINE_JIT && !t->eval_interp && jit_ready(t, -1, true, false)) {
    try {
      return jit_map_eval(t, n, d_arg, d_res, s);
    } catch (const std::runtime_error& e) {  // no segment length without scratch: the interpreter, same bits
      if (!spill_refusal(e)) throw;
      t->eval_interp = true;
      t->fallback_why = e.what();
    }
  }
  const GridShape g = grid_shape(t, n);
  const size_t lds = static_cast<size_t>(t->ws) * g.cols * sizeof(double);
  const bool use_lds = t->w_lds != 0 && lds <= static_cast<size_t>(device_lds_bytes());
  if (use_lds) {
    allow_lds(reinterpret_cast<const void*>(tape_fwd_kernel<true>), lds);
    hipLaunchKernelGGL(tape_fwd_kernel<true>, dim3(g.grid), dim3(g.bs), lds, s, device_tape(t),
                       static_cast<int>(t->alg.size()), static_cast<int>(t->ws), d_arg, d_res, n,
                       nullptr, int64_t(0), g.ipw);
  } else {
    t->scratch.reserve(static_cast<size_t>(t->ws) * g.slots * sizeof(double));
    hipLaunchKernelGGL(tape_fwd_kernel<false>, dim3(g.grid), dim3(g.bs), 0, s, device_tape(t),
                       static_cast<int>(t->alg.size()), static_cast<int>(t->ws), d_arg, d_res, n,
                       t->scratch.as<double>(), g.slots, g.ipw);
  }
  hip_check(hipGetLastError(), "tape_fwd_kernel launch");
  return 0;
}

}  // namespace sxh

extern "C" {

int cxh_map_eval_host(cxh_tape t, int64_t n, const double* const* arg, double* const* res) {
  CXH_TRY
  if (!t) return cxh_fail("null tape");
  if (n <= 0) return 0;
  const size_t ni = t->nnz_in.size(), no = t->nnz_out.size();
  std::vector<DeviceBuffer> din(ni), dout(no);
  std::vector<const double*> pa(ni, nullptr);
  std::vector<double*> pr(no, nullptr);
  for (size_t i = 0; i < ni; ++i) {
    if (arg[i] && t->nnz_in[i] > 0) {
      din[i].upload(arg[i], static_cast<size_t>(n) * t->nnz_in[i] * sizeof(double));
      pa[i] = din[i].as<double>();
    }
  }
  for (size_t i = 0; i < no; ++i) {
    if (res[i] && t->nnz_out[i] > 0) {
      dout[i].reserve(static_cast<size_t>(n) * t->nnz_out[i] * sizeof(double));
      pr[i] = dout[i].as<double>();
    }
  }
  int rc = cxh_map_eval(t, n, pa.data(), pr.data(), CXH_AOS, n, nullptr);
  if (rc) return rc;
  for (size_t i = 0; i < no; ++i)
    if (pr[i])
      hip_check(hipMemcpy(res[i], pr[i], static_cast<size_t>(n) * t->nnz_out[i] * sizeof(double),
                          hipMemcpyDeviceToHost),
                "hipMemcpy D2H");
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return 0;
  CXH_CATCH
}

}  // extern "C"

namespace sxh {

static void segments(const cxh_tape_s* t, int64_t every, int& seg_len, int& n_seg) {
  const int64_t ni = static_cast<int64_t>(t->alg.size());
  if (every <= 0 || every >= ni) {
    seg_len = static_cast<int>(std::max<int64_t>(ni, 1));
    n_seg = 1;
  } else {
    seg_len = static_cast<int>(every);
    n_seg = static_cast<int>((ni + every - 1) / every);
  }
}

static int launch_adj(cxh_tape t, int mode, int64_t n, const double* const* arg, const double* const* aseed,
                      double* const* res, double* const* asens, int32_t layout, int64_t ld,
                      int64_t every, double* ckpt_ext, hipStream_t s) {
  if (!t) return cxh_fail("null tape");
  if (n < 0) return cxh_fail("negative n");
  if (layout != CXH_SOA && layout != CXH_AOS) return cxh_fail("bad layout");
  if (layout == CXH_SOA && ld < n) return cxh_fail("ld < n");
  check_status_cap(t, n);
  const IOTable R = make_io(res, t->nnz_out, layout, ld), G = make_io(asens, t->nnz_in, layout, ld);
  const int rc = adj_io(t, mode, n, make_io(arg, t->nnz_in, layout, ld), make_io(aseed, t->nnz_out, layout, ld), R,
                        G, every, ckpt_ext, s);
  if (rc) return rc;
  // what the call wrote: outputs (modes 0 and 1), sensitivities (0 and 2)
  return regularity(t, n, mode == 2 ? IOTable{} : R, t->nnz_out, mode == 1 ? IOTable{} : G, t->nnz_in, s);
}

std::vector<uint32_t> adjoint_live(const std::vector<Instr>& alg, int64_t ws) {
  std::vector<uint32_t> bits((alg.size() + 31) / 32, 0u);
  std::vector<char> act(static_cast<size_t>(std::max<int64_t>(ws, 1)), 0);
  for (int64_t k = static_cast<int64_t>(alg.size()) - 1; k >= 0; --k) {
    const Instr& e = alg[k];
    if (e.op == OP_OUTPUT) {
      act[e.i1] = 1;
      continue;
    }
    const bool live = act[e.i0] != 0;
    act[e.i0] = 0;
    if (live) bits[k >> 5] |= 1u << (k & 31);
    if (!live || e.op == OP_CONST || e.op == OP_INPUT) continue;
    if (e.op == OP_IF_ELSE_ZERO) {
      act[e.i2] = 1;
    } else {
      if (!op_zero_partial(e.op, 0)) act[e.i1] = 1;
      if (op_is_binary(e.op) && !op_zero_partial(e.op, 1)) act[e.i2] = 1;
    }
  }
  return bits;
}

int adj_io(cxh_tape t, int mode, int64_t n, const IOTable& arg, const IOTable& aseed, const IOTable& res,
           const IOTable& asens, int64_t every, double* ckpt_ext, hipStream_t s) {
  if (n == 0 || t->alg.empty()) return 0;
  if (t->engine == CXH_ENGINE_JIT) {
    // the split sweep (modes 1 and 2) keeps one engine's checkpoint layout
    // across its two calls, so it never changes engine half way
    if (mode != 0) jit_settle(t);
    if ((mode != 0 || jit_ready(t, every, true, true)) && jit_adj_usable(t, every))
      return jit_adj(t, mode, n, arg, aseed, res, asens, every, ckpt_ext, s);
  }
  if (!t->d_live.ptr) {
    const std::vector<uint32_t> bits = adjoint_live(t->alg, t->ws);
    t->d_live.upload(bits.data(), bits.size() * sizeof(uint32_t));
  }
  AdjArgs A{};
  A.live = t->d_live.as<uint32_t>();
  A.tape = device_tape(t);
  A.n_instr = static_cast<int>(t->alg.size());
  A.ws = static_cast<int>(t->ws);
  segments(t, every, A.seg_len, A.n_seg);
  A.arg = arg;
  A.res = res;
  A.aseed = aseed;
  A.asens = asens;
  A.unused_in = t->d_unused.as<int2>();
  A.n_unused = static_cast<int>(t->unused_in.size());
  A.n = n;
  A.mode = mode;
  const GridShape g = grid_shape(t, n);
  A.ipw = g.ipw;
  A.lanes = g.slots;
  if (mode != 1) t->trail.reserve(static_cast<size_t>(A.seg_len) * A.lanes * sizeof(double));
  A.trail = t->trail.as<double>();
  if (ckpt_ext) {
    A.ckpt = ckpt_ext;
  } else if (A.n_seg > 1) {
    t->ckpt.reserve(static_cast<size_t>(A.n_seg) * A.ws * A.lanes * sizeof(double));
    A.ckpt = t->ckpt.as<double>();
  }
  const size_t lds_tr = 64 * static_cast<size_t>(g.cols) * sizeof(double);  // trail chunk staging
  const size_t lds = lds_tr + 2 * static_cast<size_t>(A.ws) * g.cols * sizeof(double);
  const bool use_lds = t->w_lds != 0 && lds <= static_cast<size_t>(device_lds_bytes());
  if (use_lds) {
    allow_lds(reinterpret_cast<const void*>(tape_adj_kernel<true>), lds);
    hipLaunchKernelGGL(tape_adj_kernel<true>, dim3(g.grid), dim3(g.bs), lds, s, A);
  } else {
    t->scratch.reserve(2 * static_cast<size_t>(A.ws) * A.lanes * sizeof(double));
    A.wglob = t->scratch.as<double>();
    allow_lds(reinterpret_cast<const void*>(tape_adj_kernel<false>), lds_tr);
    hipLaunchKernelGGL(tape_adj_kernel<false>, dim3(g.grid), dim3(g.bs), lds_tr, s, A);
  }
  hip_check(hipGetLastError(), "tape_adj_kernel launch");
  return 0;
}

}  // namespace sxh

extern "C" {

int cxh_map_fwd_adj(cxh_tape t, int64_t n, const double* const* arg, const double* const* adj_out,
                    double* const* res, double* const* adj_in, int32_t layout, int64_t ld,
                    int64_t checkpoint_every, void* stream) {
  CXH_TRY
  return launch_adj(t, 0, n, arg, adj_out, res, adj_in, layout, ld, checkpoint_every, nullptr,
                    static_cast<hipStream_t>(stream));
  CXH_CATCH
}

int cxh_ckpt_size(cxh_tape t, int64_t n, int64_t checkpoint_every, int32_t block, int64_t* n_doubles) {
  CXH_TRY
  if (!t || !n_doubles) return cxh_fail("null argument");
  (void)block;
  if (t->engine == CXH_ENGINE_JIT) {
    jit_settle(t);
    if (t->alg.empty()) {
      *n_doubles = 0;
      return 0;
    }
    // which engine the split sweep will run on is settled by building its
    // kernels (with a device present; without one the size is the plan's)
    int n_dev = 0;
    const bool device = hipGetDeviceCount(&n_dev) == hipSuccess && n_dev > 0;
    if (!device || jit_adj_usable(t, checkpoint_every)) {
      *n_doubles = jit_ckpt_rows(t, checkpoint_every) * round_up(n, 64);
      return 0;
    }
  }
  int seg_len, n_seg;
  segments(t, checkpoint_every, seg_len, n_seg);
  *n_doubles = static_cast<int64_t>(n_seg) * t->ws * grid_shape(t, n).slots;
  return 0;
  CXH_CATCH
}

int cxh_map_fwd_ckpt(cxh_tape t, int64_t n, const double* const* arg, double* const* res, int32_t layout,
                     int64_t ld, int64_t checkpoint_every, double* ckpt, void* stream) {
  CXH_TRY
  if (!ckpt) return cxh_fail("null checkpoint buffer");
  return launch_adj(t, 1, n, arg, nullptr, res, nullptr, layout, ld, checkpoint_every, ckpt,
                    static_cast<hipStream_t>(stream));
  CXH_CATCH
}

int cxh_map_rev_ckpt(cxh_tape t, int64_t n, const double* const* arg, const double* const* adj_out,
                     double* const* adj_in, int32_t layout, int64_t ld, int64_t checkpoint_every,
                     const double* ckpt, void* stream) {
  CXH_TRY
  if (!ckpt) return cxh_fail("null checkpoint buffer");
  return launch_adj(t, 2, n, arg, adj_out, nullptr, adj_in, layout, ld, checkpoint_every,
                    const_cast<double*>(ckpt), static_cast<hipStream_t>(stream));
  CXH_CATCH
}

int cxh_jac_create(cxh_tape t, int32_t forward, int64_t ndir, const int64_t* seed_ptr, const int64_t* seed_idx,
                   int64_t nnz_jac, const int64_t* src_dir, const int64_t* src_nz, cxh_jac* out) {
  CXH_TRY
  if (!t || !out) return cxh_fail("null argument");
  if (!forward) return cxh_fail("cxh_jac_create: only forward (D1) plans run on the tangent kernel");
  if (ndir < 0 || ndir > 4096) return cxh_fail("cxh_jac_create: bad number of directions");
  auto j = std::make_unique<cxh_jac_s>();
  j->tape = t;
  j->ndir = static_cast<int>(ndir);
  j->nnz_jac = nnz_jac;
  int64_t nin = 0, nout = 0;
  for (int64_t v : t->nnz_in) {
    j->in_off.push_back(static_cast<int>(nin));
    nin += v;
  }
  for (int64_t v : t->nnz_out) {
    j->out_off.push_back(static_cast<int>(nout));
    nout += v;
  }
  std::vector<double> seedval(static_cast<size_t>(std::max<int64_t>(nin, 1)) * std::max<int64_t>(ndir, 1), 0.0);
  for (int64_t d = 0; d < ndir; ++d)
    for (int64_t q = seed_ptr[d]; q < seed_ptr[d + 1]; ++q) {
      int64_t c = seed_idx[q];
      if (c < 0 || c >= nin) throw std::runtime_error("cxh_jac_create: seed index out of range");
      seedval[c * ndir + d] = 1.0;
    }
  std::vector<int2> scatter(static_cast<size_t>(std::max<int64_t>(nout, 1)) * std::max<int64_t>(ndir, 1),
                            int2{-1, -1});
  for (int64_t k = 0; k < nnz_jac; ++k) {
    if (src_dir[k] < 0) continue;
    if (src_dir[k] >= ndir || src_nz[k] < 0 || src_nz[k] >= nout)
      throw std::runtime_error("cxh_jac_create: scatter source out of range");
    int2& tg = scatter[src_nz[k] * ndir + src_dir[k]];
    if (tg.x < 0) tg.x = static_cast<int>(k);
    else if (tg.y < 0) tg.y = static_cast<int>(k);
    else throw std::runtime_error("cxh_jac_create: more than two targets per sensitivity");
  }
  j->seedval = seedval;
  j->scatter = scatter;
  *out = j.release();
  return 0;
  CXH_CATCH
}

int cxh_jac_destroy(cxh_jac j) {
  delete j;
  return 0;
}

int cxh_map_jac(cxh_jac j, int64_t n, const double* const* arg, double* const* res, double* jac,
                int32_t layout, int64_t ld, void* stream) {
  CXH_TRY
  if (!j) return cxh_fail("null jac plan");
  cxh_tape t = j->tape;
  if (n < 0) return cxh_fail("negative n");
  if (layout != CXH_SOA && layout != CXH_AOS) return cxh_fail("bad layout");
  if (layout == CXH_SOA && ld < n) return cxh_fail("ld < n");
  if (n == 0 || t->alg.empty()) return 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t jac_s_nz = layout == CXH_SOA ? ld : 1;
  const int64_t jac_s_inst = layout == CXH_SOA ? 1 : j->nnz_jac;
  if (t->engine == CXH_ENGINE_JIT)
    return jit_map_jac(j, n, make_io(arg, t->nnz_in, layout, ld), make_io(res, t->nnz_out, layout, ld), jac,
                       jac_s_nz, jac_s_inst, s);
  if (!j->d_seed.ptr) {  // device copies of the plan, at the first interpreter launch
    j->d_seed.upload(j->seedval.data(), j->seedval.size() * sizeof(double));
    j->d_scatter.upload(j->scatter.data(), j->scatter.size() * sizeof(int2));
    j->d_in_off.upload(j->in_off.data(), j->in_off.size() * sizeof(int));
    j->d_out_off.upload(j->out_off.data(), j->out_off.size() * sizeof(int));
  }
  JacArgs A{};
  A.tape = device_tape(t);
  A.n_instr = static_cast<int>(t->alg.size());
  A.ws = static_cast<int>(t->ws);
  A.ndir = j->ndir;
  A.arg = make_io(arg, t->nnz_in, layout, ld);
  A.res = make_io(res, t->nnz_out, layout, ld);
  A.in_off = j->d_in_off.as<int>();
  A.out_off = j->d_out_off.as<int>();
  A.seedval = j->d_seed.as<double>();
  A.scatter = j->d_scatter.as<int2>();
  A.jac = jac;
  A.jac_s_nz = jac_s_nz;
  A.jac_s_inst = jac_s_inst;
  A.n = n;
  const GridShape g = grid_shape(t, n);
  A.ipw = g.ipw;
  A.lanes = g.slots;
  const size_t lds = static_cast<size_t>(1 + A.ndir) * A.ws * g.cols * sizeof(double);
  const bool use_lds = t->w_lds != 0 && lds <= static_cast<size_t>(device_lds_bytes());
  if (use_lds) {
    allow_lds(reinterpret_cast<const void*>(tape_jac_kernel<true>), lds);
    hipLaunchKernelGGL(tape_jac_kernel<true>, dim3(g.grid), dim3(g.bs), lds, s, A);
  } else {
    t->scratch.reserve(static_cast<size_t>(1 + A.ndir) * A.ws * A.lanes * sizeof(double));
    A.wglob = t->scratch.as<double>();
    hipLaunchKernelGGL(tape_jac_kernel<false>, dim3(g.grid), dim3(g.bs), 0, s, A);
  }
  hip_check(hipGetLastError(), "tape_jac_kernel launch");
  return 0;
  CXH_CATCH
}

int cxh_device_info(int32_t* n_cu, int32_t* lds_per_cu_kb, char* name, int32_t name_len) {
  CXH_TRY
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hipDeviceProp_t p;
  hip_check(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
  if (n_cu) *n_cu = p.multiProcessorCount;
  if (lds_per_cu_kb) *lds_per_cu_kb = static_cast<int32_t>(p.sharedMemPerMultiprocessor / 1024);
  if (name && name_len > 0) {
    std::strncpy(name, p.gcnArchName, name_len - 1);
    name[name_len - 1] = 0;
  }
  return 0;
  CXH_CATCH
}

}  // extern "C"
