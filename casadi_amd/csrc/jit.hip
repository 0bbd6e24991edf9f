// Tape -> HIP specialiser: straight-line gfx950 kernels for one SX tape.
//
// The interpreter kernels (tape_kernels.hip) walk the instruction stream at
// run time; every instruction costs a wave-uniform fetch, a dispatch branch
// and indexed work-vector traffic.  Here the stream is instead unrolled into
// HIP source once per tape, with every tape slot value an SSA variable (held
// in VGPRs, spilled by the compiler where needed), compiled through hiprtc for
// the current device and launched with one lane per instance.  This is the GPU
// counterpart of CasADi's C code generation of an SXFunction
// (SXFunction::codegen_body, casadi/core/sx_function.cpp:344-434), with the
// numeric forward/reverse sweep semantics of SXFunction::eval
// (casadi/core/sx_function.cpp:72-127) and SXFunction::ad_reverse
// (casadi/core/sx_function.cpp:1368-1569).
//
// Long tapes are cut into checkpoint segments of `every` instructions.  The
// state crossing a segment boundary is only the set of slots live there
// (written before, read after), stored [row][instance] in HBM.  Per segment
// there are two kernels:
//   fwd s: load the live-in state, run the segment, write outputs and the
//          live-out state;
//   rev s: load the live-in state, recompute the segment, then sweep it
//          backwards: adjoints of the live-out state come in, adjoints of the
//          live-in state go out, input sensitivities are stored where their
//          INPUT instruction is met.
// A fused forward+adjoint is fwd 0..S-2 followed by rev S-1..0 (the last
// segment is computed once, inside its reverse kernel).  Each segment is its
// own hiprtc program, so compile cost grows linearly with the tape and kernels
// are only built for the modes that are used.
//
// Arithmetic is the same as the interpreter's: op_fun_full / op_der_full from
// sx_ops.h (compiled into every program), -ffp-contract=off, and the adjoint of
// each value accumulates its contributions in the same (reverse tape) order.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/casadi_hip.h"
#include "capi_common.hpp"
#include "sx_ops.h"
#include "tape_internal.hpp"

namespace sxh {

static const char kSxOpsSource[] =
#include "sx_ops_rtc.inc"
    ;

// Kernel argument block; the generated source declares the same layout.
struct CxjArgs {
  IOTable arg, res, aseed, asens;
  const double* st_in;
  double* st_out;
  const double* adj_in;
  double* adj_out;
  long long n, lds;
};
static_assert(sizeof(IOArg) == 24, "IOArg layout");
static_assert(sizeof(CxjArgs) == 4 * CXH_MAX_IO * 24 + 6 * 8, "CxjArgs layout");

static const char kPrologue[] = R"CXJ(
struct IOArg { const double* p; long long s_nz; long long s_inst; };
struct IOTable { IOArg a[16]; };
struct CxjArgs {
  IOTable arg, res, aseed, asens;
  const double* st_in; double* st_out; const double* adj_in; double* adj_out;
  long long n, lds;
};
#define CXJ_AT(t, i, j) ((t).a[i].p + (long long)(j) * (t).a[i].s_nz + k * (t).a[i].s_inst)
#define CXJ_IN(i, j) ((A.arg.a[i].p) ? *CXJ_AT(A.arg, i, j) : 0.0)
#define CXJ_OUT(i, j, v) do { if (A.res.a[i].p) *const_cast<double*>(CXJ_AT(A.res, i, j)) = (v); } while (0)
#define CXJ_SEED(i, j) ((A.aseed.a[i].p) ? *CXJ_AT(A.aseed, i, j) : 0.0)
#define CXJ_SENS(i, j, v) do { if (A.asens.a[i].p) *const_cast<double*>(CXJ_AT(A.asens, i, j)) = (v); } while (0)
#define CXJ_ST_IN(r) A.st_in[(long long)(r) * A.lds + k]
#define CXJ_ST_OUT(r) A.st_out[(long long)(r) * A.lds + k]
#define CXJ_ADJ_IN(r) A.adj_in[(long long)(r) * A.lds + k]
#define CXJ_ADJ_OUT(r) A.adj_out[(long long)(r) * A.lds + k]
)CXJ";

// ---------------------------------------------------------------- plans ---

// Segments of a tape for one checkpoint interval, with the live slots at each
// boundary and their rows in the checkpoint buffer.
struct JitPlan {
  int n_seg = 1;
  std::vector<int> t0;                // n_seg + 1 boundaries
  std::vector<std::vector<int>> live;  // live slots at each boundary
  std::vector<int64_t> ck_off;        // first checkpoint row of boundary s
  int64_t ck_rows = 0;
  int max_live = 0;
  std::vector<hipFunction_t> fwd, rev;
};

struct JitTape {
  std::map<int64_t, JitPlan> plans;  // keyed by segment length
  DeviceBuffer st, adj, ck;
};

static int64_t default_every(int64_t n_instr) { return n_instr <= 4096 ? n_instr : 2048; }

static JitPlan& plan_for(cxh_tape_s* t, int64_t every) {
  if (!t->jit) t->jit = std::make_shared<JitTape>();
  const int64_t ni = static_cast<int64_t>(t->alg.size());
  if (every <= 0 || every >= ni) every = std::max<int64_t>(ni, 1);
  auto it = t->jit->plans.find(every);
  if (it != t->jit->plans.end()) return it->second;
  JitPlan P;
  P.n_seg = static_cast<int>((ni + every - 1) / every);
  for (int s = 0; s <= P.n_seg; ++s) P.t0.push_back(static_cast<int>(std::min<int64_t>(ni, s * every)));
  // backward liveness: live before t = (live after t - written(t)) + read(t)
  P.live.assign(P.n_seg + 1, {});
  std::vector<char> live(t->ws, 0);
  int b = P.n_seg;
  auto snap = [&](int s) {
    std::vector<int>& L = P.live[s];
    L.clear();
    for (int64_t j = 0; j < t->ws; ++j)
      if (live[j]) L.push_back(static_cast<int>(j));
  };
  snap(b);
  for (int64_t k = ni - 1; k >= 0; --k) {
    const Instr& e = t->alg[k];
    if (e.op == OP_OUTPUT) {
      live[e.i1] = 1;
    } else {
      live[e.i0] = 0;
      if (op_is_arith(e.op)) {
        live[e.i1] = 1;
        if (op_is_binary(e.op)) live[e.i2] = 1;
      }
    }
    while (b > 0 && P.t0[b - 1] == k) snap(--b);
  }
  P.ck_off.assign(P.n_seg + 1, 0);
  int64_t rows = 0;
  for (int s = 1; s < P.n_seg; ++s) {
    P.ck_off[s] = rows;
    rows += static_cast<int64_t>(P.live[s].size());
  }
  P.ck_rows = rows;
  for (auto& L : P.live) P.max_live = std::max<int>(P.max_live, static_cast<int>(L.size()));
  P.fwd.assign(P.n_seg, nullptr);
  P.rev.assign(P.n_seg, nullptr);
  return t->jit->plans.emplace(every, std::move(P)).first->second;
}

// ------------------------------------------------------------ codegen ---

static std::string lit(double d) {
  if (std::isnan(d)) return "__builtin_nan(\"\")";
  if (std::isinf(d)) return d > 0 ? "__builtin_huge_val()" : "(-__builtin_huge_val())";
  char b[64];
  std::snprintf(b, sizeof b, "%a", d);
  return b;
}

// Value names: instruction t's result is v<t>; the live-in value of slot j is
// w<j>; their adjoints are a<t> / b<j>.
struct Names {
  std::vector<int64_t> cur;  // slot -> code (t >= 0, or -(j+1) for live-in)
  static std::string val(int64_t c) { return c >= 0 ? "v" + std::to_string(c) : "w" + std::to_string(-c - 1); }
  static std::string adj(int64_t c) { return c >= 0 ? "a" + std::to_string(c) : "b" + std::to_string(-c - 1); }
};

static std::string gen_segment(const cxh_tape_s* t, const JitPlan& P, int s, bool rev) {
  const int t0 = P.t0[s], t1 = P.t0[s + 1];
  std::string o;
  o.reserve(static_cast<size_t>(t1 - t0) * (rev ? 160 : 64) + 4096);
  o += kSxOpsSource;
  o += kPrologue;
  o += "extern \"C\" __global__ void __launch_bounds__(256) cxj_kernel(const CxjArgs A) {\n";
  o += "  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;\n";
  o += "  if (k >= A.n) return;\n";
  Names N;
  N.cur.assign(t->ws, 0);
  const std::vector<int>& lin = P.live[s];
  const std::vector<int>& lout = P.live[s + 1];
  for (size_t r = 0; r < lin.size(); ++r) {
    const int j = lin[r];
    N.cur[j] = -(j + 1);
    o += "  const double w" + std::to_string(j) + " = ";
    o += s > 0 ? "CXJ_ST_IN(" + std::to_string(r) + ");\n" : "0.0;\n";
  }
  // operand codes per instruction, for the reverse sweep
  std::vector<int64_t> ox(t1 - t0, 0), oy(t1 - t0, 0);
  char buf[256];
  for (int k = t0; k < t1; ++k) {
    const Instr& e = t->alg[k];
    switch (e.op) {
      case OP_CONST:
        o += "  const double v" + std::to_string(k) + " = " + lit(e.d) + ";\n";
        N.cur[e.i0] = k;
        break;
      case OP_INPUT:
        std::snprintf(buf, sizeof buf, "  const double v%d = CXJ_IN(%d, %d);\n", k, e.i1, e.i2);
        o += buf;
        N.cur[e.i0] = k;
        break;
      case OP_OUTPUT:
        ox[k - t0] = N.cur[e.i1];
        std::snprintf(buf, sizeof buf, "  CXJ_OUT(%d, %d, %s);\n", e.i0, e.i2, Names::val(N.cur[e.i1]).c_str());
        o += buf;
        break;
      default: {
        const int64_t x = N.cur[e.i1];
        const int64_t y = op_is_binary(e.op) ? N.cur[e.i2] : x;
        ox[k - t0] = x;
        oy[k - t0] = y;
        std::snprintf(buf, sizeof buf, "  const double v%d = sxh::op_fun_full(%d, %s, %s);\n", k, e.op,
                      Names::val(x).c_str(), Names::val(y).c_str());
        o += buf;
        N.cur[e.i0] = k;
      }
    }
  }
  if (!rev) {
    if (s + 1 < P.n_seg)
      for (size_t r = 0; r < lout.size(); ++r)
        o += "  CXJ_ST_OUT(" + std::to_string(r) + ") = " + Names::val(N.cur[lout[r]]) + ";\n";
    o += "}\n";
    return o;
  }
  // reverse sweep
  std::vector<char> dv(t1 - t0, 0), dw(t->ws, 0);
  auto declared = [&](int64_t c) -> char& { return c >= 0 ? dv[c - t0] : dw[-c - 1]; };
  auto acc = [&](int64_t c, const std::string& term) {
    char& d = declared(c);
    if (d) {
      o += "  " + Names::adj(c) + " += " + term + ";\n";
    } else {
      o += "  double " + Names::adj(c) + " = " + term + ";\n";
      d = 1;
    }
  };
  if (s + 1 < P.n_seg)
    for (size_t r = 0; r < lout.size(); ++r) {
      const int64_t c = N.cur[lout[r]];
      o += "  double " + Names::adj(c) + " = CXJ_ADJ_IN(" + std::to_string(r) + ");\n";
      declared(c) = 1;
    }
  for (int k = t1 - 1; k >= t0; --k) {
    const Instr& e = t->alg[k];
    switch (e.op) {
      case OP_CONST:
        break;
      case OP_OUTPUT:
        std::snprintf(buf, sizeof buf, "CXJ_SEED(%d, %d)", e.i0, e.i2);
        acc(ox[k - t0], buf);
        break;
      case OP_INPUT:
        std::snprintf(buf, sizeof buf, "  CXJ_SENS(%d, %d, %s);\n", e.i1, e.i2,
                      dv[k - t0] ? Names::adj(k).c_str() : "0.0");
        o += buf;
        break;
      default: {
        if (!dv[k - t0]) break;  // no contribution reaches this value
        const std::string ks = std::to_string(k);
        const std::string x = Names::val(ox[k - t0]), y = Names::val(oy[k - t0]);
        o += "  double d" + ks + "[2];\n  sxh::op_der_full(" + std::to_string(e.op) + ", " + x + ", " + y +
             ", v" + ks + ", d" + ks + ");\n";
        if (e.op == OP_IF_ELSE_ZERO) {
          acc(oy[k - t0], "sxh::sxh_if_else_zero(d" + ks + "[1], a" + ks + ")");
        } else if (op_is_binary(e.op)) {
          acc(ox[k - t0], "d" + ks + "[0] * a" + ks);
          acc(oy[k - t0], "d" + ks + "[1] * a" + ks);
        } else {
          acc(ox[k - t0], "d" + ks + "[0] * a" + ks);
        }
      }
    }
  }
  if (s > 0)
    for (size_t r = 0; r < lin.size(); ++r) {
      const int64_t c = -(lin[r] + 1);
      o += "  CXJ_ADJ_OUT(" + std::to_string(r) + ") = " + (dw[lin[r]] ? Names::adj(c) : std::string("0.0")) +
           ";\n";
    }
  if (s == 0)
    for (const int2& u : t->unused_in) {
      std::snprintf(buf, sizeof buf, "  CXJ_SENS(%d, %d, 0.0);\n", u.x, u.y);
      o += buf;
    }
  o += "}\n";
  return o;
}

// ------------------------------------------------------------ compile ---

static std::string device_arch() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hipDeviceProp_t p;
  hip_check(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
  std::string a = p.gcnArchName;
  return a.substr(0, a.find(':'));
}

// Compiles one program; returns the code object or throws with the log.
static std::vector<char> rtc_compile(const std::string& src, const std::string& arch) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "cxj_segment.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    throw std::runtime_error("hiprtcCreateProgram failed");
  const std::string off = "--offload-arch=" + arch;
  const char* opts[] = {off.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult r = hiprtcCompileProgram(prog, sizeof opts / sizeof opts[0], opts);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    throw std::runtime_error("hiprtc compile failed: " + log.substr(0, 4000));
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  std::vector<char> code(n);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  return code;
}

// Process-wide cache: identical sources share one loaded module.
static std::mutex g_jit_mu;
static std::unordered_map<std::string, hipFunction_t> g_jit_cache;

static hipFunction_t jit_function(const std::string& src) {
  std::lock_guard<std::mutex> lk(g_jit_mu);
  auto it = g_jit_cache.find(src);
  if (it != g_jit_cache.end()) return it->second;
  if (const char* dir = std::getenv("CXH_JIT_DUMP")) {
    const std::string path = std::string(dir) + "/cxj_" + std::to_string(g_jit_cache.size()) + ".hip";
    if (FILE* f = std::fopen(path.c_str(), "w")) {
      std::fwrite(src.data(), 1, src.size(), f);
      std::fclose(f);
    }
  }
  const std::vector<char> code = rtc_compile(src, device_arch());
  hipModule_t mod;
  hip_check(hipModuleLoadData(&mod, code.data()), "hipModuleLoadData");
  hipFunction_t fn;
  hip_check(hipModuleGetFunction(&fn, mod, "cxj_kernel"), "hipModuleGetFunction");
  g_jit_cache.emplace(src, fn);
  return fn;
}

static hipFunction_t segment_kernel(cxh_tape_s* t, JitPlan& P, int s, bool rev) {
  hipFunction_t& f = rev ? P.rev[s] : P.fwd[s];
  if (!f) f = jit_function(gen_segment(t, P, s, rev));
  return f;
}

static void launch(hipFunction_t f, int64_t n, int block, CxjArgs& A, hipStream_t s) {
  const unsigned grid = static_cast<unsigned>((n + block - 1) / block);
  void* params[] = {&A};
  hip_check(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, s, params, nullptr), "jit kernel launch");
}

// ------------------------------------------------------------ drivers ---

void jit_prepare(cxh_tape_s* t, int64_t every, bool fwd, bool rev) {
  JitPlan& P = plan_for(t, every);
  for (int s = 0; s < P.n_seg; ++s) {
    if (fwd) segment_kernel(t, P, s, false);
    if (rev) segment_kernel(t, P, s, true);
  }
}

int64_t jit_ckpt_rows(cxh_tape_s* t, int64_t every) { return plan_for(t, every).ck_rows; }

int jit_map_eval(cxh_tape_s* t, int64_t n, const IOTable& arg, const IOTable& res, hipStream_t s) {
  JitPlan& P = plan_for(t, default_every(static_cast<int64_t>(t->alg.size())));
  const int64_t lds = round_up(n, 64);
  CxjArgs A{};
  A.arg = arg;
  A.n = n;
  A.lds = lds;
  if (P.n_seg > 1) t->jit->st.reserve(2 * static_cast<size_t>(std::max(P.max_live, 1)) * lds * sizeof(double));
  double* st = t->jit->st.as<double>();
  for (int g = 0; g < P.n_seg; ++g) {
    hipFunction_t f = segment_kernel(t, P, g, false);
    A.res = res;
    A.st_in = st ? st + static_cast<int64_t>(g % 2) * P.max_live * lds : nullptr;
    A.st_out = st ? st + static_cast<int64_t>((g + 1) % 2) * P.max_live * lds : nullptr;
    launch(f, n, t->block, A, s);
  }
  return 0;
}

// mode 0: fused forward + adjoint; 1: forward storing checkpoints in ckpt_ext;
// 2: reverse from the checkpoints in ckpt_ext.
int jit_adj(cxh_tape_s* t, int mode, int64_t n, const IOTable& arg, const IOTable& aseed, const IOTable& res,
            const IOTable& asens, int64_t every, double* ckpt_ext, hipStream_t s) {
  JitPlan& P = plan_for(t, every);
  const int64_t lds = round_up(n, 64);
  double* ck = ckpt_ext;
  if (!ck && P.ck_rows) {
    t->jit->ck.reserve(static_cast<size_t>(P.ck_rows) * lds * sizeof(double));
    ck = t->jit->ck.as<double>();
  }
  auto state = [&](int b) -> double* { return (b > 0 && b < P.n_seg) ? ck + P.ck_off[b] * lds : nullptr; };
  IOTable none{};
  CxjArgs A{};
  A.arg = arg;
  A.n = n;
  A.lds = lds;
  const int last_fwd = mode == 1 ? P.n_seg : P.n_seg - 1;
  if (mode != 2) {
    for (int g = 0; g < last_fwd; ++g) {
      hipFunction_t f = segment_kernel(t, P, g, false);
      A.res = res;
      A.st_in = state(g);
      A.st_out = state(g + 1);
      launch(f, n, t->block, A, s);
    }
  }
  if (mode == 1) return 0;
  if (P.n_seg > 1) t->jit->adj.reserve(2 * static_cast<size_t>(P.max_live) * lds * sizeof(double));
  double* adj = t->jit->adj.as<double>();
  A.aseed = aseed;
  A.asens = asens;
  for (int g = P.n_seg - 1; g >= 0; --g) {
    hipFunction_t f = segment_kernel(t, P, g, true);
    A.res = (mode == 0 && g == P.n_seg - 1) ? res : none;
    A.st_in = state(g);
    A.st_out = nullptr;
    A.adj_in = adj ? adj + static_cast<int64_t>((g + 1) % 2) * P.max_live * lds : nullptr;
    A.adj_out = adj ? adj + static_cast<int64_t>(g % 2) * P.max_live * lds : nullptr;
    launch(f, n, t->block, A, s);
  }
  return 0;
}

std::string jit_source(cxh_tape_s* t, int64_t every, int seg, bool rev) {
  JitPlan& P = plan_for(t, every);
  if (seg < 0 || seg >= P.n_seg) throw std::runtime_error("segment index out of range");
  return gen_segment(t, P, seg, rev);
}

}  // namespace sxh

using namespace sxh;

extern "C" {

int cxh_tape_engine(cxh_tape t, int32_t engine) {
  if (!t) return cxh_fail("null tape");
  if (engine != CXH_ENGINE_INTERP && engine != CXH_ENGINE_JIT) return cxh_fail("engine must be 0 or 1");
  t->engine = engine;
  return 0;
}

int cxh_tape_jit_prepare(cxh_tape t, int64_t checkpoint_every, int32_t forward, int32_t reverse) {
  CXH_TRY
  if (!t) return cxh_fail("null tape");
  if (t->alg.empty()) return 0;
  jit_prepare(t, checkpoint_every, forward != 0, reverse != 0);
  return 0;
  CXH_CATCH
}

int cxh_tape_jit_source(cxh_tape t, int64_t checkpoint_every, int32_t segment, int32_t reverse, char* buf,
                        int64_t cap, int64_t* len, int32_t* n_segments) {
  CXH_TRY
  if (!t) return cxh_fail("null tape");
  if (n_segments) *n_segments = plan_for(t, checkpoint_every).n_seg;
  if (t->alg.empty()) {
    if (len) *len = 0;
    return 0;
  }
  const std::string src = jit_source(t, checkpoint_every, segment, reverse != 0);
  if (len) *len = static_cast<int64_t>(src.size());
  if (buf && cap > 0) {
    const size_t m = std::min<size_t>(src.size(), static_cast<size_t>(cap - 1));
    std::memcpy(buf, src.data(), m);
    buf[m] = '\0';
  }
  return 0;
  CXH_CATCH
}

int cxh_jit_compile_check(const char* src, const char* arch) {
  CXH_TRY
  if (!src || !arch) return cxh_fail("null argument");
  rtc_compile(src, arch);
  return 0;
  CXH_CATCH
}

}  // extern "C"
