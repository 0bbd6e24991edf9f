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

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/casadi_hip.h"
#include "capi_common.hpp"
#include "sx_ops.h"
#include "tape_internal.hpp"
#include "jit_common.hpp"

namespace sxh {

extern const char kSxOpsSource[] =
#include "sx_ops_rtc.inc"
    ;

extern const char kPrologue[] = R"CXJ(
struct IOArg { const double* p; long long s_nz; long long s_inst; };
struct IOTable { IOArg a[16]; };
struct CxjArgs {
  IOTable arg, res, aseed, asens;
  const double* st_in; double* st_out; const double* adj_in; double* adj_out;
  long long n, lds;
  double* spill;
};
#ifdef CXJ_SOA32
// every block of the launch is SoA (instance stride 1) and n * 8 < 2^32: the
// row start is wave-uniform and the instance is a 32-bit byte offset
#define CXJ_AT(t, i, j) ((const double*)((const char*)((t).a[i].p + (long long)(j) * (t).a[i].s_nz) + (unsigned)(k * 8)))
#else
#define CXJ_AT(t, i, j) ((t).a[i].p + (long long)(j) * (t).a[i].s_nz + k * (t).a[i].s_inst)
#endif
#define CXJ_IN(i, j) ((A.arg.a[i].p) ? *CXJ_AT(A.arg, i, j) : 0.0)
#define CXJ_OUT(i, j, v) do { if (A.res.a[i].p) *const_cast<double*>(CXJ_AT(A.res, i, j)) = (v); } while (0)
#define CXJ_SEED(i, j) ((A.aseed.a[i].p) ? *CXJ_AT(A.aseed, i, j) : 0.0)
#define CXJ_SENS(i, j, v) do { if (A.asens.a[i].p) *const_cast<double*>(CXJ_AT(A.asens, i, j)) = (v); } while (0)
#define CXJ_ST_IN(r) A.st_in[(long long)(r) * A.lds + k]
#define CXJ_ST_OUT(r) A.st_out[(long long)(r) * A.lds + k]
#define CXJ_ADJ_IN(r) A.adj_in[(long long)(r) * A.lds + k]
#define CXJ_ADJ_OUT(r) A.adj_out[(long long)(r) * A.lds + k]
#define CXJ_SNAP(p, r) (p)[(long long)(r) * A.lds + k]
#ifndef CXJ_KEEP
#define CXJ_KEEP(x) asm volatile("" : "+v"(x))
#define CXJ_KEEP_PTR(x) asm volatile("" : "+s"(x))
#define CXJ_AFTER(p, x) asm volatile("" : "+s"(p) : "v"(x))
#define CXJ_GLOBAL __attribute__((address_space(1)))
#define CXJ_OPAQUE(p) asm volatile("" : "+v"(p))
#define CXJ_OPAQUE_AFTER(p, x) asm volatile("" : "+v"(p) : "v"(x))
#endif
// whether any lane of the wavefront holds b
#ifndef CXJ_ANY
#define CXJ_ANY(b) (__builtin_amdgcn_ballot_w64(b) != 0)
#endif
#ifdef CXJ_TRIG_CALLS
#define CXJ_SINCOS(x_, s_, c_) do { const sxh::SinCos r_ = sxh::sxh_sincos_call(x_); s_ = r_.s; c_ = r_.c; } while (0)
#else
#define CXJ_SINCOS(x_, s_, c_) sxh::sxh_sincos(x_, &s_, &c_)
#endif
)CXJ";

// Sites of sin/cos pairs above which a program calls one out-of-line copy
// (CXH_JIT_TRIG_CALLS overrides; 0 always calls, a negative value never).
static int trig_call_sites() {
  if (const char* e = std::getenv("CXH_JIT_TRIG_CALLS")) return std::atoi(e);
  return 32;
}

// The finished program: switches its sin/cos pairs to calls when it has more
// than trig_call_sites() of them.
std::string finish_source(std::string o) {
  const int lim = trig_call_sites();
  if (lim < 0) return o;
  int64_t sites = 0;  // "  CXJ_SINCOS(" starts an emitted line; the definitions have one space
  for (size_t p = o.find("  CXJ_SINCOS("); p != std::string::npos; p = o.find("  CXJ_SINCOS(", p + 1)) ++sites;
  return sites > lim ? "#define CXJ_TRIG_CALLS\n" + o : o;
}

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
  int64_t snap_rows = 0;  // sub-block boundary states of one reverse kernel (max over segments)
  int ipt = 1;            // instances per thread of the forward kernels
  std::vector<hipFunction_t> fwd, rev;
  bool trig_cache = false;                     // pass 1 stores sin/cos pairs for the recomputation
  std::vector<int> n_sub;                      // sub-blocks of each segment's reverse kernel
  std::vector<hipFunction_t> snap, back, pair;  // a reverse kernel split: first pass, the rest, and
                                                // pair[g] = the rest of g beside the first pass of g-1
  // Forward-only evaluation keeps every value that crosses a boundary in one
  // home row for its whole life (a row is reused once its value is dead), so
  // a value that passes through a segment unread costs that segment nothing.
  // home[s][r]: the row of the value in live slot live[s][r] (boundaries
  // 1 .. n_seg-1); fresh[s][r]: that value is defined in segment s-1, which
  // stores it.  Built at the first evaluation (home_rows < 0 until then).
  std::vector<std::vector<int>> home;
  std::vector<std::vector<char>> fresh;
  int home_rows = -1;
  std::vector<hipFunction_t> fwd_home, fwd_home_soa;  // evaluation kernels on the home rows
};

// The home rows of a plan's boundaries: values are (slot, defining
// instruction); at each boundary the rows of values dead there are freed
// first, then each value new there takes a free row.
static void home_plan(const cxh_tape_s* t, JitPlan& P) {
  if (P.home_rows >= 0) return;
  P.home.assign(P.n_seg + 1, {});
  P.fresh.assign(P.n_seg + 1, {});
  std::vector<int64_t> def(static_cast<size_t>(t->ws), -1);
  std::map<std::pair<int, int64_t>, int> row_of;  // live value -> its row
  std::vector<int> free_rows;
  int rows = 0;
  int b = 1;
  for (int64_t k = 0; k <= static_cast<int64_t>(t->alg.size()) && b < P.n_seg; ++k) {
    while (b < P.n_seg && P.t0[b] == k) {
      const std::vector<int>& L = P.live[b];
      std::map<std::pair<int, int64_t>, int> next;
      for (int j : L) {
        auto it = row_of.find({j, def[j]});
        if (it != row_of.end()) next.insert(*it);
      }
      for (const auto& kv : row_of)
        if (!next.count(kv.first)) free_rows.push_back(kv.second);
      P.home[b].resize(L.size());
      P.fresh[b].resize(L.size());
      for (size_t r = 0; r < L.size(); ++r) {
        const std::pair<int, int64_t> v{L[r], def[L[r]]};
        auto it = next.find(v);
        if (it == next.end()) {
          int row;
          if (!free_rows.empty()) {
            row = free_rows.back();
            free_rows.pop_back();
          } else {
            row = rows++;
          }
          it = next.emplace(v, row).first;
          P.fresh[b][r] = 1;
        }
        P.home[b][r] = it->second;
      }
      row_of.swap(next);
      ++b;
    }
    if (k < static_cast<int64_t>(t->alg.size()) && t->alg[k].op != OP_OUTPUT) def[t->alg[k].i0] = k;
  }
  P.home_rows = rows;
  P.fwd_home.assign(P.n_seg, nullptr);
  P.fwd_home_soa.assign(P.n_seg, nullptr);
}

// Instances per thread of the forward segment kernels (CXH_JIT_IPT, 1-8):
// with several, the loads of the later instances are issued while the first
// one computes.
static int fwd_ipt() {
  if (const char* e = std::getenv("CXH_JIT_IPT")) return std::min(8, std::max(1, std::atoi(e)));
  return 1;
}

// The tape's instances per thread (cxh_tape_ipt) for a plan: forward kernels
// generated for another count are dropped and generated again (the source
// cache makes switching back cheap).
static void sync_ipt(const cxh_tape_s* t, JitPlan& P) {
  const int want = t->ipt > 0 ? t->ipt : fwd_ipt();
  if (P.ipt == want) return;
  P.ipt = want;
  std::fill(P.fwd.begin(), P.fwd.end(), nullptr);
  std::fill(P.fwd_home.begin(), P.fwd_home.end(), nullptr);
  std::fill(P.fwd_home_soa.begin(), P.fwd_home_soa.end(), nullptr);
}

struct JitTape {
  std::map<int64_t, JitPlan> plans;  // keyed by segment length
  int64_t eval_every = 0;            // segment length of forward-only evaluation (0: not chosen yet)
  DeviceBuffer st, adj, ck, snap;
};

// Whether reverse kernels run split in two, the halves of consecutive
// segments in one launch (CXH_JIT_SPLIT=0 keeps one kernel per segment).
bool split_reverse() {
  if (const char* e = std::getenv("CXH_JIT_SPLIT")) return std::atoi(e) != 0;
  return true;
}

// Whether the first pass of a reverse kernel stores the sin/cos pairs of its
// sub-blocks for their recomputation to load (CXH_JIT_TRIG_CACHE=0: computed
// again).  Fixed per plan when the plan is made.
bool trig_cache_enabled() {
  if (const char* e = std::getenv("CXH_JIT_TRIG_CACHE")) return std::atoi(e) != 0;
  return true;
}

// Instructions per sub-block of the reverse kernels (0: whole segment).
static int sub_block_len() {
  if (const char* e = std::getenv("CXH_JIT_SUBBLOCK")) return std::max(0, std::atoi(e));
  return 256;
}

// Slots live at each of the boundaries bnd (ascending, inside [t0, t1]),
// given the live set at t1.
static std::vector<std::vector<int>> live_sets(const cxh_tape_s* t, int t0, int t1, const std::vector<int>& live_end,
                                               const std::vector<int>& bnd) {
  std::vector<std::vector<int>> out(bnd.size());
  std::vector<char> live(t->ws, 0);
  for (int j : live_end) live[j] = 1;
  int b = static_cast<int>(bnd.size()) - 1;
  auto snap = [&](int i) {
    for (int64_t j = 0; j < t->ws; ++j)
      if (live[j]) out[i].push_back(static_cast<int>(j));
  };
  while (b >= 0 && bnd[b] == t1) snap(b--);
  for (int k = t1 - 1; k >= t0 && b >= 0; --k) {
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
    while (b >= 0 && bnd[b] == k) snap(b--);
  }
  return out;
}

// Sub-block boundaries of the reverse kernel of [t0, t1).
static std::vector<int> sub_bounds(int t0, int t1) {
  const int sub = sub_block_len();
  std::vector<int> c{t0};
  if (sub > 0)
    for (int x = t0 + sub; x < t1 - sub / 4; x += sub) c.push_back(x);
  c.push_back(t1);
  return c;
}

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
  P.n_sub.assign(P.n_seg, 1);
  P.trig_cache = trig_cache_enabled();
  for (int s = 0; s < P.n_seg; ++s) {
    const std::vector<int> c = sub_bounds(P.t0[s], P.t0[s + 1]);
    const auto lc = live_sets(t, P.t0[s], P.t0[s + 1], P.live[s + 1], c);
    int64_t rows = 0;
    for (size_t j = 1; j + 1 < c.size(); ++j) rows += static_cast<int64_t>(lc[j].size());
    if (P.trig_cache) {  // two rows per sincos site of sub-blocks 0..m-2 (one site per operand per sub-block)
      std::vector<int64_t> code(t->ws);
      for (int64_t q = 0; q < t->ws; ++q) code[q] = -(q + 1);
      for (size_t j = 0; j + 2 < c.size(); ++j) {
        std::unordered_set<int64_t> seen;
        for (int k = c[j]; k < c[j + 1]; ++k) {
          const Instr& e = t->alg[k];
          if ((e.op == OP_SIN || e.op == OP_COS) && seen.insert(code[e.i1]).second) rows += 2;
          if (e.op != OP_OUTPUT) code[e.i0] = k;
        }
      }
    }
    P.snap_rows = std::max(P.snap_rows, rows);
    P.n_sub[s] = static_cast<int>(c.size()) - 1;
  }
  P.fwd.assign(P.n_seg, nullptr);
  P.rev.assign(P.n_seg, nullptr);
  P.snap.assign(P.n_seg, nullptr);
  P.back.assign(P.n_seg, nullptr);
  P.pair.assign(P.n_seg, nullptr);
  P.ipt = t->ipt > 0 ? t->ipt : fwd_ipt();
  return t->jit->plans.emplace(every, std::move(P)).first->second;
}

// ------------------------------------------------------------ codegen ---

bool fast_div_enabled() {
  const char* e = std::getenv("CXH_JIT_FAST_DIV");
  return !(e && e[0] == '0');
}

std::string lit(double d) {
  if (std::isnan(d)) return "__builtin_nan(\"\")";
  if (std::isinf(d)) return d > 0 ? "__builtin_huge_val()" : "(-__builtin_huge_val())";
  char b[64];
  std::snprintf(b, sizeof b, "%a", d);
  return b;
}

// Value codes: the live-in value of slot j is -(j+1); the value instruction t
// produces in pass p is 4t+p (p = 0: first evaluation, p = 1: recomputation
// inside the reverse kernel); the copy of slot j's value taken at the start of
// recomputed sub-block b is 4(b*ws + j)+2.  Names: w<j>, v<t>, r<t>, h<b>_<j>;
// adjoints b<j>, a<t>, q<t>, g<b>_<j>.
static std::string val_name(int64_t c, int64_t ws) {
  if (c < 0) return "w" + std::to_string(-c - 1);
  const int64_t x = c >> 2;
  switch (c & 3) {
    case 0: return "v" + std::to_string(x);
    case 1: return "r" + std::to_string(x);
    default: return "h" + std::to_string(x / ws) + "_" + std::to_string(x % ws);
  }
}
static std::string adj_name(int64_t c, int64_t ws) {
  if (c < 0) return "b" + std::to_string(-c - 1);
  const int64_t x = c >> 2;
  switch (c & 3) {
    case 0: return "a" + std::to_string(x);
    case 1: return "q" + std::to_string(x);
    default: return "g" + std::to_string(x / ws) + "_" + std::to_string(x % ws);
  }
}

// Segment s of the tape as a device function `name(A, k)` evaluating
// instance k: the forward sweep (rev = false) or the recompute + reverse sweep.
//
// The reverse sweep needs every value of the segment that a partial derivative
// reads; for long segments those exceed the register file and spill to
// scratch.  The reverse kernel therefore cuts its segment into sub-blocks
// c_0 < c_1 < ... < c_m: one forward pass keeps only the slots live at each
// sub-block boundary (in registers), then for j = m-1 .. 0 sub-block j is
// recomputed from its boundary state and swept backwards, handing the
// adjoints of its live-in slots to sub-block j-1 (two-level checkpointing
// inside the kernel; the HBM checkpoints stay at the segment boundaries).  The
// boundary values enter each recomputation through an empty asm statement so
// that the compiler cannot merge the recomputation with the first pass (which
// would keep every first-pass value live again).
//
// phase splits a reverse kernel in two: 1 = its first pass only (the sub-block
// boundary states, and the outputs of every sub-block but the last), 2 = the
// rest, with the last sub-block reloaded from its stored state as well.  The
// two halves of consecutive segments then run side by side (jit_adj).
static std::string gen_body(const cxh_tape_s* t, const JitPlan& P, int s, bool rev, const std::string& name,
                            bool sub_blocks = true, int phase = 0, bool home = false) {
  const int t0 = P.t0[s], t1 = P.t0[s + 1];
  const int len = t1 - t0;
  const int64_t ws = t->ws;
  std::string o;
  o.reserve(static_cast<size_t>(len) * (rev ? 220 : 64) + 1024);
  o += "__device__ __forceinline__ void " + name + "(const CxjArgs& A, const long long k) {\n";
  std::vector<int64_t> cur(ws, 0);
  const std::vector<int>& lin = P.live[s];
  const std::vector<int>& lout = P.live[s + 1];
  // Forward segments load only the live-in values they read: the others pass
  // through to the live-out rows in copy loops (below), so that a wide live
  // set (a reverse Function keeps every forward value its sweep needs) does
  // not become thousands of values the compiler holds in registers.
  std::vector<char> read_in(static_cast<size_t>(ws), rev ? 1 : 0);
  if (!rev) {
    std::vector<char> written(static_cast<size_t>(ws), 0);
    for (int k = t0; k < t1; ++k) {
      const Instr& e = t->alg[k];
      if (e.op == OP_CONST || e.op == OP_INPUT) {
        written[e.i0] = 1;
        continue;
      }
      if (!written[e.i1]) read_in[e.i1] = 1;
      if (op_is_binary(e.op) && e.op != OP_OUTPUT && !written[e.i2]) read_in[e.i2] = 1;
      if (e.op != OP_OUTPUT) written[e.i0] = 1;
    }
  }
  std::unordered_map<int, int> lin_row;  // forward segments: live-in slot -> its row, loaded at first read
  for (size_t r = 0; r < lin.size(); ++r) {
    const int j = lin[r];
    cur[j] = -(j + 1);
    if (!read_in[j]) continue;
    if (!rev && s > 0) {
      lin_row[j] = home ? P.home[s][r] : static_cast<int>(r);
      continue;
    }
    o += "  const double w" + std::to_string(j) + " = ";
    o += s > 0 ? "CXJ_ST_IN(" + std::to_string(r) + ");\n" : "0.0;\n";
  }
  // A live-in value is loaded where the segment first reads it, through a
  // pointer made opaque after the value defined kLoadAhead definitions
  // earlier: the load can be issued no sooner, so the compiler does not
  // hoist every load of a wide live set to the top and hold them all.
  constexpr size_t kLoadAhead = 32;
  std::vector<std::string> defined;  // names of the values this segment has defined, in order
  int n_ptr = 0;
  auto load_live_in = [&](int64_t c) {
    if (c >= 0) return;
    const int j = static_cast<int>(-c - 1);
    auto it = lin_row.find(j);
    if (it == lin_row.end()) return;  // eager, or already loaded
    const std::string p = "cxj_q" + std::to_string(n_ptr++);
    o += "  CXJ_GLOBAL const double* " + p + " = (CXJ_GLOBAL const double*)A.st_in;\n";
    if (defined.size() >= kLoadAhead)
      o += "  CXJ_OPAQUE_AFTER(" + p + ", " + defined[defined.size() - kLoadAhead] + ");\n";
    o += "  const double w" + std::to_string(j) + " = " + p + "[(long long)" + std::to_string(it->second) +
         " * A.lds + k];\n";
    lin_row.erase(it);
  };
  auto V = [&](int64_t c) { return val_name(c, ws); };
  auto Adj = [&](int64_t c) { return adj_name(c, ws); };
  std::vector<int64_t> ox(len, 0), oy(len, 0), vc(len, 0);
  char buf[256];
  // sin and cos of one value: one sincos call, its names by operand code
  std::unordered_map<int64_t, std::pair<std::string, std::string>> trig;
  // With trig_io set, the sincos sites met by emit_fwd are stored to (1) or
  // loaded from (2) rows of the sub-block state buffer through pointer
  // trig_ptr: site i of sub-block trig_sub at rows trig_rows[trig_sub][i]
  // and + 1.  Pass 1 stores them, the recomputation loads them.
  int trig_io = 0, trig_sub = 0;
  size_t trig_next = 0;
  std::string trig_ptr;
  std::vector<std::vector<int64_t>> trig_rows;
  int64_t* trig_counter = nullptr;  // the next free row (pass 1)
  auto trig_pairs = [&](int a, int b, int pass) {
    std::vector<int64_t> sim = cur;
    std::unordered_map<int64_t, int> seen;
    for (int k = a; k < b; ++k) {
      const Instr& e = t->alg[k];
      if (e.op == OP_OUTPUT) continue;
      if (e.op == OP_SIN) seen[sim[e.i1]] |= 1;
      if (e.op == OP_COS) seen[sim[e.i1]] |= 2;
      sim[e.i0] = 4 * static_cast<int64_t>(k) + pass;
    }
    std::unordered_set<int64_t> both;
    for (auto& kv : seen)
      if (kv.second == 3) both.insert(kv.first);
    return both;
  };
  // all_trig: every SIN / COS through sincos (the reverse sweep of this range
  // reads the complementary value as its partial); else only pairs
  auto emit_fwd = [&](int a, int b, int pass, bool outputs, bool all_trig = false) {
    const std::unordered_set<int64_t> pairs = all_trig ? std::unordered_set<int64_t>{} : trig_pairs(a, b, pass);
    for (int k = a; k < b; ++k) {
      const Instr& e = t->alg[k];
      const int64_t code = 4 * static_cast<int64_t>(k) + pass;
      const std::string nm = V(code);
      if (!lin_row.empty() && e.op != OP_CONST && e.op != OP_INPUT) {
        load_live_in(cur[e.i1]);
        if (e.op != OP_OUTPUT && op_is_binary(e.op)) load_live_in(cur[e.i2]);
      }
      if (e.op != OP_OUTPUT) defined.push_back(nm);
      if ((e.op == OP_SIN || e.op == OP_COS) && (all_trig || pairs.count(cur[e.i1]))) {
        const int64_t x = cur[e.i1];
        auto it = trig.find(x);
        if (it == trig.end()) {
          const std::string id = std::to_string(code);
          it = trig.emplace(x, std::make_pair("ts" + id, "tc" + id)).first;
          if (trig_io == 2) {
            const int64_t r = trig_rows[trig_sub][trig_next++];
            o += "  const double ts" + id + " = CXJ_SNAP(" + trig_ptr + ", " + std::to_string(r) + "), tc" + id +
                 " = CXJ_SNAP(" + trig_ptr + ", " + std::to_string(r + 1) + ");\n";
          } else {
            o += "  double ts" + id + ", tc" + id + ";\n  CXJ_SINCOS(" + V(x) + ", ts" + id + ", tc" + id + ");\n";
          }
          if (trig_io == 1) {
            const int64_t r = (*trig_counter)++;
            ++*trig_counter;
            trig_rows[trig_sub].push_back(r);
            o += "  CXJ_SNAP(" + trig_ptr + ", " + std::to_string(r) + ") = ts" + id + ";\n  CXJ_SNAP(" + trig_ptr +
                 ", " + std::to_string(r + 1) + ") = tc" + id + ";\n";
          }
        }
        o += "  const double " + nm + " = " + (e.op == OP_SIN ? it->second.first : it->second.second) + ";\n";
        ox[k - t0] = oy[k - t0] = x;
        cur[e.i0] = vc[k - t0] = code;
        continue;
      }
      switch (e.op) {
        case OP_CONST:
          o += "  const double " + nm + " = " + lit(e.d) + ";\n";
          cur[e.i0] = vc[k - t0] = code;
          break;
        case OP_INPUT:
          std::snprintf(buf, sizeof buf, "  const double %s = CXJ_IN(%d, %d);\n", nm.c_str(), e.i1, e.i2);
          o += buf;
          cur[e.i0] = vc[k - t0] = code;
          break;
        case OP_OUTPUT:
          ox[k - t0] = cur[e.i1];
          if (outputs) {
            std::snprintf(buf, sizeof buf, "  CXJ_OUT(%d, %d, %s);\n", e.i0, e.i2, V(cur[e.i1]).c_str());
            o += buf;
          }
          break;
        default: {
          const int64_t x = cur[e.i1];
          const int64_t y = op_is_binary(e.op) ? cur[e.i2] : x;
          ox[k - t0] = x;
          oy[k - t0] = y;
          std::snprintf(buf, sizeof buf, "  const double %s = sxh::op_fun_full(%d, %s, %s);\n", nm.c_str(), e.op,
                        V(x).c_str(), V(y).c_str());
          o += buf;
          cur[e.i0] = vc[k - t0] = code;
        }
      }
    }
  };
  if (!rev && home) {  // the values defined here that cross the end go to their home rows
    emit_fwd(t0, t1, 0, true);
    if (s + 1 < P.n_seg)
      for (size_t r = 0; r < lout.size(); ++r)
        if (P.fresh[s + 1][r])
          o += "  CXJ_ST_OUT(" + std::to_string(P.home[s + 1][r]) + ") = " + V(cur[lout[r]]) + ";\n";
    o += "}\n";
    return o;
  }
  if (!rev) {
    emit_fwd(t0, t1, 0, true);
    if (s + 1 < P.n_seg) {
      std::unordered_map<int, int> row_in;
      for (size_t r = 0; r < lin.size(); ++r) row_in[lin[r]] = static_cast<int>(r);
      std::vector<std::pair<int, int>> pass;  // (live-in row, live-out row) of values passed through unread
      for (size_t r = 0; r < lout.size(); ++r) {
        const int j = lout[r];
        if (cur[j] == -(j + 1) && !read_in[j]) {
          if (s > 0) pass.push_back({row_in.at(j), static_cast<int>(r)});
          else o += "  CXJ_ST_OUT(" + std::to_string(r) + ") = 0.0;\n";
        } else {
          o += "  CXJ_ST_OUT(" + std::to_string(r) + ") = " + V(cur[j]) + ";\n";
        }
      }
      // runs of consecutive rows as loops (one copy per iteration)
      for (size_t i = 0; i < pass.size();) {
        size_t m = 1;
        while (i + m < pass.size() && pass[i + m].first == pass[i].first + static_cast<int>(m) &&
               pass[i + m].second == pass[i].second + static_cast<int>(m))
          ++m;
        const std::string ri = std::to_string(pass[i].first), ro = std::to_string(pass[i].second);
        if (m <= 2) {
          for (size_t q = 0; q < m; ++q)
            o += "  CXJ_ST_OUT(" + std::to_string(pass[i + q].second) + ") = CXJ_ST_IN(" +
                 std::to_string(pass[i + q].first) + ");\n";
        } else {
          o += "#pragma unroll 1\n  for (int i = 0; i < " + std::to_string(m) + "; ++i) CXJ_ST_OUT(" + ro +
               " + i) = CXJ_ST_IN(" + ri + " + i);\n";
        }
        i += m;
      }
    }
    o += "}\n";
    return o;
  }
  // sub-block boundaries and the live slots there
  const std::vector<int> c = sub_blocks ? sub_bounds(t0, t1) : std::vector<int>{t0, t1};
  const int m = static_cast<int>(c.size()) - 1;
  const std::vector<std::vector<int>> lc = live_sets(t, t0, t1, lout, c);
  // pass 1: forward over sub-blocks 0..m-2, storing the boundary states to the
  // snapshot buffer (A.st_out, unused otherwise by reverse kernels).  Phase 2
  // keeps its bookkeeping (names, rows) and drops its code.
  std::vector<std::vector<int64_t>> snap(m + 1), in_code(m + 1), row(m + 1);
  snap[0] = cur;
  int64_t nrow = 0;
  std::string dropped;
  if (phase == 2) std::swap(o, dropped);
  const bool tcache = P.trig_cache && sub_blocks;
  trig_rows.assign(m, {});
  trig_counter = &nrow;
  for (int j = 0; j + 1 < m; ++j) {
    if (tcache) {  // one site per operand per sub-block, as the recomputation meets them
      trig.clear();
      trig_io = 1;
      trig_sub = j;
      trig_ptr = "A.st_out";
    }
    emit_fwd(c[j], c[j + 1], 0, true, tcache);
    trig_io = 0;
    snap[j + 1] = cur;
    row[j + 1].assign(ws, -1);
    for (int slot : lc[j + 1]) {
      row[j + 1][slot] = nrow;
      o += "  CXJ_SNAP(A.st_out, " + std::to_string(nrow++) + ") = " + V(cur[slot]) + ";\n";
    }
  }
  if (phase == 2) std::swap(o, dropped);
  if (nrow > P.snap_rows)  // the buffer holds P.snap_rows rows per instance
    throw std::runtime_error("reverse kernel: sub-block state rows exceed the plan's buffer");
  if (phase == 1) {
    o += "}\n";
    return o;
  }
  std::unordered_set<int64_t> decl;
  auto acc = [&](int64_t code, const std::string& term) {
    if (decl.count(code)) {
      o += "  " + Adj(code) + " += " + term + ";\n";
    } else {
      o += "  double " + Adj(code) + " = " + term + ";\n";
      decl.insert(code);
    }
  };
  for (int j = m - 1; j >= 0; --j) {
    const bool last = j == m - 1;
    const bool cont = last && phase != 2;  // continues from pass 1's values
    cur = snap[j];
    if (!cont && !lc[j].empty()) {
      // reload the boundary state through an opaque pointer: neither the
      // values nor the loads can be taken over from the first pass.  The
      // pointer also takes the adjoints sub-block j+1 hands over as inputs, so
      // that this recomputation cannot be scheduled before that reverse sweep
      // (which would keep every recomputed value live at once).
      const std::string sp = "cxj_sp" + std::to_string(j);
      o += std::string("  const double* ") + sp + " = " + (j == 0 ? "A.st_in" : "A.st_out") + ";\n  CXJ_KEEP_PTR(" +
           sp + ");\n";
      if (!last)
        for (int slot : lc[j + 1]) {
          const int64_t from = in_code[j + 1][slot];
          if (decl.count(from)) o += "  CXJ_AFTER(" + sp + ", " + Adj(from) + ");\n";
        }
      for (int slot : lc[j]) {
        const int64_t h = 4 * (static_cast<int64_t>(j) * ws + slot) + 2;
        if (j == 0 && s == 0) {
          o += "  double " + V(h) + " = 0.0;\n  CXJ_KEEP(" + V(h) + ");\n";  // read before written
        } else {
          const int64_t r = j == 0 ? std::find(lin.begin(), lin.end(), slot) - lin.begin() : row[j][slot];
          o += "  double " + V(h) + " = CXJ_SNAP(" + sp + ", " + std::to_string(r) + ");\n";
        }
        cur[slot] = h;
      }
    }
    const bool tload = tcache && !cont && j + 1 < m && !trig_rows[j].empty();
    if (tload) {  // the sin/cos pairs pass 1 stored, through a pointer ordered like the state reload
      const std::string tp = "cxj_tp" + std::to_string(j);
      o += "  const double* " + tp + " = A.st_out;\n  CXJ_KEEP_PTR(" + tp + ");\n";
      for (int slot : lc[j + 1]) {
        const int64_t from = in_code[j + 1][slot];
        if (decl.count(from)) o += "  CXJ_AFTER(" + tp + ", " + Adj(from) + ");\n";
      }
      trig_io = 2;
      trig_sub = j;
      trig_next = 0;
      trig_ptr = tp;
    }
    in_code[j] = cur;
    emit_fwd(c[j], c[j + 1], cont ? 0 : 1, last, true);
    trig_io = 0;
    if (last) {
      if (s + 1 < P.n_seg)
        for (size_t r = 0; r < lout.size(); ++r) {
          const int64_t cd = cur[lout[r]];
          o += "  double " + Adj(cd) + " = CXJ_ADJ_IN(" + std::to_string(r) + ");\n";
          decl.insert(cd);
        }
    } else {
      // adjoints of the slots live at c[j+1], from the names sub-block j+1
      // read them by to the names this recomputation gives them
      for (int slot : lc[j + 1]) {
        const int64_t from = in_code[j + 1][slot], to = cur[slot];
        if (from == to || !decl.count(from)) continue;
        o += "  double " + Adj(to) + " = " + Adj(from) + ";\n";
        decl.insert(to);
      }
    }
    for (int k = c[j + 1] - 1; k >= c[j]; --k) {
      const Instr& e = t->alg[k];
      switch (e.op) {
        case OP_CONST:
          break;
        case OP_OUTPUT:
          std::snprintf(buf, sizeof buf, "CXJ_SEED(%d, %d)", e.i0, e.i2);
          acc(ox[k - t0], buf);
          break;
        case OP_INPUT: {
          const int64_t v = vc[k - t0];
          std::snprintf(buf, sizeof buf, "  CXJ_SENS(%d, %d, %s);\n", e.i1, e.i2,
                        decl.count(v) ? Adj(v).c_str() : "0.0");
          o += buf;
          break;
        }
        default: {
          const int64_t v = vc[k - t0];
          if (!decl.count(v)) break;  // no contribution reaches this value
          const std::string ks = std::to_string(k), an = Adj(v);
          auto tg = (e.op == OP_SIN || e.op == OP_COS) ? trig.find(ox[k - t0]) : trig.end();
          if (tg != trig.end()) {  // d sin = cos, d cos = -sin, from the same sincos
            o += "  const double d" + ks + "[2] = {" +
                 (e.op == OP_SIN ? tg->second.second : "-" + tg->second.first) + ", 0.0};\n";
          } else {
            o += "  double d" + ks + "[2];\n  sxh::op_der_full(" + std::to_string(e.op) + ", " + V(ox[k - t0]) +
                 ", " + V(oy[k - t0]) + ", " + V(v) + ", d" + ks + ");\n";
          }
          if (e.op == OP_IF_ELSE_ZERO) {
            acc(oy[k - t0], "sxh::sxh_if_else_zero(d" + ks + "[1], " + an + ")");
          } else if (op_is_binary(e.op)) {  // a constant-zero partial makes no contribution
            if (!op_zero_partial(e.op, 0)) acc(ox[k - t0], "d" + ks + "[0] * " + an);
            if (!op_zero_partial(e.op, 1)) acc(oy[k - t0], "d" + ks + "[1] * " + an);
          } else if (!op_zero_partial(e.op, 0)) {
            acc(ox[k - t0], "d" + ks + "[0] * " + an);
          }
        }
      }
    }
  }
  if (s > 0)
    for (size_t r = 0; r < lin.size(); ++r) {
      const int64_t cd = in_code[0][lin[r]];
      o += "  CXJ_ADJ_OUT(" + std::to_string(r) + ") = " + (decl.count(cd) ? Adj(cd) : std::string("0.0")) + ";\n";
    }
  if (s == 0)
    for (const int2& u : t->unused_in) {
      std::snprintf(buf, sizeof buf, "  CXJ_SENS(%d, %d, 0.0);\n", u.x, u.y);
      o += buf;
    }
  o += "}\n";
  return o;
}

// One segment kernel: one lane per instance, the arguments by value.  phase
// (reverse kernels): 0 whole, 1 first pass (cxj_snap), 2 the rest.
static std::string gen_segment(const cxh_tape_s* t, const JitPlan& P, int s, bool rev, int phase = 0,
                               bool soa32 = false, bool home = false) {
  std::string o = soa32 ? "#define CXJ_SOA32\n" : "";
  o += kSxOpsSource;
  o += kPrologue;
  o += gen_body(t, P, s, rev, "cxj_body", true, rev ? phase : 0, home && !rev);
  if (rev || P.ipt == 1) {
    o += std::string("extern \"C\" __global__ void __launch_bounds__(256) ") +
         (rev ? (phase == 1 ? "cxj_snap" : "cxj_rev") : "cxj_fwd") +
         "(const CxjArgs A) {\n"
         "  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;\n"
         "  if (k < A.n) cxj_body(A, k);\n"
         "}\n";
  } else {  // P.ipt consecutive blocks' worth of instances per block, coalesced per step
    o += "extern \"C\" __global__ void __launch_bounds__(256) cxj_fwd(const CxjArgs A) {\n"
         "  const long long k0 = (long long)blockIdx.x * blockDim.x * " + std::to_string(P.ipt) + " + threadIdx.x;\n"
         "#pragma unroll\n"
         "  for (int j = 0; j < " + std::to_string(P.ipt) + "; ++j) {\n"
         "    const long long k = k0 + (long long)j * blockDim.x;\n"
         "    if (k < A.n) cxj_body(A, k);\n"
         "  }\n"
         "}\n";
  }
  return finish_source(std::move(o));
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

hipFunction_t jit_function(const std::string& src, const char* name) {
  std::lock_guard<std::mutex> lk(g_jit_mu);
  const std::string key = std::string(name) + "\n" + src;
  auto it = g_jit_cache.find(key);
  if (it != g_jit_cache.end()) return it->second;
  if (no_alloc_active())
    throw std::runtime_error(std::string("no-allocation call: kernel ") + name +
                             " is not compiled and loaded yet; make the allocating call once first");
  if (const char* dir = std::getenv("CXH_JIT_DUMP")) {
    const std::string path = std::string(dir) + "/cxj_" + std::to_string(g_jit_cache.size()) + ".hip";
    if (FILE* f = std::fopen(path.c_str(), "w")) {
      std::fwrite(src.data(), 1, src.size(), f);
      std::fclose(f);
    }
  }
  // CXH_JIT_LOG=1: one line per compiled kernel on stderr (name, source
  // size, compile time, scratch), for programs that take long to compile
  const bool log = std::getenv("CXH_JIT_LOG") && std::atoi(std::getenv("CXH_JIT_LOG")) != 0;
  if (log) std::fprintf(stderr, "[jit] compiling %s (%zu bytes of source)\n", name, src.size());
  const auto t0 = std::chrono::steady_clock::now();
  const std::vector<char> code = rtc_compile(src, device_arch());
  hipModule_t mod;
  hip_check(hipModuleLoadData(&mod, code.data()), "hipModuleLoadData");
  hipFunction_t fn;
  hip_check(hipModuleGetFunction(&fn, mod, name), "hipModuleGetFunction");
  if (log) {
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "[jit] %s compiled in %.1f s, scratch %d B/lane\n", name, sec, scratch_bytes(fn));
  }
  g_jit_cache.emplace(key, fn);
  return fn;
}

// Bytes of scratch (private segment) per lane of a loaded kernel.
int scratch_bytes(hipFunction_t f) {
  int v = 0;
  hip_check(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, f), "hipFuncGetAttribute");
  return v;
}

// Generated kernels whose registers the compiler spills to scratch are not
// launched.  Two such kernels gave wrong results or faulted on the GPU (a
// Hessian tangent kernel spilling ~1.8 KB per lane, a rollout reverse kernel
// with 384-instruction sub-blocks spilling 204 B per lane) while the same
// generated source was correct when executed on the host, and the cause was
// not found.  The generators keep every kernel in use free of scratch: the
// tangent kernels lower their register budget until the compiler needs none,
// the others raise this error with the setting to change.
hipFunction_t jit_function_checked(const std::string& src, const char* name, const char* remedy) {
  hipFunction_t f = jit_function(src, name);
  const int sb = scratch_bytes(f);
  if (sb > 0)
    throw std::runtime_error(std::string(name) + ": the generated kernel spills " + std::to_string(sb) +
                             " bytes per lane to scratch and is not launched; " + remedy);
  return f;
}

static hipFunction_t segment_kernel(cxh_tape_s* t, JitPlan& P, int s, bool rev) {
  hipFunction_t& f = rev ? P.rev[s] : P.fwd[s];
  if (!f)
    f = jit_function_checked(gen_segment(t, P, s, rev), rev ? "cxj_rev" : "cxj_fwd",
                             rev ? "use sub-blocks of at most 256 instructions (CXH_JIT_SUBBLOCK), a shorter "
                                   "checkpoint interval or the interpreter engine"
                                 : "use a shorter checkpoint interval or the interpreter engine");
  return f;
}

// The evaluation kernel of segment s on the home rows (32-bit offsets when soa32).
static hipFunction_t segment_kernel_home(cxh_tape_s* t, JitPlan& P, int s, bool soa32) {
  home_plan(t, P);
  hipFunction_t& f = soa32 ? P.fwd_home_soa[s] : P.fwd_home[s];
  if (!f)
    f = jit_function_checked(gen_segment(t, P, s, false, 0, soa32, true), "cxj_fwd",
                             "use a shorter checkpoint interval or the interpreter engine");
  return f;
}

// Whether every block of a launch over n instances is SoA (instance stride 1)
// with byte offsets below 2^32: the launch can take the 32-bit-offset kernels.
static bool soa32_launch(const IOTable& a, size_t na, const IOTable& b, size_t nb, int64_t n) {
  if (n <= 0 || static_cast<uint64_t>(n) * 8u >= (uint64_t(1) << 32)) return false;
  for (size_t i = 0; i < na; ++i)
    if (a.a[i].p && a.a[i].s_inst != 1) return false;
  for (size_t i = 0; i < nb; ++i)
    if (b.a[i].p && b.a[i].s_inst != 1) return false;
  return true;
}

// The two halves of a split reverse kernel (phase 1 or 2).
static hipFunction_t split_kernel(cxh_tape_s* t, JitPlan& P, int s, int phase) {
  hipFunction_t& f = phase == 1 ? P.snap[s] : P.back[s];
  if (!f)
    f = jit_function_checked(gen_segment(t, P, s, true, phase), phase == 1 ? "cxj_snap" : "cxj_rev",
                             "use sub-blocks of at most 256 instructions (CXH_JIT_SUBBLOCK), a shorter checkpoint "
                             "interval or the interpreter engine");
  return f;
}

// Segment g's reverse kernel after its first pass (blockIdx.y = 0, arguments
// A) beside the first pass of segment g-1 (blockIdx.y = 1, arguments B): the
// two halves are independent, and one launch runs them on twice the SIMDs.
static std::string pair_source(const cxh_tape_s* t, const JitPlan& P, int g) {
  std::string o = kSxOpsSource;
  o += kPrologue;
  o += gen_body(t, P, g, true, "cxj_back_body", true, 2);
  o += gen_body(t, P, g - 1, true, "cxj_snap_body", true, 1);
  o += "extern \"C\" __global__ void __launch_bounds__(256) cxj_rev(const CxjArgs A, const CxjArgs B) {\n"
       "  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;\n"
       "  if (k >= A.n) return;\n"
       "  if (blockIdx.y == 0) cxj_back_body(A, k);\n"
       "  else cxj_snap_body(B, k);\n"
       "}\n";
  return finish_source(std::move(o));
}

static hipFunction_t pair_kernel(cxh_tape_s* t, JitPlan& P, int g) {
  hipFunction_t& f = P.pair[g];
  if (!f) {
    f = jit_function_checked(pair_source(t, P, g), "cxj_rev",
                             "use sub-blocks of at most 256 instructions (CXH_JIT_SUBBLOCK), a shorter checkpoint "
                             "interval or the interpreter engine");
  }
  return f;
}

void launch(hipFunction_t f, int64_t n, int block, CxjArgs& A, hipStream_t s) {
  const unsigned grid = static_cast<unsigned>((n + block - 1) / block);
  void* params[] = {&A};
  hip_check(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, s, params, nullptr), "jit kernel launch");
}

void launch_pair(hipFunction_t f, int64_t n, int block, CxjArgs& A, CxjArgs& B, hipStream_t s) {
  const unsigned grid = static_cast<unsigned>((n + block - 1) / block);
  void* params[] = {&A, &B};
  hip_check(hipModuleLaunchKernel(f, grid, 2, 1, block, 1, 1, 0, s, params, nullptr), "jit kernel launch");
}

// ------------------------------------------------------------ drivers ---

// Whether the reverse sweep of plan P runs split (jit_adj).
static bool split_plan(const JitPlan& P) { return split_reverse() && P.n_seg > 1 && P.snap_rows > 0; }

void jit_prepare(cxh_tape_s* t, int64_t every, bool fwd, bool rev) {
  if (roll_active(t)) return roll_prepare(t, every, fwd, rev);
  JitPlan& P = plan_for(t, every);
  const bool split = split_plan(P);
  for (int s = 0; s < P.n_seg; ++s) {
    if (fwd) segment_kernel(t, P, s, false);
    if (!rev) continue;
    if (!split) {
      segment_kernel(t, P, s, true);
    } else {
      if (s == P.n_seg - 1) split_kernel(t, P, s, 1);
      if (s == 0) split_kernel(t, P, s, 2);
      else pair_kernel(t, P, s);
    }
  }
}

int64_t jit_ckpt_rows(cxh_tape_s* t, int64_t every) {
  return roll_active(t) ? roll_ckpt_rows(t, every) : plan_for(t, every).ck_rows;
}

static JitPlan& eval_plan(cxh_tape_s* t);

// The kernels of forward-only evaluation, both address forms.
static void jit_prepare_eval(cxh_tape_s* t) {
  if (roll_active(t)) return roll_prepare_eval(t);
  JitPlan& P = eval_plan(t);
  sync_ipt(t, P);
  for (int g = 0; g < P.n_seg; ++g) {
    segment_kernel_home(t, P, g, false);
    segment_kernel_home(t, P, g, true);
  }
}

// ---------------------------------------------------- asynchronous compile ---
//
// A worker thread builds the kernels on its own copy of the program (the
// tape's instructions and launch settings, none of its device state), so the
// caller's tape is never touched while it runs.  When it is done, the next
// call on the tape merges what it built (plans, loop kernels) into the tape.

struct JitJob {
  cxh_tape_s clone;
  std::array<int64_t, 3> key{};  // (every, fwd, rev); every = -1: forward-only evaluation
  std::atomic<int> state{1};     // 1 running, 2 done, 3 failed
  std::string error;
  std::thread th;
  ~JitJob() {
    if (th.joinable()) th.join();
  }
};

// Workers of tapes never destroyed are waited for when the library unloads,
// before the compile cache they write to goes (this registry is created
// after that cache, so it is torn down first).
struct JitJobs {
  std::mutex mu;
  std::vector<std::weak_ptr<JitJob>> live;
  ~JitJobs() {
    for (auto& w : live)
      if (auto j = w.lock())
        if (j->th.joinable()) j->th.join();
  }
};
static JitJobs& jit_jobs() {
  static JitJobs r;
  return r;
}

struct JitAsync {
  bool on = false;
  std::shared_ptr<JitJob> job;                   // the running or finished worker
  std::vector<std::array<int64_t, 3>> done, failed;
  std::string last_error;
};

static void jit_adopt(cxh_tape_s* t, cxh_tape_s& from) {
  if (from.jit) {
    if (!t->jit) {
      t->jit = from.jit;
    } else {
      for (auto& kv : from.jit->plans) t->jit->plans.emplace(kv.first, std::move(kv.second));
      if (t->jit->eval_every == 0) t->jit->eval_every = from.jit->eval_every;
    }
  }
  roll_adopt(t, from);
}

// The finished worker's results taken over (or its error recorded).
static void jit_finish(cxh_tape_s* t) {
  JitAsync& A = *t->async;
  std::shared_ptr<JitJob> job = std::move(A.job);
  if (job->th.joinable()) job->th.join();
  if (job->state.load() == 2) {
    jit_adopt(t, job->clone);
    A.done.push_back(job->key);
  } else {
    A.failed.push_back(job->key);
    A.last_error = job->error;
  }
}

void jit_settle(cxh_tape_s* t) {
  if (t->async && t->async->job) jit_finish(t);
}

bool jit_async_pause(cxh_tape_s* t) {
  if (!t->async || !t->async->on) return false;
  jit_settle(t);
  t->async->on = false;
  return true;
}

void jit_async_set(cxh_tape_s* t, bool on) {
  if (on && !t->async) t->async = std::make_shared<JitAsync>();
  if (t->async) t->async->on = on;
}

bool jit_ready(cxh_tape_s* t, int64_t every, bool fwd, bool rev) {
  if (!t->async || !t->async->on) return true;
  JitAsync& A = *t->async;
  if (A.job) {
    if (A.job->state.load() == 1) return false;  // still compiling: the interpreter takes this call
    jit_finish(t);
  }
  const std::array<int64_t, 3> key{every, fwd ? 1 : 0, rev ? 1 : 0};
  auto covers = [&](const std::array<int64_t, 3>& d) {
    return d[0] == key[0] && d[1] >= key[1] && d[2] >= key[2];
  };
  if (std::any_of(A.done.begin(), A.done.end(), covers)) return true;
  // a failed build is left to the specialised path, which raises its error
  if (std::any_of(A.failed.begin(), A.failed.end(), covers)) return true;
  if (no_alloc_active()) return true;
  auto job = std::make_shared<JitJob>();
  job->clone.alg = t->alg;
  job->clone.ws = t->ws;
  job->clone.nnz_in = t->nnz_in;
  job->clone.nnz_out = t->nnz_out;
  job->clone.unused_in = t->unused_in;
  job->clone.block = t->block;
  job->clone.ipt = t->ipt;
  job->clone.engine = t->engine;
  job->key = key;
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  JitJob* j = job.get();  // the tape owns the job and joins the thread before letting it go
  job->th = std::thread([j, dev, every, fwd, rev]() {
    try {
      hip_check(hipSetDevice(dev), "hipSetDevice");
      if (every < 0) jit_prepare_eval(&j->clone);
      else jit_prepare(&j->clone, every, fwd, rev);
      j->state.store(2);
    } catch (const std::exception& e) {
      j->error = e.what();
      j->state.store(3);
    }
  });
  {
    JitJobs& r = jit_jobs();
    std::lock_guard<std::mutex> lk(r.mu);
    r.live.erase(std::remove_if(r.live.begin(), r.live.end(), [](const std::weak_ptr<JitJob>& w) { return w.expired(); }),
                 r.live.end());
    r.live.push_back(job);
  }
  A.job = std::move(job);
  return false;
}

// Forward-only evaluation chooses its own segment length: the default one,
// halved until no forward segment kernel needs scratch.
static JitPlan& eval_plan(cxh_tape_s* t) {
  if (t->jit && t->jit->eval_every > 0) return plan_for(t, t->jit->eval_every);
  for (int64_t every = default_every(static_cast<int64_t>(t->alg.size()));; every = (every + 1) / 2) {
    JitPlan& P = plan_for(t, every);
    sync_ipt(t, P);
    home_plan(t, P);
    bool ok = true;
    for (int g = 0; g < P.n_seg && ok; ++g) {
      if (P.fwd_home[g]) continue;
      hipFunction_t f = jit_function(gen_segment(t, P, g, false, 0, false, true), "cxj_fwd");
      if (scratch_bytes(f) > 0) {
        ok = false;
      } else {
        P.fwd_home[g] = f;
      }
    }
    if (ok) {
      t->jit->eval_every = every;
      return P;
    }
    if (every <= 64) throw std::runtime_error("cxj_fwd: no segment length avoids scratch spills");
  }
}

int jit_map_eval(cxh_tape_s* t, int64_t n, const IOTable& arg, const IOTable& res, hipStream_t s) {
  if (roll_active(t)) return roll_map_eval(t, n, arg, res, s);
  JitPlan& P = eval_plan(t);
  sync_ipt(t, P);
  home_plan(t, P);
  const int64_t lds = round_up(n, 64);
  CxjArgs A{};
  A.arg = arg;
  A.n = n;
  A.lds = lds;
  // one buffer of home rows: a segment reads and writes the values' own rows
  if (P.n_seg > 1) t->jit->st.reserve(static_cast<size_t>(std::max(P.home_rows, 1)) * lds * sizeof(double));
  double* st = P.n_seg > 1 ? t->jit->st.as<double>() : nullptr;
  const bool soa = soa32_launch(arg, t->nnz_in.size(), res, t->nnz_out.size(), n);
  for (int g = 0; g < P.n_seg; ++g) {
    hipFunction_t f = segment_kernel_home(t, P, g, soa);
    A.res = res;
    A.st_in = st;
    A.st_out = st;
    launch(f, (n + P.ipt - 1) / P.ipt, t->block, A, s);
  }
  return 0;
}

// mode 0: fused forward + adjoint; 1: forward storing checkpoints in ckpt_ext;
// 2: reverse from the checkpoints in ckpt_ext.
int jit_adj(cxh_tape_s* t, int mode, int64_t n, const IOTable& arg, const IOTable& aseed, const IOTable& res,
            const IOTable& asens, int64_t every, double* ckpt_ext, hipStream_t s) {
  if (roll_active(t)) return roll_adj(t, mode, n, arg, aseed, res, asens, every, ckpt_ext, s);
  JitPlan& P = plan_for(t, every);
  sync_ipt(t, P);
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
      launch(f, (n + P.ipt - 1) / P.ipt, t->block, A, s);
    }
  }
  if (mode == 1) return 0;
  if (P.n_seg > 1) t->jit->adj.reserve(2 * static_cast<size_t>(P.max_live) * lds * sizeof(double));
  double* adj = t->jit->adj.as<double>();
  const bool split = split_plan(P);
  const size_t snap_elems = static_cast<size_t>(P.snap_rows) * lds;
  if (P.snap_rows) t->jit->snap.reserve((split ? 2 : 1) * snap_elems * sizeof(double));
  double* snap = P.snap_rows ? t->jit->snap.as<double>() : nullptr;
  A.aseed = aseed;
  A.asens = asens;
  auto set_adj = [&](CxjArgs& B, int g) {
    B.adj_in = adj ? adj + static_cast<int64_t>((g + 1) % 2) * P.max_live * lds : nullptr;
    B.adj_out = adj ? adj + static_cast<int64_t>(g % 2) * P.max_live * lds : nullptr;
  };
  if (!split) {
    for (int g = P.n_seg - 1; g >= 0; --g) {
      hipFunction_t f = segment_kernel(t, P, g, true);
      A.res = (mode == 0 && g == P.n_seg - 1) ? res : none;
      A.st_in = state(g);
      A.st_out = snap;
      set_adj(A, g);
      launch(f, n, t->block, A, s);
    }
    return 0;
  }
  // Split: the first pass of the last segment alone, then for g = n_seg-1..1
  // the rest of segment g beside the first pass of segment g-1 in one launch,
  // then the rest of segment 0.  The boundary states of consecutive segments
  // alternate between two buffers.
  auto buf = [&](int g) { return snap + static_cast<size_t>(g % 2) * snap_elems; };
  const int last = P.n_seg - 1;
  CxjArgs B = A;
  B.res = mode == 0 ? res : none;
  B.st_in = state(last);
  B.st_out = buf(last);
  launch(split_kernel(t, P, last, 1), n, t->block, B, s);
  for (int g = last; g >= 0; --g) {
    A.res = (mode == 0 && g == last) ? res : none;
    A.st_in = state(g);
    A.st_out = buf(g);
    set_adj(A, g);
    if (g == 0) {
      launch(split_kernel(t, P, 0, 2), n, t->block, A, s);
      break;
    }
    B.res = none;
    B.st_in = state(g - 1);
    B.st_out = buf(g - 1);
    launch_pair(pair_kernel(t, P, g), n, t->block, A, B, s);
  }
  return 0;
}

std::string jit_source(cxh_tape_s* t, int64_t every, int seg, int rev) {
  JitPlan& P = plan_for(t, every);
  if (seg < 0 || seg >= P.n_seg) throw std::runtime_error("segment index out of range");
  if (rev < 0 || rev > 5) throw std::runtime_error("reverse must be 0 .. 5");
  if (rev == 5) {  // forward-only evaluation on the home rows
    home_plan(t, P);
    return gen_segment(t, P, seg, false, 0, false, true);
  }
  if (rev == 4) {
    if (seg < 1) throw std::runtime_error("a pair kernel pairs segment g >= 1 with g-1");
    return pair_source(t, P, seg);
  }
  return gen_segment(t, P, seg, rev != 0, rev >= 2 ? rev - 1 : 0);
}

void jit_plan_info(cxh_tape_s* t, int64_t every, int64_t* info, int64_t* ck_off) {
  JitPlan& P = plan_for(t, every);
  info[0] = P.n_seg;
  info[1] = P.ck_rows;
  info[2] = P.snap_rows;
  info[3] = P.max_live;
  info[4] = split_plan(P) ? 1 : 0;
  info[5] = P.ipt;
  home_plan(t, P);
  info[6] = P.home_rows;
  if (ck_off)
    for (int s = 0; s <= P.n_seg; ++s) ck_off[s] = P.ck_off[s];
}

// Operand codes (instruction indices) of the values whose sin and cos are
// both taken in the tape: computed by one sincos call.
std::unordered_set<int64_t> step_trig_pairs(const cxh_tape_s* t) {
  std::vector<int64_t> cur(t->ws, 0);
  std::unordered_map<int64_t, int> seen;
  for (size_t k = 0; k < t->alg.size(); ++k) {
    const Instr& e = t->alg[k];
    if (e.op == OP_OUTPUT) continue;
    if (e.op == OP_SIN) seen[cur[e.i1]] |= 1;
    if (e.op == OP_COS) seen[cur[e.i1]] |= 2;
    cur[e.i0] = static_cast<int64_t>(k);
  }
  std::unordered_set<int64_t> both;
  for (auto& kv : seen)
    if (kv.second == 3) both.insert(kv.first);
  return both;
}

// Emits `const double v<k> = sin/cos(v<x>)` through sincos (one call per
// operand, names remembered in trig) and returns true, for SIN / COS when
// all_trig or the operand is in pairs.
bool emit_trig(std::string& o, const Instr& e, int64_t k, int64_t x, const std::string& xname, bool all_trig,
               const std::unordered_set<int64_t>& pairs,
               std::unordered_map<int64_t, std::pair<std::string, std::string>>& trig) {
  if ((e.op != OP_SIN && e.op != OP_COS) || !(all_trig || pairs.count(x))) return false;
  auto it = trig.find(x);
  if (it == trig.end()) {
    const std::string id = std::to_string(k);
    it = trig.emplace(x, std::make_pair("ts" + id, "tc" + id)).first;
    o += "  double ts" + id + ", tc" + id + ";\n  CXJ_SINCOS(" + xname + ", ts" + id + ", tc" + id + ");\n";
  }
  o += "  const double v" + std::to_string(k) + " = " + (e.op == OP_SIN ? it->second.first : it->second.second) +
       ";\n";
  return true;
}

// ------------------------------------------------------ tangent sweeps ---
//
// The compressed-Jacobian sweep of XFunction::jac
// (casadi/core/x_function.hpp:442-840) as straight-line code: the primal and,
// per seed direction, the tangent of every value (forward mode,
// SXFunction::ad_forward numerically, casadi/core/sx_function.cpp:1194-1366),
// with the tangents scattered into the Jacobian nonzeros at the OUTPUT
// instructions.  Tangents that are structurally zero (no seeded input reaches
// the value) are known at generation time and never computed, as CasADi's
// symbolic sweep drops structural zeros.  Directions are processed in groups,
// each group a body of its own that recomputes the primal; the groups of one
// segment share a kernel (blockIdx.y picks the body) and run side by side.
// Long tapes are cut into segments whose state is the live values plus their
// nonzero tangents.

struct JitJac {
  int group = 1, n_group = 1, n_seg = 1;
  std::vector<int> t0;                                         // segment boundaries
  std::vector<std::vector<int>> live;                          // live slots per boundary
  std::vector<std::vector<int>> def;                           // [b][i]: instruction that wrote live[b][i] (-1: none)
  std::vector<std::vector<std::vector<std::pair<int, int>>>> tnz;  // [g][b]: nonzero (slot, dir)
  std::vector<std::unordered_map<int64_t, int>> home;          // [g]: home row of a value crossing a boundary
  int64_t home_rows = 0;                                       // home rows per group
  int64_t spill_rows = 0;                                      // spill rows (max over kernels)
  std::vector<hipFunction_t> fn;                               // [s], all groups
  DeviceBuffer st, spill, zero;
};

// Key of a value in the home map: its defining instruction and the direction
// within the group (0: the primal).
static int64_t home_key(int def, int dir_in_group) { return static_cast<int64_t>(def) * 128 + dir_in_group; }

// Directions per group.  Each group is a body of its own over all instances
// (n/64 waves) and the kernels hold about one wave per SIMD (register-bound),
// so by default the directions are split into as many groups as fill the
// device's SIMDs once; CXH_JIT_DIRGROUP overrides.  n <= 0: no batch size
// known (source introspection), 8 directions per group.
static int jac_dir_group(int ndir, int64_t n) {
  if (const char* e = std::getenv("CXH_JIT_DIRGROUP")) return std::max(1, std::atoi(e));
  if (n <= 0) return 8;
  int dev = 0, cus = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "device CU count");
  const int64_t simds = 4 * static_cast<int64_t>(std::max(cus, 1));
  const int64_t waves = std::max<int64_t>(1, (n + 63) / 64);
  const int64_t groups = std::min<int64_t>(std::max<int64_t>(1, (simds + waves / 2) / waves), std::max(ndir, 1));
  return static_cast<int>((ndir + groups - 1) / groups);
}

static JitJac& jac_plan(cxh_jac_s* j, int64_t n) {
  const int want = j->ndir == 0 ? 1 : std::min(std::max(j->ndir, 1), std::min(jac_dir_group(j->ndir, n), 64));
  if (j->jit && j->jit->group == want) return *j->jit;
  const cxh_tape_s* t = j->tape;
  auto J = std::make_shared<JitJac>();
  const int nd = j->ndir;
  const int64_t ni = static_cast<int64_t>(t->alg.size());
  J->group = want;
  J->n_group = std::max(1, (nd + J->group - 1) / J->group);
  int64_t every = j->seg_len > 0 ? j->seg_len : default_every(ni);
  if (const char* e = std::getenv("CXH_JIT_TANSEG")) every = std::max(1, std::atoi(e));
  every = std::max<int64_t>(1, std::min(every, ni));
  J->n_seg = static_cast<int>((ni + every - 1) / every);
  for (int s = 0; s <= J->n_seg; ++s) J->t0.push_back(static_cast<int>(std::min<int64_t>(ni, s * every)));
  J->live = live_sets(t, 0, static_cast<int>(ni), {}, J->t0);
  J->tnz.assign(J->n_group, std::vector<std::vector<std::pair<int, int>>>(J->n_seg + 1));
  for (int g = 0; g < J->n_group; ++g) {
    const int d0 = g * J->group, d1 = std::min(nd, d0 + J->group);
    std::vector<uint64_t> mask(t->ws, 0);
    int b = 1;
    for (int64_t k = 0; k < ni; ++k) {
      const Instr& e = t->alg[k];
      switch (e.op) {
        case OP_CONST: mask[e.i0] = 0; break;
        case OP_INPUT: {
          uint64_t m = 0;
          const int64_t flat = j->in_off[e.i1] + e.i2;
          for (int d = d0; d < d1; ++d)
            if (j->seedval[flat * nd + d] != 0.0) m |= uint64_t(1) << (d - d0);
          mask[e.i0] = m;
        } break;
        case OP_OUTPUT: break;
        default:
          mask[e.i0] = e.op == OP_IF_ELSE_ZERO ? mask[e.i2]
                       : op_is_binary(e.op)    ? (mask[e.i1] | mask[e.i2])
                                               : mask[e.i1];
      }
      while (b < J->n_seg && J->t0[b] == k + 1) {
        for (int slot : J->live[b])
          for (int d = d0; d < d1; ++d)
            if (mask[slot] >> (d - d0) & 1) J->tnz[g][b].push_back({slot, d});
        ++b;
      }
    }
  }
  // the instruction that wrote each live slot at each boundary
  J->def.assign(J->n_seg + 1, {});
  {
    std::vector<int> defk(t->ws, -1);
    int b = 1;
    for (int64_t k = 0; k < ni; ++k) {
      const Instr& e = t->alg[k];
      if (e.op != OP_OUTPUT) defk[e.i0] = static_cast<int>(k);
      while (b < J->n_seg && J->t0[b] == k + 1) {
        for (int slot : J->live[b]) J->def[b].push_back(defk[slot]);
        ++b;
      }
    }
  }
  // Home rows.  A value that crosses a boundary keeps one row from the
  // segment that defines it to the last segment that reads it: it is stored
  // once (when evicted or at the end of its segment) and later segments load
  // it where they use it, instead of every boundary copying every live value.
  // Rows are reused between values whose segment spans do not overlap.
  auto seg_of = [&](int k) {
    return static_cast<int>(std::upper_bound(J->t0.begin(), J->t0.end(), k) - J->t0.begin()) - 1;
  };
  J->home.assign(J->n_group, {});
  for (int g = 0; g < J->n_group; ++g) {
    const int d0 = g * J->group;
    std::map<int64_t, std::pair<int, int>> span;  // key -> (defining segment, last reading segment)
    for (int b = 1; b < J->n_seg; ++b) {
      std::unordered_map<int, int> slot_def;
      for (size_t i = 0; i < J->live[b].size(); ++i) {
        const int def = J->def[b][i];
        slot_def[J->live[b][i]] = def;
        if (def < 0) continue;
        auto it = span.emplace(home_key(def, 0), std::make_pair(seg_of(def), b)).first;
        it->second.second = std::max(it->second.second, b);
      }
      for (const auto& sd : J->tnz[g][b]) {
        const int def = slot_def.at(sd.first);
        auto it = span.emplace(home_key(def, sd.second - d0 + 1), std::make_pair(seg_of(def), b)).first;
        it->second.second = std::max(it->second.second, b);
      }
    }
    std::vector<std::vector<int64_t>> starts(J->n_seg + 1), ends(J->n_seg + 1);
    for (const auto& kv : span) {
      starts[kv.second.first].push_back(kv.first);
      ends[kv.second.second].push_back(kv.first);
    }
    std::set<int> free_rows;
    int rows = 0;
    for (int sg = 0; sg < J->n_seg; ++sg) {
      if (sg > 0)
        for (int64_t key : ends[sg - 1]) free_rows.insert(J->home[g].at(key));
      for (int64_t key : starts[sg]) {
        int r;
        if (free_rows.empty()) {
          r = rows++;
        } else {
          r = *free_rows.begin();
          free_rows.erase(free_rows.begin());
        }
        J->home[g][key] = r;
      }
    }
    J->home_rows = std::max<int64_t>(J->home_rows, rows);
  }
  J->fn.assign(J->n_seg, nullptr);
  j->jit = J;
  return *J;
}

// Starting register budget of a tangent kernel, in doubles held at once.
static int jac_reg_budget(const cxh_jac_s* j) {
  if (const char* e = std::getenv("CXH_JIT_TANREGS")) return std::max(8, std::atoi(e));
  return j->reg_budget > 0 ? j->reg_budget : 200;
}

// Reloads are issued this many instructions ahead of their use.  Kernels
// with few wavefronts (nothing else to run while a reload is in flight) need
// a longer lookahead than those that fill the device.
static int jac_prefetch(const cxh_jac_s* j) {
  if (const char* e = std::getenv("CXH_JIT_PREFETCH")) return std::max(0, std::atoi(e));
  return j->prefetch >= 0 ? j->prefetch : 16;
}

// Kernel of direction group g, segment s.  Block 0 of the asens table carries
// the Jacobian nonzeros: nonzero q of instance k at p[q*s_nz + k*s_inst].
//
// The straight-line body of a long gradient tape keeps hundreds of values
// live (a reverse sweep reads forward values from far back), more than the
// register file holds; left to the compiler they spill to scratch at one wave
// per SIMD.  The generator therefore allocates registers itself: it walks the
// segment with at most R values resident and, when over budget, evicts the
// value whose next use is furthest away (Belady), storing it once to a spill
// row (or straight to its boundary row when it is live out) and reloading it
// a few instructions before its next use.  Segment live-ins and inputs are
// loaded next to their first use instead of at the top.  Each group of loads
// or stores goes through its own opaque copy of the per-instance base pointer
// (reloads: made after the most recently computed value), so the compiler can
// neither fold a reload into its store nor hoist it back up, and no row offset
// stays live across the kernel.
//
// State and spill rows are tiled [group][instance/64][row][64]: a row's offset
// is a compile-time constant and accesses of a wave are 512 B contiguous.
static std::string gen_tangent(const cxh_jac_s* j, const JitJac& J, int g, int s, int R, int* spill_rows) {
  const cxh_tape_s* t = j->tape;
  const int nd = j->ndir;
  const int d0 = g * J.group, d1 = std::min(nd, d0 + J.group);
  const int t0 = J.t0[s], t1 = J.t0[s + 1];
  const bool last_seg = s + 1 >= J.n_seg;

  // ---- values and the reads of every instruction
  struct Val {
    std::string base, name;  // SSA base name, current name (literal text for literals)
    int kind = 0;            // 0 computed, 1 live-in (state row), 2 input, 3 literal
    int in_row = -1, ai = 0, aj = 0;
    int home = -1;           // spill row holding the value once stored
    int out_row = -1;        // boundary row if live out (its home when stored)
    std::vector<int> uses;
    size_t nu = 0;
    bool resident = false, dirty = false, stored = false;
    int reloads = 0;
  };
  std::vector<Val> val;
  std::unordered_map<int64_t, int> vid;
  auto key = [](int64_t code, int d) { return (code + (int64_t(1) << 40)) * 128 + (d + 1); };
  auto mk = [&](int kind, std::string nm) {
    Val v;
    v.kind = kind;
    v.base = v.name = std::move(nm);
    val.push_back(std::move(v));
    return static_cast<int>(val.size()) - 1;
  };
  auto add = [&](int64_t code, int d, int kind, std::string nm) {
    const int id = mk(kind, std::move(nm));
    vid[key(code, d)] = id;
    return id;
  };
  auto Vn = [](int64_t c) { return c >= 0 ? "v" + std::to_string(c) : "w" + std::to_string(-c - 1); };
  auto Tn = [](int64_t c, int d) {
    return (c >= 0 ? "u" + std::to_string(c) : "z" + std::to_string(-c - 1)) + "_" + std::to_string(d);
  };
  struct Op {
    int k;                    // instruction; t1 marks the segment end
    std::vector<int> reads;   // deduplicated
    std::vector<int> defs;    // values defined (with uses or not)
    int x = -1, y = -1;
    int trig = 0, s_id = -1, c_id = -1;  // 1: sincos here, 2: value of an earlier sincos
    std::vector<int> tx, ty, td;  // per direction of the group
  };
  std::vector<Op> ops;
  const int G = d1 - d0;
  std::vector<int64_t> cur(t->ws, 0);
  std::vector<uint64_t> mask(t->ws, 0);
  auto vof = [&](int64_t c, int d) { return vid.at(key(c, d)); };
  // live-ins: read from their home rows where used (never-written slots: 0)
  const std::vector<int>& lin = J.live[s];
  const auto& home = J.home[g];
  std::unordered_map<int, int> in_def;
  for (size_t r = 0; r < lin.size(); ++r) {
    const int64_t c = -(lin[r] + 1);
    cur[lin[r]] = c;
    const int def = s > 0 ? J.def[s][r] : -1;
    in_def[lin[r]] = def;
    const int v = add(c, -1, def >= 0 ? 1 : 3, def >= 0 ? Vn(c) : std::string("0.0"));
    if (def >= 0) val[v].in_row = home.at(home_key(def, 0));
  }
  if (s > 0)
    for (const auto& sd : J.tnz[g][s]) {
      mask[sd.first] |= uint64_t(1) << (sd.second - d0);
      const int v = add(-(sd.first + 1), sd.second, 1, Tn(-(sd.first + 1), sd.second));
      val[v].in_row = home.at(home_key(in_def.at(sd.first), sd.second - d0 + 1));
    }
  // sin and cos of one operand in this segment: one sincos gives both
  std::unordered_set<int64_t> pair_ops;
  {
    std::vector<int64_t> c2(cur);
    std::unordered_map<int64_t, int> seen;
    for (int k = t0; k < t1; ++k) {
      const Instr& e = t->alg[k];
      if (e.op == OP_OUTPUT) continue;
      if (e.op == OP_SIN) seen[c2[e.i1]] |= 1;
      if (e.op == OP_COS) seen[c2[e.i1]] |= 2;
      c2[e.i0] = k;
    }
    for (const auto& kv : seen)
      if (kv.second == 3) pair_ops.insert(kv.first);
  }
  std::unordered_map<int64_t, std::pair<int, int>> sib;  // operand -> (sin, cos) values
  for (int k = t0; k < t1; ++k) {
    const Instr& e = t->alg[k];
    switch (e.op) {
      case OP_CONST:
        add(k, -1, 3, lit(e.d));
        cur[e.i0] = k;
        mask[e.i0] = 0;
        break;
      case OP_INPUT: {
        const int v = add(k, -1, 2, Vn(k));
        val[v].ai = e.i1;
        val[v].aj = e.i2;
        uint64_t m = 0;
        const int64_t flat = j->in_off[e.i1] + e.i2;
        for (int d = d0; d < d1; ++d) {
          const double sv = j->seedval[flat * nd + d];
          if (sv != 0.0) {
            m |= uint64_t(1) << (d - d0);
            add(k, d, 3, lit(sv));
          }
        }
        cur[e.i0] = k;
        mask[e.i0] = m;
      } break;
      case OP_OUTPUT: {
        Op op;
        op.k = k;
        const int64_t c = cur[e.i1];
        op.x = vof(c, -1);
        if (g == 0) op.reads.push_back(op.x);
        op.tx.assign(G, -1);
        const int64_t flat = j->out_off[e.i0] + e.i2;
        for (int d = d0; d < d1; ++d) {
          const int2 tg = j->scatter[flat * nd + d];
          if ((tg.x >= 0 || tg.y >= 0) && (mask[e.i1] >> (d - d0) & 1)) {
            op.tx[d - d0] = vof(c, d);
            op.reads.push_back(op.tx[d - d0]);
          }
        }
        ops.push_back(std::move(op));
      } break;
      default: {
        Op op;
        op.k = k;
        const bool bin = op_is_binary(e.op);
        const int64_t x = cur[e.i1], y = bin ? cur[e.i2] : x;
        const uint64_t mx = mask[e.i1], my = bin ? mask[e.i2] : 0;
        const uint64_t m = e.op == OP_IF_ELSE_ZERO ? my : (mx | my);
        op.x = vof(x, -1);
        op.y = vof(y, -1);
        const bool trig = e.op == OP_SIN || e.op == OP_COS;
        const auto sit = trig ? sib.find(x) : sib.end();
        if (sit != sib.end()) {  // the value (and partial) come from an earlier sincos
          op.trig = 2;
          op.s_id = sit->second.first;
          op.c_id = sit->second.second;
          vid[key(k, -1)] = e.op == OP_SIN ? op.s_id : op.c_id;
          if (m) op.reads.push_back(e.op == OP_SIN ? op.c_id : op.s_id);
        } else if (trig && (m || pair_ops.count(x))) {
          op.trig = 1;
          op.reads.push_back(op.x);
          op.s_id = mk(0, "ts" + std::to_string(k));
          op.c_id = mk(0, "tc" + std::to_string(k));
          vid[key(k, -1)] = e.op == OP_SIN ? op.s_id : op.c_id;
          if (pair_ops.count(x)) sib[x] = {op.s_id, op.c_id};
          op.defs = {op.s_id, op.c_id};
        } else {
          op.reads.push_back(op.x);
          if (op.y != op.x) op.reads.push_back(op.y);
          op.defs.push_back(add(k, -1, 0, Vn(k)));
        }
        op.tx.assign(G, -1);
        op.ty.assign(G, -1);
        op.td.assign(G, -1);
        for (int d = d0; d < d1; ++d) {
          const uint64_t bit = uint64_t(1) << (d - d0);
          if (!(m & bit)) continue;
          if (e.op != OP_IF_ELSE_ZERO && (mx & bit)) op.tx[d - d0] = vof(x, d);
          if ((e.op == OP_IF_ELSE_ZERO || bin) && (my & bit)) op.ty[d - d0] = vof(y, d);
          for (int q : {op.tx[d - d0], op.ty[d - d0]})
            if (q >= 0 && std::find(op.reads.begin(), op.reads.end(), q) == op.reads.end()) op.reads.push_back(q);
          op.td[d - d0] = add(k, d, 0, Tn(k, d));
          op.defs.push_back(op.td[d - d0]);
        }
        cur[e.i0] = k;
        mask[e.i0] = m;
        ops.push_back(std::move(op));
      }
    }
  }
  // the end: live-out values (state rows of the next segment)
  // values this segment defines that later segments read: to their home rows
  std::vector<std::pair<int, int>> outs;  // (row, value id or -1 for a structural zero)
  if (!last_seg) {
    Op op;
    op.k = t1;
    std::unordered_map<int, int> out_def;
    for (size_t i = 0; i < J.live[s + 1].size(); ++i) {
      const int slot = J.live[s + 1][i], def = J.def[s + 1][i];
      out_def[slot] = def;
      if (def >= t0 && def < t1) outs.push_back({home.at(home_key(def, 0)), vof(cur[slot], -1)});
    }
    for (const auto& sd : J.tnz[g][s + 1]) {
      const int def = out_def.at(sd.first);
      if (def < t0 || def >= t1) continue;  // a live-in: already home
      const bool nzv = mask[sd.first] >> (sd.second - d0) & 1;
      outs.push_back({home.at(home_key(def, sd.second - d0 + 1)), nzv ? vof(cur[sd.first], sd.second) : -1});
    }
    for (auto& ro : outs)
      if (ro.second >= 0 && val[ro.second].kind != 3) {
        if (val[ro.second].out_row < 0) val[ro.second].out_row = ro.first;
        if (std::find(op.reads.begin(), op.reads.end(), ro.second) == op.reads.end()) op.reads.push_back(ro.second);
      }
    ops.push_back(std::move(op));
  }
  for (size_t i = 0; i < ops.size(); ++i)
    for (int v : ops[i].reads)
      if (val[v].kind != 3) val[v].uses.push_back(static_cast<int>(i));

  // ---- allocation and emission
  const int LA = jac_prefetch(j);
  std::string b;  // body
  b.reserve(static_cast<size_t>(t1 - t0) * (90 + 70 * G));
  std::set<std::pair<int, int>> res;  // (next use, value) of resident values
  std::vector<int> free_rows;
  int n_spill = 0;
  std::string prev;  // most recently computed value, orders the reloads after it
  auto next_use = [&](const Val& v) { return v.nu < v.uses.size() ? v.uses[v.nu] : INT32_MAX; };
  int n_ptr = 0;
  bool use_si = false, use_so = false, use_ss = false;
  std::set<int> use_arg;
  // Every group of accesses goes through its own opaque copy of the
  // per-instance base pointer: the row offset is then an immediate or a
  // short-lived add, never an index shared by (and kept live between) all
  // accesses of one row.
  auto opaque = [&](const char* base, bool is_const, const std::string& after) {
    const std::string p = "cxj_p" + std::to_string(n_ptr++);
    b += std::string("  CXJ_GLOBAL ") + (is_const ? "const " : "") + "double* " + p + " = " + base + ";\n";
    b += after.empty() ? "  CXJ_OPAQUE(" + p + ");\n" : "  CXJ_OPAQUE_AFTER(" + p + ", " + after + ");\n";
    return p;
  };
  std::vector<std::pair<int, std::string>> st_rows, sp_rows;  // pending stores
  auto flush = [&]() {
    if (!st_rows.empty()) {
      use_so = true;
      const std::string p = opaque("cxj_so", false, "");
      for (const auto& x : st_rows) b += "  " + p + "[" + std::to_string(64 * x.first) + "] = " + x.second + ";\n";
      st_rows.clear();
    }
    if (!sp_rows.empty()) {
      use_ss = true;
      const std::string p = opaque("cxj_ss", false, "");
      for (const auto& x : sp_rows) b += "  " + p + "[" + std::to_string(64 * x.first) + "] = " + x.second + ";\n";
      sp_rows.clear();
    }
  };
  auto evict = [&](int id) {
    Val& v = val[id];
    res.erase({next_use(v), id});
    v.resident = false;
    if (!v.dirty) return;
    v.dirty = false;
    v.stored = true;
    if (v.out_row >= 0) {
      st_rows.push_back({v.out_row, v.name});
    } else {
      if (free_rows.empty()) free_rows.push_back(n_spill++);
      v.home = free_rows.back();
      free_rows.pop_back();
      sp_rows.push_back({v.home, v.name});
    }
  };
  auto over_budget = [&](int not_before) {
    while (static_cast<int>(res.size()) > R) {
      const auto it = std::prev(res.end());
      if (it->first <= not_before) break;
      evict(it->second);
    }
    flush();
  };
  // bring the values of `ids` into registers, loads ordered after `prev`
  auto reload = [&](const std::vector<int>& ids) {
    std::string p_st, p_sp, p_in;
    std::map<int, std::string> p_arg;
    auto ptr = [&](std::string& p, const std::string& base) {
      if (p.empty()) p = opaque(base.c_str(), true, prev);
      return p;
    };
    for (int id : ids) {
      Val& v = val[id];
      if (v.kind == 3 || v.resident) continue;
      const std::string nm = v.base + "_" + std::to_string(++v.reloads);
      std::string load;
      if (v.stored && v.out_row >= 0) {
        use_so = true;
        load = ptr(p_st, "cxj_so") + "[" + std::to_string(64 * v.out_row) + "]";
      } else if (v.stored) {
        use_ss = true;
        load = ptr(p_sp, "cxj_ss") + "[" + std::to_string(64 * v.home) + "]";
      } else if (v.kind == 1) {
        use_si = true;
        load = ptr(p_in, "cxj_si") + "[" + std::to_string(64 * v.in_row) + "]";
      } else if (v.kind == 2) {
        use_arg.insert(v.ai);
        load = ptr(p_arg[v.ai], "cxj_a" + std::to_string(v.ai)) + "[" + std::to_string(v.aj) + " * A.arg.a[" +
               std::to_string(v.ai) + "].s_nz]";
      } else {
        throw std::logic_error("tangent codegen: value neither resident nor stored");
      }
      b += "  const double " + nm + " = " + load + ";\n";
      v.name = nm;
      v.resident = true;
      res.insert({next_use(v), id});
    }
  };
  char buf[256];
  for (size_t i = 0; i < ops.size(); ++i) {
    const Op& op = ops[i];
    const int ii = static_cast<int>(i);
    if (op.k == t1) {  // values already in their boundary row stay there
      std::vector<int> need;
      for (const auto& ro : outs)
        if (ro.second >= 0 && !(val[ro.second].stored && val[ro.second].out_row == ro.first)) need.push_back(ro.second);
      reload(need);
    } else {
      // this instruction's operands, and those of the one LA ahead (the
      // reload is then in flight while the instructions between execute)
      std::vector<int> need(op.reads);
      const size_t ahead = i + static_cast<size_t>(LA);
      if (LA > 0 && ahead < ops.size() && ops[ahead].k != t1)
        for (int id : ops[ahead].reads)  // only what memory already holds
          if (val[id].kind == 1 || val[id].kind == 2 || val[id].stored) need.push_back(id);
      reload(need);
    }
    over_budget(ii + LA);
    if (op.k == t1) {  // segment end: boundary rows
      for (const auto& ro : outs) {
        const std::string rr = std::to_string(ro.first);
        if (ro.second < 0) {
          st_rows.push_back({ro.first, "0.0"});
          continue;
        }
        const Val& v = val[ro.second];
        if (v.stored && v.out_row == ro.first) continue;  // already there
        st_rows.push_back({ro.first, v.name});
      }
      flush();
    } else {
      const Instr& e = t->alg[op.k];
      const std::string ks = std::to_string(op.k);
      if (e.op == OP_OUTPUT) {
        if (g == 0) {
          std::snprintf(buf, sizeof buf, "  CXJ_OUT(%d, %d, %s);\n", e.i0, e.i2, val[op.x].name.c_str());
          b += buf;
        }
        const int64_t flat = j->out_off[e.i0] + e.i2;
        for (int d = d0; d < d1; ++d) {
          const int2 tg = j->scatter[flat * nd + d];
          const std::string v = op.tx[d - d0] >= 0 ? val[op.tx[d - d0]].name : std::string("0.0");
          if (tg.x >= 0) b += "  CXJ_SENS(0, " + std::to_string(tg.x) + ", " + v + ");\n";
          if (tg.y >= 0) b += "  CXJ_SENS(0, " + std::to_string(tg.y) + ", " + v + ");\n";
        }
      } else {
        const std::string& X = val[op.x].name;
        const std::string& Y = val[op.y].name;
        if (op.trig == 1) {
          b += "  double ts" + ks + ", tc" + ks + ";\n  CXJ_SINCOS(" + X + ", ts" + ks + ", tc" + ks + ");\n";
        } else if (op.trig == 0) {
          b += "  const double v" + ks + " = sxh::op_fun_full(" + std::to_string(e.op) + ", " + X + ", " + Y + ");\n";
        }
        bool any = false;
        for (int q : op.td) any |= q >= 0;
        if (any && op.trig) {  // d sin = cos, d cos = -sin, from the sincos
          b += "  const double d" + ks + "[2] = {" +
               (e.op == OP_SIN ? val[op.c_id].name : "-" + val[op.s_id].name) + ", 0.0};\n";
        } else if (any) {
          b += "  double d" + ks + "[2];\n  sxh::op_der_full(" + std::to_string(e.op) + ", " + X + ", " + Y + ", v" +
               ks + ", d" + ks + ");\n";
        }
        if (any) {
          for (int d = d0; d < d1; ++d) {
            const int q = op.td[d - d0];
            if (q < 0) continue;
            const int a = op.tx[d - d0], c = op.ty[d - d0];
            std::string rhs;
            if (e.op == OP_IF_ELSE_ZERO) {
              rhs = "sxh::sxh_if_else_zero(d" + ks + "[1], " + val[c].name + ")";
            } else if (a >= 0 && c >= 0) {
              rhs = "d" + ks + "[0] * " + val[a].name + " + d" + ks + "[1] * " + val[c].name;
            } else if (a >= 0) {
              rhs = "d" + ks + "[0] * " + val[a].name;
            } else {
              rhs = "d" + ks + "[1] * " + val[c].name;
            }
            b += "  const double " + val[q].name + " = " + rhs + ";\n";
          }
        }
      }
    }
    // reads advance; dead values leave the registers and free their rows
    for (int id : op.reads) {
      Val& v = val[id];
      if (v.kind == 3) continue;
      res.erase({next_use(v), id});
      ++v.nu;
      if (v.nu < v.uses.size()) {
        res.insert({next_use(v), id});
      } else {
        v.resident = false;
        if (v.home >= 0) free_rows.push_back(v.home);
      }
    }
    // definitions enter the registers
    {
      for (int id : op.defs) {
        Val& v = val[id];
        if (v.uses.empty()) continue;
        prev = v.name;
        v.resident = true;
        v.dirty = true;
        res.insert({next_use(v), id});
      }
    }
    over_budget(ii);
  }
  if (spill_rows) *spill_rows = n_spill;
  // group g's rows follow those of groups 0..g-1: [group][instance/64][row][64]
  const std::string gs = std::to_string(g);
  std::string o;
  o.reserve(b.size() + 1024);
  o += "__device__ __forceinline__ void cxj_body_g" + gs + "(const CxjArgs& A, const long long k) {\n"
       "  const long long cxj_tb = (k >> 6) * (CXJ_TROWS * 64LL) + (k & 63) + " + gs + "LL * A.lds * CXJ_TROWS;\n"
       "  const long long cxj_sb = (k >> 6) * (CXJ_SROWS * 64LL) + (k & 63) + " + gs + "LL * A.lds * CXJ_SROWS;\n";
  if (use_si) o += "  CXJ_GLOBAL const double* const cxj_si = (CXJ_GLOBAL const double*)A.st_in + cxj_tb;\n";
  if (use_so) o += "  CXJ_GLOBAL double* const cxj_so = (CXJ_GLOBAL double*)A.st_out + cxj_tb;\n";
  if (use_ss) o += "  CXJ_GLOBAL double* const cxj_ss = (CXJ_GLOBAL double*)A.spill + cxj_sb;\n";
  for (int i : use_arg) {
    const std::string is = std::to_string(i);
    o += "  CXJ_GLOBAL const double* const cxj_a" + is + " = (CXJ_GLOBAL const double*)A.arg.a[" + is +
         "].p + k * A.arg.a[" + is + "].s_inst;\n";
  }
  o += "  (void)cxj_tb;\n  (void)cxj_sb;\n";
  o += b;
  o += "}\n";
  return o;
}

// Kernel of segment s: the bodies of all direction groups, blockIdx.y picking
// the group, so that the groups run side by side (one wave per 64 instances
// and group) rather than one after another.
static std::string gen_tangent_kernel(const cxh_jac_s* j, const JitJac& J, int s, int R, int* spill_rows) {
  std::string bodies;
  int sp = 0;
  for (int g = 0; g < J.n_group; ++g) {
    int ns = 0;
    bodies += gen_tangent(j, J, g, s, R, &ns);
    sp = std::max(sp, ns);
  }
  if (spill_rows) *spill_rows = sp;
  std::string o = kSxOpsSource;
  o += kPrologue;
  o.reserve(o.size() + bodies.size() + 1024);
  o += "#define CXJ_TROWS " + std::to_string(J.home_rows) + "\n#define CXJ_SROWS " + std::to_string(sp) + "\n";
  o += bodies;
  o += "__device__ __forceinline__ void cxj_body(const CxjArgs& A, const long long k, const int grp) {\n"
       "  switch (grp) {\n";
  for (int g = 0; g < J.n_group; ++g)
    o += "    case " + std::to_string(g) + ": cxj_body_g" + std::to_string(g) + "(A, k); break;\n";
  o += "  }\n}\n"
       "extern \"C\" __global__ void __launch_bounds__(256) " + std::string(j->kernel) + "(const CxjArgs A) {\n"
       "  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;\n"
       "  if (k < A.n) cxj_body(A, k, (int)blockIdx.y);\n"
       "}\n";
  return finish_source(std::move(o));
}

int jit_map_jac(cxh_jac_s* j, int64_t n, const IOTable& arg, const IOTable& res, double* jac, int64_t jac_s_nz,
                int64_t jac_s_inst, hipStream_t s) {
  JitJac& J = jac_plan(j, n);
  for (int sg = 0; sg < J.n_seg; ++sg) {
    hipFunction_t& f = J.fn[sg];
    // the budget counts values, not the compiler's temporaries: where the
    // compiler still needs scratch, regenerate with fewer values resident
    for (int R = jac_reg_budget(j); !f; R -= 24) {
      if (R < 24) throw std::runtime_error(std::string(j->kernel) + ": no register budget avoids scratch spills");
      int ns = 0;
      hipFunction_t g = jit_function(gen_tangent_kernel(j, J, sg, R, &ns), j->kernel);
      if (scratch_bytes(g) > 0) continue;
      f = g;
      J.spill_rows = std::max<int64_t>(J.spill_rows, ns);
    }
  }
  const int64_t lds = round_up(n, 64);  // instances, padded to whole tiles of 64
  const int64_t homes = static_cast<int64_t>(J.n_group) * J.home_rows * lds;
  if (homes > 0) J.st.reserve(static_cast<size_t>(homes) * sizeof(double));
  if (J.spill_rows > 0) J.spill.reserve(static_cast<size_t>(J.n_group) * J.spill_rows * lds * sizeof(double));
  double* st = J.st.as<double>();
  CxjArgs A{};
  A.arg = arg;
  for (auto& a : A.arg.a)  // an absent input reads zeros: one device zero, strides 0
    if (!a.p) {
      if (!J.zero.ptr) {
        static const double zeros[8] = {};
        J.zero.upload(zeros, sizeof zeros);
      }
      a.p = J.zero.as<double>();
      a.s_nz = 0;
      a.s_inst = 0;
    }
  A.res = res;
  A.asens.a[0].p = jac;
  A.asens.a[0].s_nz = jac_s_nz;
  A.asens.a[0].s_inst = jac_s_inst;
  A.n = n;
  A.lds = lds;
  A.spill = J.spill_rows > 0 ? J.spill.as<double>() : nullptr;
  const int block = j->tape->block;
  const unsigned grid = static_cast<unsigned>((n + block - 1) / block);
  A.st_in = homes > 0 ? st : nullptr;  // home rows: read and written in place
  A.st_out = homes > 0 ? st : nullptr;
  for (int sg = 0; sg < J.n_seg; ++sg) {
    void* params[] = {&A};
    hip_check(hipModuleLaunchKernel(J.fn[sg], grid, static_cast<unsigned>(J.n_group), 1, block, 1, 1, 0, s, params,
                                    nullptr),
              "jit tangent kernel launch");
  }
  return 0;
}

std::string jit_tangent_source(cxh_jac_s* j, int s, int* n_group, int* n_seg) {
  JitJac& J = jac_plan(j, 0);
  if (n_group) *n_group = J.n_group;
  if (n_seg) *n_seg = J.n_seg;
  if (s < 0 || s >= J.n_seg) throw std::runtime_error("segment out of range");
  return gen_tangent_kernel(j, J, s, jac_reg_budget(j), nullptr);
}

// ------------------------------------------------------ specialised LDL ---
//
// casadi_ldl + casadi_ldl_solve (casadi/core/runtime/casadi_ldl.hpp:24-109)
// for one sparsity pattern, as a straight-line program over the instance's
// own values: every operation of the batched LDL kernel (linsol_kernels.hip),
// in its order, with the pattern's indices resolved at generation time --
// including the products with entries of the work vector that are
// structurally zero (v - l*0 is v except for the sign of a zero result, so
// dropping them would not be bit-exact).  The program is evaluated by the
// register-budgeted generator of the tangent sweeps with no seed directions,
// as one kernel: the factor keeps most of L^T live until the solve, so a
// segment boundary would carry all of it through memory.
//
// Inputs: block 0 = A's nonzeros, block 1 = the right-hand sides (nrhs blocks
// of n).  Outputs: block 0 = the solutions (same layout as block 1; the same
// memory may hold both), block 1 = the number of zero or non-finite pivots.

struct LinsolJit {
  cxh_tape_s tape;
  cxh_jac_s jac;
  int nrhs = 0;
  DeviceBuffer cnt;
};

namespace {

struct TapeEmit {
  std::vector<Instr> alg;
  int ws = 0;
  // Deferred inputs: a negative handle names an input that is read afresh
  // at every use.  A solve program that named every factor entry at its top,
  // or kept one for its second use, would hold them live across thousands of
  // instructions (and across segment boundaries, through memory).
  std::vector<std::pair<int, int>> pend;  // (block, nonzero) of each handle
  int deferred(int blk, int nz) {
    pend.push_back({blk, nz});
    return -static_cast<int>(pend.size());
  }
  int use(int x) {
    if (x >= 0) return x;
    const auto& b = pend[static_cast<size_t>(-x - 1)];
    return input(b.first, b.second);
  }
  int op(int code, int x, int y) {
    x = use(x);
    y = use(y);
    Instr e{};
    e.op = code;
    e.i0 = ws;
    e.i1 = x;
    e.i2 = y;
    alg.push_back(e);
    return ws++;
  }
  int konst(double d) {
    Instr e{};
    e.op = OP_CONST;
    e.i0 = ws;
    e.d = d;
    alg.push_back(e);
    return ws++;
  }
  int un(int code, int x) { return op(code, x, x); }
  // c ? a : b for c in {0, 1}, exactly (a selected -0 stays -0): the sum of
  // the two if_else_zero terms is right up to the sign of a zero, and the
  // sign comes from the same selection applied to copysign(1, .).
  int select(int c, int a, int b, int one) {
    const int nc = un(OP_NOT, c);
    const int mag = op(OP_ADD, op(OP_IF_ELSE_ZERO, c, a), op(OP_IF_ELSE_ZERO, nc, b));
    const int sgn = op(OP_ADD, op(OP_IF_ELSE_ZERO, c, op(OP_COPYSIGN, one, a)),
                       op(OP_IF_ELSE_ZERO, nc, op(OP_COPYSIGN, one, b)));
    return op(OP_COPYSIGN, mag, sgn);
  }
  int input(int blk, int nz) {
    Instr e{};
    e.op = OP_INPUT;
    e.i0 = ws;
    e.i1 = blk;
    e.i2 = nz;
    alg.push_back(e);
    return ws++;
  }
  void output(int blk, int nz, int x) {
    x = use(x);
    Instr e{};
    e.op = OP_OUTPUT;
    e.i0 = blk;
    e.i1 = x;
    e.i2 = nz;
    alg.push_back(e);
  }
};

// Offsets of the input / output blocks in the flattened nonzero numbering.
void set_io_offsets(LinsolJit& L) {
  L.jac.in_off.clear();
  L.jac.out_off.clear();
  int acc = 0;
  for (int64_t v : L.tape.nnz_in) {
    L.jac.in_off.push_back(acc);
    acc += static_cast<int>(v);
  }
  acc = 0;
  for (int64_t v : L.tape.nnz_out) {
    L.jac.out_off.push_back(acc);
    acc += static_cast<int>(v);
  }
}

__global__ void count_to_status(const double* __restrict__ cnt, int64_t n, int32_t* __restrict__ status) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) status[k] = static_cast<int32_t>(cnt[k]);
}

}  // namespace

std::shared_ptr<LinsolJit> jit_ldl_create(int n, const std::vector<int>& a_colind, const std::vector<int>& a_row,
                                       const std::vector<int>& lt_colind, const std::vector<int>& lt_row,
                                       const std::vector<int>& p, int nrhs, int part) {
  TapeEmit e;
  const int Z = e.konst(0.0);
  const int nnz_a = a_colind[n], nnz_lt = lt_colind[n];
  std::vector<int> w(n, Z), lt(nnz_lt, Z), d(n, Z);
  if (part != LS_SOLVE) {
    // scatter the permuted A into L^T and D
    for (int c = 0; c < n; ++c) {
      const int c1 = p[c];
      for (int k = a_colind[c1]; k < a_colind[c1 + 1]; ++k) w[a_row[k]] = e.input(0, k);
      for (int k = lt_colind[c]; k < lt_colind[c + 1]; ++k) lt[k] = w[p[lt_row[k]]];
      d[c] = w[p[c]];
      for (int k = a_colind[c1]; k < a_colind[c1 + 1]; ++k) w[a_row[k]] = Z;
    }
    // up-looking factorisation, column by column
    for (int c = 0; c < n; ++c) {
      int dc = d[c];
      for (int k = lt_colind[c]; k < lt_colind[c + 1]; ++k) {
        const int r = lt_row[k];
        int v = lt[k];
        for (int k2 = lt_colind[r]; k2 < lt_colind[r + 1]; ++k2)
          v = e.op(OP_SUB, v, e.op(OP_MUL, lt[k2], w[lt_row[k2]]));
        w[r] = v;
        v = e.op(OP_DIV, v, d[r]);
        lt[k] = v;
        dc = e.op(OP_SUB, dc, e.op(OP_MUL, w[r], v));
      }
      d[c] = dc;
      for (int k = lt_colind[c]; k < lt_colind[c + 1]; ++k) w[lt_row[k]] = Z;
    }
    // zero or non-finite pivots: !(d != 0 && d - d == 0)
    int cnt = Z;
    for (int c = 0; c < n; ++c) {
      const int nz = e.op(OP_NE, d[c], Z);
      const int fin = e.op(OP_EQ, e.op(OP_SUB, d[c], d[c]), Z);
      const int bad = e.un(OP_NOT, e.op(OP_AND, nz, fin));
      cnt = e.op(OP_ADD, cnt, bad);
    }
    if (part == LS_FACTOR) {  // L^T and D leave the program: the solve program reads them back
      for (int k = 0; k < nnz_lt; ++k) e.output(0, k, lt[k]);
      for (int c = 0; c < n; ++c) e.output(1, c, d[c]);
      e.output(2, 0, cnt);
    } else {
      e.output(1, 0, cnt);
    }
  } else {
    for (int k = 0; k < nnz_lt; ++k) lt[k] = e.deferred(0, k);
    for (int c = 0; c < n; ++c) d[c] = e.deferred(1, c);
  }
  // x = P' L'^-1 D^-1 L^-1 P b for every right-hand side
  const int rhs_blk = part == LS_SOLVE ? 2 : 1;
  if (part != LS_FACTOR)
    for (int q = 0; q < nrhs; ++q) {
      for (int i = 0; i < n; ++i) w[i] = e.input(rhs_blk, q * n + p[i]);
      for (int c = 0; c < n; ++c) {
        int v = w[c];
        for (int k = lt_colind[c]; k < lt_colind[c + 1]; ++k) v = e.op(OP_SUB, v, e.op(OP_MUL, lt[k], w[lt_row[k]]));
        w[c] = v;
      }
      for (int i = 0; i < n; ++i) w[i] = e.op(OP_DIV, w[i], d[i]);
      for (int c = n - 1; c >= 0; --c) {
        const int wc = w[c];
        for (int k = lt_colind[c]; k < lt_colind[c + 1]; ++k)
          w[lt_row[k]] = e.op(OP_SUB, w[lt_row[k]], e.op(OP_MUL, wc, lt[k]));
      }
      for (int i = 0; i < n; ++i) e.output(0, q * n + p[i], w[i]);
    }
  auto L = std::make_shared<LinsolJit>();
  L->nrhs = nrhs;
  L->tape.alg = std::move(e.alg);
  L->tape.ws = e.ws;
  const int64_t nx = static_cast<int64_t>(n) * nrhs;
  if (part == LS_FUSED) {
    L->tape.nnz_in = {nnz_a, nx};
    L->tape.nnz_out = {nx, 1};
  } else if (part == LS_FACTOR) {
    L->tape.nnz_in = {nnz_a};
    L->tape.nnz_out = {nnz_lt, n, 1};
  } else {
    L->tape.nnz_in = {nnz_lt, n, nx};
    L->tape.nnz_out = {nx};
  }
  L->tape.engine = CXH_ENGINE_JIT;
  L->jac.tape = &L->tape;
  L->jac.ndir = 0;
  set_io_offsets(*L);
  // measured on the 274 x 274 KKT batch (8192 instances = 128 wavefronts):
  // budget 160 with reloads 64 instructions ahead, 0.34 ms against 0.80 ms
  // at the tangent sweeps' 200 / 8 (whose kernel also needed scratch).  The
  // solve alone runs in stretches of 2048 operations: as one straight line
  // its kernel took the compiler close to four minutes (register allocation
  // over one basic block), in 2048s a few seconds each.
  L->jac.seg_len = part == LS_SOLVE ? 2048 : static_cast<int64_t>(L->tape.alg.size());
  L->jac.reg_budget = 160;
  L->jac.prefetch = 64;
  L->jac.kernel = part == LS_FUSED ? "cxj_ldl" : part == LS_FACTOR ? "cxj_ldl_fact" : "cxj_ldl_solve";
  return L;
}

int jit_linsol_run(LinsolJit& L, int64_t n_inst, const double* a, int64_t a_s, int64_t a_t, double* x, int64_t x_s,
                int64_t x_t, int32_t* status, hipStream_t s) {
  const int64_t ld = round_up(std::max<int64_t>(n_inst, 1), 64);
  L.cnt.reserve(static_cast<size_t>(ld) * sizeof(double));
  IOTable arg{}, res{};
  arg.a[0] = {a, a_s, a_t};
  arg.a[1] = {x, x_s, x_t};
  res.a[0] = {x, x_s, x_t};
  res.a[1] = {L.cnt.as<double>(), ld, 1};
  const int rc = jit_map_jac(&L.jac, n_inst, arg, res, nullptr, 0, 0, s);
  if (rc) return rc;
  if (status) {
    hipLaunchKernelGGL(count_to_status, dim3(static_cast<unsigned>((n_inst + 255) / 256)), dim3(256), 0, s,
                       L.cnt.as<double>(), n_inst, status);
    hip_check(hipGetLastError(), "count_to_status launch");
  }
  return 0;
}

int jit_linsol_factor(LinsolJit& L, int64_t n_inst, const double* a, int64_t a_s, int64_t a_t, double* const* fac,
                      int n_fac, int64_t ld, int32_t* status, hipStream_t s) {
  if (n_fac + 1 > CXH_MAX_IO) throw std::runtime_error("too many factor blocks");
  L.cnt.reserve(static_cast<size_t>(ld) * sizeof(double));
  IOTable arg{}, res{};
  arg.a[0] = {a, a_s, a_t};
  for (int b = 0; b < n_fac; ++b) res.a[b] = {fac[b], ld, 1};
  res.a[n_fac] = {L.cnt.as<double>(), ld, 1};
  const int rc = jit_map_jac(&L.jac, n_inst, arg, res, nullptr, 0, 0, s);
  if (rc) return rc;
  if (status) {
    hipLaunchKernelGGL(count_to_status, dim3(static_cast<unsigned>((n_inst + 255) / 256)), dim3(256), 0, s,
                       L.cnt.as<double>(), n_inst, status);
    hip_check(hipGetLastError(), "count_to_status launch");
  }
  return 0;
}

int jit_linsol_solve(LinsolJit& L, int64_t n_inst, const double* const* fac, int n_fac, int64_t ld, double* x,
                     int64_t x_s, int64_t x_t, hipStream_t s) {
  if (n_fac + 1 > CXH_MAX_IO) throw std::runtime_error("too many factor blocks");
  IOTable arg{}, res{};
  for (int b = 0; b < n_fac; ++b) arg.a[b] = {fac[b], ld, 1};
  arg.a[n_fac] = {x, x_s, x_t};
  res.a[0] = {x, x_s, x_t};
  return jit_map_jac(&L.jac, n_inst, arg, res, nullptr, 0, 0, s);
}

std::string jit_linsol_source(LinsolJit& L, int segment) {
  return jit_tangent_source(&L.jac, segment, nullptr, nullptr);
}

int jit_linsol_segments(LinsolJit& L) {
  int ns = 0;
  jit_tangent_source(&L.jac, 0, nullptr, &ns);
  return ns;
}

// casadi_qr + casadi_qr_singular + casadi_qr_solve
// (casadi/core/runtime/casadi_qr.hpp:24-230) for one sparsity pattern, the
// same way: every operation of the batched QR kernel (linsol_kernels.hip) in
// its order, the selections of casadi_house exact (LinsolJit::select).  The
// program is about four times the LDL's, so it runs in segments whose values
// that cross a boundary keep home rows.
std::shared_ptr<LinsolJit> jit_qr_create(int ncol, int nrow, const std::vector<int>& a_colind,
                                         const std::vector<int>& a_row, const std::vector<int>& v_colind,
                                         const std::vector<int>& v_row, const std::vector<int>& r_colind,
                                         const std::vector<int>& r_row, const std::vector<int>& prinv,
                                         const std::vector<int>& pc, double eps, int nrhs, bool tr, int part) {
  TapeEmit e;
  const int Z = e.konst(0.0), ONE = e.konst(1.0), TWO = e.konst(2.0), M1 = e.konst(-1.0), EPS = e.konst(eps);
  const int nnz_a = a_colind[ncol], nnz_v = v_colind[ncol], nnz_r = r_colind[ncol];
  std::vector<int> x(nrow, Z), v(nnz_v, Z), r(nnz_r, Z), beta(ncol, Z);
  auto add_dot = [&](const std::vector<int>& w, int c) {  // 0.0 + sum v[k] * w[v_row[k]], in order
    int alpha = Z;
    for (int k = v_colind[c]; k < v_colind[c + 1]; ++k) alpha = e.op(OP_ADD, alpha, e.op(OP_MUL, v[k], w[v_row[k]]));
    return alpha;
  };
  auto sub_axpy = [&](int alpha, std::vector<int>& w, int c) {
    for (int k = v_colind[c]; k < v_colind[c + 1]; ++k)
      w[v_row[k]] = e.op(OP_SUB, w[v_row[k]], e.op(OP_MUL, alpha, v[k]));
  };
  if (part != LS_SOLVE) {
    int nz_r = 0;
    for (int c = 0; c < ncol; ++c) {
      const int pcc = pc[c];
      for (int k = a_colind[pcc]; k < a_colind[pcc + 1]; ++k) x[prinv[a_row[k]]] = e.input(0, k);
      for (int k = r_colind[c]; k < r_colind[c + 1] && r_row[k] < c; ++k) {
        const int rr = r_row[k];
        int alpha = add_dot(x, rr);
        alpha = e.op(OP_MUL, alpha, beta[rr]);
        sub_axpy(alpha, x, rr);
        r[nz_r++] = x[rr];
        x[rr] = Z;
      }
      for (int k = v_colind[c]; k < v_colind[c + 1]; ++k) {
        v[k] = x[v_row[k]];
        x[v_row[k]] = Z;
      }
      // casadi_house on v[v_colind[c] ..]
      const int off = v_colind[c], nv = v_colind[c + 1] - v_colind[c];
      const int v0 = v[off];
      int sigma = Z;
      for (int i = 1; i < nv; ++i) sigma = e.op(OP_ADD, sigma, e.op(OP_MUL, v[off + i], v[off + i]));
      const int sq = e.un(OP_SQRT, e.op(OP_ADD, e.op(OP_MUL, v0, v0), sigma));
      const int sz = e.op(OP_EQ, sigma, Z), np = e.op(OP_LE, v0, Z);
      const int inner =
          e.select(np, e.op(OP_SUB, v0, sq), e.op(OP_DIV, e.un(OP_NEG, sigma), e.op(OP_ADD, v0, sq)), ONE);
      const int v0n = e.select(sz, ONE, inner, ONE);
      v[off] = v0n;
      beta[c] = e.select(sz, e.op(OP_MUL, TWO, np), e.op(OP_DIV, M1, e.op(OP_MUL, sq, v0n)), ONE);
      r[nz_r++] = sq;
    }
    // casadi_qr_singular: number of !(|R_cc| >= eps)
    int cnt = Z;
    for (int c = 0; c < ncol; ++c)
      cnt = e.op(OP_ADD, cnt, e.un(OP_NOT, e.op(OP_LE, EPS, e.un(OP_FABS, r[r_colind[c + 1] - 1]))));
    if (part == LS_FACTOR) {  // V, R and beta leave the program: the solve program reads them back
      for (int k = 0; k < nnz_v; ++k) e.output(0, k, v[k]);
      for (int k = 0; k < nnz_r; ++k) e.output(1, k, r[k]);
      for (int c = 0; c < ncol; ++c) e.output(2, c, beta[c]);
      e.output(3, 0, cnt);
    } else {
      e.output(1, 0, cnt);
    }
  } else {
    // the factor leaves its work vector all zero (every entry it fills is
    // cleared again), which is the state the solve starts from
    for (int k = 0; k < nnz_v; ++k) v[k] = e.deferred(0, k);
    for (int k = 0; k < nnz_r; ++k) r[k] = e.deferred(1, k);
    for (int c = 0; c < ncol; ++c) beta[c] = e.deferred(2, c);
  }
  std::vector<int>& w = x;  // the solve reuses the factor's work vector, as the kernel does
  const int rhs_blk = part == LS_SOLVE ? 3 : 1;
  if (part != LS_FACTOR)
    for (int q = 0; q < nrhs; ++q) {
      if (tr) {
        for (int c = 0; c < ncol; ++c) w[c] = e.input(rhs_blk, q * ncol + pc[c]);
        for (int c = 0; c < ncol; ++c)
          for (int k = r_colind[c]; k < r_colind[c + 1]; ++k) {
            const int rr = r_row[k];
            w[c] = rr == c ? e.op(OP_DIV, w[c], r[k]) : e.op(OP_SUB, w[c], e.op(OP_MUL, r[k], w[rr]));
          }
        for (int c1 = 0; c1 < ncol; ++c1) {
          const int c = ncol - 1 - c1;
          sub_axpy(e.op(OP_MUL, add_dot(w, c), beta[c]), w, c);
        }
        for (int c = 0; c < ncol; ++c) e.output(0, q * ncol + c, w[prinv[c]]);
      } else {
        for (int c = 0; c < nrow; ++c) w[c] = Z;
        for (int c = 0; c < ncol; ++c) w[prinv[c]] = e.input(rhs_blk, q * ncol + c);
        for (int c = 0; c < ncol; ++c) sub_axpy(e.op(OP_MUL, add_dot(w, c), beta[c]), w, c);
        for (int c = ncol - 1; c >= 0; --c)
          for (int k = r_colind[c + 1] - 1; k >= r_colind[c]; --k) {
            const int rr = r_row[k];
            w[rr] = rr == c ? e.op(OP_DIV, w[rr], r[k]) : e.op(OP_SUB, w[rr], e.op(OP_MUL, r[k], w[c]));
          }
        for (int c = 0; c < ncol; ++c) e.output(0, q * ncol + pc[c], w[c]);
      }
    }
  auto L = std::make_shared<LinsolJit>();
  L->nrhs = nrhs;
  L->tape.alg = std::move(e.alg);
  L->tape.ws = e.ws;
  const int64_t nx = static_cast<int64_t>(ncol) * nrhs;
  if (part == LS_FUSED) {
    L->tape.nnz_in = {nnz_a, nx};
    L->tape.nnz_out = {nx, 1};
  } else if (part == LS_FACTOR) {
    L->tape.nnz_in = {nnz_a};
    L->tape.nnz_out = {nnz_v, nnz_r, ncol, 1};
  } else {
    L->tape.nnz_in = {nnz_v, nnz_r, ncol, nx};
    L->tape.nnz_out = {nx};
  }
  L->tape.engine = CXH_ENGINE_JIT;
  L->jac.tape = &L->tape;
  L->jac.ndir = 0;
  set_io_offsets(*L);
  // the solve alone: 4096 operations a segment compile in seconds, 16384 took
  // the compiler close to four minutes (register allocation over one block)
  L->jac.seg_len = part == LS_SOLVE ? 4096 : 16384;
  L->jac.reg_budget = 160;
  L->jac.prefetch = 64;
  L->jac.kernel = part == LS_FUSED ? "cxj_qr" : part == LS_FACTOR ? "cxj_qr_fact" : "cxj_qr_solve";
  return L;
}

// ------------------------------------------------------------ mapaccum ---
//
// Function::mapaccum (casadi/core/function.cpp:661-787): N chained calls of a
// step Function, the first n_accum outputs fed back as the first n_accum
// inputs; the other inputs and all outputs are stacked horizontally, step t's
// block after step t-1's.  On the GPU one lane runs the whole chain: a loop
// over steps around ONE specialised step body, the state in registers, so code
// size and compile time do not grow with N.  The adjoint checkpoints the state
// every K steps (HBM, [row][instance]), then for each K-step segment from the
// last: recompute and store the states of its steps, and sweep the steps
// backwards with a specialised reverse step body, the adjoint of the state
// carried in registers.

struct MaccPlan {
  int n_in = 0, n_out = 0, n_accum = 0;
  int64_t N = 0, ns = 0;                 // steps, state size (nonzeros)
  std::vector<int64_t> soff;             // state offset of accumulated block i
  std::vector<int64_t> acc_input_instr;  // per state entry: INPUT instruction index or -1
};

static MaccPlan macc_plan(const cxh_tape_s* t, int64_t N, int n_accum) {
  MaccPlan M;
  M.n_in = static_cast<int>(t->nnz_in.size());
  M.n_out = static_cast<int>(t->nnz_out.size());
  M.n_accum = n_accum;
  M.N = N;
  if (n_accum < 0 || n_accum > std::min(M.n_in, M.n_out)) throw std::runtime_error("mapaccum: bad n_accum");
  for (int i = 0; i < n_accum; ++i) {
    if (t->nnz_in[i] != t->nnz_out[i])
      throw std::runtime_error("mapaccum: accumulated input and output " + std::to_string(i) + " differ in size");
    M.soff.push_back(M.ns);
    M.ns += t->nnz_in[i];
  }
  M.acc_input_instr.assign(M.ns, -1);
  std::vector<int64_t> out_seen(M.ns, 0);
  for (size_t k = 0; k < t->alg.size(); ++k) {
    const Instr& e = t->alg[k];
    if (e.op == OP_INPUT && e.i1 < n_accum) M.acc_input_instr[M.soff[e.i1] + e.i2] = static_cast<int64_t>(k);
    if (e.op == OP_OUTPUT && e.i0 < n_accum) ++out_seen[M.soff[e.i0] + e.i2];
  }
  for (int64_t j = 0; j < M.ns; ++j)
    if (out_seen[j] != 1) throw std::runtime_error("mapaccum: every accumulated output nonzero needs one OUTPUT");
  return M;
}

// `const double v<k> = v<x> / c` through sxh_div_c (flagging into `wide`)
// when y is a tape constant the fast division takes; empty otherwise.
static std::string macc_fast_div(const cxh_tape_s* t, int k, const Instr& e, int64_t x, int64_t y) {
  if (e.op != OP_DIV || t->alg[y].op != OP_CONST || !fast_divisor(t->alg[y].d)) return std::string();
  const double c = t->alg[y].d;
  return "  const double v" + std::to_string(k) + " = sxh::sxh_div_c(v" + std::to_string(x) + ", " + lit(c) + ", " +
         lit(1.0 / c) + ", wide);\n";
}

// Forward step body: state s -> sn, per-step inputs/outputs at column block
// `step`.  Outputs are written when `outs` is true.  fast: the variant
// `cxj_step_fwd_f` dividing by constants through sxh_div_c (*ndiv counts
// them), else `cxj_step_fwd_x` with IEEE division.
static std::string gen_macc_step_fwd(const cxh_tape_s* t, const MaccPlan& M, bool fast, int* ndiv = nullptr) {
  std::string o = std::string("__device__ __forceinline__ void cxj_step_fwd") + (fast ? "_f" : "_x") +
                  "(const CxjArgs& A, const long long k, const long long step, const double* s, double* sn, "
                  "const bool outs" + (fast ? ", bool& wide" : "") + ") {\n";
  if (ndiv) *ndiv = 0;
  std::vector<int64_t> cur(t->ws, 0);
  char buf[320];
  const std::unordered_set<int64_t> pairs = step_trig_pairs(t);
  std::unordered_map<int64_t, std::pair<std::string, std::string>> trig;
  for (size_t kk = 0; kk < t->alg.size(); ++kk) {
    const Instr& e = t->alg[kk];
    const int k = static_cast<int>(kk);
    const bool is_trig = e.op == OP_SIN || e.op == OP_COS;
    if (is_trig && emit_trig(o, e, k, cur[e.i1], "v" + std::to_string(cur[e.i1]), false, pairs, trig)) {
      cur[e.i0] = k;
      continue;
    }
    switch (e.op) {
      case OP_CONST:
        o += "  const double v" + std::to_string(k) + " = " + lit(e.d) + ";\n";
        cur[e.i0] = k;
        break;
      case OP_INPUT:
        if (e.i1 < M.n_accum)
          std::snprintf(buf, sizeof buf, "  const double v%d = s[%lld];\n", k,
                        static_cast<long long>(M.soff[e.i1] + e.i2));
        else
          std::snprintf(buf, sizeof buf, "  const double v%d = CXJ_IN(%d, step * %lld + %d);\n", k, e.i1,
                        static_cast<long long>(t->nnz_in[e.i1]), e.i2);
        o += buf;
        cur[e.i0] = k;
        break;
      case OP_OUTPUT:
        std::snprintf(buf, sizeof buf, "  if (outs) CXJ_OUT(%d, step * %lld + %d, v%lld);\n", e.i0,
                      static_cast<long long>(t->nnz_out[e.i0]), e.i2, static_cast<long long>(cur[e.i1]));
        o += buf;
        if (e.i0 < M.n_accum)
          o += "  sn[" + std::to_string(M.soff[e.i0] + e.i2) + "] = v" + std::to_string(cur[e.i1]) + ";\n";
        break;
      default: {
        const int64_t x = cur[e.i1], y = op_is_binary(e.op) ? cur[e.i2] : x;
        const std::string fd = fast ? macc_fast_div(t, k, e, x, y) : std::string();
        if (!fd.empty()) {
          o += fd;
          if (ndiv) ++*ndiv;
        } else {
          std::snprintf(buf, sizeof buf, "  const double v%d = sxh::op_fun_full(%d, v%lld, v%lld);\n", k, e.op,
                        static_cast<long long>(x), static_cast<long long>(y));
          o += buf;
        }
        cur[e.i0] = k;
      }
    }
  }
  o += "}\n";
  return o;
}

// Reverse step body: recompute the step from s, then sweep back.  as holds
// the adjoint of the step's output state on entry and of its input state on
// exit; seeds of step t's outputs and sensitivities of its inputs are at
// column block `step`.
static std::string gen_macc_step_rev(const cxh_tape_s* t, const MaccPlan& M, bool fast) {
  const int ni = static_cast<int>(t->alg.size());
  std::string o = std::string("__device__ __forceinline__ void cxj_step_rev") + (fast ? "_f" : "_x") +
                  "(const CxjArgs& A, const long long k, const long long step, const double* s, double* as" +
                  (fast ? ", bool& wide" : "") + ") {\n";
  std::vector<int64_t> cur(t->ws, 0), ox(ni, 0), oy(ni, 0);
  char buf[320];
  std::unordered_map<int64_t, std::pair<std::string, std::string>> trig;
  for (int k = 0; k < ni; ++k) {
    const Instr& e = t->alg[k];
    const bool is_trig = e.op == OP_SIN || e.op == OP_COS;
    if (is_trig && emit_trig(o, e, k, cur[e.i1], "v" + std::to_string(cur[e.i1]), true, {}, trig)) {
      ox[k] = oy[k] = cur[e.i1];
      cur[e.i0] = k;
      continue;
    }
    switch (e.op) {
      case OP_CONST:
        o += "  const double v" + std::to_string(k) + " = " + lit(e.d) + ";\n";
        cur[e.i0] = k;
        break;
      case OP_INPUT:
        if (e.i1 < M.n_accum)
          std::snprintf(buf, sizeof buf, "  const double v%d = s[%lld];\n", k,
                        static_cast<long long>(M.soff[e.i1] + e.i2));
        else
          std::snprintf(buf, sizeof buf, "  const double v%d = CXJ_IN(%d, step * %lld + %d);\n", k, e.i1,
                        static_cast<long long>(t->nnz_in[e.i1]), e.i2);
        o += buf;
        cur[e.i0] = k;
        break;
      case OP_OUTPUT:
        ox[k] = cur[e.i1];
        break;
      default:
        ox[k] = cur[e.i1];
        oy[k] = op_is_binary(e.op) ? cur[e.i2] : ox[k];
        if (fast && !macc_fast_div(t, k, e, ox[k], oy[k]).empty()) {
          o += macc_fast_div(t, k, e, ox[k], oy[k]);
        } else {
          std::snprintf(buf, sizeof buf, "  const double v%d = sxh::op_fun_full(%d, v%lld, v%lld);\n", k, e.op,
                        static_cast<long long>(ox[k]), static_cast<long long>(oy[k]));
          o += buf;
        }
        cur[e.i0] = k;
    }
  }
  std::vector<char> decl(ni, 0);
  auto acc = [&](int64_t c, const std::string& term) {
    if (decl[c]) {
      o += "  a" + std::to_string(c) + " += " + term + ";\n";
    } else {
      o += "  double a" + std::to_string(c) + " = " + term + ";\n";
      decl[c] = 1;
    }
  };
  for (int k = ni - 1; k >= 0; --k) {
    const Instr& e = t->alg[k];
    switch (e.op) {
      case OP_CONST:
        break;
      case OP_OUTPUT:
        std::snprintf(buf, sizeof buf, "CXJ_SEED(%d, step * %lld + %d)", e.i0, static_cast<long long>(t->nnz_out[e.i0]),
                      e.i2);
        acc(ox[k], buf);
        if (e.i0 < M.n_accum) acc(ox[k], "as[" + std::to_string(M.soff[e.i0] + e.i2) + "]");
        break;
      case OP_INPUT:
        if (e.i1 >= M.n_accum) {
          std::snprintf(buf, sizeof buf, "  CXJ_SENS(%d, step * %lld + %d, %s);\n", e.i1,
                        static_cast<long long>(t->nnz_in[e.i1]), e.i2,
                        decl[k] ? ("a" + std::to_string(k)).c_str() : "0.0");
          o += buf;
        }
        break;
      default: {
        if (!decl[k]) break;
        const std::string ks = std::to_string(k);
        auto tg = (e.op == OP_SIN || e.op == OP_COS) ? trig.find(ox[k]) : trig.end();
        if (tg != trig.end()) {
          o += "  const double d" + ks + "[2] = {" +
               (e.op == OP_SIN ? tg->second.second : "-" + tg->second.first) + ", 0.0};\n";
        } else {
          o += "  double d" + ks + "[2];\n  sxh::op_der_full(" + std::to_string(e.op) + ", v" + std::to_string(ox[k]) +
               ", v" + std::to_string(oy[k]) + ", v" + ks + ", d" + ks + ");\n";
        }
        if (e.op == OP_IF_ELSE_ZERO) {
          acc(oy[k], "sxh::sxh_if_else_zero(d" + ks + "[1], a" + ks + ")");
        } else if (op_is_binary(e.op)) {  // a constant-zero partial makes no contribution
          if (!op_zero_partial(e.op, 0)) acc(ox[k], "d" + ks + "[0] * a" + ks);
          if (!op_zero_partial(e.op, 1)) acc(oy[k], "d" + ks + "[1] * a" + ks);
        } else if (!op_zero_partial(e.op, 0)) {
          acc(ox[k], "d" + ks + "[0] * a" + ks);
        }
      }
    }
  }
  // adjoint of the input state (all reads of `as` above happen first)
  for (int64_t j = 0; j < M.ns; ++j) {
    const int64_t c = M.acc_input_instr[j];
    o += "  as[" + std::to_string(j) + "] = " + (c >= 0 && decl[c] ? "a" + std::to_string(c) : std::string("0.0")) +
         ";\n";
  }
  // unread non-accumulated inputs of this step
  for (const int2& u : t->unused_in)
    if (u.x >= M.n_accum) {
      std::snprintf(buf, sizeof buf, "  CXJ_SENS(%d, step * %lld + %d, 0.0);\n", u.x,
                    static_cast<long long>(t->nnz_in[u.x]), u.y);
      o += buf;
    }
  o += "}\n";
  return o;
}

// Kernels over n instances.  mode 0: forward chain; 1: fused forward +
// adjoint; 2: forward storing the checkpoints; 3: adjoint from stored
// checkpoints.  Checkpoints (S = ceil(N/K) rows of the state, [row][lds]) at
// A.st_in; the states of one segment's steps (K rows) at A.st_out.
// 4: one segment's reverse steps beside the next-earlier segment's recompute
// (cxj_macc_pair, launched once per segment after mode 2).
static const char* const kMaccKernel[] = {"cxj_macc_fwd", "cxj_macc_adj", "cxj_macc_fwd_ckpt", "cxj_macc_rev_ckpt",
                                          "cxj_macc_pair"};

static std::string gen_macc(const cxh_tape_s* t, const MaccPlan& M, int mode, int64_t K) {
  std::string o = kSxOpsSource;
  o += kPrologue;
  // Modes 0, 2 and 4 (evaluation, checkpointing forward, paired reverse)
  // divide by tape constants through sxh_div_c in a fast pass and run a loop
  // again with IEEE division when a lane of the wavefront reported a dividend
  // outside its exact range (each pass rewrites every row and output it
  // stores).  The one-launch sweeps (1, 3) divide as IEEE 754 does.
  int nd = 0;
  const std::string fwd_f = (mode == 0 || mode == 2 || mode == 4) && fast_div_enabled()
                                ? gen_macc_step_fwd(t, M, true, &nd) : std::string();
  const bool fastv = nd > 0;
  if (fastv) o += fwd_f;
  o += gen_macc_step_fwd(t, M, false);
  if (mode == 1 || mode == 3 || mode == 4) {
    if (fastv) o += gen_macc_step_rev(t, M, true);
    o += gen_macc_step_rev(t, M, false);
  }
  auto fn = [&](const char* base, bool exact) { return std::string(base) + (fastv && !exact ? "_f" : "_x"); };
  auto flag = [&](bool exact) { return fastv && !exact ? std::string(", wide") : std::string(); };
  auto passes = [&](const std::function<std::string(bool)>& loops, const std::string& restart) {
    if (!fastv) return loops(true);
    return "  bool wide = false;\n" + loops(false) + "  if (CXJ_ANY(wide)) {\n" + restart + loops(true) + "  }\n";
  };
  const std::string ns = std::to_string(std::max<int64_t>(M.ns, 1)), nsi = std::to_string(M.ns);
  const std::string N = std::to_string(M.N), Ks = std::to_string(K);
  const int64_t S = (M.N + K - 1) / K;
  std::string load_x0, store_sens;
  for (int i = 0; i < M.n_accum; ++i)
    for (int64_t j = 0; j < t->nnz_in[i]; ++j) {
      const std::string sj = std::to_string(M.soff[i] + j);
      load_x0 += "  s[" + sj + "] = CXJ_IN(" + std::to_string(i) + ", " + std::to_string(j) + ");\n";
      store_sens += "  CXJ_SENS(" + std::to_string(i) + ", " + std::to_string(j) + ", as[" + sj + "]);\n";
    }
  if (mode == 4) {
    // Segment seg's reverse steps from its stored step states (A.st_out), the
    // state adjoint handed from launch to launch in A.adj_in / A.adj_out; and
    // the recompute of a segment from its checkpoint (A.st_in) storing its
    // step states.  Same arithmetic, in the same order, as mode 1.
    o += "__device__ __forceinline__ void cxj_back(const CxjArgs& A, const long long k, const long long seg) {\n"
         "  double s[" + ns + "], as[" + ns + "], as0[" + ns + "];\n"
         "  const long long t0 = seg * " + Ks + ", t1 = (t0 + " + Ks + " < " + N + ") ? t0 + " + Ks + " : " + N + ";\n"
         "  if (seg == " + std::to_string(S) + " - 1)\n"
         "    for (int j = 0; j < " + nsi + "; ++j) as[j] = 0.0;\n"
         "  else\n"
         "    for (int j = 0; j < " + nsi + "; ++j) as[j] = CXJ_ADJ_IN(j);\n"
         "  for (int j = 0; j < " + nsi + "; ++j) as0[j] = as[j];\n" +
         passes(
             [&](bool x) {
               return "#pragma unroll 1\n"
                      "  for (long long step = t1 - 1; step >= t0; --step) {\n"
                      "    for (int j = 0; j < " + nsi + "; ++j) s[j] = CXJ_SNAP(A.st_out, (step - t0) * " + nsi +
                      " + j);\n"
                      "    " + fn("cxj_step_rev", x) + "(A, k, step, s, as" + flag(x) + ");\n"
                      "  }\n";
             },
             "    for (int j = 0; j < " + nsi + "; ++j) as[j] = as0[j];\n") +
         "  if (seg == 0) {\n" + store_sens + "  } else {\n"
         "    for (int j = 0; j < " + nsi + "; ++j) CXJ_ADJ_OUT(j) = as[j];\n"
         "  }\n"
         "}\n"
         "__device__ __forceinline__ void cxj_recompute(const CxjArgs& A, const long long k, const long long seg) {\n"
         "  double s[" + ns + "], sn[" + ns + "];\n"
         "  const long long t0 = seg * " + Ks + ", t1 = (t0 + " + Ks + " < " + N + ") ? t0 + " + Ks + " : " + N + ";\n"
         "  for (int j = 0; j < " + nsi + "; ++j) s[j] = CXJ_SNAP(A.st_in, seg * " + nsi + " + j);\n" +
         passes(
             [&](bool x) {
               return "#pragma unroll 1\n"
                      "  for (long long step = t0; step < t1; ++step) {\n"
                      "    for (int j = 0; j < " + nsi + "; ++j) CXJ_SNAP(A.st_out, (step - t0) * " + nsi +
                      " + j) = s[j];\n"
                      "    if (step + 1 < t1) {\n"
                      "      " + fn("cxj_step_fwd", x) + "(A, k, step, s, sn, false" + flag(x) + ");\n"
                      "      for (int j = 0; j < " + nsi + "; ++j) s[j] = sn[j];\n"
                      "    }\n"
                      "  }\n";
             },
             "    for (int j = 0; j < " + nsi + "; ++j) s[j] = CXJ_SNAP(A.st_in, seg * " + nsi + " + j);\n") +
         "}\n"
         "extern \"C\" __global__ void __launch_bounds__(256) cxj_macc_pair(const CxjArgs A, const CxjArgs B, "
         "const long long seg) {\n"
         "  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;\n"
         "  if (k >= A.n) return;\n"
         "  if (blockIdx.y == 0) {\n"
         "    if (seg < " + std::to_string(S) + ") cxj_back(A, k, seg);\n"
         "  } else if (seg >= 1) {\n"
         "    cxj_recompute(B, k, seg - 1);\n"
         "  }\n"
         "}\n";
    return finish_source(std::move(o));
  }
  o += "__device__ __forceinline__ void cxj_body(const CxjArgs& A, const long long k) {\n"
       "  double s[" + ns + "], sn[" + ns + "], as[" + ns + "];\n"
       "  double* ck = const_cast<double*>(A.st_in);\n";
  if (mode != 3) {  // forward chain
    o += load_x0 + passes(
                       [&](bool x) {
                         std::string l = "#pragma unroll 1\n"
                                         "  for (long long step = 0; step < " + N + "; ++step) {\n";
                         if (mode != 0)
                           l += "    if (step % " + Ks + " == 0)\n"
                                "      for (int j = 0; j < " + nsi + "; ++j) CXJ_SNAP(ck, (step / " + Ks + ") * " +
                                nsi + " + j) = s[j];\n";
                         if (mode == 2)  // with A.st_out, also the step states of the last segment (what its recompute would store)
                           l += "    if (A.st_out && step >= " + std::to_string((S - 1) * K) + ")\n"
                                "      for (int j = 0; j < " + nsi + "; ++j) CXJ_SNAP(A.st_out, (step - " +
                                std::to_string((S - 1) * K) + ") * " + nsi + " + j) = s[j];\n";
                         return l + "    " + fn("cxj_step_fwd", x) + "(A, k, step, s, sn, true" + flag(x) + ");\n"
                                    "    for (int j = 0; j < " + nsi + "; ++j) s[j] = sn[j];\n"
                                    "  }\n";
                       },
                       load_x0);
  }
  if (mode == 1 || mode == 3) {  // reverse over the segments
    o += "  for (int j = 0; j < " + nsi + "; ++j) as[j] = 0.0;\n"
         "#pragma unroll 1\n"
         "  for (long long seg = " + std::to_string(S) + " - 1; seg >= 0; --seg) {\n"
         "    const long long t0 = seg * " + Ks + ", t1 = (t0 + " + Ks + " < " + N + ") ? t0 + " + Ks + " : " + N + ";\n"
         "    for (int j = 0; j < " + nsi + "; ++j) s[j] = CXJ_SNAP(ck, seg * " + nsi + " + j);\n"
         "#pragma unroll 1\n"
         "    for (long long step = t0; step < t1; ++step) {\n"
         "      for (int j = 0; j < " + nsi + "; ++j) CXJ_SNAP(A.st_out, (step - t0) * " + nsi + " + j) = s[j];\n"
         "      if (step + 1 < t1) {\n"
         "        cxj_step_fwd_x(A, k, step, s, sn, false);\n"
         "        for (int j = 0; j < " + nsi + "; ++j) s[j] = sn[j];\n"
         "      }\n"
         "    }\n"
         "#pragma unroll 1\n"
         "    for (long long step = t1 - 1; step >= t0; --step) {\n"
         "      for (int j = 0; j < " + nsi + "; ++j) s[j] = CXJ_SNAP(A.st_out, (step - t0) * " + nsi + " + j);\n"
         "      cxj_step_rev_x(A, k, step, s, as);\n"
         "    }\n"
         "  }\n" + store_sens;
  }
  o += std::string("}\nextern \"C\" __global__ void __launch_bounds__(256) ") + kMaccKernel[mode] +
       "(const CxjArgs A) {\n"
       "  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;\n"
       "  if (k < A.n) cxj_body(A, k);\n"
       "}\n";
  return finish_source(std::move(o));
}

// ----------------------------------------------------- heterogeneous batch ---

// A tile is `block` consecutive instances of one Function.
struct CxjTile {
  long long start;
  int fun;
  int pad;
};

static const char kBatchKernel[] = R"CXJ(
struct CxjTile { long long start; int fun; int pad; };
)CXJ";

// One kernel holding the fused forward + reverse body of every Function of the
// batch.  Workgroups stride over the tile list (persistent when the grid is
// smaller than the list); the Function of a tile is wave-uniform.
static std::string gen_batch(const std::vector<cxh_tape_s*>& tapes) {
  std::string o = kSxOpsSource;
  o += kPrologue;
  o += kBatchKernel;
  for (size_t f = 0; f < tapes.size(); ++f) {
    JitPlan& P = plan_for(tapes[f], 0);
    if (P.n_seg != 1) throw std::runtime_error("batch: tape is not a single segment");
    o += gen_body(tapes[f], P, 0, true, "cxj_f" + std::to_string(f), false);
  }
  o += "extern \"C\" __global__ void __launch_bounds__(256) cxj_batch(const CxjArgs* __restrict__ args, "
       "const CxjTile* __restrict__ tiles, const long long n_tiles) {\n"
       "  for (long long i = blockIdx.x; i < n_tiles; i += gridDim.x) {\n"
       "    const CxjTile T = tiles[i];\n"
       "    const long long k = T.start + threadIdx.x;\n"
       "    switch (T.fun) {\n";
  for (size_t f = 0; f < tapes.size(); ++f) {
    const std::string fs = std::to_string(f);
    o += "      case " + fs + ": { const CxjArgs& A = args[" + fs + "]; if (k < A.n) cxj_f" + fs + "(A, k); } break;\n";
  }
  o += "    }\n  }\n}\n";
  return finish_source(std::move(o));
}

}  // namespace sxh

using namespace sxh;

struct cxh_mapaccum_s {
  cxh_tape_s* tape = nullptr;
  sxh::MaccPlan plan;
  std::vector<int64_t> nnz_in, nnz_out;  // of the mapaccum Function
  std::map<std::pair<int, int64_t>, hipFunction_t> fn;  // by (mode, checkpoint interval)
  sxh::DeviceBuffer scratch;
};

struct cxh_batch_s {
  std::vector<cxh_tape_s*> tapes;
  std::vector<int64_t> counts;
  std::vector<CxjArgs> args;
  std::vector<char> bound;
  DeviceBuffer d_args, d_tiles;
  hipFunction_t fn = nullptr;
  int n_cu = 256;
  int block = 256, grid_mult = 0, order = 0;  // grid_mult 0: one workgroup per tile
  int64_t n_tiles = 0;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool dirty = true;
  int prefer_graph = -1;  // from the tuner: 1 the graph replay was faster, 0 the direct launch, -1 not tuned
  ~cxh_batch_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
  }
};

namespace sxh {

// Tile orders: 0 grouped by Function; 1 Functions interleaved in proportion
// to their tile counts; 2 grouped, most expensive Function (tape length) first.
static void build_tiles(cxh_batch_s* b) {
  std::vector<CxjTile> tiles;
  const int nf = static_cast<int>(b->tapes.size());
  std::vector<int64_t> nt(nf);
  for (int f = 0; f < nf; ++f) nt[f] = (b->counts[f] + b->block - 1) / b->block;
  auto tile = [&](int f, int64_t i) { return CxjTile{static_cast<long long>(i * b->block), f, 0}; };
  if (b->order == 1) {
    std::vector<std::pair<double, CxjTile>> key;
    for (int f = 0; f < nf; ++f)
      for (int64_t i = 0; i < nt[f]; ++i) key.push_back({(i + 0.5) / static_cast<double>(nt[f]), tile(f, i)});
    std::stable_sort(key.begin(), key.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    for (auto& kv : key) tiles.push_back(kv.second);
  } else {
    std::vector<int> fo(nf);
    for (int f = 0; f < nf; ++f) fo[f] = f;
    if (b->order == 2)
      std::stable_sort(fo.begin(), fo.end(),
                       [&](int x, int y) { return b->tapes[x]->alg.size() > b->tapes[y]->alg.size(); });
    for (int f : fo)
      for (int64_t i = 0; i < nt[f]; ++i) tiles.push_back(tile(f, i));
  }
  b->n_tiles = static_cast<int64_t>(tiles.size());
  b->d_tiles.upload(tiles.data(), tiles.size() * sizeof(CxjTile));
  b->dirty = true;
}

static void batch_launch_direct(cxh_batch_s* b, hipStream_t s) {
  if (b->n_tiles == 0) return;
  const int64_t g = b->grid_mult > 0 ? std::min<int64_t>(b->n_tiles, static_cast<int64_t>(b->grid_mult) * b->n_cu)
                                     : b->n_tiles;
  const CxjArgs* a = b->d_args.as<CxjArgs>();
  const CxjTile* tl = b->d_tiles.as<CxjTile>();
  long long nt = b->n_tiles;
  void* params[] = {&a, &tl, &nt};
  hip_check(hipModuleLaunchKernel(b->fn, static_cast<unsigned>(g), 1, 1, b->block, 1, 1, 0, s, params, nullptr),
            "cxj_batch launch");
}

// The launch captured into a graph (again after a change of launch shape).
static void batch_graph(cxh_batch_s* b) {
  if (!b->dirty) return;
  if (b->exec) hip_check(hipGraphExecDestroy(b->exec), "hipGraphExecDestroy");
  if (b->graph) hip_check(hipGraphDestroy(b->graph), "hipGraphDestroy");
  b->exec = nullptr;
  b->graph = nullptr;
  hip_check(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  batch_launch_direct(b, b->cap);
  hip_check(hipStreamEndCapture(b->cap, &b->graph), "hipStreamEndCapture");
  hip_check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  b->dirty = false;
}

static void batch_check_bound(const cxh_batch_s* b) {
  for (size_t f = 0; f < b->bound.size(); ++f)
    if (!b->bound[f] && b->counts[f] > 0)
      throw std::runtime_error("cxh_batch: Function " + std::to_string(f) + " has no bound buffers");
}

}  // namespace sxh

extern "C" {

int cxh_mapaccum_create(cxh_tape step, int64_t n_steps, int32_t n_accum, cxh_mapaccum* out) {
  CXH_TRY
  if (!step || !out) return cxh_fail("cxh_mapaccum_create: null argument");
  if (n_steps <= 0) return cxh_fail("cxh_mapaccum_create: n_steps must be positive");
  if (step->alg.empty()) return cxh_fail("cxh_mapaccum_create: empty tape");
  auto m = std::make_unique<cxh_mapaccum_s>();
  m->tape = step;
  m->plan = macc_plan(step, n_steps, n_accum);
  for (size_t i = 0; i < step->nnz_in.size(); ++i)
    m->nnz_in.push_back(static_cast<int>(i) < n_accum ? step->nnz_in[i] : n_steps * step->nnz_in[i]);
  for (size_t i = 0; i < step->nnz_out.size(); ++i) m->nnz_out.push_back(n_steps * step->nnz_out[i]);
  *out = m.release();
  return 0;
  CXH_CATCH
}

int cxh_mapaccum_destroy(cxh_mapaccum m) {
  delete m;
  return 0;
}

int cxh_mapaccum_source(cxh_mapaccum m, int32_t adjoint, int64_t checkpoint_steps, char* buf, int64_t cap,
                        int64_t* len) {
  CXH_TRY
  if (!m) return cxh_fail("null mapaccum");
  const int64_t K = checkpoint_steps > 0 ? std::min(checkpoint_steps, m->plan.N) : m->plan.N;
  if (adjoint < 0 || adjoint > 4) return cxh_fail("cxh_mapaccum_source: mode must be 0..4");
  const std::string src = gen_macc(m->tape, m->plan, adjoint, K);
  if (len) *len = static_cast<int64_t>(src.size());
  if (buf && cap > 0) {
    const size_t n = std::min<size_t>(src.size(), static_cast<size_t>(cap - 1));
    std::memcpy(buf, src.data(), n);
    buf[n] = '\0';
  }
  return 0;
  CXH_CATCH
}

static int macc_launch(cxh_mapaccum m, int mode, int64_t n, const double* const* arg,
                       const double* const* adj_out, double* const* res, double* const* adj_in, int32_t layout,
                       int64_t ld, int64_t checkpoint_steps, double* ckpt_ext, void* stream) {
  if (!m) return cxh_fail("null mapaccum");
  if (n < 0) return cxh_fail("negative n");
  if (layout != CXH_SOA && layout != CXH_AOS) return cxh_fail("bad layout");
  if (layout == CXH_SOA && ld < n) return cxh_fail("ld < n");
  if (n == 0) return 0;
  const int64_t K = checkpoint_steps > 0 ? std::min(checkpoint_steps, m->plan.N) : m->plan.N;
  const int64_t lds = round_up(n, 64);
  const int64_t S = (m->plan.N + K - 1) / K;
  const int64_t ns = std::max<int64_t>(m->plan.ns, 1);
  CxjArgs A{};
  A.arg = make_io(arg, m->nnz_in, layout, ld);
  A.res = make_io(res, m->nnz_out, layout, ld);
  A.aseed = make_io(adj_out, m->nnz_out, layout, ld);
  A.asens = make_io(adj_in, m->nnz_in, layout, ld);
  A.n = n;
  A.lds = lds;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if ((mode == 1 || mode == 3) && split_reverse() && S > 1) {
    // The reverse chain per segment, each launch running one segment's reverse
    // steps beside the recompute of the segment before it: checkpoints (S
    // states), the step states of two segments, the state adjoint twice.
    const int64_t rows = (ckpt_ext ? 0 : S * ns) + 2 * K * ns + 2 * ns;
    m->scratch.reserve(static_cast<size_t>(rows) * lds * sizeof(double));
    double* base = m->scratch.as<double>();
    double* ck = ckpt_ext ? ckpt_ext : base;
    double* states = base + (ckpt_ext ? 0 : S * ns * lds);
    double* adjb = states + 2 * K * ns * lds;
    long long first = S;  // the first pair launch: seg = S recomputes segment S - 1 only
    if (mode == 1) {      // forward storing the checkpoints, the outputs and the last segment's step states
      hipFunction_t& ff = m->fn[{2, K}];
      if (!ff)
        ff = jit_function_checked(gen_macc(m->tape, m->plan, 2, K), kMaccKernel[2],
                                  "evaluate the unrolled chain with the interpreter engine instead");
      A.st_in = ck;
      A.st_out = states + ((S - 1) % 2) * K * ns * lds;
      launch(ff, n, m->tape->block, A, st);
      first = S - 1;
    }
    hipFunction_t& fp = m->fn[{4, K}];
    if (!fp)
      fp = jit_function_checked(gen_macc(m->tape, m->plan, 4, K), kMaccKernel[4],
                                "evaluate the unrolled chain with the interpreter engine instead");
    A.res = IOTable{};
    A.st_in = ck;
    CxjArgs B = A;
    const unsigned grid = static_cast<unsigned>((n + m->tape->block - 1) / m->tape->block);
    for (long long seg = first; seg >= 0; --seg) {
      A.st_out = states + (seg % 2) * K * ns * lds;
      A.adj_in = adjb + ((seg + 1) % 2) * ns * lds;
      A.adj_out = adjb + (seg % 2) * ns * lds;
      B.st_out = states + ((seg + 1) % 2) * K * ns * lds;  // the buffer of seg - 1
      void* params[] = {&A, &B, &seg};
      hip_check(hipModuleLaunchKernel(fp, grid, 2, 1, m->tape->block, 1, 1, 0, st, params, nullptr),
                "jit kernel launch");
    }
    return 0;
  }
  hipFunction_t& f = m->fn[{mode, mode == 0 ? 0 : K}];
  if (!f)
    f = jit_function_checked(gen_macc(m->tape, m->plan, mode, K), kMaccKernel[mode],
                             "evaluate the unrolled chain with the interpreter engine instead");
  if (mode != 0) {
    const int64_t rows = (ckpt_ext ? 0 : S * ns) + (mode == 2 ? 0 : K * ns);
    m->scratch.reserve(static_cast<size_t>(std::max<int64_t>(rows, 1)) * lds * sizeof(double));
    double* base = m->scratch.as<double>();
    A.st_in = ckpt_ext ? ckpt_ext : base;
    A.st_out = mode == 2 ? nullptr : ckpt_ext ? base : base + S * ns * lds;
  }
  launch(f, n, m->tape->block, A, st);
  return 0;
}

int cxh_mapaccum_eval(cxh_mapaccum m, int64_t n, const double* const* arg, double* const* res, int32_t layout,
                      int64_t ld, void* stream) {
  CXH_TRY
  return macc_launch(m, 0, n, arg, nullptr, res, nullptr, layout, ld, 0, nullptr, stream);
  CXH_CATCH
}

int cxh_mapaccum_fwd_adj(cxh_mapaccum m, int64_t n, const double* const* arg, const double* const* adj_out,
                         double* const* res, double* const* adj_in, int32_t layout, int64_t ld,
                         int64_t checkpoint_steps, void* stream) {
  CXH_TRY
  return macc_launch(m, 1, n, arg, adj_out, res, adj_in, layout, ld, checkpoint_steps, nullptr, stream);
  CXH_CATCH
}

int cxh_mapaccum_ckpt_size(cxh_mapaccum m, int64_t n, int64_t checkpoint_steps, int64_t* n_doubles) {
  if (!m || !n_doubles) return cxh_fail("null argument");
  const int64_t K = checkpoint_steps > 0 ? std::min(checkpoint_steps, m->plan.N) : m->plan.N;
  *n_doubles = (m->plan.N + K - 1) / K * std::max<int64_t>(m->plan.ns, 1) * round_up(std::max<int64_t>(n, 1), 64);
  return 0;
}

int cxh_mapaccum_fwd_ckpt(cxh_mapaccum m, int64_t n, const double* const* arg, double* const* res,
                          int32_t layout, int64_t ld, int64_t checkpoint_steps, double* ckpt, void* stream) {
  CXH_TRY
  if (!ckpt) return cxh_fail("null checkpoint buffer");
  return macc_launch(m, 2, n, arg, nullptr, res, nullptr, layout, ld, checkpoint_steps, ckpt, stream);
  CXH_CATCH
}

int cxh_mapaccum_rev_ckpt(cxh_mapaccum m, int64_t n, const double* const* arg, const double* const* adj_out,
                          double* const* adj_in, int32_t layout, int64_t ld, int64_t checkpoint_steps,
                          const double* ckpt, void* stream) {
  CXH_TRY
  if (!ckpt) return cxh_fail("null checkpoint buffer");
  return macc_launch(m, 3, n, arg, adj_out, nullptr, adj_in, layout, ld, checkpoint_steps, const_cast<double*>(ckpt),
                     stream);
  CXH_CATCH
}

int cxh_batch_create(int32_t n_fun, const cxh_tape* tapes, const int64_t* counts, cxh_batch* out) {
  CXH_TRY
  if (!out || !tapes || !counts || n_fun <= 0) return cxh_fail("cxh_batch_create: bad argument");
  auto b = std::make_unique<cxh_batch_s>();
  for (int f = 0; f < n_fun; ++f) {
    if (!tapes[f]) return cxh_fail("cxh_batch_create: null tape");
    if (counts[f] < 0) return cxh_fail("cxh_batch_create: negative count");
    if (tapes[f]->alg.empty()) return cxh_fail("cxh_batch_create: empty tape");
    b->tapes.push_back(tapes[f]);
    b->counts.push_back(counts[f]);
  }
  b->args.assign(n_fun, CxjArgs{});
  b->bound.assign(n_fun, 0);
  b->fn = jit_function_checked(gen_batch(b->tapes), "cxj_batch",
                               "batch fewer or smaller Functions, or launch them one by one");
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hip_check(hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
  hip_check(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking), "hipStreamCreateWithFlags");
  b->d_args.reserve(sizeof(CxjArgs) * n_fun);
  build_tiles(b.get());
  *out = b.release();
  return 0;
  CXH_CATCH
}

int cxh_batch_bind(cxh_batch b, int32_t fun, const double* const* arg, const double* const* adj_out,
                   double* const* res, double* const* adj_in, int32_t layout, int64_t ld) {
  CXH_TRY
  if (!b || fun < 0 || fun >= static_cast<int32_t>(b->tapes.size())) return cxh_fail("cxh_batch_bind: bad Function");
  if (layout != CXH_SOA && layout != CXH_AOS) return cxh_fail("bad layout");
  const cxh_tape_s* t = b->tapes[fun];
  const int64_t n = b->counts[fun];
  if (layout == CXH_SOA && ld < n) return cxh_fail("ld < n");
  CxjArgs& A = b->args[fun];
  A = CxjArgs{};
  A.arg = make_io(arg, t->nnz_in, layout, ld);
  A.aseed = make_io(adj_out, t->nnz_out, layout, ld);
  A.res = make_io(res, t->nnz_out, layout, ld);
  A.asens = make_io(adj_in, t->nnz_in, layout, ld);
  A.n = n;
  A.lds = round_up(std::max<int64_t>(n, 1), 64);
  hip_check(hipMemcpy(static_cast<char*>(b->d_args.ptr) + sizeof(CxjArgs) * fun, &A, sizeof(CxjArgs),
                      hipMemcpyHostToDevice),
            "cxh_batch_bind upload");
  b->bound[fun] = 1;
  return 0;
  CXH_CATCH
}

int cxh_batch_config(cxh_batch b, int32_t block, int32_t grid_per_cu, int32_t order) {
  CXH_TRY
  if (!b) return cxh_fail("null batch");
  if (block != -1) {
    if (block != 64 && block != 128 && block != 256) return cxh_fail("block must be 64, 128 or 256");
    b->block = block;
  }
  if (grid_per_cu != -1) {
    if (grid_per_cu < 0 || grid_per_cu > 64) return cxh_fail("grid_per_cu must be in [0, 64]");
    b->grid_mult = grid_per_cu;
  }
  if (order != -1) {
    if (order < 0 || order > 2) return cxh_fail("order must be 0, 1 or 2");
    b->order = order;
  }
  b->prefer_graph = -1;  // a tuned preference holds for the tuned shape only
  build_tiles(b);
  return 0;
  CXH_CATCH
}

int cxh_batch_tune(cxh_batch b, void* stream, int32_t reps, int32_t* best, double* best_ms) {
  CXH_TRY
  if (!b) return cxh_fail("null batch");
  batch_check_bound(b);
  (void)stream;
  reps = std::max(reps, 1);
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  float best_t = 1e30f;
  int bb = b->block, bg = b->grid_mult, bo = b->order;
  const int blocks[] = {64, 128, 256};
  const int grids[] = {0, 1, 2, 4, 8};
  for (int blk : blocks)
    for (int ord = 0; ord < 3; ++ord) {
      b->block = blk;
      b->order = ord;
      build_tiles(b);
      for (int gm : grids) {
        b->grid_mult = gm;
        batch_launch_direct(b, b->cap);  // warm-up
        hip_check(hipEventRecord(e0, b->cap), "hipEventRecord");
        for (int r = 0; r < reps; ++r) batch_launch_direct(b, b->cap);
        hip_check(hipEventRecord(e1, b->cap), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        ms /= reps;
        if (ms < best_t) {
          best_t = ms;
          bb = blk;
          bg = gm;
          bo = ord;
        }
      }
    }
  b->block = bb;
  b->grid_mult = bg;
  b->order = bo;
  build_tiles(b);
  // the winning shape replayed from its graph against launched directly
  {
    batch_graph(b);
    float t_dir = 0, t_graph = 0;
    for (int pass = 0; pass < 2; ++pass) {
      if (pass == 0) batch_launch_direct(b, b->cap);
      else hip_check(hipGraphLaunch(b->exec, b->cap), "hipGraphLaunch");  // warm-up
      hip_check(hipEventRecord(e0, b->cap), "hipEventRecord");
      for (int r = 0; r < reps; ++r) {
        if (pass == 0) batch_launch_direct(b, b->cap);
        else hip_check(hipGraphLaunch(b->exec, b->cap), "hipGraphLaunch");
      }
      hip_check(hipEventRecord(e1, b->cap), "hipEventRecord");
      hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
      hip_check(hipEventElapsedTime(pass == 0 ? &t_dir : &t_graph, e0, e1), "hipEventElapsedTime");
    }
    b->prefer_graph = t_graph <= t_dir ? 1 : 0;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (best) {
    best[0] = bb;
    best[1] = bg;
    best[2] = bo;
  }
  if (best_ms) *best_ms = best_t;
  return 0;
  CXH_CATCH
}

int cxh_batch_launch(cxh_batch b, void* stream) {
  CXH_TRY
  if (!b) return cxh_fail("null batch");
  batch_check_bound(b);
  batch_graph(b);
  hip_check(hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream)), "hipGraphLaunch");
  return 0;
  CXH_CATCH
}

int cxh_batch_launch_best(cxh_batch b, void* stream) {
  CXH_TRY
  if (!b) return cxh_fail("null batch");
  batch_check_bound(b);
  if (b->prefer_graph == 0) {
    batch_launch_direct(b, static_cast<hipStream_t>(stream));
    return 0;
  }
  batch_graph(b);
  hip_check(hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream)), "hipGraphLaunch");
  return 0;
  CXH_CATCH
}

int cxh_batch_prefers_graph(cxh_batch b) {
  if (!b) return cxh_fail("null batch");
  return b->prefer_graph;
}

int cxh_batch_launch_direct(cxh_batch b, void* stream) {
  CXH_TRY
  if (!b) return cxh_fail("null batch");
  batch_check_bound(b);
  batch_launch_direct(b, static_cast<hipStream_t>(stream));
  return 0;
  CXH_CATCH
}

int cxh_batch_info(cxh_batch b, int32_t* config, int64_t* n_tiles) {
  if (!b) return cxh_fail("null batch");
  if (config) {
    config[0] = b->block;
    config[1] = b->grid_mult;
    config[2] = b->order;
  }
  if (n_tiles) *n_tiles = b->n_tiles;
  return 0;
}

int cxh_batch_destroy(cxh_batch b) {
  delete b;
  return 0;
}

int cxh_batch_source(int32_t n_fun, const cxh_tape* tapes, char* buf, int64_t cap, int64_t* len) {
  CXH_TRY
  if (!tapes || n_fun <= 0) return cxh_fail("cxh_batch_source: bad argument");
  std::vector<cxh_tape_s*> v;
  for (int f = 0; f < n_fun; ++f) {
    if (!tapes[f] || tapes[f]->alg.empty()) return cxh_fail("cxh_batch_source: null or empty tape");
    v.push_back(tapes[f]);
  }
  const std::string src = gen_batch(v);
  if (len) *len = static_cast<int64_t>(src.size());
  if (buf && cap > 0) {
    const size_t m = std::min<size_t>(src.size(), static_cast<size_t>(cap - 1));
    std::memcpy(buf, src.data(), m);
    buf[m] = '\0';
  }
  return 0;
  CXH_CATCH
}

int cxh_tape_engine(cxh_tape t, int32_t engine) {
  if (!t) return cxh_fail("null tape");
  if (engine != CXH_ENGINE_INTERP && engine != CXH_ENGINE_JIT) return cxh_fail("engine must be 0 or 1");
  t->engine = engine;
  return 0;
}

int cxh_tape_jit_async(cxh_tape t, int32_t on) {
  CXH_TRY
  if (!t) return cxh_fail("null tape");
  if (on) {
    if (!t->async) t->async = std::make_shared<JitAsync>();
    t->async->on = true;
  } else if (t->async) {
    jit_settle(t);
    t->async->on = false;
  }
  return 0;
  CXH_CATCH
}

int cxh_tape_jit_pending(cxh_tape t) {
  if (!t) return cxh_fail("null tape");
  return t->async && t->async->job && t->async->job->state.load() == 1 ? 1 : 0;
}

int cxh_tape_jit_prepare(cxh_tape t, int64_t checkpoint_every, int32_t forward, int32_t reverse) {
  CXH_TRY
  if (!t) return cxh_fail("null tape");
  if (t->alg.empty()) return 0;
  jit_settle(t);
  jit_prepare(t, checkpoint_every, forward != 0, reverse != 0);
  return 0;
  CXH_CATCH
}

int cxh_tape_jit_source(cxh_tape t, int64_t checkpoint_every, int32_t segment, int32_t reverse, char* buf,
                        int64_t cap, int64_t* len, int32_t* n_segments) {
  CXH_TRY
  if (!t) return cxh_fail("null tape");
  jit_settle(t);
  if (n_segments) *n_segments = plan_for(t, checkpoint_every).n_seg;
  if (t->alg.empty()) {
    if (len) *len = 0;
    return 0;
  }
  const std::string src = jit_source(t, checkpoint_every, segment, reverse);
  if (len) *len = static_cast<int64_t>(src.size());
  if (buf && cap > 0) {
    const size_t m = std::min<size_t>(src.size(), static_cast<size_t>(cap - 1));
    std::memcpy(buf, src.data(), m);
    buf[m] = '\0';
  }
  return 0;
  CXH_CATCH
}

int cxh_tape_jit_plan(cxh_tape t, int64_t checkpoint_every, int64_t* info, int64_t* ck_off) {
  CXH_TRY
  if (!t || !info) return cxh_fail("null argument");
  if (t->alg.empty()) return cxh_fail("empty tape");
  jit_settle(t);
  jit_plan_info(t, checkpoint_every, info, ck_off);
  return 0;
  CXH_CATCH
}

int cxh_jac_jit_source(cxh_jac j, int32_t segment, char* buf, int64_t cap, int64_t* len, int32_t* n_groups,
                       int32_t* n_segments) {
  CXH_TRY
  if (!j) return cxh_fail("null jac plan");
  int ng = 0, ns = 0;
  const std::string src = jit_tangent_source(j, segment, &ng, &ns);
  if (n_groups) *n_groups = ng;
  if (n_segments) *n_segments = ns;
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
