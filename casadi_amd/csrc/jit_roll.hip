// Flat tapes rerolled into loops.
//
// A tape built by unrolling a recurrence (a rollout written as a Python loop
// over steps, the way CasADi users write single- and multiple-shooting
// problems) is one long straight line whose middle repeats: after a prelude,
// the same P instructions come back R times, each copy reading the values of
// the copy before it where the previous copy read its predecessor's, the same
// constants and prelude values, and input / output nonzeros whose index
// advances by a fixed stride.  SXFunction::init (casadi/core/sx_function.cpp:
// 476-841) sorts the expression graph depth first, so such a tape comes out
// periodic except for the first copy, which also holds the constants at
// their first use.
//
// Instead of straight-line segments of the whole tape (jit.hip), the
// specialiser then emits ONE body of P instructions and runs it R times in a
// loop, the values that cross from one iteration to the next in registers:
// the program stays small (it fits the instruction cache and compiles in about
// a second whatever R is) and each lane evaluates its instance exactly as the
// flat tape does -- the same operations on the same operands, the adjoint of
// every value accumulated in the same (reverse tape) order.
//
// Reverse sweep: the loop state (the values an iteration hands on, plus the
// non-constant prelude values every iteration reads) is checkpointed every K
// iterations.  Then, from the last checkpoint segment to the first, one
// launch runs a segment's reverse iterations from its stored iteration states
// (blockIdx.y = 0) beside the recompute of the segment before it, storing its
// iteration states (blockIdx.y = 1), as the mapaccum kernels do
// (Function::mapaccum, casadi/core/function.cpp:661-787).  The epilogue's
// reverse sweep runs before the last segment's, the prelude's after the first
// segment's.
//
// Which adjoints are structurally nonzero (adjoint_live) must be the same in
// every iteration, or the tape is not rerolled: a handed-over adjoint is then
// declared exactly where the flat sweep declares one.  The loop adds to the
// handed-over value where the flat sweep starts from its first term, so only
// the sign of a zero adjoint may differ; every other result is bit-identical
// to the flat kernels and the interpreter.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/casadi_hip.h"
#include "capi_common.hpp"
#include "jit_common.hpp"
#include "sx_ops.h"
#include "tape_internal.hpp"

namespace sxh {

// The loop structure of a tape: prelude [0, a), R copies of the body
// [a + bP, a + (b+1)P), epilogue [a + RP, n).
struct RollPlan {
  bool ok = false;
  int64_t a = 0, P = 0, R = 0, n = 0;
  std::vector<int64_t> d1, d2;  // operands of each instruction: the instruction that defined them (-1: none)
  // Operand classes of the body (block 0, instruction a + j): kind 0 a value
  // of the same iteration, 1 a value of the previous iteration (state entry
  // ref), 2 a prelude constant (instruction ref), 3 a non-constant prelude
  // value (state entry ref).
  std::vector<int8_t> k1, k2;
  std::vector<int64_t> r1, r2;
  std::vector<int64_t> stride;  // per body position: row step of its INPUT / OUTPUT
  // State: entries carry the value of a body position out of an iteration
  // (st_pos >= 0) or a prelude value through it (st_pos = -1); st_init is the
  // prelude instruction giving the entry's value before the loop (-1: none).
  int64_t ns = 0;
  std::vector<int64_t> st_pos, st_init;
  std::vector<int64_t> pos_st;                  // body position -> entry (-1)
  std::unordered_map<int64_t, int64_t> inv_st;  // prelude instruction -> entry of a passed-through value
  // Trig rows: the operands (Emitter keys) whose sin and cos the body both
  // takes, in order of first use.  The forward passes that store iteration
  // states store each pair beside them, and the reverse iterations load the
  // pairs instead of calling sincos again.  Empty when the cache is off or
  // the body takes only one of the two for some operand.
  std::vector<int64_t> trig_keys;
  // Rows per iteration in the iteration-state buffers: the state, then two
  // per trig key.
  int64_t state_rows() const { return std::max<int64_t>(ns, 1); }
  int64_t step_rows() const { return state_rows() + 2 * static_cast<int64_t>(trig_keys.size()); }
  // Loads one iteration ahead: each loop issues the next iteration's loads
  // (its inputs; in the reverse also its state and trig rows and seeds)
  // before it computes the current one, so that their latency overlaps the
  // arithmetic instead of stalling each iteration's start.
  bool prefetch = false;
  std::vector<uint32_t> live;                   // adjoint_live bits of the flat tape
  bool is_live(int64_t k) const { return (live[k >> 5] >> (k & 31)) & 1u; }
};

struct RollTape {
  RollPlan plan;
  bool analysed = false;
  bool use = false;  // fixed at the analysis, so that checkpoint sizes and kernels never disagree
  std::unordered_map<int64_t, hipFunction_t> fn;  // (mode, K) -> kernel
  DeviceBuffer scratch;
};

namespace {

// Whether flat tapes are rerolled (CXH_JIT_ROLL=0: never).
bool roll_enabled() {
  const char* e = std::getenv("CXH_JIT_ROLL");
  return !(e && e[0] == '0');
}

// Whether the loops load one iteration ahead (CXH_JIT_ROLL_PREFETCH=0:
// never), and the most values an iteration may load for it: the loop then
// holds two sets in registers.
bool roll_prefetch_enabled() {
  const char* e = std::getenv("CXH_JIT_ROLL_PREFETCH");
  return !(e && e[0] == '0');
}
constexpr int64_t kMaxPrefetch = 48;


void ssa_operands(const cxh_tape_s* t, std::vector<int64_t>& d1, std::vector<int64_t>& d2) {
  const int64_t n = static_cast<int64_t>(t->alg.size());
  std::vector<int64_t> cur(t->ws, -1);
  d1.assign(n, -1);
  d2.assign(n, -1);
  for (int64_t k = 0; k < n; ++k) {
    const Instr& e = t->alg[k];
    switch (e.op) {
      case OP_CONST:
      case OP_INPUT:
        cur[e.i0] = k;
        break;
      case OP_OUTPUT:
        d1[k] = cur[e.i1];
        break;
      default:
        d1[k] = cur[e.i1];
        if (op_is_binary(e.op)) d2[k] = cur[e.i2];
        cur[e.i0] = k;
    }
  }
}

// Whether instruction k + P repeats instruction k: the same operation, the
// same constant, the same input / output block, and operands that are either
// the same prelude value (same instruction, before `lim`) or shifted by P.
bool repeats(const cxh_tape_s* t, const std::vector<int64_t>& d1, const std::vector<int64_t>& d2, int64_t k,
             int64_t P, int64_t lim) {
  const Instr& x = t->alg[k];
  const Instr& y = t->alg[k + P];
  if (x.op != y.op) return false;
  if (x.op == OP_CONST) return x.i1 == y.i1 && x.i2 == y.i2;  // the same bits
  if (x.op == OP_INPUT) return x.i1 == y.i1;
  if (x.op == OP_OUTPUT && x.i0 != y.i0) return false;
  for (const std::vector<int64_t>* d : {&d1, &d2}) {
    const int64_t u = (*d)[k], v = (*d)[k + P];
    if (u < 0 && v < 0) continue;
    if (u < 0 || v < 0) return false;
    if (v - u == P) continue;
    if (v == u && u < lim) continue;
    return false;
  }
  return true;
}

// Loop of period P: the longest run of repeating instructions; its start a
// is moved forward until every operand that stays the same refers to the
// prelude.  Returns the body count R (0: none).
int64_t find_run(const cxh_tape_s* t, const std::vector<int64_t>& d1, const std::vector<int64_t>& d2, int64_t P,
                 int64_t& a_out) {
  const int64_t n = static_cast<int64_t>(t->alg.size());
  int64_t best_len = 0, best_a = 0;
  for (int64_t k = 0; k + P < n;) {
    if (!repeats(t, d1, d2, k, P, n)) {
      ++k;
      continue;
    }
    int64_t e = k;
    while (e + P < n && repeats(t, d1, d2, e, P, n)) ++e;
    if (e - k > best_len) {
      best_len = e - k;
      best_a = k;
    }
    k = e;
  }
  int64_t a = best_a, end = best_a + best_len;  // repeats(k) for k in [a, end)
  for (int iter = 0; iter < 64 && end - a >= P; ++iter) {
    // operands that stay the same must lie before a
    int64_t need = a;
    for (int64_t k = a; k < end; ++k)
      for (const std::vector<int64_t>* d : {&d1, &d2}) {
        const int64_t u = (*d)[k];
        if (u >= 0 && (*d)[k + P] == u && u >= a) need = std::max(need, u + 1);
      }
    if (need == a) break;
    a += (need - a + P - 1) / P * P;
  }
  if (end - a < P) return 0;
  a_out = a;
  return (end - a) / P + 1;
}

// Classifies the operand u of body position j (block 0), given the operand v
// of the same position in block 1 (R >= 2).
bool classify(const cxh_tape_s* t, const RollPlan& L, int64_t u, int64_t v, int8_t& kind, int64_t& ref) {
  if (u < 0) {
    kind = -1;
    ref = -1;
    return true;
  }
  if (u >= L.a) {  // the same iteration
    kind = 0;
    ref = u - L.a;
    return u < L.a + L.P;
  }
  if (v == u + L.P) {  // the previous iteration's value (block -1: the prelude)
    if (u < L.a - L.P) return false;  // two iterations back
    kind = 1;
    ref = u + L.P - L.a;  // body position
    return true;
  }
  if (v == u) {
    kind = t->alg[u].op == OP_CONST ? 2 : 3;
    ref = u;
    return true;
  }
  return false;
}

// The loop of period P, when the tape has one: prelude, body count, operand
// classes and state.  False when the repeating part is short or its
// structure is not one a loop body can hold.
bool try_period(const cxh_tape_s* t, int64_t P, RollPlan& L) {
  const int64_t n = L.n;
  int64_t a = 0;
  L.R = find_run(t, L.d1, L.d2, P, a);
  L.a = a;
  L.P = P;
  if (L.R < 3 || L.R * P * 2 < n) return false;  // less than half the tape repeats
  // the input / output rows of each body position advance by a fixed stride
  L.stride.assign(P, 0);
  for (int64_t j = 0; j < P; ++j) {
    const Instr& e = t->alg[a + j];
    if (e.op != OP_INPUT && e.op != OP_OUTPUT) continue;
    const int64_t s = static_cast<int64_t>(t->alg[a + P + j].i2) - e.i2;
    for (int64_t b = 1; b < L.R; ++b)
      if (static_cast<int64_t>(t->alg[a + b * P + j].i2) != e.i2 + b * s) {
        L.R = b;  // the stride breaks: keep the blocks before
        break;
      }
    L.stride[j] = s;
  }
  if (L.R < 3 || L.R * P * 2 < n) return false;
  // which adjoints are structurally nonzero must be the same in every
  // iteration: the one body's reverse code serves them all
  for (int64_t j = 0; j < P; ++j)
    for (int64_t b = 1; b < L.R; ++b)
      if (L.is_live(a + b * P + j) != L.is_live(a + j)) return false;
  // operand classes of the body
  L.k1.assign(P, -1);
  L.k2.assign(P, -1);
  L.r1.assign(P, -1);
  L.r2.assign(P, -1);
  std::vector<char> carried(P, 0), exported(P, 0);
  std::vector<int64_t> invariants;
  std::unordered_set<int64_t> inv_seen;
  auto note = [&](int8_t kind, int64_t ref) {
    if (kind == 1) carried[ref] = 1;
    if (kind == 3 && inv_seen.insert(ref).second) invariants.push_back(ref);
  };
  for (int64_t j = 0; j < P; ++j) {
    const int64_t k = a + j;
    if (!classify(t, L, L.d1[k], L.d1[k + P], L.k1[j], L.r1[j])) return false;
    if (!classify(t, L, L.d2[k], L.d2[k + P], L.k2[j], L.r2[j])) return false;
    note(L.k1[j], L.r1[j]);
    note(L.k2[j], L.r2[j]);
  }
  // epilogue operands: its own values, the last iteration's, prelude values
  const int64_t last0 = a + (L.R - 1) * P, epi0 = a + L.R * P;
  for (int64_t k = epi0; k < n; ++k)
    for (int64_t u : {L.d1[k], L.d2[k]}) {
      if (u < 0 || u >= epi0) continue;
      if (u >= last0) {
        exported[u - last0] = 1;
      } else if (u < a) {
        if (t->alg[u].op != OP_CONST && inv_seen.insert(u).second) invariants.push_back(u);
      } else {
        return false;  // an earlier iteration's value
      }
    }
  // state entries: carried or exported positions, then passed-through values
  L.ns = 0;
  L.st_pos.clear();
  L.st_init.clear();
  L.inv_st.clear();
  L.pos_st.assign(P, -1);
  std::unordered_set<int64_t> init_used;
  for (int64_t j = 0; j < P; ++j) {
    if (!carried[j] && !exported[j]) continue;
    L.pos_st[j] = L.ns++;
    L.st_pos.push_back(j);
    const int64_t init = carried[j] ? a - P + j : -1;
    if (init >= 0 && !init_used.insert(init).second) return false;
    L.st_init.push_back(init);
  }
  for (int64_t u : invariants) {
    // one state entry per prelude value: a value both handed to the first
    // iteration and read by every iteration would split its adjoint in two
    if (!init_used.insert(u).second) return false;
    L.inv_st[u] = L.ns++;
    L.st_pos.push_back(-1);
    L.st_init.push_back(u);
  }
  return L.ns > 0;
}

void analyse(const cxh_tape_s* t, RollPlan& L) {
  L = RollPlan{};
  const int64_t n = static_cast<int64_t>(t->alg.size());
  L.n = n;
  if (n < 256) return;
  ssa_operands(t, L.d1, L.d2);
  // candidate periods: the shifts at which the operation sequence agrees most
  const int64_t pmax = std::min<int64_t>(n / 3, std::max<int64_t>(64, 200000000 / n));
  std::vector<std::pair<int64_t, int64_t>> score;
  for (int64_t P = 8; P <= pmax; ++P) {
    int64_t c = 0;
    for (int64_t k = 0; k + P < n; ++k) c += t->alg[k].op == t->alg[k + P].op;
    score.push_back({c, P});
  }
  std::sort(score.begin(), score.end(), [](const auto& x, const auto& y) {
    return x.first != y.first ? x.first > y.first : x.second < y.second;
  });
  RollPlan best;
  const std::vector<uint32_t> live = adjoint_live(t->alg, t->ws);
  for (size_t c = 0; c < std::min<size_t>(6, score.size()); ++c) {
    RollPlan cand;
    cand.n = n;
    cand.d1 = L.d1;
    cand.d2 = L.d2;
    cand.live = live;
    if (try_period(t, score[c].second, cand) && (!best.ok || cand.R * cand.P > best.R * best.P)) {
      cand.ok = true;
      best = std::move(cand);
    }
  }
  if (best.ok) L = std::move(best);
}

// ------------------------------------------------------------ codegen ---

// Names: v<k> the value of instruction k (body positions by their block-0
// instruction a + j), c<e> state entry e read by the body or the epilogue;
// adjoints a<k>, ac<e>.
struct Emitter {
  const cxh_tape_s* t;
  const RollPlan& L;
  std::string& o;
  enum Region { PRE, BODY, EPI } region;
  std::unordered_set<std::string> decl;
  std::unordered_map<int64_t, std::pair<std::string, std::string>> trig;
  // Trig rows (RollPlan::trig_keys) in the body: 1 stores each pair after
  // its sincos (when store_trig, a constant at every call site), 2 loads it
  // instead; pair i sits in rows row0 + 2i (sin) and row0 + 2i + 1 (cos) of
  // `st`.
  int trig_io = 0;
  // Iteration loads: when set, every load of the body (inputs, trig rows,
  // seeds) is recorded here and read from in_[i] instead, the caller issuing
  // the recorded expressions itself (cxj_load_fwd / cxj_load_rev).
  std::vector<std::string>* pre = nullptr;
  // Divisions by a constant of moderate magnitude through sxh_div_c (the
  // function declares `bool& wide`); div_c counts them.
  bool fast_div = false;
  int64_t div_c = 0;
  char buf[320];

  Emitter(const cxh_tape_s* t_, const RollPlan& L_, std::string& o_, Region r) : t(t_), L(L_), o(o_), region(r) {}

  std::string load(const std::string& expr) const {
    if (!pre) return expr;
    pre->push_back(expr);
    return "in_[" + std::to_string(pre->size() - 1) + "]";
  }

  // Operand w (1 or 2) of instruction k: its value name, adjoint name ("" for
  // a constant: no adjoint) and a key for the sin/cos pairing.
  void operand(int64_t k, int w, std::string& val, std::string& adj, int64_t& key) const {
    const int64_t u = w == 1 ? L.d1[k] : L.d2[k];
    if (region == BODY) {
      const int64_t j = k - L.a;
      const int8_t kind = w == 1 ? L.k1[j] : L.k2[j];
      const int64_t ref = w == 1 ? L.r1[j] : L.r2[j];
      if (kind == 0) {
        val = "v" + std::to_string(L.a + ref);
        adj = "a" + std::to_string(L.a + ref);
        key = L.a + ref;
      } else if (kind == 2) {
        val = "v" + std::to_string(ref);
        adj = "";
        key = ref;
      } else {
        const int64_t e = kind == 1 ? L.pos_st[ref] : L.inv_st.at(ref);
        val = "c" + std::to_string(e);
        adj = "ac" + std::to_string(e);
        key = -(e + 1);
      }
      return;
    }
    if (region == EPI) {
      const int64_t last0 = L.a + (L.R - 1) * L.P, epi0 = L.a + L.R * L.P;
      if (u >= epi0) {
        val = "v" + std::to_string(u);
        adj = "a" + std::to_string(u);
        key = u;
      } else if (u >= last0) {
        const int64_t e = L.pos_st[u - last0];
        val = "c" + std::to_string(e);
        adj = "ac" + std::to_string(e);
        key = -(e + 1);
      } else if (t->alg[u].op == OP_CONST) {
        val = "v" + std::to_string(u);
        adj = "";
        key = u;
      } else {
        const int64_t e = L.inv_st.at(u);
        val = "c" + std::to_string(e);
        adj = "ac" + std::to_string(e);
        key = -(e + 1);
      }
      return;
    }
    val = "v" + std::to_string(u);
    adj = t->alg[u].op == OP_CONST ? "" : "a" + std::to_string(u);
    key = u;
  }

  // Row expression of the INPUT / OUTPUT at k.
  std::string row(int64_t k, int64_t i2) const {
    if (region != BODY) return std::to_string(i2);
    const int64_t s = L.stride[k - L.a];
    return s == 0 ? std::to_string(i2) : std::to_string(i2) + " + step * " + std::to_string(s);
  }

  // Constants of other regions read here, declared as literals.
  void declare_constants(int64_t k0, int64_t k1) {
    std::unordered_set<int64_t> done;
    for (int64_t k = k0; k < k1; ++k)
      for (int w = 1; w <= 2; ++w) {
        const int64_t u = w == 1 ? L.d1[k] : L.d2[k];
        if (u < 0 || u >= k0 || t->alg[u].op != OP_CONST || !done.insert(u).second) continue;
        o += "  const double v" + std::to_string(u) + " = " + lit(t->alg[u].d) + ";\n";
      }
  }

  // Forward sweep of [k0, k1); all_trig: every sin / cos through sincos (a
  // reverse sweep follows and reads the complementary value), else only the
  // operands whose sin and cos are both taken.
  void forward(int64_t k0, int64_t k1, bool outs, bool all_trig) {
    std::unordered_set<int64_t> pairs;
    if (!all_trig) {
      std::unordered_map<int64_t, int> seen;
      for (int64_t k = k0; k < k1; ++k) {
        const Instr& e = t->alg[k];
        if (e.op != OP_SIN && e.op != OP_COS) continue;
        std::string v, a;
        int64_t key;
        operand(k, 1, v, a, key);
        seen[key] |= e.op == OP_SIN ? 1 : 2;
      }
      for (auto& kv : seen)
        if (kv.second == 3) pairs.insert(kv.first);
    }
    for (int64_t k = k0; k < k1; ++k) {
      const Instr& e = t->alg[k];
      const std::string ks = std::to_string(k);
      std::string x, ax, y, ay;
      int64_t kx = 0, ky = 0;
      if (e.op != OP_CONST && e.op != OP_INPUT) operand(k, 1, x, ax, kx);
      if (trig_io && (e.op == OP_SIN || e.op == OP_COS) && trig_row(e, k, kx, x)) continue;
      if (emit_trig(o, e, k, kx, x, all_trig, pairs, trig)) continue;
      switch (e.op) {
        case OP_CONST:
          o += "  const double v" + ks + " = " + lit(e.d) + ";\n";
          break;
        case OP_INPUT:
          o += "  const double v" + ks + " = " + load("CXJ_IN(" + std::to_string(e.i1) + ", " + row(k, e.i2) + ")") +
               ";\n";
          break;
        case OP_OUTPUT:
          if (outs) o += "  if (outs) CXJ_OUT(" + std::to_string(e.i0) + ", " + row(k, e.i2) + ", " + x + ");\n";
          break;
        default:
          if (op_is_binary(e.op)) operand(k, 2, y, ay, ky);
          else y = x;
          if (fast_div && e.op == OP_DIV && ay.empty() && fast_divisor(t->alg[ky].d)) {
            const double c = t->alg[ky].d;
            o += "  const double v" + ks + " = sxh::sxh_div_c(" + x + ", " + lit(c) + ", " + lit(1.0 / c) + ", wide);\n";
            ++div_c;
            break;
          }
          std::snprintf(buf, sizeof buf, "  const double v%s = sxh::op_fun_full(%d, %s, %s);\n", ks.c_str(), e.op,
                        x.c_str(), y.c_str());
          o += buf;
      }
    }
  }

  // Sin or cos of a cached operand: the pair from its rows (trig_io 2), or
  // from sincos and stored to them (trig_io 1).
  bool trig_row(const Instr& e, int64_t k, int64_t key, const std::string& xname) {
    const auto pos = std::find(L.trig_keys.begin(), L.trig_keys.end(), key);
    if (pos == L.trig_keys.end()) return false;
    auto it = trig.find(key);
    if (it == trig.end()) {
      const std::string id = std::to_string(k), i = std::to_string(2 * (pos - L.trig_keys.begin()));
      it = trig.emplace(key, std::make_pair("ts" + id, "tc" + id)).first;
      if (trig_io == 2) {
        o += "  const double ts" + id + " = " + load("CXJ_SNAP(st, row0 + " + i + ")") + ";\n  const double tc" + id +
             " = " + load("CXJ_SNAP(st, row0 + " + i + " + 1)") + ";\n";
      } else {
        o += "  double ts" + id + ", tc" + id + ";\n  CXJ_SINCOS(" + xname + ", ts" + id + ", tc" + id + ");\n";
        o += "  if (store_trig) {\n    CXJ_SNAP(st, row0 + " + i + ") = ts" + id + ";\n    CXJ_SNAP(st, row0 + " + i +
             " + 1) = tc" + id + ";\n  }\n";
      }
    }
    o += "  const double v" + std::to_string(k) + " = " + (e.op == OP_SIN ? it->second.first : it->second.second) +
         ";\n";
    return true;
  }

  void acc(const std::string& target, const std::string& term) {
    if (target.empty()) return;  // a constant
    if (decl.count(target)) {
      o += "  " + target + " += " + term + ";\n";
    } else {
      o += "  double " + target + " = " + term + ";\n";
      decl.insert(target);
    }
  }

  // Reverse sweep of [k0, k1) after forward(k0, k1, ..., true).
  void reverse(int64_t k0, int64_t k1) {
    for (int64_t k = k1 - 1; k >= k0; --k) {
      const Instr& e = t->alg[k];
      const std::string ks = std::to_string(k), an = "a" + ks;
      std::string x, ax, y, ay;
      int64_t kx = 0, ky = 0;
      switch (e.op) {
        case OP_CONST:
          break;
        case OP_OUTPUT:
          operand(k, 1, x, ax, kx);
          acc(ax, load("CXJ_SEED(" + std::to_string(e.i0) + ", " + row(k, e.i2) + ")"));
          break;
        case OP_INPUT:
          o += "  CXJ_SENS(" + std::to_string(e.i1) + ", " + row(k, e.i2) + ", " + (decl.count(an) ? an : "0.0") +
               ");\n";
          break;
        default: {
          if (!decl.count(an)) break;  // no contribution reaches this value
          operand(k, 1, x, ax, kx);
          if (op_is_binary(e.op)) operand(k, 2, y, ay, ky);
          else { y = x; ay = ax; }
          auto tg = (e.op == OP_SIN || e.op == OP_COS) ? trig.find(kx) : trig.end();
          if (tg != trig.end()) {  // d sin = cos, d cos = -sin, from the same sincos
            o += "  const double d" + ks + "[2] = {" + (e.op == OP_SIN ? tg->second.second : "-" + tg->second.first) +
                 ", 0.0};\n";
          } else {
            o += "  double d" + ks + "[2];\n  sxh::op_der_full(" + std::to_string(e.op) + ", " + x + ", " + y +
                 ", v" + ks + ", d" + ks + ");\n";
          }
          if (e.op == OP_IF_ELSE_ZERO) {
            acc(ay, "sxh::sxh_if_else_zero(d" + ks + "[1], " + an + ")");
          } else if (op_is_binary(e.op)) {  // a constant-zero partial makes no contribution
            if (!op_zero_partial(e.op, 0)) acc(ax, "d" + ks + "[0] * " + an);
            if (!op_zero_partial(e.op, 1)) acc(ay, "d" + ks + "[1] * " + an);
          } else if (!op_zero_partial(e.op, 0)) {
            acc(ax, "d" + ks + "[0] * " + an);
          }
        }
      }
    }
  }
};

// The operands whose sin and cos the body both takes, in order of first use
// (the trig rows); empty if the body takes only one of the two for some
// operand, as then its forward and reverse would not share the sincos.
std::vector<int64_t> trig_keys(const cxh_tape_s* t, const RollPlan& L) {
  std::string unused;
  const Emitter E(t, L, unused, Emitter::BODY);
  std::vector<int64_t> order;
  std::unordered_map<int64_t, int> seen;
  for (int64_t k = L.a; k < L.a + L.P; ++k) {
    const Instr& e = t->alg[k];
    if (e.op != OP_SIN && e.op != OP_COS) continue;
    std::string v, a;
    int64_t key = 0;
    E.operand(k, 1, v, a, key);
    if (!seen.count(key)) order.push_back(key);
    seen[key] |= e.op == OP_SIN ? 1 : 2;
  }
  for (const auto& kv : seen)
    if (kv.second != 3) return {};
  return order;
}

// The state entries read by the body (carried in or passed through).
std::string load_state(const RollPlan& L, const char* arr) {
  std::string o;
  for (int64_t e = 0; e < L.ns; ++e)
    o += "  const double c" + std::to_string(e) + " = " + arr + "[" + std::to_string(e) + "];\n";
  return o;
}

// A load function: in_[i] = the i-th recorded load expression of an iteration.
std::string load_function(const char* name, const char* params, const std::string& head,
                          const std::vector<std::string>& loads) {
  std::string o = std::string("__device__ __forceinline__ void ") + name + "(const CxjArgs& A, const long long k, " +
                  params + ", double* in_) {\n" + head;
  for (size_t i = 0; i < loads.size(); ++i) o += "  in_[" + std::to_string(i) + "] = " + loads[i] + ";\n";
  return o + "}\n";
}

// The prelude, iteration and epilogue functions.  The iteration functions
// read what they load from in_ (cxj_load_fwd / cxj_load_rev fill it); *nf and
// *nr receive how many values each loads, *fast whether they come in the two
// variants below.
std::string gen_functions(const cxh_tape_s* t, const RollPlan& L, bool rev, int64_t* nf, int64_t* nr, bool* fast,
                          bool allow_fast) {
  const int64_t a = L.a, P = L.P, epi0 = a + L.R * P, n = L.n;
  std::string o;
  {  // prelude forward: outputs, the state before the loop
    o += "__device__ __forceinline__ void cxj_pre_fwd(const CxjArgs& A, const long long k, double* s, const bool "
         "outs) {\n";
    Emitter E(t, L, o, Emitter::PRE);
    E.forward(0, a, true, false);
    for (int64_t e = 0; e < L.ns; ++e)
      o += "  s[" + std::to_string(e) + "] = " + (L.st_init[e] >= 0 ? "v" + std::to_string(L.st_init[e]) : "0.0") +
           ";\n";
    o += "}\n";
  }
  // The iteration functions: `_x` divides as IEEE 754 does; when the body
  // divides by constants, `_f` divides through sxh_div_c and reports in
  // `wide` a dividend outside its exact range (the loops then run again
  // with `_x`, see gen_roll).
  *fast = false;
  const bool fast_ok = allow_fast && fast_div_enabled();
  {  // one iteration forward
    const std::string params = "const CxjArgs& A, const long long k, const long long step, const double* s, double* sn, "
                               "const bool outs, double* st, const long long row0, const bool store_trig, "
                               "const double* in_";
    std::vector<std::string> loads;
    auto body = [&](const std::string& name, bool fast_div) {
      std::string f = "__device__ __forceinline__ void " + name + "(" + params + (fast_div ? ", bool& wide" : "") +
                      ") {\n" + load_state(L, "s");
      loads.clear();
      Emitter E(t, L, f, Emitter::BODY);
      E.trig_io = L.trig_keys.empty() ? 0 : 1;
      E.pre = &loads;
      E.fast_div = fast_div;
      E.declare_constants(a, a + P);
      E.forward(a, a + P, true, false);
      for (int64_t e = 0; e < L.ns; ++e)
        f += "  sn[" + std::to_string(e) + "] = " +
             (L.st_pos[e] >= 0 ? "v" + std::to_string(a + L.st_pos[e]) : "c" + std::to_string(e)) + ";\n";
      return std::make_pair(f + "}\n", E.div_c);
    };
    const auto fv = body("cxj_step_fwd_f", true);
    if (fast_ok && fv.second > 0) {
      o += fv.first;
      *fast = true;
    }
    o += body("cxj_step_fwd_x", false).first;
    o += load_function("cxj_load_fwd", "const long long step", "", loads);
    *nf = static_cast<int64_t>(loads.size());
  }
  {  // epilogue forward: outputs from the state after the loop
    o += "__device__ __forceinline__ void cxj_epi_fwd(const CxjArgs& A, const long long k, const double* s, const "
         "bool outs) {\n";
    o += load_state(L, "s");
    Emitter E(t, L, o, Emitter::EPI);
    E.declare_constants(epi0, n);
    E.forward(epi0, n, true, false);
    o += "}\n";
  }
  *nr = 0;
  if (!rev) return o;
  {  // epilogue reverse: the adjoint of the state after the loop
    o += "__device__ __forceinline__ void cxj_epi_rev(const CxjArgs& A, const long long k, const double* s, double* "
         "as) {\n";
    o += load_state(L, "s");
    Emitter E(t, L, o, Emitter::EPI);
    E.declare_constants(epi0, n);
    E.forward(epi0, n, false, true);
    E.reverse(epi0, n);
    for (int64_t e = 0; e < L.ns; ++e) {
      const std::string an = "ac" + std::to_string(e);
      o += "  as[" + std::to_string(e) + "] = " + (E.decl.count(an) ? an : "0.0") + ";\n";
    }
    o += "}\n";
  }
  {  // one iteration reverse: as holds the adjoint of the state the iteration
     // hands on, on return that of the state it received; its stored state
     // (rows srow + e of st) comes first in in_
    const std::string params = "const CxjArgs& A, const long long k, const long long step, double* as, const double* in_";
    std::vector<std::string> loads;
    auto body = [&](const std::string& name, bool fast_div) {
      std::string f = "__device__ __forceinline__ void " + name + "(" + params + (fast_div ? ", bool& wide" : "") +
                      ") {\n" + load_state(L, "in_");
      loads.clear();
      for (int64_t e = 0; e < L.ns; ++e) loads.push_back("CXJ_SNAP(st, srow + " + std::to_string(e) + ")");
      Emitter E(t, L, f, Emitter::BODY);
      E.trig_io = L.trig_keys.empty() ? 0 : 2;
      E.pre = &loads;
      E.fast_div = fast_div;
      E.declare_constants(a, a + P);
      E.forward(a, a + P, false, true);
      for (int64_t e = 0; e < L.ns; ++e) {  // adjoints handed over, where they are structurally nonzero
        const std::string es = std::to_string(e);
        const int64_t def = L.st_pos[e] >= 0 ? a + L.st_pos[e] : L.st_init[e];
        if (!L.is_live(def)) continue;
        const std::string nm = L.st_pos[e] >= 0 ? "a" + std::to_string(def) : "ac" + es;
        f += "  double " + nm + " = as[" + es + "];\n";
        E.decl.insert(nm);
      }
      E.reverse(a, a + P);
      for (int64_t e = 0; e < L.ns; ++e) {
        const std::string an = "ac" + std::to_string(e);
        f += "  as[" + std::to_string(e) + "] = " + (E.decl.count(an) ? an : "0.0") + ";\n";
      }
      return std::make_pair(f + "}\n", E.div_c);
    };
    const auto fv = body("cxj_step_rev_f", true);
    if (*fast) o += fv.first;  // the same divisions as the forward iteration
    o += body("cxj_step_rev_x", false).first;
    // the iteration's trig pairs sit after its state, from row srow + state_rows
    o += load_function("cxj_load_rev", "const long long step, const double* st, const long long srow",
                       "  const long long row0 = srow + " + std::to_string(L.state_rows()) + ";\n  (void)row0;\n",
                       loads);
    *nr = static_cast<int64_t>(loads.size());
  }
  {  // prelude reverse: input sensitivities from the adjoint of the state before the loop
    o += "__device__ __forceinline__ void cxj_pre_rev(const CxjArgs& A, const long long k, const double* as) {\n";
    Emitter E(t, L, o, Emitter::PRE);
    E.forward(0, a, false, true);
    for (int64_t e = 0; e < L.ns; ++e) {
      const int64_t u = L.st_init[e];
      if (u < 0 || t->alg[u].op == OP_CONST || !L.is_live(u)) continue;
      const std::string nm = "a" + std::to_string(u);
      o += "  double " + nm + " = as[" + std::to_string(e) + "];\n";
      E.decl.insert(nm);
    }
    E.reverse(0, a);
    char buf[96];
    for (const int2& u : t->unused_in) {
      std::snprintf(buf, sizeof buf, "  CXJ_SENS(%d, %d, 0.0);\n", u.x, u.y);
      o += buf;
    }
    o += "}\n";
  }
  return o;
}

// Kernels: mode 0 the forward chain (evaluation); 2 the forward chain storing
// the state every K iterations (checkpoint rows [seg][ns] at A.st_in, the
// state after the loop at row block S) and, with A.st_out, the iteration
// states of the last segment; 4 one segment's reverse iterations beside the
// recompute of the segment before it.
const char* const kRollKernel[] = {"cxj_roll_fwd", "", "cxj_roll_ckpt", "", "cxj_roll_pair"};

// lean: the variant without the one-iteration-ahead loads and the fast
// constant divisions (fewer registers: more wavefronts per SIMD for large
// batches); the same bits.
std::string gen_roll(const cxh_tape_s* t, const RollPlan& L, int mode, int64_t K, bool lean = false) {
  if (K <= 0) K = L.R;  // evaluation: no checkpoints
  std::string o = kSxOpsSource;
  o += kPrologue;
  int64_t nf = 0, nr = 0;
  bool fastv = false;
  o += gen_functions(t, L, mode == 4, &nf, &nr, &fastv, !lean);
  const int64_t S = (L.R + K - 1) / K;
  const std::string ns = std::to_string(std::max<int64_t>(L.ns, 1)), nsi = std::to_string(L.ns);
  const std::string R = std::to_string(L.R), Ks = std::to_string(K), Ss = std::to_string(S);
  // iteration-state buffers: rs rows per iteration, the trig pairs from row ns1
  const std::string rs = std::to_string(L.step_rows()), ns1 = std::to_string(L.state_rows());
  const bool trig = !L.trig_keys.empty();
  const bool pf = !lean && L.prefetch && nf <= kMaxPrefetch && nr <= kMaxPrefetch;
  const std::string nfa = std::to_string(std::max<int64_t>(nf, 1)), nra = std::to_string(std::max<int64_t>(nr, 1));
  // A loop over `step` from `first` to `last` (inclusive; down: decreasing)
  // running body(buffer), text that uses `step` and reads its iteration's
  // loads from the buffer that load(step expression, buffer) fills.  With
  // prefetch the first iteration's loads are issued before the loop and each
  // iteration issues the next one's into the other of two buffers before its
  // arithmetic, the next step clamped to `last` so that every load stays
  // within the loop's rows.  A trip holds two iterations, so the buffers
  // alternate without copies (a copy would wait for the loads in flight).
  // Without prefetch each iteration loads into b0 at its start.
  using Text = std::function<std::string(const std::string&, const std::string&)>;
  using Body = std::function<std::string(const std::string&)>;
  auto iterate = [&](bool down, const std::string& first, const std::string& last, const Text& load, const Body& body,
                     bool prefetch) {
    if (!prefetch)
      return "#pragma unroll 1\n  for (long long step = " + first + "; " + (down ? "step >= " : "step <= ") + last +
             "; " + (down ? "--step" : "++step") + ") {\n    " + load("step", "b0") + ";\n" + body("b0") + "  }\n";
    const std::string sg = down ? " - " : " + ", more = down ? " > " : " < ";
    auto next = [&](const std::string& i) { return "(" + i + more + last + " ? " + i + sg + "1 : " + last + ")"; };
    return "  {\n  long long i = " + first + ";\n  " + load("i", "b0") + ";\n#pragma unroll 1\n  for (;; i " +
           (down ? "-" : "+") + "= 2) {\n    " + load(next("i"), "b1") + ";\n    {\n    const long long step = i;\n" +
           body("b0") + "    }\n    if (i == " + last + ") break;\n    " + load(next("(i" + sg + "1)"), "b0") +
           ";\n    {\n    const long long step = i" + sg + "1;\n" + body("b1") + "    }\n    if (i" + sg + "1 == " +
           last + ") break;\n  }\n  }\n";
  };
  const Text load_fwd = [](const std::string& x, const std::string& b) {
    return "cxj_load_fwd(A, k, " + x + ", " + b + ")";
  };
  const Text load_rev = [&](const std::string& x, const std::string& b) {
    return "cxj_load_rev(A, k, " + x + ", A.st_out, (" + x + " - t0) * " + rs + ", " + b + ")";
  };
  // The iteration function of a pass and its trailing argument: the fast
  // variant, flagging into `wide`, or the IEEE one.
  auto step_fn = [&](const char* base, bool exact) { return std::string(base) + (fastv && !exact ? "_f" : "_x"); };
  auto flag = [&](bool exact) { return fastv && !exact ? std::string(", wide") : std::string(); };
  // loops(exact) once; with fast variants, the fast pass (prefetching, no
  // branch in the loop body) and, when any lane of the wavefront reported a
  // dividend outside sxh_div_c's exact range, `restart` and the loops again
  // with IEEE division: each pass writes every row and output it stores.
  auto passes = [&](const std::function<std::string(bool)>& loops, const std::string& restart) {
    if (!fastv) return loops(true);
    return "  bool wide = false;\n" + loops(false) + "  if (CXJ_ANY(wide)) {\n" + restart + loops(true) + "  }\n";
  };
  if (mode == 4) {
    o += "__device__ __forceinline__ void cxj_back(const CxjArgs& A, const long long k, const long long seg) {\n"
         "  double s[" + ns + "], as[" + ns + "], as0[" + ns + "], b0[" + nra + "], b1[" + nra + "];\n"
         "  const long long t0 = seg * " + Ks + ", t1 = (t0 + " + Ks + " < " + R + ") ? t0 + " + Ks + " : " + R + ";\n"
         "  if (seg == " + Ss + " - 1) {\n"
         "    for (int j = 0; j < " + nsi + "; ++j) s[j] = CXJ_SNAP(A.st_in, " + Ss + " * " + nsi + " + j);\n"
         "    cxj_epi_rev(A, k, s, as);\n"
         "  } else {\n"
         "    for (int j = 0; j < " + nsi + "; ++j) as[j] = CXJ_ADJ_IN(j);\n"
         "  }\n"
         "  for (int j = 0; j < " + nsi + "; ++j) as0[j] = as[j];\n" +
         passes(
             [&](bool exact) {
               return iterate(true, "t1 - 1", "t0", load_rev,
                              [&](const std::string& b) {
                                return "    " + step_fn("cxj_step_rev", exact) + "(A, k, step, as, " + b +
                                       flag(exact) + ");\n";
                              },
                              pf && !exact);
             },
             "    for (int j = 0; j < " + nsi + "; ++j) as[j] = as0[j];\n") +
         "  if (seg == 0) cxj_pre_rev(A, k, as);\n"
         "  else for (int j = 0; j < " + nsi + "; ++j) CXJ_ADJ_OUT(j) = as[j];\n"
         "}\n"
         "__device__ __forceinline__ void cxj_recompute(const CxjArgs& A, const long long k, const long long seg) {\n"
         "  double s[" + ns + "], sn[" + ns + "], b0[" + nfa + "], b1[" + nfa + "];\n"
         "  const long long t0 = seg * " + Ks + ", t1 = (t0 + " + Ks + " < " + R + ") ? t0 + " + Ks + " : " + R + ";\n"
         "  for (int j = 0; j < " + nsi + "; ++j) s[j] = CXJ_SNAP(A.st_in, seg * " + nsi + " + j);\n" +
         passes(
             [&](bool exact) {
               return iterate(false, "t0", "t1 - 1", load_fwd, [&](const std::string& b) {
                 return "    for (int j = 0; j < " + nsi + "; ++j) CXJ_SNAP(A.st_out, (step - t0) * " + rs +
                        " + j) = s[j];\n" +
                        // with trig rows the last iteration runs too: its reverse reads its pairs
                        (trig ? std::string("    {\n") : std::string("    if (step + 1 < t1) {\n")) + "      " +
                        step_fn("cxj_step_fwd", exact) + "(A, k, step, s, sn, false, A.st_out, " +
                        (trig ? "(step - t0) * " + rs + " + " + ns1 + ", true" : std::string("0, false")) + ", " + b +
                        flag(exact) + ");\n"
                        "      for (int j = 0; j < " + nsi + "; ++j) s[j] = sn[j];\n"
                        "    }\n";
               }, pf && !exact);
             },
             "    for (int j = 0; j < " + nsi + "; ++j) s[j] = CXJ_SNAP(A.st_in, seg * " + nsi + " + j);\n") +
         "}\n"
         "extern \"C\" __global__ void __launch_bounds__(256) cxj_roll_pair(const CxjArgs A, const CxjArgs B, "
         "const long long seg) {\n"
         "  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;\n"
         "  if (k >= A.n) return;\n"
         "  if (blockIdx.y == 0) {\n"
         "    if (seg < " + Ss + ") cxj_back(A, k, seg);\n"
         "  } else if (seg >= 1) {\n"
         "    cxj_recompute(B, k, seg - 1);\n"
         "  }\n"
         "}\n";
    return finish_source(std::move(o));
  }
  o += "__device__ __forceinline__ void cxj_body(const CxjArgs& A, const long long k) {\n"
       "  double s[" + ns + "], s0[" + ns + "], sn[" + ns + "], b0[" + nfa + "], b1[" + nfa + "];\n"
       "  cxj_pre_fwd(A, k, s, true);\n"
       "  for (int j = 0; j < " + nsi + "; ++j) s0[j] = s[j];\n";
  if (mode == 2) o += "  double* ck = const_cast<double*>(A.st_in);\n";
  // one loop over [from, to) (not empty): the checkpoint every K iterations
  // (mode 2), with states the iteration states of the last segment and, with
  // trig, their trig rows
  const std::string last0 = std::to_string((S - 1) * K);
  auto loop = [&](bool exact, const std::string& from, const std::string& to, bool states, bool trig_rows) {
    return iterate(false, from, to + " - 1", load_fwd, [&](const std::string& b) {
      std::string x;
      if (mode == 2)
        x += "    if (step % " + Ks + " == 0)\n"
             "      for (int j = 0; j < " + nsi + "; ++j) CXJ_SNAP(ck, (step / " + Ks + ") * " + nsi + " + j) = s[j];\n";
      if (states)
        x += "    if (A.st_out && step >= " + last0 + ")\n"
             "      for (int j = 0; j < " + nsi + "; ++j) CXJ_SNAP(A.st_out, (step - " + last0 + ") * " + rs +
             " + j) = s[j];\n";
      return x + "    " + step_fn("cxj_step_fwd", exact) + "(A, k, step, s, sn, true, " +
             (trig_rows ? "A.st_out, (step - " + last0 + ") * " + rs + " + " + ns1 + ", true"
                        : std::string("nullptr, 0, false")) +
             ", " + b + flag(exact) + ");\n"
             "    for (int j = 0; j < " + nsi + "; ++j) s[j] = sn[j];\n";
    }, pf && !exact);
  };
  o += passes(
      [&](bool exact) {
        if (mode != 2) return loop(exact, "0", R, false, false);
        if (!trig) return loop(exact, "0", R, true, false);
        // the iterations before the last segment without a branch on the trig stores
        return (S > 1 ? loop(exact, "0", last0, false, false) : std::string()) + "  if (A.st_out) {\n" +
               loop(exact, last0, R, true, true) + "  } else {\n" + loop(exact, last0, R, false, false) + "  }\n";
      },
      "    for (int j = 0; j < " + nsi + "; ++j) s[j] = s0[j];\n");
  if (mode == 2)
    o += "  for (int j = 0; j < " + nsi + "; ++j) CXJ_SNAP(ck, " + Ss + " * " + nsi + " + j) = s[j];\n";
  o += "  cxj_epi_fwd(A, k, s, true);\n"
       "}\n"
       "extern \"C\" __global__ void __launch_bounds__(256) " + std::string(kRollKernel[mode]) +
       "(const CxjArgs A) {\n"
       "  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;\n"
       "  if (k < A.n) cxj_body(A, k);\n"
       "}\n";
  return finish_source(std::move(o));
}

RollTape& roll_state(cxh_tape_s* t) {
  if (!t->roll) t->roll = std::make_shared<RollTape>();
  RollTape& r = *t->roll;
  if (!r.analysed) {
    analyse(t, r.plan);
    if (r.plan.ok && trig_cache_enabled()) r.plan.trig_keys = trig_keys(t, r.plan);
    r.plan.prefetch = roll_prefetch_enabled();
    r.use = r.plan.ok && roll_enabled();
    r.analysed = true;
  }
  return r;
}

// Iterations per checkpoint segment for a checkpoint interval of `every`
// instructions of the flat tape (<= 0: segments of about 2048 instructions).
int64_t roll_K(const RollPlan& L, int64_t every) {
  const int64_t e = every > 0 ? every : 2048;
  return std::max<int64_t>(1, std::min<int64_t>(L.R, e / L.P));
}

hipFunction_t roll_kernel(cxh_tape_s* t, int mode, int64_t K, bool lean = false) {
  RollTape& r = roll_state(t);
  hipFunction_t& f = r.fn[(mode * 2 + (lean ? 1 : 0)) * 1000003 + K];
  if (!f)
    f = jit_function_checked(gen_roll(t, r.plan, mode, K, lean), kRollKernel[mode],
                             "set CXH_JIT_ROLL=0 for the straight-line segment kernels or use the interpreter engine");
  return f;
}

// Whether a batch of n instances runs the lean kernels.  The full ones hold
// two iterations' loads and both division variants in registers (up to 190
// VGPRs for the cart-pole pair kernel, against about 64 lean): the best
// choice while a launch has at most one or two wavefronts per SIMD, and
// 2-3% slower on the MI355X once it has four or more (262 144 instances and
// up), where the lean kernels keep more wavefronts resident.  The limit is
// 128 instances per SIMD (131 072 on 1024 SIMDs); CXH_JIT_ROLL_FAT_MAX sets it
// in instances.
bool roll_lean(int64_t n) {
  if (const char* e = std::getenv("CXH_JIT_ROLL_FAT_MAX")) return n > std::atoll(e);
  static const int64_t fat_max = [] {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
    return static_cast<int64_t>(cus) * 4 * 128;
  }();
  return n > fat_max;
}

}  // namespace

// Whether the tape runs as a loop: decided once per tape, when it is first
// analysed (CXH_JIT_ROLL at that time).
bool roll_active(cxh_tape_s* t) { return !t->alg.empty() && roll_state(t).use; }

int64_t roll_ckpt_rows(cxh_tape_s* t, int64_t every) {
  const RollPlan& L = roll_state(t).plan;
  const int64_t K = roll_K(L, every);
  return ((L.R + K - 1) / K + 1) * std::max<int64_t>(L.ns, 1);
}

// Both variants (full and lean): the batch size picks one at launch, and an
// asynchronous build should leave no compile to the first call of either.
void roll_prepare(cxh_tape_s* t, int64_t every, bool fwd, bool rev) {
  const int64_t K = roll_K(roll_state(t).plan, every);
  for (const bool lean : {false, true}) {
    if (fwd) roll_kernel(t, 0, K, lean);
    if (rev) {
      roll_kernel(t, 2, K, lean);
      roll_kernel(t, 4, K, lean);
    }
  }
}

void roll_prepare_eval(cxh_tape_s* t) {
  roll_kernel(t, 0, 0, false);
  roll_kernel(t, 0, 0, true);
}

void roll_adopt(cxh_tape_s* t, cxh_tape_s& from) {
  if (!from.roll) return;
  if (!t->roll || !t->roll->analysed) {
    t->roll = from.roll;
    return;
  }
  for (const auto& kv : from.roll->fn) t->roll->fn.emplace(kv.first, kv.second);
}

int roll_map_eval(cxh_tape_s* t, int64_t n, const IOTable& arg, const IOTable& res, hipStream_t s) {
  CxjArgs A{};
  A.arg = arg;
  A.res = res;
  A.n = n;
  A.lds = round_up(n, 64);
  launch(roll_kernel(t, 0, 0, roll_lean(n)), n, t->block, A, s);
  return 0;
}

// mode 0: fused forward + adjoint; 1: forward storing checkpoints in
// ckpt_ext; 2: reverse from the checkpoints in ckpt_ext.
int roll_adj(cxh_tape_s* t, int mode, int64_t n, const IOTable& arg, const IOTable& aseed, const IOTable& res,
             const IOTable& asens, int64_t every, double* ckpt_ext, hipStream_t s) {
  RollTape& r = roll_state(t);
  const RollPlan& L = r.plan;
  const int64_t K = roll_K(L, every), S = (L.R + K - 1) / K, ns = std::max<int64_t>(L.ns, 1), rs = L.step_rows();
  const int64_t lds = round_up(n, 64);
  CxjArgs A{};
  A.arg = arg;
  A.n = n;
  A.lds = lds;
  if (mode == 1) {
    A.res = res;
    A.st_in = ckpt_ext;
    A.st_out = nullptr;
    launch(roll_kernel(t, 2, K, roll_lean(n)), n, t->block, A, s);
    return 0;
  }
  // checkpoints (unless given), the iteration states (and trig rows) of two
  // segments, the state adjoint twice
  const int64_t ck_rows = mode == 0 ? (S + 1) * ns : 0;
  r.scratch.reserve(static_cast<size_t>(ck_rows + 2 * K * rs + 2 * ns) * lds * sizeof(double));
  double* base = r.scratch.as<double>();
  double* ck = mode == 0 ? base : ckpt_ext;
  double* states = base + ck_rows * lds;
  double* adjb = states + 2 * K * rs * lds;
  long long first = S;  // seg = S: recompute of the last segment only
  if (mode == 0) {
    A.res = res;
    A.st_in = ck;
    A.st_out = states + ((S - 1) % 2) * K * rs * lds;
    launch(roll_kernel(t, 2, K, roll_lean(n)), n, t->block, A, s);
    first = S - 1;
  }
  hipFunction_t fp = roll_kernel(t, 4, K, roll_lean(n));
  A.res = IOTable{};
  A.aseed = aseed;
  A.asens = asens;
  A.st_in = ck;
  CxjArgs B = A;
  const unsigned grid = static_cast<unsigned>((n + t->block - 1) / t->block);
  for (long long seg = first; seg >= 0; --seg) {
    A.st_out = states + (seg % 2) * K * rs * lds;
    A.adj_in = adjb + ((seg + 1) % 2) * ns * lds;
    A.adj_out = adjb + (seg % 2) * ns * lds;
    B.st_out = states + ((seg + 1) % 2) * K * rs * lds;  // the buffer of seg - 1
    void* params[] = {&A, &B, &seg};
    hip_check(hipModuleLaunchKernel(fp, grid, 2, 1, t->block, 1, 1, 0, s, params, nullptr), "cxj_roll_pair launch");
  }
  return 0;
}

}  // namespace sxh

using namespace sxh;

extern "C" {

int cxh_tape_roll_plan(cxh_tape t, int64_t checkpoint_every, int64_t* info) {
  CXH_TRY
  if (!t || !info) return cxh_fail("null argument");
  jit_settle(t);
  const RollTape& r = roll_state(t);
  const RollPlan& L = r.plan;
  const int64_t K = L.ok ? roll_K(L, checkpoint_every) : 0;
  const int64_t v[9] = {L.ok ? 1 : 0, L.a, L.P, L.R, L.ns, K, K ? (L.R + K - 1) / K : 0, r.use ? 1 : 0,
                        L.ok ? L.step_rows() : 0};
  for (int i = 0; i < 9; ++i) info[i] = v[i];
  return 0;
  CXH_CATCH
}

int cxh_tape_roll_source(cxh_tape t, int64_t checkpoint_every, int32_t mode, char* buf, int64_t cap,
                         int64_t* len) {
  CXH_TRY
  if (!t) return cxh_fail("null tape");
  if (mode != 0 && mode != 2 && mode != 4) return cxh_fail("cxh_tape_roll_source: mode must be 0, 2 or 4");
  jit_settle(t);
  const RollPlan& L = roll_state(t).plan;
  if (!L.ok) return cxh_fail("cxh_tape_roll_source: the tape has no loop structure");
  const std::string src = gen_roll(t, L, mode, mode == 0 ? 0 : roll_K(L, checkpoint_every));
  if (len) *len = static_cast<int64_t>(src.size());
  if (buf && cap > 0) {
    const size_t m = std::min<size_t>(src.size(), static_cast<size_t>(cap - 1));
    std::memcpy(buf, src.data(), m);
    buf[m] = '\0';
  }
  return 0;
  CXH_CATCH
}

}  // extern "C"
