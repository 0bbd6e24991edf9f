/*
 * casadi_hip.h -- C ABI of the MI355X (gfx950) batched evaluator for CasADi
 * SX tapes and sparse LDL^T / QR solves.
 *
 * Conventions (in the style of casadi/casadi_c.h:109-131):
 *  - every function returns int: 0 = success, negative = invalid argument or
 *    HIP error (text from cxh_last_error()), positive = numerical failure
 *    count (e.g. instances with a zero pivot);
 *  - no C++ exception crosses this boundary;
 *  - device pointers are caller-owned; handles own their device copies of the
 *    tape / sparsity structure;
 *  - a handle is not re-entrant: use one handle per host thread and stream.
 *
 * Two groups of entry points:
 *  cxh_*  the GPU boundary: tapes, batched evaluation, derivative sweeps,
 *         heterogeneous batches, batched sparse LDL^T / QR.
 *  cxs_*  the host symbolic layer (SX graphs, Function construction, tape
 *         compilation, derivative Functions, colourings, orderings) that
 *         produces what the GPU boundary consumes when CasADi itself is not
 *         present.
 */
#ifndef CASADI_HIP_H
#define CASADI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Last error message of the calling thread ("" if none). */
const char* cxh_last_error(void);

/* ------------------------------------------------------------------------ */
/* GPU boundary                                                              */
/* ------------------------------------------------------------------------ */

typedef struct cxh_tape_s* cxh_tape;
typedef struct cxh_jac_s* cxh_jac;
typedef struct cxh_batch_s* cxh_batch;
typedef struct cxh_mapaccum_s* cxh_mapaccum;
typedef struct cxh_ldl_s* cxh_ldl;
typedef struct cxh_qr_s* cxh_qr;

/* Layout of per-instance data.
 *  CXH_SOA: nonzero j of instance k at p[j*ld + k]   (ld >= n; coalesced)
 *  CXH_AOS: nonzero j of instance k at p[k*nnz + j]  (CasADi Map layout,
 *           casadi/core/map.cpp:141-157)                                    */
#define CXH_SOA 0
#define CXH_AOS 1

/* Create a device tape from a CasADi SXFunction algorithm.
 * Replaces: the numeric state SXFunction holds for eval --
 *   algorithm_ (std::vector<ScalarAtomic>, casadi/core/sx_function.hpp:37-44,
 *   245-261), worksize_ and the nnz_in / nnz_out of the Function.
 * instr: n_instr records of 16 bytes laid out exactly as ScalarAtomic
 *   {int op; int i0; union {double d; struct {int i1, i2;};}}, with
 *   casadi::Operation codes (casadi/core/calculus.hpp:60-218).
 * The tape is validated (slot and I/O indices in range, supported ops);
 * Free parameters are rejected; a tape with OP_CALL records goes to
 * cxh_tape_create_nested.                                                   */
int cxh_tape_create(const void* instr, int64_t n_instr, int64_t worksize,
                    int32_t n_in, const int64_t* nnz_in,
                    int32_t n_out, const int64_t* nnz_out, cxh_tape* out);
int cxh_tape_destroy(cxh_tape t);

/* A tape holding nested Function calls.  An OP_CALL record's i1 is the index
 * k of its call: callees[k] is the called Function's tape, call_dep[k] its
 * argument slots (one per input nonzero of the callee, inputs in order) and
 * call_res[k] its result slots (one per output nonzero; -1: not needed).
 * Every call is inlined -- its arguments copied to slots of their own, the
 * callee's tape run on slots appended to the work vector, its results copied
 * out after it ends -- so every kernel sees one flat tape computing what the
 * call computes.  A callee may itself have been created nested.
 * Replaces: OP_CALL with its ExtendedAlgEl record
 *   (casadi/core/sx_function.hpp:287-313), evaluated by SXFunction::call_fwd
 *   (casadi/core/sx_function.cpp:1596-1620); the tape without calls is the
 *   same as cxh_tape_create's.                                               */
int cxh_tape_create_nested(const void* instr, int64_t n_instr, int64_t worksize,
                           int32_t n_in, const int64_t* nnz_in, int32_t n_out,
                           const int64_t* nnz_out, int32_t n_calls, const cxh_tape* callees,
                           const int32_t* const* call_dep, const int32_t* const* call_res,
                           cxh_tape* out);

/* The tape the kernels run (after inlining): *n_instr records of 16 bytes
 * (at most cap copied to instr) and its work-vector size.
 * Replaces: SXFunction::instructions_sx / instruction_* accessors
 *   (casadi/core/sx_function.hpp:187-231) for a tape held by the library.   */
int cxh_tape_instructions(cxh_tape t, void* instr, int64_t cap, int64_t* n_instr, int64_t* worksize);
/* The tape's input and output blocks: their counts and nonzeros (arrays of
 * at least *n_in / *n_out entries, at most 16 each; NULL skips).
 * Replaces: FunctionInternal::n_in / n_out / nnz_in / nnz_out
 *   (casadi/core/function_internal.hpp) for a tape held by the library.     */
int cxh_tape_io(cxh_tape t, int32_t* n_in, int32_t* n_out, int64_t* nnz_in, int64_t* nnz_out);
/* A tape from CasADi's serialized SXFunction: the text Function::serialize()
 * returns (len characters, debug mode or not).  Every field of the
 * ProtoFunction / FunctionInternal / XFunction / SXFunction bodies is read in
 * the reference's order with its version branches, the SX graphs of the
 * inputs and outputs are parsed, and the algorithm, worksize and the
 * nonzeros of the inputs and outputs go to cxh_tape_create.  Refused: other
 * Function classes, free variables, OP_CALL records, jit libraries, and any
 * stream that does not parse to its end.
 * Replaces: Function::deserialize for an SXFunction
 *   (casadi/core/function.cpp:1491-1560, casadi/core/sx_function.cpp:1926-1990,
 *   casadi/core/serializing_stream.cpp), feeding the GPU boundary.          */
int cxh_tape_deserialize(const char* data, int64_t len, cxh_tape* out);
/* Kernel configuration: threads per block (64/128/256), whether the
 * interpreter's work vector lives in LDS (1) or in HBM scratch (0), and
 * instances per wavefront of the interpreter (8/16/32/64; the 64/ipw lanes of
 * an instance group share one wavefront's instruction stream); -1 keeps the
 * current value.  Replaces nothing in the reference: launch tuning of the
 * "hip" Map.                                                                */
int cxh_tape_config(cxh_tape t, int32_t block, int32_t w_in_lds, int32_t ipw);
/* Instances per thread of the specialised forward kernels (1/2/4/8; 0: the
 * CXH_JIT_IPT setting, else 1): a thread evaluates that many instances, the
 * loads of the later ones issued while the first computes.  Replaces nothing
 * in the reference: launch tuning of the "hip" Map.                         */
int cxh_tape_ipt(cxh_tape t, int32_t ipt);
/* Autotuner of one map evaluation (cxh_map_eval's arguments; the outputs are
 * written, reps + 1 times per setting).  It times every setting the tape's
 * engine reads -- threads per block (64/128/256), then for the interpreter
 * the work vector in LDS or HBM and the lanes per instance (64/ipw for ipw in
 * 64/32/16/8), for the specialised kernels the instances per thread (1/2/4)
 * -- with events on `stream`, keeps the fastest on the tape, and returns it
 * in best[4] = {block, w_in_lds, ipw, ipt} with its time in ms per call.
 * Every setting gives the same bits.  A specialised setting whose kernel
 * would need scratch is skipped.  Replaces nothing in the reference: CasADi's
 * OmpMap leaves the split of instances over threads to OpenMP
 * (casadi/core/map.cpp:340-386).                                            */
int cxh_tape_tune(cxh_tape t, int64_t n, const double* const* arg, double* const* res, int32_t layout,
                  int64_t ld, int32_t reps, void* stream, int32_t* best, double* best_ms);
/* Regularity check of batched evaluation, per instance.  With on = 1, every
 * later cxh_map_eval, cxh_map_fwd_adj, cxh_map_fwd_ckpt and cxh_map_rev_ckpt
 * call on this tape scans what it wrote (the outputs; the input
 * sensitivities of the adjoint calls), instance by instance, into status (a
 * device array of cap int32): status[k] = 1 when instance k wrote a NaN,
 * | 2 when it wrote an infinity, 0 otherwise.  The call then synchronizes its
 * stream and returns the number of flagged instances: positive is the ABI's
 * numerical failure count, 0 means every instance is regular.  A call over
 * more than cap instances fails.  on = 0 turns the check off (status and
 * cap ignored).
 * Replaces: the error flag OmpMap::eval ORs over its instances
 *   (casadi/core/map.cpp:371-380) and the NaN / Inf scan of the outputs under
 *   the regularity_check option (casadi/core/function_internal.cpp:1016-1030),
 *   here reported per instance.                                            */
int cxh_tape_regularity_check(cxh_tape t, int32_t on, int32_t* status, int64_t cap);

/* Execution engine of a tape for cxh_map_eval / cxh_map_fwd_adj /
 * cxh_ckpt_size / cxh_map_fwd_ckpt / cxh_map_rev_ckpt:
 *  CXH_ENGINE_INTERP: the tape interpreter kernels (no compilation);
 *  CXH_ENGINE_JIT: straight-line kernels generated from the tape and compiled
 *    at first use through hiprtc, one per checkpoint segment and direction
 *    (the GPU counterpart of CasADi's C code generation of an SXFunction,
 *    SXFunction::codegen_body, casadi/core/sx_function.cpp:344-434).
 * Results agree between engines; checkpoint buffers do not carry over.      */
#define CXH_ENGINE_INTERP 0
#define CXH_ENGINE_JIT 1
int cxh_tape_engine(cxh_tape t, int32_t engine);
/* Compile (ahead of the first launch) the specialised kernels for a
 * checkpoint interval: forward segments and/or reverse segments.            */
int cxh_tape_jit_prepare(cxh_tape t, int64_t checkpoint_every, int32_t forward, int32_t reverse);
/* Asynchronous compile (on = 1) for a tape on the specialised engine: a
 * cxh_map_eval or fused cxh_map_fwd_adj call whose kernels are not built yet
 * starts a worker thread that builds them on a copy of the program, and runs
 * on the interpreter meanwhile (the two engines give the same bits); the
 * first call after the worker is done takes the specialised kernels.  Other
 * calls on the tape (the split checkpointed sweep, cxh_ckpt_size, the
 * introspection calls, cxh_tape_destroy) wait for a running worker first.
 * A build that fails is left to the specialised path, which then raises its
 * error.  on = 0 waits for a running worker and compiles in the calling
 * thread again.  cxh_tape_jit_pending: 1 while a worker runs, else 0.
 * Replaces nothing in the reference: CasADi's Map evaluates at once; this
 * hides the first call's compile of map(n, "hip").                         */
int cxh_tape_jit_async(cxh_tape t, int32_t on);
int cxh_tape_jit_pending(cxh_tape t);
/* A specialised kernel that would need scratch is not launched.  When that
 * refusal meets a tape's evaluation (no segment length avoids scratch) or its
 * adjoint at some checkpoint interval (decided before anything is launched,
 * and the same for cxh_ckpt_size, so a split sweep keeps one layout), those
 * calls run on the interpreter from then on, with the same bits.  Returns
 * bit 1 if evaluation, bit 2 if the adjoint at some interval has fallen back,
 * and the refusal's message in why (NUL-terminated, truncated to cap; may be
 * NULL).  Replaces nothing in the reference.                                  */
int cxh_tape_jit_fallback(cxh_tape t, char* why, int64_t cap);
/* Introspection: generated HIP source of one segment kernel into buf
 * (NUL-terminated, truncated to cap), its full length in *len and the number
 * of segments.  reverse: 0 the forward kernel, 1 the whole recompute +
 * reverse-sweep kernel, 2 and 3 its two halves when split (2: the first pass
 * storing the sub-block boundary states, 3: the rest), 5 the kernel of
 * forward-only evaluation (each value crossing a boundary in its home row). */
int cxh_tape_jit_source(cxh_tape t, int64_t checkpoint_every, int32_t segment, int32_t reverse,
                        char* buf, int64_t cap, int64_t* len, int32_t* n_segments);
/* Introspection: the plan of a checkpoint interval as the drivers size their
 * buffers from it.  info[7] = {segments, checkpoint rows, sub-block state rows
 * of one reverse kernel, largest live set at a segment boundary, whether the
 * reverse sweep runs split (1) or one kernel per segment (0), instances per
 * thread of the forward kernels, home rows of forward-only evaluation (one
 * buffer that every evaluation segment reads and writes)}; ck_off (may be
 * NULL; segments + 1 entries):
 * first checkpoint row of each boundary.  reverse = 4 in cxh_tape_jit_source
 * gives the source of the launch pairing segment `segment`'s reverse half
 * with the first pass of the segment before it.                            */
int cxh_tape_jit_plan(cxh_tape t, int64_t checkpoint_every, int64_t* info, int64_t* ck_off);
/* Introspection: the loop structure of a tape whose middle repeats one body
 * (a recurrence unrolled into the graph: rollouts, shooting problems), which
 * the specialised engine runs as a loop instead of straight-line segments.
 * info[9] = {found (1/0), prelude length, body length, iterations, state
 * entries crossing an iteration, iterations per checkpoint segment for
 * checkpoint_every, checkpoint segments, whether the tape runs as the loop
 * (CXH_JIT_ROLL when the tape was first analysed, and a loop was found),
 * rows per iteration in the reverse sweep's iteration-state buffers (the
 * state, then a sin and a cos row per sincos operand of the body, which the
 * recompute stores and the reverse reads; CXH_JIT_TRIG_CACHE=0: none)}.
 * Reference: the graph order of SXFunction::init
 *   (casadi/core/sx_function.cpp:476-841) is what makes such tapes periodic;
 *   evaluation as in SXFunction::eval (casadi/core/sx_function.cpp:72-127).   */
int cxh_tape_roll_plan(cxh_tape t, int64_t checkpoint_every, int64_t* info);
/* Introspection: generated source of the loop kernels; mode 0 the forward
 * chain, 2 the forward chain storing checkpoints, 4 the paired reverse
 * launch (one segment's reverse iterations beside the recompute of the
 * segment before it).                                                       */
int cxh_tape_roll_source(cxh_tape t, int64_t checkpoint_every, int32_t mode, char* buf, int64_t cap,
                         int64_t* len);
/* Introspection: generated source of the specialised tangent kernel of one
 * segment of a Jacobian plan (the bodies of all direction groups, selected by
 * blockIdx.y), with the group / segment counts.  cxh_map_jac runs these
 * kernels when the plan's tape uses CXH_ENGINE_JIT.                         */
int cxh_jac_jit_source(cxh_jac j, int32_t segment, char* buf, int64_t cap, int64_t* len,
                       int32_t* n_groups, int32_t* n_segments);
/* Compile a generated source for arch (e.g. "gfx950") without loading it;
 * 0 when it compiles, the compiler log in cxh_last_error otherwise.        */
int cxh_jit_compile_check(const char* src, const char* arch);

/* Batched forward evaluation of one tape over n instances, device pointers.
 * Replaces: Map::eval / Map::eval_gen (casadi/core/map.cpp:141-157, 327-334)
 *   looping SXFunction::eval (casadi/core/sx_function.cpp:72-127).
 * arg[i] / res[i]: per-input / per-output device arrays (NULL arg = zeros,
 *   NULL res = not requested, as in SXFunction::eval).
 * ld: leading dimension for CXH_SOA (ignored for CXH_AOS).
 * stream: hipStream_t (NULL = default stream). Asynchronous.               */
int cxh_map_eval(cxh_tape t, int64_t n, const double* const* arg, double* const* res,
                 int32_t layout, int64_t ld, void* stream);

/* Same as cxh_map_eval on HOST arrays in CasADi's AoS Map layout
 * (the drop-in for Map::eval's arg/res); synchronous.                       */
int cxh_map_eval_host(cxh_tape t, int64_t n, const double* const* arg, double* const* res);

/* Fused forward + adjoint sweep (one launch): primal outputs res and the
 * gradient adj_in = (d out / d in)^T adj_out for every instance.
 * Replaces: evaluating f and then its reverse Function adj1_<f>
 *   (XFunction::get_reverse, casadi/core/x_function.hpp:892-935, built by
 *   SXFunction::ad_reverse, casadi/core/sx_function.cpp:1368-1569).
 * checkpoint_every: instructions per checkpointed segment of the reverse
 *   sweep (0 = keep the whole trail).  res or adj_in entries may be NULL.   */
int cxh_map_fwd_adj(cxh_tape t, int64_t n, const double* const* arg,
                    const double* const* adj_out, double* const* res,
                    double* const* adj_in, int32_t layout, int64_t ld,
                    int64_t checkpoint_every, void* stream);

/* Reduced map: the n instances share the inputs flagged in reduce_in (arg[i]
 * is then nnz_in(i) doubles, read by every instance) and the outputs flagged
 * in reduce_out are summed over the instances (res[j] receives nnz_out(j)
 * doubles).  Other blocks are SoA [nz][ld] as in cxh_map_eval.  NULL flag
 * arrays mean no reduction.  The sum runs in a fixed order that depends on n
 * only (bitwise reproducible), not instance by instance.
 * Replaces: Function::map(name, parallelization, n, reduce_in, reduce_out)
 *   (casadi/core/function.cpp:796-817: repmat of the reduced inputs, repsum of
 *   the reduced outputs around a parallel map) and MapSum::eval_gen
 *   (casadi/core/mapsum.cpp:155-186) for the "hip" parallelization.         */
int cxh_map_eval_sum(cxh_tape t, int64_t n, const double* const* arg, double* const* res,
                     const int32_t* reduce_in, const int32_t* reduce_out, int64_t ld, void* stream);

/* cxh_map_eval_sum without any allocation: same arguments, same results.
 * No counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The summed instance outputs and the
 * partial sums go to one static device array of 64 MiB (8 Mi doubles) held
 * by each GPU the library runs on until the process ends, instead of the
 * tape's own buffers: rows * (round_up(n, 64) + min(1024, ceil(n / 2048)))
 * doubles must fit (rows = the summed output nonzeros), else the call fails
 * with a message.  It compiles and allocates nothing: make one
 * cxh_map_eval_sum call with the same tape and a batch at least as large
 * first (it loads the kernels and sizes the tape's work buffers), else it
 * fails saying so.  Every call on one GPU shares that GPU's array and the
 * array keeps the last call's contents: calls that overlap on one GPU
 * (other streams or threads) race.  Nothing in it needs clearing, as every
 * call writes all that it reads.                                           */
int cxh_map_eval_sum_noalloc(cxh_tape t, int64_t n, const double* const* arg, double* const* res,
                             const int32_t* reduce_in, const int32_t* reduce_out, int64_t ld, void* stream);

/* Fused forward + adjoint of a reduced map: the seed adj_out[j] of a reduced
 * output is nnz_out(j) doubles applied to every instance, the sensitivity
 * adj_in[i] of a reduced input is the sum over instances (nnz_in(i) doubles).
 * Replaces: the reverse Function of a MapSum (MapSum::get_reverse,
 *   casadi/core/mapsum.cpp:354-415) evaluated after the MapSum.              */
int cxh_map_fwd_adj_sum(cxh_tape t, int64_t n, const double* const* arg,
                        const double* const* adj_out, double* const* res, double* const* adj_in,
                        const int32_t* reduce_in, const int32_t* reduce_out, int64_t ld,
                        int64_t checkpoint_every, void* stream);

/* Dependency propagation, forward: n_words independent 64-bit seed words
 * (64 directions each) through the tape.  arg[i] holds nnz_in(i) rows of
 * words, res[j] receives nnz_out(j) rows; SoA, word k of nonzero r at
 * [r * ld + k].  A null arg[i] reads as zero, a null res[j] is skipped.
 * Replaces: SXFunction::sp_forward (casadi/core/sx_function.cpp:1657-1681),
 *   called once per 64 columns by FunctionInternal::get_jac_sparsity
 *   (casadi/core/function_internal.cpp:1928-2066).                          */
int cxh_sp_forward(cxh_tape t, int64_t n_words, const uint64_t* const* arg, uint64_t* const* res,
                   int64_t ld, void* stream);

/* Dependency propagation, reverse: the words of res[j] are seeds (cleared on
 * return, as the reference does), and each arg[i] word is ORed with the
 * dependencies found (the caller zeroes it first for a fresh sweep).
 * Replaces: SXFunction::sp_reverse (casadi/core/sx_function.cpp:1733-1782).  */
int cxh_sp_reverse(cxh_tape t, int64_t n_words, uint64_t* const* arg, uint64_t* const* res, int64_t ld,
                   void* stream);

/* Split adjoint for autograd: a forward pass that stores the work vector at
 * the start of every segment into a caller-owned device buffer `ckpt`
 * (cxh_ckpt_size doubles), then the reverse sweep restarting from those
 * checkpoints.  Same results as cxh_map_fwd_adj with the same
 * checkpoint_every.                                                         */
int cxh_ckpt_size(cxh_tape t, int64_t n, int64_t checkpoint_every, int32_t block,
                  int64_t* n_doubles);
int cxh_map_fwd_ckpt(cxh_tape t, int64_t n, const double* const* arg, double* const* res,
                     int32_t layout, int64_t ld, int64_t checkpoint_every, double* ckpt,
                     void* stream);
int cxh_map_rev_ckpt(cxh_tape t, int64_t n, const double* const* arg,
                     const double* const* adj_out, double* const* adj_in,
                     int32_t layout, int64_t ld, int64_t checkpoint_every,
                     const double* ckpt, void* stream);

/* Compressed-Jacobian sweep plan (colouring + scatter) for a tape with one
 * output block: forward (1) or adjoint (0) directions, ndir seeds given as
 * CSR (seed_ptr[ndir+1], seed_idx) over input (fwd) / output (adj) nonzeros
 * of input/output block 0, and for each of nnz_jac Jacobian nonzeros the
 * (direction, sensitivity nonzero) it is read from (-1 = structural zero).
 * Replaces: the seed / scatter bookkeeping of XFunction::jac
 *   (casadi/core/x_function.hpp:567-826) with colourings from
 *   FunctionInternal::get_partition (casadi/core/function_internal.cpp:2068-2165). */
int cxh_jac_create(cxh_tape t, int32_t forward, int64_t ndir, const int64_t* seed_ptr,
                   const int64_t* seed_idx, int64_t nnz_jac, const int64_t* src_dir,
                   const int64_t* src_nz, cxh_jac* out);
int cxh_jac_destroy(cxh_jac j);
/* Fused primal + all seed directions + scatter into the Jacobian nonzeros
 * jac (device, nnz_jac per instance) for n instances; res may be NULL.      */
int cxh_map_jac(cxh_jac j, int64_t n, const double* const* arg, double* const* res,
                double* jac, int32_t layout, int64_t ld, void* stream);

/* Heterogeneous batch: instances of several Functions, forward + adjoint each,
 * in one persistent launch captured once into a hipGraph and replayed.
 * Replaces: one Map::eval of f plus one of its reverse Function adj1_f per
 *   Function (casadi/core/map.cpp:141-157; XFunction::get_reverse,
 *   casadi/core/x_function.hpp:892-935).
 * Every tape is compiled (specialised, single segment) into one kernel whose
 * workgroups stride over a list of tiles (block consecutive instances of one
 * Function).  counts[f]: instances of tape f.  The tapes must outlive the
 * batch.                                                                     */
int cxh_batch_create(int32_t n_fun, const cxh_tape* tapes, const int64_t* counts, cxh_batch* out);
/* Device buffers of Function f (as cxh_map_fwd_adj): arg, adj_out (output
 * cotangents; NULL blocks = 0), res and adj_in (NULL = not requested).       */
int cxh_batch_bind(cxh_batch b, int32_t fun, const double* const* arg, const double* const* adj_out,
                   double* const* res, double* const* adj_in, int32_t layout, int64_t ld);
/* Launch shape: threads per block (64/128/256), persistent workgroups per
 * CU (0 = one workgroup per tile), tile order (0 grouped by Function,
 * 1 interleaved, 2 longest tape first); -1 keeps a value.                    */
int cxh_batch_config(cxh_batch b, int32_t block, int32_t grid_per_cu, int32_t order);
/* Autotuner: times every launch shape (reps launches each, hipEvents on an
 * internal stream) and keeps the fastest; best[3] = {block, grid_per_cu,
 * order}, *best_ms its time.  Then times that shape replayed from its
 * captured graph against launched directly, and remembers which was faster
 * (cxh_batch_prefers_graph; cxh_batch_launch_best takes it).  Runs the
 * batch, so buffers must be bound.                                          */
int cxh_batch_tune(cxh_batch b, void* stream, int32_t reps, int32_t* best, double* best_ms);
/* Replay the captured graph on stream (captured again after a config change). */
int cxh_batch_launch(cxh_batch b, void* stream);
/* The same kernel launched without the graph. */
int cxh_batch_launch_direct(cxh_batch b, void* stream);
/* The graph replay or the direct launch, whichever the tuner found faster
 * (the graph before any tuning or after a cxh_batch_config change).
 * cxh_batch_prefers_graph: 1 graph, 0 direct, -1 not tuned.                  */
int cxh_batch_launch_best(cxh_batch b, void* stream);
int cxh_batch_prefers_graph(cxh_batch b);
int cxh_batch_info(cxh_batch b, int32_t* config, int64_t* n_tiles);
int cxh_batch_destroy(cxh_batch b);
/* Introspection: the generated source of a batch kernel over these tapes
 * (NUL-terminated into buf, truncated to cap; full length in *len).         */
int cxh_batch_source(int32_t n_fun, const cxh_tape* tapes, char* buf, int64_t cap, int64_t* len);

/* mapaccum over a batch: N chained evaluations of a step tape per instance,
 * its first n_accum outputs fed back as its first n_accum inputs.
 * Replaces: Function::mapaccum / fold (casadi/core/function.cpp:661-787)
 *   evaluated by Map::eval per instance.
 * The mapaccum Function's blocks: inputs i < n_accum are the initial state
 *   (nnz_in(i)); the other inputs and every output are the N step blocks
 *   stacked (N*nnz, step t's block after step t-1's), as CasADi's horzcat.
 * One lane runs the chain: a loop around one specialised step body (hiprtc,
 * compiled at first use), so code size does not grow with N.
 * cxh_mapaccum_fwd_adj: outputs and the adjoint of every input for output
 *   cotangents adj_out, checkpointing the state every checkpoint_steps steps
 *   (0 = keep only the initial state and recompute).                        */
int cxh_mapaccum_create(cxh_tape step, int64_t n_steps, int32_t n_accum, cxh_mapaccum* out);
int cxh_mapaccum_destroy(cxh_mapaccum m);
int cxh_mapaccum_eval(cxh_mapaccum m, int64_t n, const double* const* arg, double* const* res,
                      int32_t layout, int64_t ld, void* stream);
int cxh_mapaccum_fwd_adj(cxh_mapaccum m, int64_t n, const double* const* arg,
                         const double* const* adj_out, double* const* res, double* const* adj_in,
                         int32_t layout, int64_t ld, int64_t checkpoint_steps, void* stream);
/* Split adjoint (for autograd): the forward chain storing the state every
 * checkpoint_steps steps into a caller-owned device buffer of
 * cxh_mapaccum_ckpt_size doubles, then the adjoint restarting from it; same
 * results as cxh_mapaccum_fwd_adj.                                          */
int cxh_mapaccum_ckpt_size(cxh_mapaccum m, int64_t n, int64_t checkpoint_steps, int64_t* n_doubles);
int cxh_mapaccum_fwd_ckpt(cxh_mapaccum m, int64_t n, const double* const* arg, double* const* res,
                          int32_t layout, int64_t ld, int64_t checkpoint_steps, double* ckpt, void* stream);
int cxh_mapaccum_rev_ckpt(cxh_mapaccum m, int64_t n, const double* const* arg,
                          const double* const* adj_out, double* const* adj_in, int32_t layout,
                          int64_t ld, int64_t checkpoint_steps, const double* ckpt, void* stream);
/* Introspection: generated source of a kernel: mode 0 forward chain, 1 fused
 * forward + adjoint, 2 forward storing checkpoints, 3 adjoint from them,
 * 4 one segment's reverse steps beside the recompute of the segment before
 * it (what modes 1 and 3 launch once per segment after the checkpoints).   */
int cxh_mapaccum_source(cxh_mapaccum m, int32_t mode, int64_t checkpoint_steps, char* buf,
                        int64_t cap, int64_t* len);

/* Batched sparse LDL^T on CasADi's symbolic factorisation.
 * sp_a, sp_lt: CasADi compact sparsity [nrow, ncol, colind..., row...] of A
 *   and of L^T (strictly upper), p: AMD permutation (Sparsity::ldl,
 *   casadi/core/sparsity.cpp:622-652).
 * Replaces: LinsolLdl::nfact / solve (casadi/solvers/linsol_ldl.cpp:119-132)
 *   = casadi_ldl + casadi_ldl_solve (casadi/core/runtime/casadi_ldl.hpp). */
int cxh_ldl_create(const int64_t* sp_a, const int64_t* sp_lt, const int64_t* p, cxh_ldl* out);
int cxh_ldl_destroy(cxh_ldl l);
/* a: nnz(A) values per instance, x: n*nrhs rhs values per instance,
 * overwritten by the solution; both in `layout` (SOA: [nz][n_inst]).
 * status (device, may be NULL): per instance, number of zero pivots.
 * Returns 0; use cxh_ldl_count_failures for the host-side count.           */
int cxh_ldl_factor_solve(cxh_ldl l, int64_t n_inst, const double* a, double* x, int32_t nrhs,
                         int32_t* status, int32_t layout, void* stream);
/* Numeric factorisation alone: L^T and D of n_inst instances stay in device
 * memory owned by the handle until the next nfact (or factor + solve), so one
 * factorisation serves any number of cxh_ldl_solve calls.
 * status (device, may be NULL): per instance, the number of zero or
 * non-finite pivots.
 * Replaces: LinsolLdl::nfact (casadi/solvers/linsol_ldl.cpp:119-126,
 *   casadi_ldl) behind Linsol::nfact (casadi/core/linsol.cpp:138-169).      */
int cxh_ldl_nfact(cxh_ldl l, int64_t n_inst, const double* a, int32_t* status, int32_t layout, void* stream);
/* Solve with the factors of the last cxh_ldl_nfact (same n_inst; an error if
 * there is none, as Linsol::solve asserts is_nfact): x holds nrhs right-hand
 * sides per instance and is overwritten by the solutions.  tr is accepted and
 * ignored (A is symmetric), as LinsolLdl::solve does.
 * Replaces: LinsolLdl::solve (casadi/solvers/linsol_ldl.cpp:128-132,
 *   casadi_ldl_solve) behind Linsol::solve (casadi/core/linsol.cpp:211-218). */
int cxh_ldl_solve(cxh_ldl l, int64_t n_inst, double* x, int32_t nrhs, int32_t tr, int32_t layout, void* stream);
/* Per instance (device int32 arrays of the last nfact's n_inst), from the D
 * of the last nfact: the number of negative entries (the inertia an
 * interior-point method checks) and of nonzero entries.
 * Replaces: LinsolLdl::neig / LinsolLdl::rank
 *   (casadi/solvers/linsol_ldl.cpp:134-150).                                */
int cxh_ldl_neig(cxh_ldl l, int32_t* neig, void* stream);
int cxh_ldl_rank(cxh_ldl l, int32_t* rank, void* stream);
/* Engine of cxh_ldl_factor_solve: CXH_ENGINE_JIT (default) compiles, at the
 * first call per nrhs, the factor + solve of this pattern into one
 * straight-line kernel (every operation of casadi_ldl / casadi_ldl_solve in
 * its order, indices resolved: bit-identical results); CXH_ENGINE_INTERP
 * runs the generic kernel that reads the pattern from memory.  Replaces
 * nothing in the reference (a launch choice of the batched plugin).        */
int cxh_ldl_engine(cxh_ldl l, int32_t engine);
/* Introspection: generated source of the specialised kernel for nrhs.     */
int cxh_ldl_jit_source(cxh_ldl l, int32_t nrhs, char* buf, int64_t cap, int64_t* len);

/* Batched sparse Householder QR (casadi_qr / casadi_qr_solve,
 * casadi/core/runtime/casadi_qr.hpp:53-200, LinsolQr
 * casadi/solvers/linsol_qr.cpp:126-178) on Sparsity::qr_sparse output.
 * status: per instance, number of |R_ii| < eps (casadi_qr_singular).        */
int cxh_qr_create(const int64_t* sp_a, const int64_t* sp_v, const int64_t* sp_r,
                  const int64_t* prinv, const int64_t* pc, double eps, cxh_qr* out);
int cxh_qr_destroy(cxh_qr q);
int cxh_qr_factor_solve(cxh_qr q, int64_t n_inst, const double* a, double* x, int32_t nrhs,
                        int32_t tr, int32_t* status, int32_t layout, void* stream);
/* Numeric factorisation alone, as cxh_ldl_nfact: V, R and beta stay in the
 * handle's device memory.  status: per instance, the number of |R_ii| < eps
 * (casadi_qr_singular; LinsolQr::nfact fails on a nonzero count).
 * Replaces: LinsolQr::nfact (casadi/solvers/linsol_qr.cpp:126-170).         */
int cxh_qr_nfact(cxh_qr q, int64_t n_inst, const double* a, int32_t* status, int32_t layout, void* stream);
/* Solve with the factors of the last cxh_qr_nfact (tr: with A^T).
 * Replaces: LinsolQr::solve (casadi/solvers/linsol_qr.cpp:172-178,
 *   casadi_qr_solve).                                                        */
int cxh_qr_solve(cxh_qr q, int64_t n_inst, double* x, int32_t nrhs, int32_t tr, int32_t layout, void* stream);
/* Per instance (device doubles): the determinant of A from the factors of the
 * last nfact, casadi_det times the sign of the row and column permutations.
 * Replaces: LinsolQr::det (casadi/solvers/linsol_qr.cpp, casadi_det in
 *   casadi/core/runtime/casadi_det.hpp:26-45).                               */
int cxh_qr_det(cxh_qr q, double* det, void* stream);
/* Remember the factorisations of the last n_cache batches (0 = none, the
 * default): an nfact whose batch of A values equals a remembered one (same
 * instance count and layout, every value ==) copies its factors back instead
 * of factoring.  Only batches without a singular instance are remembered.
 * Replaces: LinsolQr's "cache" option (casadi/solvers/linsol_qr.cpp:61-78,
 *   126-170, casadi_cache_check), per batch instead of per matrix.           */
int cxh_qr_cache(cxh_qr q, int32_t n_cache);
/* Engine of cxh_qr_factor_solve, as cxh_ldl_engine: CXH_ENGINE_JIT (default)
 * compiles the factor + solve of this pattern per (nrhs, tr) into straight-
 * line kernels (every operation of casadi_qr / casadi_qr_solve in its order:
 * bit-identical results), CXH_ENGINE_INTERP runs the generic kernel.
 * Replaces nothing in the reference (a launch choice of the batched plugin). */
int cxh_qr_engine(cxh_qr q, int32_t engine);
/* Introspection: generated source of segment `segment` of the specialised
 * kernels for (nrhs, tr), with the number of segments.                      */
int cxh_qr_jit_source(cxh_qr q, int32_t nrhs, int32_t tr, int32_t segment, char* buf, int64_t cap,
                      int64_t* len, int32_t* n_segments);

/* Facts of the current device: compute units, LDS per compute unit in KiB,
 * the device name (truncated to name_len).  Replaces nothing in the
 * reference (launch shaping of the batched kernels).                       */
int cxh_device_info(int32_t* n_cu, int32_t* lds_per_cu_kb, char* name, int32_t name_len);

/* ------------------------------------------------------------------------ */
/* Host symbolic layer                                                       */
/* ------------------------------------------------------------------------ */

typedef uint64_t cxs_sx;                 /* scalar expression node          */
typedef struct cxs_fun_s* cxs_fun;       /* SX Function                     */
typedef struct cxs_buf_s* cxs_buf;       /* result buffer (int64 or nodes)  */

/* SXElem construction (casadi/core/sx_elem.cpp:71-263) */
cxs_sx cxs_sym(const char* name);
cxs_sx cxs_const(double v);
cxs_sx cxs_unary(int32_t op, cxs_sx x);
cxs_sx cxs_binary(int32_t op, cxs_sx x, cxs_sx y);
cxs_sx cxs_neg(cxs_sx x);
cxs_sx cxs_if_else(cxs_sx c, cxs_sx a, cxs_sx b);
/* elementwise over arrays; incx/incy = 0 broadcasts a scalar               */
int cxs_unary_n(int32_t op, int64_t n, const cxs_sx* x, cxs_sx* out);
int cxs_binary_n(int32_t op, int64_t n, const cxs_sx* x, int64_t incx, const cxs_sx* y,
                 int64_t incy, cxs_sx* out);
int cxs_node_info(cxs_sx x, int32_t* op, double* value, cxs_sx* dep0, cxs_sx* dep1);
const char* cxs_sym_name(cxs_sx x);
int cxs_is_equal(cxs_sx x, cxs_sx y, int32_t depth);
int64_t cxs_node_count(void);

/* Function(name, ex_in, ex_out) (casadi/core/function.cpp:96-160 ->
 * SXFunction::init, casadi/core/sx_function.cpp:476-841).
 * sp_in[i]: compact sparsity of input i, nz_in[i]: its nonzero nodes.      */
int cxs_function_new(const char* name, int32_t n_in, const int64_t* const* sp_in,
                     const cxs_sx* const* nz_in, const char* const* names_in,
                     int32_t n_out, const int64_t* const* sp_out, const cxs_sx* const* nz_out,
                     const char* const* names_out, int32_t allow_free, cxs_fun* out);
int cxs_function_free(cxs_fun f);
int cxs_function_info(cxs_fun f, int32_t* n_in, int32_t* n_out, int64_t* n_instr,
                      int64_t* worksize, int64_t* n_free);
const char* cxs_function_name(cxs_fun f);
const char* cxs_function_name_in(cxs_fun f, int32_t i);
const char* cxs_function_name_out(cxs_fun f, int32_t i);
/* sparsity (compact form) of input / output i, as a buffer                 */
int cxs_function_sparsity_in(cxs_fun f, int32_t i, cxs_buf* out);
int cxs_function_sparsity_out(cxs_fun f, int32_t i, cxs_buf* out);
/* expressions of output i / input i (nodes)                                 */
int cxs_function_nz_out(cxs_fun f, int32_t i, cxs_buf* out);
int cxs_function_nz_in(cxs_fun f, int32_t i, cxs_buf* out);
/* the compiled tape as 16-byte ScalarAtomic records                         */
int cxs_function_tape(cxs_fun f, void* instr_out);
/* numeric evaluation on the host VM, one instance                           */
int cxs_function_eval(cxs_fun f, const double* const* arg, double* const* res);
/* Map serial / openmp on host AoS arrays (casadi/core/map.cpp:141-157, 340-386) */
int cxs_function_map_eval(cxs_fun f, int64_t n, const double* const* arg, double* const* res,
                          int32_t n_threads);
/* symbolic call: arg_nz[i] nodes for input i (pattern of input i)          */
int cxs_function_call(cxs_fun f, const cxs_sx* const* arg_nz, cxs_sx* const* res_nz);
/* derivative Functions (casadi/core/function_internal.cpp:2220-2440)       */
int cxs_function_jacobian(cxs_fun f, cxs_fun* out);
int cxs_function_forward(cxs_fun f, int64_t nfwd, cxs_fun* out);
int cxs_function_reverse(cxs_fun f, int64_t nadj, cxs_fun* out);
/* Jacobian / Hessian plan for the numeric sweeps: jsp (compact pattern),
 * direction seeds CSR and scatter sources; symmetric selects star colouring
 * (XFunction::jac / get_partition, casadi/core/x_function.hpp:442-840,
 * casadi/core/function_internal.cpp:2068-2165).  first_block != 0: the
 * Jacobian of output 0 w.r.t. input 0 (the other inputs are data, as the
 * free symbols of SX::jacobian), else of all outputs w.r.t. all inputs
 * flattened (Function::jacobian).                                           */
int cxs_function_jac_plan(cxs_fun f, int32_t symmetric, int32_t first_block, int32_t* forward, cxs_buf* jsp,
                          cxs_buf* seed_ptr, cxs_buf* seed_idx, cxs_buf* src_dir,
                          cxs_buf* src_nz);
/* Structural Jacobian pattern of output oind w.r.t. input iind             */
int cxs_function_jac_sparsity(cxs_fun f, int32_t oind, int32_t iind, int32_t compact,
                              int32_t symmetric, cxs_buf* out);

/* SX-level calculus (casadi/core/sx_instantiator.cpp:1155-1177). f and x
 * are (compact sparsity, nodes) pairs; the result is a sparsity buffer plus
 * a node buffer.                                                            */
int cxs_jacobian(const int64_t* sp_f, const cxs_sx* nz_f, const int64_t* sp_x,
                 const cxs_sx* nz_x, int32_t symmetric, cxs_buf* sp_out, cxs_buf* nz_out);
int cxs_gradient(const int64_t* sp_f, const cxs_sx* nz_f, const int64_t* sp_x,
                 const cxs_sx* nz_x, cxs_buf* sp_out, cxs_buf* nz_out);
int cxs_hessian(const int64_t* sp_f, const cxs_sx* nz_f, const int64_t* sp_x,
                const cxs_sx* nz_x, cxs_buf* sp_out, cxs_buf* nz_out);

/* Sparsity algorithms on compact patterns (casadi/core/sparsity_internal.cpp) */
int cxs_sp_uni_coloring(const int64_t* sp, int64_t cutoff, cxs_buf* out);
int cxs_sp_star_coloring(const int64_t* sp, int64_t ordering, cxs_buf* out);
int cxs_sp_largest_first(const int64_t* sp, cxs_buf* out);
int cxs_sp_amd(const int64_t* sp, cxs_buf* out);
int cxs_sp_etree(const int64_t* sp, int32_t ata, cxs_buf* out);
int cxs_sp_ldl(const int64_t* sp, int32_t amd, cxs_buf* sp_lt, cxs_buf* p);
int cxs_sp_qr(const int64_t* sp, int32_t amd, cxs_buf* sp_v, cxs_buf* sp_r, cxs_buf* prinv,
              cxs_buf* pc);
int cxs_sp_mtimes(const int64_t* sp_a, const int64_t* sp_b, cxs_buf* out);

/* result buffers: element count (int64 or node entries) and copy-out        */
int64_t cxs_buf_size(cxs_buf b);
int cxs_buf_copy(cxs_buf b, void* dst);
int cxs_buf_free(cxs_buf b);

#ifdef __cplusplus
}
#endif

#endif /* CASADI_HIP_H */
