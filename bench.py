"""Headline benchmark: instances/s of forward + adjoint, fused vs plain GPU tape.

Workload (N=1): cart-pole 100-step RK4 rollout cost J(x0, u[0:100]) =
sum ||x_t||^2 + 0.01 u_t^2, batch 32768, FP64, gradient by the fused forward +
checkpointed reverse sweep (checkpoints every 10 steps) on the specialised
straight-line kernels.  A step = one pass over the batch with inputs resident
in HBM.

"plain" = CasADi's way on a GPU tape interpreter: one launch of the f tape,
one launch of its reverse Function adj1_f (casadi reverse(1)), each re-reading
the inputs and re-evaluating the primal.

Prints one JSON line.  --gpus N > 1 is not supported (one GPU).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_BATCH = 32768
N_STEPS = 100
CKPT_STEPS = 10
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--n", type=int, default=N_BATCH)
    p.add_argument("--cpu-sample", type=int, default=16384, help="instances timed on the CPU oracle")
    p.add_argument("--no-cpu", action="store_true")
    return p.parse_args()


def tape_stats(f, every):
    """Instruction counts and the algorithmic FP64 operations of the reverse
    segment kernels: per arithmetic instruction 1 (recompute) + per operand 1
    (partial) + 2 (multiply and accumulate the adjoint)."""
    from casadi_amd.ops import BINARY, OP
    a = f.instructions()
    op = a["op"]
    arith = (op != OP.CONST) & (op != OP.INPUT) & (op != OP.OUTPUT)
    nops = np.where(np.isin(op, [int(b) for b in BINARY]), 2, 1)
    rev_flops = int(np.sum(np.where(arith, 1 + 3 * nops, 0)))
    n_seg = (a.size + every - 1) // every
    return dict(n_instr=int(a.size), n_arith=int(arith.sum()), worksize=f.sz_w(), rev_flops=rev_flops,
                n_seg=int(n_seg))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; sharding the instance batch "
                                   "across GPUs is designed but not launched"}))
        sys.exit(2)
    import torch
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    n = args.n
    f = W.cartpole_rollout(N_STEPS)
    R = f.reverse(1)
    every = (f.n_instructions() + (N_STEPS // CKPT_STEPS) - 1) // (N_STEPS // CKPT_STEPS)
    st_f, st_r = tape_stats(f, every), tape_stats(R, R.n_instructions())

    rng = np.random.default_rng(20261202 + 2)
    x0 = rng.uniform(-1, 1, (4, n))
    us = rng.uniform(-10, 10, (N_STEPS, n))
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    X0, U, SEED = t(x0), t(us), t(np.ones((1, n)))

    tape_f = HipTape(f, engine="jit")
    t0 = time.perf_counter()
    tape_f.jit_prepare(every, forward=True, reverse=True)
    compile_s = time.perf_counter() - t0
    tape_i = HipTape(f)            # interpreter, for the plain path
    tape_r = HipTape(R)

    def fused():
        return tape_f.fwd_adj([X0, U], [SEED], n, checkpoint_every=every)

    def plain():
        tape_i.eval([X0, U], n)
        return tape_r.eval([X0, U, None, SEED], n)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps  # ms per step

    ms_fused = timed(fused, args.steps, args.warmup)
    ms_plain = timed(plain, max(2, args.steps // 4), 1)
    ms_fused_interp = timed(lambda: tape_i.fwd_adj([X0, U], [SEED], n, checkpoint_every=every),
                            max(2, args.steps // 4), 1)

    # the same cost and gradient written as CasADi users write rollouts: a
    # mapaccum of the cost-carrying step (one compiled step body, loops),
    # checkpoints every 10 steps -- reported beside the headline, not as it
    from casadi_amd.hip import HipMapAccum
    macc = HipMapAccum(W.cartpole_cost_step(), N_STEPS)
    XC0 = torch.cat([X0, torch.zeros((1, n), dtype=torch.float64, device=dev)])
    (ym,) = macc.eval([XC0, U], n)
    SM = torch.zeros((5 * N_STEPS, n), dtype=torch.float64, device=dev)
    SM[5 * (N_STEPS - 1):5 * (N_STEPS - 1) + 4] = 2 * ym[5 * (N_STEPS - 1):5 * (N_STEPS - 1) + 4]
    SM[5 * (N_STEPS - 1) + 4] = 1.0   # the final |x_N|^2 term is outside the chain; its cotangent is 2 x_N

    def fused_macc():
        return macc.fwd_adj([XC0, U], [SM], n, checkpoint_steps=CKPT_STEPS)
    ms_macc = timed(fused_macc, args.steps, args.warmup)

    # dominant kernel: the reverse segment kernels (cxj_rev), timed live with
    # events on the stream they are launched on (torch's current stream).  The
    # sweep is one cxj_rev per segment -- each but the last holding the
    # boundary-state pass of the segment before it beside its own sub-block
    # recompute + adjoint sweep -- plus one lone cxj_snap for the last
    # segment; its time divided by the segment count is the per-launch time.
    outs, ck = tape_f.fwd_ckpt([X0, U], n, every)
    ms_rev = timed(lambda: tape_f.rev_ckpt([X0, U], [SEED], n, every, ck), args.steps, 1)
    rev_launch_ms = ms_rev / st_f["n_seg"]
    flops_per_launch = st_f["rev_flops"] / st_f["n_seg"] * n
    achieved_tf = flops_per_launch / (rev_launch_ms * 1e-3) / 1e12

    # correctness of the timed path on a sample: per instance, scaled by the
    # instance's gradient norm, against casadi reverse(1) on the CPU oracle
    import oracle_ref as O
    (J,), (gx, gu) = fused()
    m = min(n, 256)
    rx, ru = O.map_eval(R, m, [x0[:, :m].T, us[:, :m].T, None, np.ones((m, 1))])
    gr = np.concatenate([rx, ru], axis=1)
    g = np.concatenate([gx[:, :m].cpu().numpy(), gu[:, :m].cpu().numpy()]).T
    per = np.max(np.abs(g - gr) / np.maximum(np.max(np.abs(gr), axis=1, keepdims=True), 1.0), axis=1)
    (Jr,) = O.map_eval(f, m, [x0[:, :m].T, us[:, :m].T])
    err_J = O.rel_err(J[:, :m].cpu().numpy(), Jr.T)
    _, (mg0, mgu) = fused_macc()
    gm = np.concatenate([mg0[:4, :m].cpu().numpy(), mgu[:, :m].cpu().numpy()]).T
    per_m = np.max(np.abs(gm - gr) / np.maximum(np.max(np.abs(gr), axis=1, keepdims=True), 1.0), axis=1)

    # HBM bytes per cxj_rev launch and its hardware FP64 work, from the
    # rocprofv3 --pmc runs summarised under profiles/ (same workload)
    traffic, hw_tflops = None, None
    prof = os.path.join(ROOT, "profiles", "cxj_rev_rollout_traffic.json")
    if os.path.exists(prof) and n == N_BATCH:
        with open(prof) as fh:
            pj = json.load(fh)
        traffic = pj["hbm_bytes_per_launch"]
        hw_tflops = pj["fp64_flops_per_launch_hw"] / (rev_launch_ms * 1e-3) / 1e12

    value = n / (ms_fused * 1e-3)
    plain_value = n / (ms_plain * 1e-3)
    bytes_per_inst = (4 + N_STEPS + 1 + 1 + 4 + N_STEPS) * 8   # x0, u, seed in; J, gradient out

    # CPU baseline: CasADi's path on the host -- map serial of f, then of its
    # reverse(1) Function -- on the C oracle VM, one thread, over the batch's
    # first instances repeated until about 10 s of work
    cpu, cpu_omp = None, None
    if not args.no_cpu:
        mc = min(args.cpu_sample, n)
        a_f = [x0[:, :mc].T.copy(), us[:, :mc].T.copy()]
        a_r = a_f + [None, np.ones((mc, 1))]
        reps, dt = 0, 0.0
        while dt < 10.0 and reps < 20:
            t0 = time.perf_counter()
            O.map_eval(f, mc, a_f, n_threads=1)
            O.map_eval(R, mc, a_r, n_threads=1)
            dt += time.perf_counter() - t0
            reps += 1
        cpu = {"value": reps * mc / dt, "unit": "instances/s", "cores": 1, "kind": "port",
               "sample": f"{reps} x {mc} instances: the rollout f ({st_f['n_instr']} instructions) then its "
                         f"reverse(1) Function ({st_r['n_instr']} instructions) on the C oracle VM, map serial, "
                         f"1 thread, {dt:.1f} s"}
        # the same map over the host's cores (map openmp): OpenMP threads over
        # instances, as many as the process may use (OMP_NUM_THREADS caps it)
        nt = len(os.sched_getaffinity(0))
        if os.environ.get("OMP_NUM_THREADS", "").isdigit():
            nt = min(nt, int(os.environ["OMP_NUM_THREADS"]))
        reps_p, dt_p = 0, 0.0
        while dt_p < 3.0 and reps_p < 200:
            t0 = time.perf_counter()
            O.map_eval(f, mc, a_f, n_threads=nt)
            O.map_eval(R, mc, a_r, n_threads=nt)
            dt_p += time.perf_counter() - t0
            reps_p += 1
        model = ""
        try:
            with open("/proc/cpuinfo") as fh:
                model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), "")
        except OSError:
            pass
        cpu_omp = {"value": reps_p * mc / dt_p, "unit": "instances/s", "cores": nt, "kind": "port",
                   "cpu_model": model, "nproc": os.cpu_count(),
                   "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
                   "sample": f"{reps_p} x {mc} instances, same two maps, map openmp, {nt} threads, {dt_p:.1f} s"}

    line = {
        "metric": "instances/sec fwd+adjoint, fused batched vs plain GPU tape @1e-13 rel err",
        "value": value,
        "unit": "instances/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_fused,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (x0~U[-1,1]^4, u~U[-10,10]^100, seed 20261202+2)",
        "config": {"workload": f"cart-pole {N_STEPS}-step RK4 rollout cost + gradient, batch {n}, "
                               f"checkpoint every {CKPT_STEPS} steps ({every} instructions, "
                               f"{st_f['n_seg']} segments)",
                   "batch": n, "tape_instructions": st_f["n_instr"], "worksize": st_f["worksize"],
                   "engine": "specialised straight-line kernels (hiprtc)", "parallelization": "hip, one GPU"},
        "plain": {"value": plain_value, "ms_per_step": ms_plain, "engine": "interpreter",
                  "tapes": [st_f["n_instr"], st_r["n_instr"]]},
        "speedup_vs_plain": value / plain_value,
        "fused_interpreter": {"value": n / (ms_fused_interp * 1e-3), "ms_per_step": ms_fused_interp},
        "fused_mapaccum": {"value": n / (ms_macc * 1e-3), "ms_per_step": ms_macc,
                           "speedup_vs_plain": ms_plain / ms_macc,
                           "err_grad_median": float(np.median(per_m)), "err_grad_max": float(per_m.max()),
                           "err_grad_p99": float(np.percentile(per_m, 99)),
                           "note": "same cost and gradient as mapaccum(100) of the cost-carrying RK4 step: "
                                   "loop kernel around one specialised step, checkpoints every 10 steps"},
        "jit_compile_s": compile_s,
        "err_grad_median": float(np.median(per)),
        "err_grad_max": float(per.max()),
        "err_grad_p99": float(np.percentile(per, 99)),
        "err_cost": err_J,
        "roofline": {"bound": "mfma", "achieved": achieved_tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tf / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "hbm_gbs_from_traffic": traffic / (rev_launch_ms * 1e-3) / 1e9 if traffic else None,
                     "hw_fp64_tflops_from_counters": hw_tflops,
                     "kernel": "cxj_rev (reverse segment: sub-block recompute + adjoint sweep, beside the "
                               "boundary-state pass of the segment before it)",
                     "kernel_ms": rev_launch_ms, "flops_per_launch": flops_per_launch,
                     "hbm_algorithmic_gbs": bytes_per_inst * n / (ms_fused * 1e-3) / 1e9,
                     "note": "FP64 vector peak; algorithmic tape ops, transcendentals counted once"},
        "cpu_baseline": cpu,
        "cpu_baseline_openmp": cpu_omp,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
