"""Headline benchmark: instances/s of forward + adjoint, fused vs plain GPU tape.

Workload (N=1): cart-pole 100-step RK4 rollout cost J(x0, u[0:100]) =
sum ||x_t||^2 + 0.01 u_t^2, batch 32768, FP64, gradient by the fused forward +
checkpointed reverse sweep (checkpoints every 10 steps) on the specialised
kernels: the flat tape of the rollout repeats one step, so it runs as one
loop body per step.  A step of the benchmark = one pass over the batch with
inputs resident in HBM.  The line also carries the cart-pole step map over
65536 instances and its dense Jacobian (f + jac), both CPU baselines, and
(unless --no-configs) the other workloads at their stated sizes: the
quadrotor OCP Hessian, the heterogeneous batch and the KKT LDL / QR solves.

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
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=N_BATCH)
    p.add_argument("--cpu-sample", type=int, default=16384, help="instances timed on the CPU oracle")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-configs", action="store_true",
                   help="skip the other workloads (Hessian, heterogeneous batch, KKT solves)")
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

    # the same steps timed one by one (events around each): the median step
    def per_step(fn, steps):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        for a, b in ev:
            a.record()
            fn()
            b.record()
        torch.cuda.synchronize()
        return sorted(a.elapsed_time(b) for a, b in ev)
    steps_ms = per_step(fused, args.steps)
    ms_median = steps_ms[len(steps_ms) // 2]

    # the same fused call captured once in a HIP graph and replayed (the
    # library launches on the capturing stream and allocates nothing after
    # its first call)
    ms_graph = None
    try:
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            fused()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        with torch.cuda.graph(g):
            fused()
        ms_graph = timed(g.replay, args.steps, args.warmup)
    except RuntimeError as e:   # capture refused: report it, keep the direct figure
        ms_graph = f"capture failed: {e}"
    ms_plain = timed(plain, max(2, args.steps // 4), 1)
    # the same two launches on the specialised engine: what fusion itself
    # gains there (the plain path above also pays for the interpreter)
    tape_fj, tape_rj = HipTape(f, engine="jit"), HipTape(R, engine="jit")
    t0 = time.perf_counter()
    tape_fj.eval([X0, U], n)
    tape_rj.eval([X0, U, None, SEED], n)
    torch.cuda.synchronize()
    compile_plain_s = time.perf_counter() - t0

    def plain_jit():
        tape_fj.eval([X0, U], n)
        return tape_rj.eval([X0, U, None, SEED], n)
    ms_plain_jit = timed(plain_jit, max(2, args.steps // 4), 1)
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

    # dominant kernel: the reverse sweep's launches, timed live with events on
    # the stream they are launched on (torch's current stream).  The rollout
    # tape repeats one RK4 step + cost term, so the specialiser runs it as a
    # loop body (rerolled): the reverse sweep from stored checkpoints is S + 1
    # launches of cxj_roll_pair -- the first recomputes the last segment's
    # iteration states, each later one runs segment g's reverse iterations
    # beside the recompute of segment g-1.  Without a loop (CXH_JIT_ROLL=0) it
    # is one cxj_rev per segment plus one lone cxj_snap.  Per-launch time: the
    # sweep's time over its launch count; algorithmic flops per launch: the
    # tape's reverse flops over the launches that sweep backwards.
    roll = tape_f.roll_plan(every)
    if roll["enabled"]:
        kernel, n_launch, n_work = "cxj_roll_pair", roll["segments"] + 1, roll["segments"]
    else:
        kernel, n_launch, n_work = "cxj_rev", st_f["n_seg"], st_f["n_seg"]
    outs, ck = tape_f.fwd_ckpt([X0, U], n, every)
    ms_rev = timed(lambda: tape_f.rev_ckpt([X0, U], [SEED], n, every, ck), args.steps, 1)
    rev_launch_ms = ms_rev / n_launch
    flops_per_launch = st_f["rev_flops"] / n_work * n
    achieved_tf = st_f["rev_flops"] * n / (ms_rev * 1e-3) / 1e12

    # correctness of the timed path on a sample of instances, against casadi
    # reverse(1) on the CPU oracle: per nonzero |g - g_ref| / max(|g_ref|, 1)
    # (the 1e-13 acceptance metric), and per instance scaled by the
    # instance's gradient norm.  The rollout is chaotic (|u| <= 10): a 1-ulp
    # difference of one sin/cos (OCML against glibc) grows along the
    # trajectory, so the bar held is the median at 1e-13 and the maximum
    # inside the rounding envelope: how far the oracle's own results move
    # when every transcendental result is nudged by one ulp (the numeric
    # sweep's envelope plus reverse(1)'s, as the two round independently).
    import oracle_ref as O
    (J,), (gx, gu) = fused()
    m = min(n, 256)
    a_s = [x0[:, :m].T, us[:, :m].T, None, np.ones((m, 1))]
    rx, ru = O.map_eval(R, m, a_s)
    gr = np.concatenate([rx, ru], axis=1)
    g = np.concatenate([gx[:, :m].cpu().numpy(), gu[:, :m].cpu().numpy()]).T
    scale = np.maximum(np.max(np.abs(gr), axis=1, keepdims=True), 1.0)
    per = np.max(np.abs(g - gr) / scale, axis=1)
    per_nz = (np.abs(g - gr) / np.maximum(np.abs(gr), 1.0)).ravel()
    _, (ox, ou) = O.adj_map(f, m, [x0[:, :m].T, us[:, :m].T], [np.ones((m, 1))])
    go = np.concatenate([ox, ou], axis=1)
    with O.perturbed_transcendentals():
        _, (px, pu) = O.adj_map(f, m, [x0[:, :m].T, us[:, :m].T], [np.ones((m, 1))])
        qx, qu = O.map_eval(R, m, a_s)
    env = float(np.max(np.abs(np.concatenate([px, pu], axis=1) - go) / scale))
    env_r = float(np.max(np.abs(np.concatenate([qx, qu], axis=1) - gr) / scale))
    (Jr,) = O.map_eval(f, m, [x0[:, :m].T, us[:, :m].T])
    err_J = O.rel_err(J[:, :m].cpu().numpy(), Jr.T)
    _, (mg0, mgu) = fused_macc()
    gm = np.concatenate([mg0[:4, :m].cpu().numpy(), mgu[:, :m].cpu().numpy()]).T
    per_m = np.max(np.abs(gm - gr) / scale, axis=1)

    # HBM bytes per launch of the dominant kernel and its hardware FP64 work,
    # from the rocprofv3 --pmc runs summarised under profiles/ (same workload)
    traffic, hw_tflops = None, None
    prof = os.path.join(ROOT, "profiles", f"{kernel}_rollout_traffic.json")
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

    c1 = cartpole_map_figures(timed, args, nt if not args.no_cpu else 1)
    others = None if args.no_configs else other_config_figures(timed)

    line = {
        "metric": "instances/sec fwd+adjoint, fused batched vs plain GPU tape @1e-13 rel err",
        "value": value,
        "unit": "instances/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_fused,
        "ms_per_step_median": ms_median,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (x0~U[-1,1]^4, u~U[-10,10]^100, seed 20261202+2)",
        "config": {"workload": f"cart-pole {N_STEPS}-step RK4 rollout cost + gradient, batch {n}, "
                               f"checkpoint every {CKPT_STEPS} steps ({every} instructions, "
                               f"{st_f['n_seg']} segments)",
                   "batch": n, "tape_instructions": st_f["n_instr"], "worksize": st_f["worksize"],
                   "engine": ("specialised kernels (hiprtc), the tape rerolled into a loop over its "
                              f"{roll['iterations']} repeated steps" if roll["enabled"] else
                              "specialised straight-line kernels (hiprtc)"),
                   "parallelization": "hip, one GPU"},
        "plain": {"value": plain_value, "ms_per_step": ms_plain, "engine": "interpreter",
                  "tapes": [st_f["n_instr"], st_r["n_instr"]]},
        "speedup_vs_plain": value / plain_value,
        "plain_specialised": {"value": n / (ms_plain_jit * 1e-3), "ms_per_step": ms_plain_jit,
                              "first_call_s": compile_plain_s,
                              "fusion_gain": ms_plain_jit / ms_fused,
                              "note": "f and its reverse(1) Function as two specialised launches (the reverse "
                                      "Function's tape does not reroll: straight-line segments)"},
        "fused_graph_replay": ({"ms_per_step": ms_graph, "value": n / (ms_graph * 1e-3),
                                "vs_direct": ms_fused / ms_graph}
                               if isinstance(ms_graph, float) else {"error": ms_graph}),
        "fused_interpreter": {"value": n / (ms_fused_interp * 1e-3), "ms_per_step": ms_fused_interp},
        "fused_mapaccum": {"value": n / (ms_macc * 1e-3), "ms_per_step": ms_macc,
                           "speedup_vs_plain": ms_plain / ms_macc,
                           "err_grad_median": float(np.median(per_m)), "err_grad_max": float(per_m.max()),
                           "err_grad_p99": float(np.percentile(per_m, 99)),
                           "note": "same cost and gradient as mapaccum(100) of the cost-carrying RK4 step: "
                                   "loop kernel around one specialised step, checkpoints every 10 steps"},
        "jit_compile_s": compile_s,
        "parity": {
            "sample": f"first {m} instances, against reverse(1) on the CPU oracle",
            "per_nonzero": {"median": float(np.median(per_nz)), "p99": float(np.percentile(per_nz, 99)),
                            "max": float(per_nz.max()), "frac_within_1e-13": float(np.mean(per_nz <= 1e-13))},
            "per_instance_scaled": {"median": float(np.median(per)), "p99": float(np.percentile(per, 99)),
                                    "max": float(per.max())},
            "envelope_1ulp": env + env_r,
            "bar_held": "median per nonzero <= 1e-13; maximum within the 1-ulp rounding envelope "
                        "(chaotic 100-step trajectories, OCML vs glibc sin/cos)",
            "met": bool(np.median(per_nz) <= 1e-13 and per.max() <= max(1e-13, env + env_r)),
            "err_cost": err_J},
        "roofline": {"bound": "fp64_valu", "achieved": achieved_tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tf / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "hbm_gbs_from_traffic": traffic / (rev_launch_ms * 1e-3) / 1e9 if traffic else None,
                     "hw_fp64_tflops_from_counters": hw_tflops,
                     "kernel": kernel, "launches_per_sweep": n_launch,
                     "kernel_what": ("rerolled reverse sweep: one checkpoint segment's reverse iterations beside "
                                     "the recompute of the segment before it" if kernel == "cxj_roll_pair" else
                                     "reverse segment: sub-block recompute + adjoint sweep, beside the "
                                     "boundary-state pass of the segment before it"),
                     "loop": roll,
                     "kernel_ms": rev_launch_ms, "flops_per_launch": flops_per_launch,
                     "hbm_algorithmic_gbs": bytes_per_inst * n / (ms_fused * 1e-3) / 1e9,
                     "traffic_source": f"stored: profiles/{kernel}_rollout_traffic.json (rocprofv3 --pmc "
                                       "FETCH_SIZE and WRITE_SIZE of this workload, calibrated), not measured "
                                       "in this run" if traffic else None,
                     "note": "FP64 vector peak (no MFMA is used: scalar FP64 tape); algorithmic tape ops, "
                             "transcendentals counted once"},
        "cpu_baseline": cpu,
        "cpu_baseline_openmp": cpu_omp,
        "cartpole_map": c1,
        "other_configs": others,
    }
    print(json.dumps(line))


def cartpole_map_figures(timed, args, n_threads):
    """The cart-pole RK4 step mapped over 65536 instances (f.map(65536)) and
    its dense 4x5 Jacobian over the same batch, on the specialised kernels:
    HIP-event times, algorithmic HBM rates (f: 5 in + 4 out = 72 B per
    instance; f + jac: 72 + 20 Jacobian nonzeros = 232 B), the same map over
    8M instances, and the oracle's map serial / openmp on the host."""
    import torch
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipJacobian, HipTape
    import oracle_ref as O
    f = W.cartpole_step()
    out = {}
    rng = np.random.default_rng(20261202)
    dev = torch.device("cuda:0")
    for n in (65536, 8 * 1024 * 1024):
        x = rng.uniform(-1, 1, (4, n))
        u = rng.uniform(-10, 10, (1, n))
        X = torch.tensor(x, dtype=torch.float64, device=dev)
        Uc = torch.tensor(u, dtype=torch.float64, device=dev)
        tp = HipTape(f, engine="jit")
        ms = timed(lambda: tp.eval([X, Uc], n), 50, 5)
        gbs = 72 * n / (ms * 1e-3) / 1e9
        key = "map_65536" if n == 65536 else "map_8M"
        out[key] = {"n": n, "ms": ms, "instances_per_s": n / (ms * 1e-3), "hbm_gbs": gbs,
                    "hbm_frac": gbs / HBM_PEAK_GBS, "bytes_per_instance": 72}
        if n != 65536:
            continue
        ti = HipTape(f)
        ms_i = timed(lambda: ti.eval([X, Uc], n), 50, 5)
        out["map_65536_interpreter"] = {"ms": ms_i, "hbm_gbs": 72 * n / (ms_i * 1e-3) / 1e9}
        hj = HipJacobian(f, engine="jit")
        ms_j = timed(lambda: hj([X, Uc], n), 50, 5)
        out["map_jac_65536"] = {"ms": ms_j, "hbm_gbs": 232 * n / (ms_j * 1e-3) / 1e9,
                                "hbm_frac": 232 * n / (ms_j * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                "bytes_per_instance": 232, "colours": hj.plan["ndir"]}
        # At this size a call's host side (argument checks, output allocation,
        # the ctypes call) takes longer than its kernel, so the figures above
        # are host-bound.  Fifty calls captured in one HIP graph and replayed
        # give the device's rate for back-to-back calls.
        for key2, fn, bpi in (("map_65536", lambda: tp.eval([X, Uc], n), 72),
                              ("map_jac_65536", lambda: hj([X, Uc], n), 232)):
            g = torch.cuda.CUDAGraph()
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                fn()
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            with torch.cuda.graph(g):
                for _ in range(50):
                    fn()
            ms_g = timed(g.replay, 5, 1) / 50
            gb = bpi * n / (ms_g * 1e-3) / 1e9
            out[key2].update({"ms_back_to_back": ms_g, "hbm_gbs_back_to_back": gb,
                              "hbm_frac_back_to_back": gb / HBM_PEAK_GBS})
        (y,) = tp.eval([X, Uc], n)
        (yr,) = O.map_eval(f, 4096, [x[:, :4096].T, u[:, :4096].T])
        out["map_65536"]["err_max_4096"] = O.rel_err(y[:, :4096].cpu().numpy(), yr.T)
        if not args.no_cpu:
            a = [x.T.copy(), u.T.copy()]
            for key, nt in (("cpu_map_serial", 1), ("cpu_map_openmp", n_threads)):
                reps, dt = 0, 0.0
                while dt < 2.0 and reps < 100:
                    t0 = time.perf_counter()
                    O.map_eval(f, n, a, n_threads=nt)
                    dt += time.perf_counter() - t0
                    reps += 1
                out[key] = {"value": reps * n / dt, "unit": "instances/s", "cores": nt, "kind": "port",
                            "sample": f"{reps} x {n} instances, oracle VM map, {dt:.1f} s"}
    return out


def other_config_figures(timed):
    """The other workloads of the benchmark set, each at its stated size, on
    the specialised kernels with the generic path beside it: the star-coloured
    Hessian of the 20-step quadrotor OCP Lagrangian (16384 instances), the
    heterogeneous batch (40000 cart-pole + 12000 quadrotor + 25000 arm
    forward + adjoint, autotuned launch) and the batched KKT solve (n = 274,
    8192 instances, ldl and qr) and the 100-step rollout's cost + gradient at
    262144 and 1048576 instances.  first_call_s includes compiling the kernels;
    ms are HIP-event times after it.  max_abs_diff compares the two engines
    (the tests hold both to the oracle)."""
    import torch
    from casadi_amd import workloads as W
    from casadi_amd.function import Function
    from casadi_amd.hip import HipBatch, HipJacobian, HipTape
    from casadi_amd.linsol import Linsol
    from casadi_amd.sx import gradient
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    out = {}

    def first(fn):
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    # config 2: Hessian of the quadrotor OCP Lagrangian
    n = 16384
    w, lam, Lag = W.quadrotor_ocp_lagrangian()
    G = Function("grad_lag", [w, lam], [gradient(Lag, w)])
    rng = np.random.default_rng(20261203)
    args = [t(rng.uniform(-1, 1, (w.nnz(), n))), t(rng.standard_normal((lam.nnz(), n)))]
    hj = HipJacobian(G, symmetric=True, first_block=True, engine="jit")
    (_, hs), c_s = first(lambda: hj(args, n))
    ms = timed(lambda: hj(args, n), 10, 1)
    hi = HipJacobian(G, symmetric=True, first_block=True)
    (_, hi_v), _ = first(lambda: hi(args, n))
    ms_i = timed(lambda: hi(args, n), 2, 0)
    bpi = (w.nnz() + lam.nnz() + hj.nnz_jac) * 8   # primal and multipliers in, Hessian nonzeros out
    gbs = bpi * n / (ms * 1e-3) / 1e9
    out["hessian_16384"] = {"n": n, "ms": ms, "instances_per_s": n / (ms * 1e-3), "first_call_s": c_s,
                            "colours": hj.plan["ndir"], "nnz": hj.nnz_jac, "interpreter_ms": ms_i,
                            "bytes_per_instance": bpi, "hbm_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
                            "max_abs_diff": float((hs - hi_v).abs().max())}
    # config 4: heterogeneous batch
    fs = [W.cartpole_step(), W.quadrotor_step(), W.arm_dynamics()]
    counts = [40000, 12000, 25000]
    rng = np.random.default_rng(20261205)
    (B, c_s) = first(lambda: HipBatch(fs, counts))
    data = []
    for i, (f, m) in enumerate(zip(fs, counts)):
        A = [t(rng.uniform(-1, 1, (f.nnz_in(j), m))) for j in range(f.n_in())]
        S = [t(rng.standard_normal((f.nnz_out(j), m))) for j in range(f.n_out())]
        B.bind(i, A, S)
        data.append((A, S))
    cfg, _ = B.tune(reps=3)
    ms_g = timed(lambda: B.launch(), 50, 5)
    ms_d = timed(lambda: B.launch(graph=False), 50, 5)
    ms_b = timed(lambda: B.launch(graph=None), 50, 5)
    tapes = [(HipTape(f), HipTape(f.reverse(1))) for f in fs]

    def plain():
        for (tf, tr), (A, S), m in zip(tapes, data, counts):
            tf.eval(A, m)
            tr.eval(A + [None] * len(S) + S, m)
    ms_p = timed(plain, 10, 1)
    nbytes = sum(m * 16 * (sum(f.nnz_in(j) for j in range(f.n_in())) + sum(f.nnz_out(j) for j in range(f.n_out())))
                 for f, m in zip(fs, counts))   # inputs + seeds in, outputs + sensitivities out
    gbs = nbytes / (ms_b * 1e-3) / 1e9
    out["batch_77000"] = {"instances": sum(counts), "ms": ms_b, "ms_graph": ms_g, "ms_direct": ms_d,
                          "first_call_s": c_s, "instances_per_s": sum(counts) / (ms_b * 1e-3), "tuned": cfg,
                          "bytes": nbytes, "hbm_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
                          "plain_interpreter_ms": ms_p}
    # config 5: batched KKT solves
    n = 8192
    sp, _, _ = W.ocp_kkt_sparsity()
    rng = np.random.default_rng(20261206)
    _, A = W.ocp_kkt_batch(rng, n)
    At, Bt = t(A), t(rng.standard_normal((sp.size1(), n)))
    for solver in ("ldl", "qr"):
        Lj = Linsol("kkt", solver, sp, {"engine": "jit"})
        (xj, _), c_s = first(lambda: Lj.solve_batch(At, Bt))
        ms = timed(lambda: Lj.solve_batch(At, Bt), 10, 1)
        Lg = Linsol("kkt", solver, sp, {"engine": "interp"})
        (xg, _), _ = first(lambda: Lg.solve_batch(At, Bt))
        ms_g = timed(lambda: Lg.solve_batch(At, Bt), 3, 0)
        bpi = (sp.nnz() + 2 * sp.size1()) * 8   # A's nonzeros and b in, x out
        gbs = bpi * n / (ms * 1e-3) / 1e9
        out[f"kkt_{solver}_8192"] = {"n": n, "ms": ms, "instances_per_s": n / (ms * 1e-3), "first_call_s": c_s,
                                     "bytes_per_instance": bpi, "hbm_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
                                     "generic_ms": ms_g, "max_abs_diff": float((xj - xg).abs().max())}
    # the same rollout cost + gradient at larger batches: above 128 instances
    # per SIMD the lean loop kernels run (no loads ahead, IEEE division: more
    # resident wavefronts), same bits
    f100 = W.cartpole_rollout(100)
    every = (f100.n_instructions() + 9) // 10
    tape = HipTape(f100, engine="jit")
    fat_max = int(os.environ.get("CXH_JIT_ROLL_FAT_MAX", torch.cuda.get_device_properties(0).multi_processor_count * 512))
    for n in (262144, 1048576):
        x0, us = t(rng.uniform(-1, 1, (4, n))), t(rng.uniform(-10, 10, (100, n)))
        seed = torch.ones((1, n), dtype=torch.float64, device=dev)
        call = lambda: tape.fwd_adj([x0, us], [seed], checkpoint_every=every)
        _, t1 = first(call)
        ms = timed(call, 5, 1)
        out[f"rollout_{n}"] = {"n": n, "ms": ms, "instances_per_s": n / (ms * 1e-3), "first_call_s": t1,
                               "kernels": "lean" if n > fat_max else "full"}
        del x0, us, seed
    return out


if __name__ == "__main__":
    main()
