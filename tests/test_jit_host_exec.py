"""The specialiser's generated code, executed on the host.

The segment bodies the specialiser emits are plain C++ apart from a few
qualifiers, so they are compiled here with g++ (the device qualifiers defined
away, the register barrier empty) and run instance by instance on the CPU,
driven through the same segment sequence the GPU driver launches: forward
segments storing the live state, then reverse segments from the last to the
first handing adjoints of the live slots backwards.  This checks the code
generation itself -- liveness at segment and sub-block boundaries,
recomputation, the adjoint hand-offs -- without a GPU, against the oracle's
reverse sweep (same host libm: bit-exact up to the sign of zero).
"""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_ref as O

MAX_IO = 16
# g++ turns sin(x) and cos(x) of one argument into one glibc sincos call,
# whose results differ from separate sin and cos in the last bit for some
# arguments (the 12-step rollout's instance 9 of seed 4: cos differs by one
# ulp).  The oracle calls sin and cos separately, so the host builds of the
# generated code must too.
NO_SINCOS_FUSION = ("-fno-builtin-sin", "-fno-builtin-cos")


class IOArg(C.Structure):
    _fields_ = [("p", C.c_void_p), ("s_nz", C.c_longlong), ("s_inst", C.c_longlong)]


class IOTable(C.Structure):
    _fields_ = [("a", IOArg * MAX_IO)]


class CxjArgs(C.Structure):
    _fields_ = [("arg", IOTable), ("res", IOTable), ("aseed", IOTable), ("asens", IOTable),
                ("st_in", C.c_void_p), ("st_out", C.c_void_p), ("adj_in", C.c_void_p), ("adj_out", C.c_void_p),
                ("n", C.c_longlong), ("lds", C.c_longlong), ("spill", C.c_void_p)]


def _table(arrs, n):
    t = IOTable()
    for i, a in enumerate(arrs):
        if a is not None:
            t.a[i].p = a.ctypes.data
            t.a[i].s_nz = n
            t.a[i].s_inst = 1
    return t


def _tiled(srcs, n, groups):
    """State (two halves) and spill buffers of the tangent kernels, tiled
    [group][instance/64][row][64] with the row counts each source declares,
    and the padded instance count the kernels take as lds."""
    rows = max(int(re.search(r"#define CXJ_TROWS (\d+)", x).group(1)) for x in srcs)
    spill = max(int(re.search(r"#define CXJ_SROWS (\d+)", x).group(1)) for x in srcs)
    lds = (n + 63) // 64 * 64
    return ([np.zeros(groups * lds * max(rows, 1)) for _ in range(2)], np.zeros(groups * lds * max(spill, 1)), spill,
            lds)


def _compile(sources, workdir, groups=None):
    """One shared library exposing run_<i>(const CxjArgs*, n) per source; with
    groups, the body runs for every direction group in turn."""
    libs = []
    for i, src in enumerate(sources):
        cut = src.index('extern "C" __global__')
        loop = ("  for (long long k = 0; k < n; ++k) cxj_body(*A, k);\n" if groups is None else
                f"  for (int g = 0; g < {groups}; ++g)\n    for (long long k = 0; k < n; ++k) cxj_body(*A, k, g);\n")
        code = ("#define __device__\n#define __forceinline__ inline\n#define CXJ_KEEP(x) ((void)0)\n#define CXJ_KEEP_PTR(x) ((void)0)\n#define CXJ_AFTER(p, x) ((void)0)\n"
                "#define CXJ_GLOBAL\n#define CXJ_OPAQUE(p) ((void)0)\n#define CXJ_OPAQUE_AFTER(p, x) ((void)0)\n#define CXJ_ANY(b) (b)\n" + src[:cut] +
                f'extern "C" void run_{i}(const CxjArgs* A, long long n) {{\n' + loop + "}\n")
        path = os.path.join(workdir, f"seg{i}.cpp")
        with open(path, "w") as fh:
            fh.write(code)
        so = os.path.join(workdir, f"seg{i}.so")
        subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", *NO_SINCOS_FUSION, "-shared", "-fPIC", "-w", path, "-o", so],
                       check=True)
        lib = C.CDLL(so)
        fn = getattr(lib, f"run_{i}")
        fn.argtypes = [C.POINTER(CxjArgs), C.c_longlong]
        fn.restype = None
        libs.append((lib, fn))
    return libs


@pytest.mark.parametrize("trig_rows", ["1", "0"])
@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("sub", ["0", "40"])
def test_generated_segments_run_on_host(monkeypatch, sub, split, trig_rows):
    """split: each reverse kernel as its two halves -- the first pass storing
    the sub-block boundary states, then the rest reloading every sub-block.
    trig_rows: the first pass also stores its sin/cos pairs for the
    recomputation to load (CXH_JIT_TRIG_CACHE)."""
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    monkeypatch.setenv("CXH_JIT_SUBBLOCK", sub)
    monkeypatch.setenv("CXH_JIT_TRIG_CACHE", trig_rows)
    f = W.cartpole_rollout(4)
    every = 170
    tape = HipTape(f, engine="jit")
    nseg = (f.n_instructions() + every - 1) // every
    fwd = [tape.jit_source(every, s, False)[0] for s in range(nseg)]
    rev = [tape.jit_source(every, s, True)[0] for s in range(nseg)]
    first = [tape.jit_source(every, s, 2)[0] for s in range(nseg)]
    rest = [tape.jit_source(every, s, 3)[0] for s in range(nseg)]
    assert all("__global__ void __launch_bounds__(256) cxj_snap(" in x for x in first)
    body = [x.split("void cxj_body(")[1] for x in first + rest]
    assert not any("op_der_full(" in x or "CXJ_ADJ" in x for x in body[:nseg])   # no adjoint work in the first pass
    assert not any(re.search(r"CXJ_SNAP\(A\.st_out, \d+\) =", x) for x in body[nseg:])   # no states stored
    if sub != "0":
        assert any("CXJ_SNAP(" in r for r in rev)       # sub-blocks were generated
    rows = [len(re.findall(r"CXJ_ST_IN\(", r)) for r in rev]  # live slots at each segment start
    n = 37
    rng = np.random.default_rng(4)
    x0, us = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (4, n))
    seed = rng.standard_normal((1, n))
    with tempfile.TemporaryDirectory() as d:
        libs = _compile(fwd[:-1] + (first + rest if split else rev), d)
        fk, rk = [fn for _, fn in libs[:nseg - 1]], [fn for _, fn in libs[nseg - 1:]]
        if split:
            fp, rk = rk[:nseg], rk[nseg:]
        state = [np.zeros((max(r, 1), n)) for r in rows] + [np.zeros((1, n))]
        J = np.zeros((1, n))
        gx, gu = np.zeros((4, n)), np.zeros((4, n))
        A = CxjArgs()
        A.n, A.lds = n, n
        A.arg = _table([x0, us], n)
        A.res = _table([J], n)
        for s in range(nseg - 1):
            A.st_in = state[s].ctypes.data
            A.st_out = state[s + 1].ctypes.data
            fk[s](C.byref(A), n)
        A.aseed = _table([seed], n)
        A.asens = _table([gx, gu], n)
        adj = [np.zeros((max(rows + [1]), n)) for _ in range(2)]
        snap = np.zeros((4096, n))          # sub-block boundary states
        A.st_out = snap.ctypes.data
        for s in range(nseg - 1, -1, -1):
            A.res = _table([J] if s == nseg - 1 else [None], n)
            A.st_in = state[s].ctypes.data
            A.adj_in = adj[(s + 1) % 2].ctypes.data
            A.adj_out = adj[s % 2].ctypes.data
            if split:
                fp[s](C.byref(A), n)
            rk[s](C.byref(A), n)
    (Jo,), (ox, ou) = O.adj_map(f, n, [x0.T, us.T], [seed.T])
    assert np.array_equal(J, Jo.T)
    assert np.array_equal(gx, ox.T) and np.array_equal(gu, ou.T)


def test_generated_tangent_sweeps_run_on_host(monkeypatch):
    """Specialised tangent kernels (cart-pole Jacobian, 5 forward colours in
    direction groups of 2, one segment; the star-coloured Hessian of a 2-step
    quadrotor OCP Lagrangian in direction groups and segments) against the
    oracle evaluating the symbolic jacobian() / hessian() Functions."""
    import casadi_amd as ca
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipJacobian
    monkeypatch.setenv("CXH_JIT_DIRGROUP", "2")
    f = W.cartpole_step()
    hj = HipJacobian(f, engine="jit")
    n = 29
    rng = np.random.default_rng(9)
    x, u = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (1, n))
    src, ng, ns = hj.jit_source(0)
    srcs = [src]
    assert ng == 3 and ns == 1
    jac = np.full((hj.nnz_jac, n), np.nan)
    y = np.zeros((4, n))
    _, sp, _, lds = _tiled(srcs, n, ng)
    with tempfile.TemporaryDirectory() as d:
        libs = _compile(srcs, d, ng)
        A = CxjArgs()
        A.n, A.lds = n, lds
        A.arg = _table([x, u], n)
        A.res = _table([y], n)
        A.asens = _table([jac], n)
        A.spill = sp.ctypes.data
        for _, fn in libs:
            fn(C.byref(A), n)
    ref = np.concatenate(O.map_eval(f.jacobian(), n, [x.T, u.T, None]), axis=1).T
    assert O.rel_err(jac, ref) <= 1e-14
    assert O.rel_err(y, O.map_eval(f, n, [x.T, u.T])[0].T) == 0.0
    # star-coloured Hessian of a 2-step quadrotor OCP Lagrangian, 4 segments
    monkeypatch.setenv("CXH_JIT_DIRGROUP", "4")
    monkeypatch.setenv("CXH_JIT_TANSEG", "400")
    monkeypatch.setenv("CXH_JIT_TANREGS", "24")   # small register budget: spills and reloads
    w, lam, Lag = W.quadrotor_ocp_lagrangian(N=2)
    G = ca.Function("grad_lag", [w, lam], [ca.gradient(Lag, w)])
    H, _ = ca.hessian(Lag, w)
    hh = HipJacobian(G, symmetric=True, first_block=True, engine="jit")
    _, ng, ns = hh.jit_source(0)
    assert ns > 1 and ng > 1
    wv, lv = rng.uniform(-1, 1, (w.nnz(), n)), rng.standard_normal((lam.nnz(), n))
    hnz = np.full((hh.nnz_jac, n), np.nan)
    gv = np.zeros((w.nnz(), n))
    srcs = [hh.jit_source(s)[0] for s in range(ns)]
    assert all(f"cxj_body_g{g}(" in x for x in srcs for g in range(ng))
    st, sp, n_spill, lds = _tiled(srcs, n, ng)
    assert n_spill > 0 and any("cxj_ss" in x for x in srcs)
    with tempfile.TemporaryDirectory() as d:
        libs = _compile(srcs, d, ng)
        A = CxjArgs()
        A.n, A.lds = n, lds
        A.arg = _table([wv, lv], n)
        A.res = _table([gv], n)
        A.asens = _table([hnz], n)
        A.spill = sp.ctypes.data
        A.st_in = A.st_out = st[0].ctypes.data    # home rows: read and written in place
        for _, fn in libs:
            fn(C.byref(A), n)
    Hf = ca.Function("hess_lag", [w, lam], [H])
    (href,) = O.map_eval(Hf, n, [wv.T, lv.T])
    assert hh.jsp == H.sparsity
    assert O.rel_err(hnz, href.T) <= 1e-13
    (gref,) = O.map_eval(G, n, [wv.T, lv.T])
    assert O.rel_err(gv, gref.T) == 0.0


@pytest.mark.parametrize("K", [0, 3, 4])
def test_mapaccum_loop_kernels_run_on_host(K):
    """The mapaccum loop kernels (one step body, checkpoints every K steps)
    against the oracle on the unrolled chain."""
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipMapAccum
    step = W.cartpole_cost_step()
    N = 10
    F = step.mapaccum("roll", N)
    m = HipMapAccum.from_function(F)
    n = 23
    rng = np.random.default_rng(11)
    x0 = np.vstack([rng.uniform(-1, 1, (4, n)), np.zeros((1, n))])
    U = rng.uniform(-5, 5, (N, n))
    seed = rng.standard_normal((5 * N, n))
    out = np.zeros((5 * N, n))
    g0, gU = np.zeros((5, n)), np.zeros((N, n))
    y = np.zeros((5 * N, n))
    out2 = np.zeros((5 * N, n))
    h0, hU = np.zeros((5, n)), np.zeros((N, n))
    with tempfile.TemporaryDirectory() as d:
        (_, fwd), (_, adj), (_, fck), (_, rck) = _compile([m.source(0), m.source(1, K), m.source(2, K),
                                                           m.source(3, K)], d)
        ckpt, steps = np.zeros((512, n)), np.zeros((512, n))
        A = CxjArgs()
        A.n, A.lds = n, n
        A.arg = _table([x0, U], n)
        A.res = _table([y], n)
        fwd(C.byref(A), n)
        A.res = _table([out], n)
        A.aseed = _table([seed], n)
        A.asens = _table([g0, gU], n)
        A.st_in, A.st_out = ckpt.ctypes.data, steps.ctypes.data
        adj(C.byref(A), n)
        # split: forward storing checkpoints, then the adjoint from them
        ckpt2 = np.zeros((512, n))
        B = CxjArgs()
        B.n, B.lds = n, n
        B.arg = _table([x0, U], n)
        B.res = _table([out2], n)
        B.st_in, B.st_out = ckpt2.ctypes.data, steps.ctypes.data
        fck(C.byref(B), n)
        B.res = _table([], n)
        B.aseed = _table([seed], n)
        B.asens = _table([h0, hU], n)
        rck(C.byref(B), n)
    (ref,) = O.map_eval(F, n, [x0.T, U.T])
    assert np.array_equal(y, ref.T) and np.array_equal(out, ref.T) and np.array_equal(out2, ref.T)
    (_,), (r0, rU) = O.adj_map(F, n, [x0.T, U.T], [seed.T])
    assert O.rel_err(g0, r0.T) <= 1e-13 and O.rel_err(gU, rU.T) <= 1e-13
    assert np.array_equal(h0, g0) and np.array_equal(hU, gU)
    # the reverse per segment (mode 4): checkpoints and the last segment's step
    # states, then for seg = S-1 .. 0 the reverse steps of seg beside the
    # recompute of seg - 1 -- bit for bit the single-kernel adjoint
    S = (N + K - 1) // K if K else 1
    Kq = K if K else N
    src = m.source(4, K)
    cut = src.index('extern "C" __global__')
    code = ("#define __device__\n#define __forceinline__ inline\n#define CXJ_KEEP(x) ((void)0)\n"
            "#define CXJ_KEEP_PTR(x) ((void)0)\n#define CXJ_AFTER(p, x) ((void)0)\n#define CXJ_GLOBAL\n"
            "#define CXJ_OPAQUE(p) ((void)0)\n#define CXJ_OPAQUE_AFTER(p, x) ((void)0)\n#define CXJ_ANY(b) (b)\n" + src[:cut] +
            "extern \"C\" void run_pair(const CxjArgs* A, const CxjArgs* B, long long n, long long seg) {\n"
            f"  if (seg < {S}) for (long long k = 0; k < n; ++k) cxj_back(*A, k, seg);\n"
            "  if (seg >= 1) for (long long k = 0; k < n; ++k) cxj_recompute(*B, k, seg - 1);\n}\n")
    with tempfile.TemporaryDirectory() as d:
        path, so = os.path.join(d, "pair.cpp"), os.path.join(d, "pair.so")
        with open(path, "w") as fh:
            fh.write(code)
        subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", *NO_SINCOS_FUSION, "-shared", "-fPIC", "-w", path, "-o", so],
                       check=True)
        run = C.CDLL(so).run_pair
        run.argtypes = [C.POINTER(CxjArgs), C.POINTER(CxjArgs), C.c_longlong, C.c_longlong]
        (_, fck2), = _compile([m.source(2, K)], d)
        ck3 = np.zeros((S * 5, n))
        states = [np.zeros((Kq * 5, n)) for _ in range(2)]
        adjb = [np.zeros((5, n)) for _ in range(2)]
        k0, kU, out3 = np.zeros((5, n)), np.zeros((N, n)), np.zeros((5 * N, n))
        A3 = CxjArgs()
        A3.n, A3.lds = n, n
        A3.arg = _table([x0, U], n)
        A3.res = _table([out3], n)
        A3.st_in = ck3.ctypes.data
        A3.st_out = states[(S - 1) % 2].ctypes.data   # the forward also stores the last segment's step states
        fck2(C.byref(A3), n)
        A3.res = _table([], n)
        A3.aseed = _table([seed], n)
        A3.asens = _table([k0, kU], n)
        B3 = CxjArgs()
        C.memmove(C.byref(B3), C.byref(A3), C.sizeof(CxjArgs))
        for seg in range(S - 1, -1, -1):
            A3.st_out = states[seg % 2].ctypes.data
            A3.adj_in = adjb[(seg + 1) % 2].ctypes.data
            A3.adj_out = adjb[seg % 2].ctypes.data
            B3.st_out = states[(seg + 1) % 2].ctypes.data
            run(C.byref(A3), C.byref(B3), n, seg)
    assert np.array_equal(out3, ref.T)
    assert np.array_equal(k0, g0) and np.array_equal(kU, gU)


def test_specialised_ldl_runs_on_host():
    """The straight-line LDL factor + solve generated for a KKT pattern (5
    stages, two right-hand sides), executed on the host, against the oracle's
    casadi_ldl / casadi_ldl_solve (itself bit-exact against the reference
    runtime): bit for bit, a zero matrix included (every pivot counted)."""
    from casadi_amd import workloads as W
    from casadi_amd.linsol import Linsol
    sp, _, _ = W.ocp_kkt_sparsity(N=5)
    Lt, p = sp.ldl(amd=True)
    src = Linsol("k", "ldl", sp).jit_source(2)
    n, nn = 5, sp.size1()
    a = np.zeros((sp.nnz(), n))
    for k in range(n - 1):
        K = W.ocp_kkt_values(np.random.default_rng(k), N=5)
        a[:, k] = K[sp.row, sp.get_col()]
    b = np.random.default_rng(9).standard_normal((2 * nn, n))
    x = np.zeros((2 * nn, n))
    cnt = np.zeros((1, n))
    st, spill, _, lds = _tiled([src], n, 1)
    with tempfile.TemporaryDirectory() as d:
        (_, fn), = _compile([src], d, 1)
        A = CxjArgs()
        A.n, A.lds = n, lds
        A.arg = _table([a, b], n)
        A.res = _table([x, cnt], n)
        A.spill = spill.ctypes.data
        A.st_in, A.st_out = st[0].ctypes.data, st[1].ctypes.data
        fn(C.byref(A), n)
    for k in range(n):
        lt, dd = O.ldl(sp.compressed(), a[:, k], Lt.compressed(), p)
        xr = O.ldl_solve(b[:, k], Lt.compressed(), lt, dd, p)
        assert np.array_equal(x[:, k], xr, equal_nan=True), k
        assert cnt[0, k] == np.sum(~((dd != 0) & np.isfinite(dd)))
    assert cnt[0, n - 1] == nn


@pytest.mark.parametrize("tr", [False, True])
def test_specialised_qr_runs_on_host(tr):
    """The straight-line QR factor + solve for a KKT pattern (5 stages, two
    right-hand sides, in segments whose crossing values keep home rows),
    executed on the host, against the oracle's casadi_qr / casadi_qr_solve:
    bit for bit, a zero matrix included (every diagonal counted)."""
    from casadi_amd import workloads as W
    from casadi_amd.linsol import Linsol
    sp, _, _ = W.ocp_kkt_sparsity(N=5)
    ls = Linsol("q", "qr", sp)
    src0, ns = ls.jit_source(2, tr, 0)
    assert ns > 1
    srcs = [src0] + [ls.jit_source(2, tr, s)[0] for s in range(1, ns)]
    n, nn = 5, sp.size1()
    a = np.zeros((sp.nnz(), n))
    for k in range(n - 1):
        K = W.ocp_kkt_values(np.random.default_rng(k), N=5)
        a[:, k] = K[sp.row, sp.get_col()]
    b = np.random.default_rng(9).standard_normal((2 * nn, n))
    x = np.zeros((2 * nn, n))
    cnt = np.zeros((1, n))
    st, spill, _, lds = _tiled(srcs, n, 1)
    with tempfile.TemporaryDirectory() as d:
        libs = _compile(srcs, d, 1)
        A = CxjArgs()
        A.n, A.lds = n, lds
        A.arg = _table([a, b], n)
        A.res = _table([x, cnt], n)
        A.spill = spill.ctypes.data
        A.st_in = A.st_out = st[0].ctypes.data
        for _, fn in libs:
            fn(C.byref(A), n)
    svc, src_ = ls.sp_v.compressed(), ls.sp_r.compressed()
    for k in range(n):
        v, r, beta = O.qr(sp.compressed(), a[:, k], svc, src_, ls.prinv, ls.pc)
        xr = O.qr_solve(b[:, k], int(tr), svc, v, src_, r, beta, ls.prinv, ls.pc)
        assert np.array_equal(x[:, k], xr, equal_nan=True), k
        assert cnt[0, k] == O.qr_singular(r, src_, ls.pc, ls.eps)[0]
    assert cnt[0, n - 1] == nn


def test_soa32_forward_segments_run_on_host(monkeypatch):
    """The forward segment kernels built for all-SoA launches (CXJ_SOA32:
    the row start wave-uniform, the instance a 32-bit byte offset) compute
    the same outputs and boundary states as the general ones, on a cart-pole
    step map and the segments of a short rollout."""
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    monkeypatch.setenv("CXH_JIT_ROLL", "0")
    f = W.cartpole_rollout(3)
    every = 150
    tape = HipTape(f, engine="jit")
    nseg = (f.n_instructions() + every - 1) // every
    src = [tape.jit_source(every, s, False)[0] for s in range(nseg)]
    n = 77
    rng = np.random.default_rng(12)
    x0, us = rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (3, n))
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for soa in (False, True):
            sub = os.path.join(d, "soa32" if soa else "general")
            os.makedirs(sub)
            libs = _compile([("#define CXJ_SOA32\n" if soa else "") + x for x in src], sub)
            J = np.full((1, n), np.nan)
            state = [np.zeros((64, n)) for _ in range(nseg + 1)]
            A = CxjArgs()
            A.n, A.lds = n, n
            A.arg = _table([x0, us], n)
            A.res = _table([J], n)
            for s, (_, fn) in enumerate(libs):
                A.st_in = state[s].ctypes.data
                A.st_out = state[s + 1].ctypes.data
                fn(C.byref(A), n)
            out[soa] = (J.copy(), [x.copy() for x in state])
    assert np.array_equal(out[True][0], out[False][0])
    assert all(np.array_equal(a, b) for a, b in zip(out[True][1], out[False][1]))
    (Jo,) = O.map_eval(f, n, [x0.T, us.T])
    assert np.array_equal(out[True][0], Jo.T)
