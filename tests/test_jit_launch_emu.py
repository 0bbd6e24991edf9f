"""Generated kernels run on the host at the device's exact launch geometry,
every memory access bounds- and race-checked.

The host-execution tests (test_jit_host_exec.py) check the generated
arithmetic.  This file checks where the kernels touch memory, as the GPU
driver launches them:

* the grid the driver launches (ceil(n / 64) blocks of 64 threads, the padded
  last wavefront included; blockIdx.y for the paired reverse launches and the
  side-by-side direction groups), calling the kernel entry points unchanged;
* the buffers the driver allocates, at their exact sizes: checkpoint rows
  `ck_rows * lds`, the two alternating sub-block state buffers
  `snap_rows * lds` each, the two adjoint hand-over buffers
  `max_live * lds` each, with `lds = round_up(n, 64)`, and the tangent
  kernels' state and spill rows tiled by 64 instances per direction group;
* every load and store checked against the buffer its base pointer came
  from, and, within one launch, against every other lane's accesses: an
  address one lane writes is touched by no other lane (lanes of a launch run
  in no fixed order on the GPU, so such an overlap would be a race that a
  sequential host run cannot show).

The header `jit_launch_emu.h` supplies the checked memory macros.  The
results are also compared with the oracle, so a launch sequence that skipped
work would fail too.
"""
import ctypes as C
import os
import re
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_ref as O
from test_jit_host_exec import NO_SINCOS_FUSION, CxjArgs, _table

HERE = os.path.dirname(os.path.abspath(__file__))
BLOCK = 64


def _round64(n):
    return (n + 63) // 64 * 64


_MEM_MACROS = re.compile(r"^#define CXJ_(AT|IN|OUT|SEED|SENS|ST_IN|ST_OUT|ADJ_IN|ADJ_OUT|SNAP)\(.*$", re.M)
_KERNEL = re.compile(r'extern "C" __global__ void __launch_bounds__\(\d+\) (\w+)\(([^)]*)\)')


def _emu_source(src):
    """The generated source with its memory macros and global pointers
    replaced by the checked ones, plus a launch entry point."""
    src = _MEM_MACROS.sub("", src)
    src = src.replace("CXJ_GLOBAL const double*", "cxj_emu::Ptr<const double>")
    src = src.replace("CXJ_GLOBAL double*", "cxj_emu::Ptr<double>")
    m = _KERNEL.search(src)
    assert m, "no kernel entry point"
    params = m.group(2)
    call = (f"{m.group(1)}(*A" + (", *B" if params.count("CxjArgs") == 2 else "") +
            (", seg" if "long long seg" in params else "") + ")")
    return ('#include "jit_launch_emu.h"\n' + src +
            '\nextern "C" void emu_launch(const CxjArgs* A, const CxjArgs* B, long long seg, unsigned gx, '
            f"unsigned gy, unsigned bx) {{\n  (void)seg;\n  CXJ_EMU_LAUNCH_LOOP({call});\n}}\n")


class Emu:
    """The compiled kernels of one test and the device buffers they see."""

    def __init__(self, sources, workdir):
        def build(i):
            path = os.path.join(workdir, f"emu{i}.cpp")
            with open(path, "w") as fh:
                fh.write(_emu_source(sources[i]))
            so = os.path.join(workdir, f"emu{i}.so")
            subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", *NO_SINCOS_FUSION, "-shared", "-fPIC", "-w", "-I", HERE,
                            path, "-o", so], check=True)
            return so
        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
            sos = list(ex.map(build, range(len(sources))))
        self.libs = []
        for so in sos:
            lib = C.CDLL(so)
            lib.emu_launch.argtypes = [C.POINTER(CxjArgs), C.POINTER(CxjArgs), C.c_longlong, C.c_uint, C.c_uint,
                                       C.c_uint]
            lib.emu_add_region.argtypes = [C.c_void_p, C.c_longlong, C.c_int]
            lib.emu_errors.restype = C.c_longlong
            lib.emu_accesses.restype = C.c_longlong
            self.libs.append(lib)
        self.buffers = []

    def buffer(self, n_doubles):
        """A device buffer of exactly n_doubles doubles (None when empty, as
        the driver passes a null pointer for an empty buffer)."""
        if n_doubles <= 0:
            return None
        b = np.zeros(int(n_doubles))
        self.buffers.append(b)
        return b

    def launch(self, i, A, n, B=None, gy=1, seg=0):
        lib = self.libs[i]
        lib.emu_clear_regions()
        for r, b in enumerate(self.buffers):
            lib.emu_add_region(b.ctypes.data, b.nbytes, r)
        lib.emu_reset_errors()
        lib.emu_launch(C.byref(A), C.byref(B if B is not None else A), seg, (n + BLOCK - 1) // BLOCK, gy, BLOCK)
        assert lib.emu_accesses() > 0
        return lib.emu_errors()


def _ptr(b):
    return None if b is None else b.ctypes.data


def _off(b, doubles):
    return None if b is None else b.ctypes.data + 8 * int(doubles)


def _rollout_sweep(f, every, n, seed_val):
    """The fused forward + adjoint of the split sweep, launched as the driver
    launches it (mode 0), on the emulator.  Returns (errors, J, grads)."""
    from casadi_amd.hip import HipTape
    tape = HipTape(f, engine="jit")
    P = tape.jit_plan(every)
    S = P["segments"]
    assert P["split"] and S > 1 and P["ipt"] == 1
    srcs = [tape.jit_source(every, g, 0)[0] for g in range(S - 1)]           # forward segments
    srcs.append(tape.jit_source(every, S - 1, 2)[0])                        # first pass of the last one
    srcs += [tape.jit_source(every, g, 4)[0] for g in range(S - 1, 0, -1)]  # pairs g, g-1
    srcs.append(tape.jit_source(every, 0, 3)[0])                            # the rest of segment 0
    rng = np.random.default_rng(seed_val)
    x0 = rng.uniform(-1, 1, (4, n))
    us = rng.uniform(-10, 10, (f.nnz_in(1), n))
    seed = rng.standard_normal((1, n))
    errors = 0
    with tempfile.TemporaryDirectory() as d:
        emu = Emu(srcs, d)
        lds = _round64(n)
        # inputs and outputs: SoA [nz][n] with ld = n, as HipTape passes them
        xb, ub, Jb, sb = emu.buffer(4 * n), emu.buffer(us.size), emu.buffer(n), emu.buffer(n)
        gxb, gub = emu.buffer(4 * n), emu.buffer(us.size)
        xb[:] = x0.ravel()
        ub[:] = us.ravel()
        sb[:] = seed.ravel()
        ck = emu.buffer(P["ck_rows"] * lds)
        adj = emu.buffer(2 * P["max_live"] * lds)
        snap = emu.buffer(2 * P["snap_rows"] * lds)

        def state(b):
            return _off(ck, P["ck_off"][b] * lds) if 0 < b < S else None

        def buf(g):
            return _off(snap, (g % 2) * P["snap_rows"] * lds)

        arg = _table([xb, ub], n)
        res = _table([Jb], n)
        none = _table([], n)
        A = CxjArgs()
        A.arg, A.n, A.lds = arg, n, lds
        for g in range(S - 1):
            A.res, A.st_in, A.st_out = res, state(g), state(g + 1)
            errors += emu.launch(g, A, n)
        A.aseed = _table([sb], n)
        A.asens = _table([gxb, gub], n)
        B = CxjArgs.from_buffer_copy(A)
        B.res, B.st_in, B.st_out = res, state(S - 1), buf(S - 1)
        errors += emu.launch(S - 1, B, n)
        for i, g in enumerate(range(S - 1, -1, -1)):
            A.res = res if g == S - 1 else none
            A.st_in, A.st_out = state(g), buf(g)
            A.adj_in = _off(adj, ((g + 1) % 2) * P["max_live"] * lds)
            A.adj_out = _off(adj, (g % 2) * P["max_live"] * lds)
            if g == 0:
                errors += emu.launch(len(srcs) - 1, A, n)
                break
            B.res, B.st_in, B.st_out = none, state(g - 1), buf(g - 1)
            errors += emu.launch(S + i, A, n, B, gy=2)
        return errors, Jb.reshape(1, n).copy(), (gxb.reshape(4, n).copy(), gub.reshape(-1, n).copy()), (x0, us, seed)


@pytest.mark.parametrize("steps,every,n,sub", [
    (20, 900, 1000, None),   # segments 900, 900, 900, 81: the padded last wavefront, a short last segment
    (4, 170, 130, "40"),     # several sub-blocks per segment, three tiles of 64 (the last ragged)
])
@pytest.mark.parametrize("trig_cache", ["1", "0"])
def test_split_reverse_sweep_in_bounds_and_race_free(monkeypatch, steps, every, n, sub, trig_cache):
    from casadi_amd import workloads as W
    if sub is not None:
        monkeypatch.setenv("CXH_JIT_SUBBLOCK", sub)
    monkeypatch.setenv("CXH_JIT_TRIG_CACHE", trig_cache)
    f = W.cartpole_rollout(steps)
    errors, J, (gx, gu), (x0, us, seed) = _rollout_sweep(f, every, n, 20261205 + steps)
    assert errors == 0
    # and the sweep computed the oracle's cost and gradient: its numeric
    # reverse sweep of the same tape, same host libm, so bit for bit
    (Jr,), (rx, ru) = O.adj_map(f, n, [x0.T, us.T], [seed.T])
    assert np.array_equal(J, Jr.T)
    assert np.array_equal(gx, rx.T) and np.array_equal(gu, ru.T)


def test_emulator_catches_planted_mistakes(monkeypatch):
    """A sub-block state buffer one row short, and the two alternating
    buffers collapsed into one, are both reported: the first as out of
    bounds, the second as a race between the two halves of a paired launch."""
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_rollout(4)
    every, n = 170, 130
    monkeypatch.setenv("CXH_JIT_SUBBLOCK", "40")
    tape = HipTape(f, engine="jit")
    P = tape.jit_plan(every)
    S = P["segments"]
    g = S - 1
    srcs = [tape.jit_source(every, g, 2)[0], tape.jit_source(every, g, 4)[0]]
    used = 1 + max(int(r) for r in re.findall(r"CXJ_SNAP\(A\.st_out, (\d+)\) =", srcs[0]))
    assert used <= P["snap_rows"]
    rng = np.random.default_rng(3)
    with tempfile.TemporaryDirectory() as d:
        emu = Emu(srcs, d)
        lds = _round64(n)
        xb, ub, sb = emu.buffer(4 * n), emu.buffer(4 * n), emu.buffer(n)
        gxb, gub = emu.buffer(4 * n), emu.buffer(4 * n)
        xb[:] = rng.uniform(-1, 1, 4 * n)
        ub[:] = rng.uniform(-10, 10, 4 * n)
        ck = emu.buffer(P["ck_rows"] * lds)
        adj = emu.buffer(2 * P["max_live"] * lds)
        short = emu.buffer((used - 1) * lds)                 # planted: one row short
        A = CxjArgs()
        A.arg, A.n, A.lds = _table([xb, ub], n), n, lds
        A.st_in, A.st_out = _off(ck, P["ck_off"][g] * lds), _ptr(short)
        assert emu.launch(0, A, n) > 0
        snap = emu.buffer(P["snap_rows"] * lds)              # planted: one buffer for both halves
        A.aseed, A.asens = _table([sb], n), _table([gxb, gub], n)
        A.st_out = _ptr(snap)
        A.adj_in, A.adj_out = _ptr(adj), _off(adj, P["max_live"] * lds)
        assert emu.launch(0, A, n) == 0                      # the first pass alone is fine
        B = CxjArgs.from_buffer_copy(A)
        B.st_in = _off(ck, P["ck_off"][g - 1] * lds)
        assert emu.launch(1, A, n, B, gy=2) > 0


def _tangent_launch(G, n, settings, monkeypatch):
    """The specialised tangent kernels of a Jacobian plan launched as
    cxh_map_jac launches them: every segment one launch, the direction
    groups side by side (blockIdx.y), home and spill rows tiled by 64
    instances per group at their exact sizes.  The sources are those of the
    first register budget (the driver lowers it where the compiler would
    still spill; the host cannot tell)."""
    from casadi_amd.hip import HipJacobian
    for k, v in settings.items():
        monkeypatch.setenv(k, v)
    hh = HipJacobian(G, symmetric=True, first_block=True, engine="jit")
    src0, ng, ns = hh.jit_source(0)
    srcs = [src0] + [hh.jit_source(s)[0] for s in range(1, ns)]
    rows = max(int(re.search(r"#define CXJ_TROWS (\d+)", x).group(1)) for x in srcs)
    spill = max(int(re.search(r"#define CXJ_SROWS (\d+)", x).group(1)) for x in srcs)
    return hh, srcs, ng, ns, rows, spill


@pytest.mark.parametrize("settings", [
    {"CXH_JIT_DIRGROUP": "4", "CXH_JIT_TANSEG": "400", "CXH_JIT_TANREGS": "24"},
    {"CXH_JIT_DIRGROUP": "3"}])
def test_tangent_kernels_in_bounds_and_race_free(monkeypatch, settings):
    import casadi_amd as ca
    from casadi_amd import workloads as W
    w, lam, Lag = W.quadrotor_ocp_lagrangian(N=2)
    G = ca.Function("grad_lag", [w, lam], [ca.gradient(Lag, w)])
    n = 130
    hh, srcs, ng, ns, rows, spill = _tangent_launch(G, n, settings, monkeypatch)
    assert ng >= 2
    rng = np.random.default_rng(9)
    wv, lv = rng.uniform(-1, 1, (w.nnz(), n)), rng.standard_normal((lam.nnz(), n))
    nnz_jac = hh.jsp.nnz()
    with tempfile.TemporaryDirectory() as d:
        emu = Emu(srcs, d)
        lds = _round64(n)
        wb, lb, gb = emu.buffer(wv.size), emu.buffer(lv.size), emu.buffer(w.nnz() * n)
        jb = emu.buffer(nnz_jac * n)
        wb[:] = wv.ravel()
        lb[:] = lv.ravel()
        st = emu.buffer(ng * lds * rows)     # home rows: read and written in place
        sp = emu.buffer(ng * lds * spill)
        errors = 0
        for s in range(ns):
            A = CxjArgs()
            A.arg, A.n, A.lds = _table([wb, lb], n), n, lds
            A.res = _table([gb], n)
            A.st_in = A.st_out = _ptr(st)
            A.spill = _ptr(sp)
            A.asens = _table([jb], n)
            errors += emu.launch(s, A, n, gy=ng)
    assert errors == 0
    H, _ = ca.hessian(Lag, w)
    (ref,) = O.map_eval(ca.Function("h", [w, lam], [H]), n, [wv.T, lv.T])
    assert O.rel_err(jb.reshape(nnz_jac, n), ref.T) <= 1e-13


@pytest.mark.parametrize("steps,every,n", [(4, 150, 130), (6, 701, 70)])
def test_evaluation_home_rows_in_bounds_and_exact(monkeypatch, steps, every, n):
    """Forward-only evaluation on the home rows: every value that crosses a
    segment boundary keeps one row of one buffer for its whole life, rows of
    dead values are reused, values passed through a segment unread are not
    touched.  The reverse Function of a rollout (its forward values stay live
    through its reverse half: wide live sets, many pass-throughs), launched
    segment by segment as the driver does: in bounds, race-free, and the
    oracle's outputs bit for bit."""
    monkeypatch.setenv("CXH_JIT_ROLL", "0")
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    f = W.cartpole_rollout(steps)
    R = f.reverse(1)
    tape = HipTape(R, engine="jit")
    P = tape.jit_plan(every)
    S = P["segments"]
    assert S > 2 and P["ipt"] == 1
    assert 0 < P["home_rows"] <= P["max_live"]      # rows reused: never more than the widest live set
    srcs = [tape.jit_source(every, g, 5)[0] for g in range(S)]
    rng = np.random.default_rng(20261211 + n)
    ins = [rng.uniform(-1, 1, (4, n)), rng.uniform(-10, 10, (steps, n)), None, rng.standard_normal((1, n))]
    with tempfile.TemporaryDirectory() as d:
        emu = Emu(srcs, d)
        lds = _round64(n)
        bufs = []
        for a in ins:
            if a is None:
                bufs.append(None)
                continue
            b = emu.buffer(a.size)
            b[:] = a.ravel()
            bufs.append(b)
        outs = [emu.buffer(R.nnz_out(j) * n) for j in range(R.n_out())]
        home = emu.buffer(P["home_rows"] * lds)
        A = CxjArgs()
        A.arg, A.res, A.n, A.lds = _table(bufs, n), _table(outs, n), n, lds
        A.st_in = A.st_out = _ptr(home)
        errors = sum(emu.launch(g, A, n) for g in range(S))
        assert errors == 0
        ref = O.map_eval(R, n, [None if a is None else a.T for a in ins])
        for y, r in zip(outs, ref):
            assert np.array_equal(y.reshape(-1, n), r.T)
