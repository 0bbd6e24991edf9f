"""GPU execution of SX tapes through the C ABI (``cxh_*``).

Per-instance data is structure-of-arrays: input ``i`` of a batch of ``n``
instances is a float64 CUDA tensor of shape ``(nnz_in(i), n)`` (nonzero-major,
row stride ``ld``), which is exactly CasADi's map argument ``repmat(x, 1, n)``
stored row-major.  PyTorch provides device memory and streams only; every
evaluation runs the hand-written kernels in ``libcasadi_hip.so``.
"""
import ctypes as C

import numpy as np

from . import _lib as L

CXH_SOA, CXH_AOS = 0, 1
ENGINES = {"interp": 0, "jit": 1}


def _torch():
    import torch
    return torch


def _stream_ptr(stream):
    """The HIP stream to launch on: the given torch stream's handle, else the
    current stream of the current device.  The raw handle query avoids
    building a Stream object per call (about 3 us of the host side of a
    small launch); torch builds without it take the object's handle."""
    if stream is not None:
        return stream.cuda_stream
    torch = _torch()
    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if raw is not None:
        return raw(torch.cuda.current_device())
    return torch.cuda.current_stream().cuda_stream


def _ptrs(tensors):
    return L.ptr_array([t.data_ptr() if t is not None else None for t in tensors])


def _check_soa(tensors, nnz, n, what):
    """The leading dimension of SoA blocks (nnz[i], >= n) after checking type,
    shape and strides (each tensor attribute read once: this runs on every
    call's host path)."""
    torch = _torch()
    Tensor, f64 = torch.Tensor, torch.float64
    ld = None
    for i, t in enumerate(tensors):
        if t is None:
            continue
        if not isinstance(t, Tensor) or t.dtype is not f64 or not t.is_cuda:
            raise TypeError(f"{what}[{i}] must be a float64 CUDA tensor")
        sh, k = t.shape, nnz[i]
        if len(sh) != 2 or sh[0] != k or sh[1] < n:
            raise ValueError(f"{what}[{i}] has shape {tuple(sh)}, expected ({k}, {n})")
        if k > 1:
            st = t.stride()
            if sh[1] > 1 and st[1] != 1:   # one instance: its stride is never used
                raise ValueError(f"{what}[{i}] must have unit stride along instances")
            if ld is None:
                ld = st[0]
            elif st[0] != ld:
                raise ValueError(f"{what}: all blocks must share the leading dimension")
        elif ld is None:
            ld = n
    return ld if ld is not None else n


class HipTape:
    """A Function's tape resident on the GPU (cxh_tape)."""

    def __init__(self, f, block=64, w_in_lds=True, engine="interp"):
        self.f = f
        self.nnz_in = [f.nnz_in(i) for i in range(f.n_in())]
        self.nnz_out = [f.nnz_out(i) for i in range(f.n_out())]
        instr = f.instructions()
        ni = np.array(self.nnz_in, np.int64)
        no = np.array(self.nnz_out, np.int64)
        h = L.P()
        L.check(L.cxh_tape_create(instr.ctypes.data if instr.size else None, instr.size, f.sz_w(),
                                  len(ni), ni.ctypes.data_as(L.PI64), len(no), no.ctypes.data_as(L.PI64),
                                  C.byref(h)), f"cxh_tape_create({f.name()})")
        self._h = h
        self.config(block, w_in_lds)
        self.set_engine(engine)

    @classmethod
    def nested(cls, instr, worksize, nnz_in, nnz_out, calls, block=64, w_in_lds=True, engine="interp"):
        """A tape holding nested Function calls (cxh_tape_create_nested): an
        OP_CALL record's i1 is the index k of calls[k] = (callee HipTape,
        argument slots, result slots), the slots as CasADi's call records hold
        them (one per callee input / output nonzero; -1 = result not needed).
        instr is a structured array of Function.instructions()'s layout."""
        from .function import INSTR_DTYPE
        self = cls.__new__(cls)
        self.f = None
        self.nnz_in, self.nnz_out = [int(k) for k in nnz_in], [int(k) for k in nnz_out]
        instr = np.ascontiguousarray(instr, INSTR_DTYPE)
        ni = np.array(self.nnz_in, np.int64)
        no = np.array(self.nnz_out, np.int64)
        dep = [np.ascontiguousarray(c[1], np.int32) for c in calls]
        res = [np.ascontiguousarray(c[2], np.int32) for c in calls]
        h = L.P()
        L.check(L.cxh_tape_create_nested(instr.ctypes.data if instr.size else None, instr.size, int(worksize),
                                         len(ni), ni.ctypes.data_as(L.PI64), len(no), no.ctypes.data_as(L.PI64),
                                         len(calls), L.ptr_array([c[0]._h.value for c in calls]),
                                         L.ptr_array([d.ctypes.data if d.size else None for d in dep]),
                                         L.ptr_array([r.ctypes.data if r.size else None for r in res]),
                                         C.byref(h)), "cxh_tape_create_nested")
        self._h = h
        self.config(block, w_in_lds)
        self.set_engine(engine)
        return self

    @classmethod
    def deserialize(cls, text, block=64, w_in_lds=True, engine="interp"):
        """A tape from CasADi's serialized SXFunction (the string
        Function.serialize() returns), read by cxh_tape_deserialize."""
        data = text.encode() if isinstance(text, str) else bytes(text)
        self = cls.__new__(cls)
        self.f = None
        h = L.P()
        L.check(L.cxh_tape_deserialize(data, len(data), C.byref(h)), "cxh_tape_deserialize")
        self._h = h
        ni, no = L.I32(), L.I32()
        a, b = (L.I64 * 16)(), (L.I64 * 16)()
        L.check(L.cxh_tape_io(h, C.byref(ni), C.byref(no), a, b), "cxh_tape_io")
        self.nnz_in, self.nnz_out = [int(a[i]) for i in range(ni.value)], [int(b[i]) for i in range(no.value)]
        self.config(block, w_in_lds)
        self.set_engine(engine)
        return self

    def instructions(self):
        """The tape the kernels run (calls inlined) and its work-vector size."""
        from .function import INSTR_DTYPE
        n, ws = L.I64(), L.I64()
        L.check(L.cxh_tape_instructions(self._h, None, 0, C.byref(n), C.byref(ws)), "cxh_tape_instructions")
        a = np.zeros(n.value, INSTR_DTYPE)
        if n.value:
            L.check(L.cxh_tape_instructions(self._h, a.ctypes.data, n.value, C.byref(n), C.byref(ws)),
                    "cxh_tape_instructions")
        return a, ws.value

    def regularity_check(self, capacity=None, on=True):
        """Per-instance regularity status (cxh_tape_regularity_check): later
        eval / fwd_adj / fwd_ckpt / rev_ckpt calls scan what they wrote,
        instance by instance, into ``self.status`` (int32 on the GPU: 1 NaN,
        | 2 infinity), and raise CasadiHipError naming the flagged instances
        when any is, as CasADi's regularity_check option raises on a NaN or
        Inf output.  capacity: the largest batch to be checked."""
        torch = _torch()
        if not on:
            L.check(L.cxh_tape_regularity_check(self._h, 0, None, 0), "cxh_tape_regularity_check")
            self.status = None
            return None
        self.status = torch.zeros(int(capacity), dtype=torch.int32, device="cuda")
        L.check(L.cxh_tape_regularity_check(self._h, 1, self.status.data_ptr(), int(capacity)),
                "cxh_tape_regularity_check")
        return self.status

    def _flagged(self, rc, n, what):
        """A positive return code is the count of instances the regularity
        check flagged."""
        if rc <= 0:
            return
        st = self.status[:n].cpu().numpy()
        bad = np.flatnonzero(st)
        kinds = {1: "NaN", 2: "Inf", 3: "NaN and Inf"}
        head = ", ".join(f"{k} ({kinds[int(st[k])]})" for k in bad[:8])
        raise L.CasadiHipError(f"{what}: {rc} of {n} instances wrote NaN or Inf: {head}"
                               + (" ..." if len(bad) > 8 else ""))

    def set_engine(self, engine):
        """'interp': tape interpreter kernels; 'jit': straight-line kernels
        generated from the tape and compiled at first use (hiprtc)."""
        if engine not in ENGINES:
            raise ValueError(f"engine must be one of {sorted(ENGINES)}")
        L.check(L.cxh_tape_engine(self._h, ENGINES[engine]), "cxh_tape_engine")
        self.engine = engine

    def jit_prepare(self, checkpoint_every=0, forward=True, reverse=True):
        """Compile the specialised kernels now instead of at first launch."""
        L.check(L.cxh_tape_jit_prepare(self._h, int(checkpoint_every), int(forward), int(reverse)),
                "cxh_tape_jit_prepare")

    def jit_source(self, checkpoint_every=0, segment=0, reverse=False):
        """Generated HIP source of one segment kernel and the segment count."""
        n, ns = L.I64(), L.I32()
        L.check(L.cxh_tape_jit_source(self._h, int(checkpoint_every), int(segment), int(reverse), None, 0,
                                      C.byref(n), C.byref(ns)), "cxh_tape_jit_source")
        buf = C.create_string_buffer(n.value + 1)
        L.check(L.cxh_tape_jit_source(self._h, int(checkpoint_every), int(segment), int(reverse), buf,
                                      n.value + 1, C.byref(n), C.byref(ns)), "cxh_tape_jit_source")
        return buf.value.decode(), ns.value

    def roll_plan(self, checkpoint_every=0):
        """The loop structure the specialiser found in the tape: dict of ok,
        prelude (instructions before the loop), period (body length),
        iterations, state (values crossing an iteration), K (iterations per
        checkpoint segment), segments, enabled (the tape runs as the loop: CXH_JIT_ROLL
        when the tape was first analysed, and a loop was found), step_rows
        (rows per iteration in the reverse's iteration-state buffers: the
        state, then the sin / cos rows the recompute stores)."""
        info = (L.I64 * 9)()
        L.check(L.cxh_tape_roll_plan(self._h, int(checkpoint_every), info), "cxh_tape_roll_plan")
        keys = ("ok", "prelude", "period", "iterations", "state", "K", "segments", "enabled", "step_rows")
        return {k: int(v) for k, v in zip(keys, info)}

    def roll_source(self, checkpoint_every=0, mode=0):
        """Generated source of the rerolled kernels: mode 0 the forward chain,
        2 the forward chain storing checkpoints, 4 the paired reverse."""
        n = L.I64()
        L.check(L.cxh_tape_roll_source(self._h, int(checkpoint_every), int(mode), None, 0, C.byref(n)),
                "cxh_tape_roll_source")
        buf = C.create_string_buffer(n.value + 1)
        L.check(L.cxh_tape_roll_source(self._h, int(checkpoint_every), int(mode), buf, n.value + 1, C.byref(n)),
                "cxh_tape_roll_source")
        return buf.value.decode()

    def jit_plan(self, checkpoint_every=0):
        """The plan of a checkpoint interval as the drivers size their buffers
        from it: dict of segments, ck_rows, snap_rows, max_live, split, ipt,
        home_rows (forward-only evaluation) and ck_off (first checkpoint row of
        each boundary)."""
        info = (L.I64 * 7)()
        L.check(L.cxh_tape_jit_plan(self._h, int(checkpoint_every), info, None), "cxh_tape_jit_plan")
        off = (L.I64 * (info[0] + 1))()
        L.check(L.cxh_tape_jit_plan(self._h, int(checkpoint_every), info, off), "cxh_tape_jit_plan")
        keys = ("segments", "ck_rows", "snap_rows", "max_live", "split", "ipt", "home_rows")
        d = {k: int(v) for k, v in zip(keys, info)}
        d["ck_off"] = [int(v) for v in off]
        return d

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and L.cxh_tape_destroy is not None:
            L.cxh_tape_destroy(h)
            self._h = None

    def config(self, block=-1, w_in_lds=-1, ipw=-1):
        """block: threads per workgroup (64/128/256); w_in_lds: work vector in
        LDS (True) or HBM (False); ipw: instances per wavefront (8/16/32/64)."""
        L.check(L.cxh_tape_config(self._h, int(block), int(w_in_lds) if w_in_lds != -1 else -1, int(ipw)))
        if block != -1:
            self.block = int(block)
        if ipw != -1:
            self.ipw = int(ipw)

    def jit_async(self, on=True):
        """Compile the specialised kernels on a worker thread (cxh_tape_jit_async):
        calls made before they are ready run on the interpreter, same bits."""
        L.check(L.cxh_tape_jit_async(self._h, int(bool(on))), "cxh_tape_jit_async")

    def jit_pending(self):
        """Whether a worker thread is still compiling kernels for this tape."""
        rc = L.cxh_tape_jit_pending(self._h)
        L.check(rc, "cxh_tape_jit_pending")
        return rc == 1

    def jit_fallback(self):
        """Which calls of this specialised tape run on the interpreter because a
        kernel they need would spill to scratch (cxh_tape_jit_fallback)."""
        buf = C.create_string_buffer(512)
        rc = L.cxh_tape_jit_fallback(self._h, buf, 512)
        L.check(rc, "cxh_tape_jit_fallback")
        return {"eval": bool(rc & 1), "adjoint": bool(rc & 2), "why": buf.value.decode()}

    def ipt(self, ipt):
        """Instances per thread of the specialised forward kernels (1/2/4/8;
        0: the CXH_JIT_IPT setting)."""
        L.check(L.cxh_tape_ipt(self._h, int(ipt)), "cxh_tape_ipt")

    def tune(self, args, n=None, reps=5, stream=None):
        """Autotune the map launch over these inputs (cxh_tape_tune): the
        fastest setting stays on the tape.  Returns (setting dict, ms per
        call).  The outputs of the timed calls are discarded."""
        like = self._like(args)
        n = n if n is not None else like.shape[1]
        ld = _check_soa(args, self.nnz_in, n, "arg")
        out = self._alloc(self.nnz_out, n, like)
        best = (L.I32 * 4)()
        ms = C.c_double()
        L.check(L.cxh_tape_tune(self._h, n, _ptrs(args), _ptrs(out), CXH_SOA, ld, int(reps), _stream_ptr(stream),
                                best, C.byref(ms)), "cxh_tape_tune")
        self.block, self.ipw = int(best[0]), int(best[2])
        cfg = {"block": int(best[0]), "w_in_lds": bool(best[1]), "lanes_per_instance": 64 // int(best[2]),
               "instances_per_thread": int(best[3]) or 1}
        return cfg, ms.value

    def _alloc(self, nnz, n, like):
        torch = _torch()
        return [torch.empty((k, n), dtype=torch.float64, device=like.device) if k else None for k in nnz]

    def _like(self, args):
        for a in args:
            if a is not None:
                return a
        torch = _torch()
        return torch.empty(0, dtype=torch.float64, device="cuda")

    # -- forward ---------------------------------------------------------------
    def eval(self, args, n=None, out=None, stream=None):
        """Outputs of n instances (list of (nnz_out(i), n) tensors)."""
        like = self._like(args)
        n = n if n is not None else like.shape[1]
        ld = _check_soa(args, self.nnz_in, n, "arg")
        out = out if out is not None else self._alloc(self.nnz_out, n, like)
        ld_o = _check_soa(out, self.nnz_out, n, "res")
        if ld_o != ld and any(k > 1 for k in self.nnz_out):
            raise ValueError("inputs and outputs must share the leading dimension")
        rc = L.cxh_map_eval(self._h, n, _ptrs(args), _ptrs(out), CXH_SOA, ld, _stream_ptr(stream))
        L.check(rc, "cxh_map_eval")
        self._flagged(rc, n, "cxh_map_eval")
        return out

    def eval_host(self, args_aos, n):
        """Map layout on the host: instance k of input i at args_aos[i][k*nnz:(k+1)*nnz]."""
        arg = [np.ascontiguousarray(a, dtype=float) if a is not None else None for a in args_aos]
        res = [np.zeros(n * k) for k in self.nnz_out]
        pa = L.ptr_array([a.ctypes.data if a is not None and a.size else None for a in arg])
        pr = L.ptr_array([r.ctypes.data if r.size else None for r in res])
        L.check(L.cxh_map_eval_host(self._h, n, pa, pr), "cxh_map_eval_host")
        return res

    # -- adjoint ---------------------------------------------------------------
    def fwd_adj(self, args, seeds, n=None, checkpoint_every=0, res=True, stream=None):
        """Primal outputs and adj_in = J^T seeds for every instance, one launch."""
        like = self._like(args)
        n = n if n is not None else like.shape[1]
        ld = _check_soa(args, self.nnz_in, n, "arg")
        _check_soa(seeds, self.nnz_out, n, "adj_out")
        outs = self._alloc(self.nnz_out, n, like) if res else [None] * len(self.nnz_out)
        adj = self._alloc(self.nnz_in, n, like)
        rc = L.cxh_map_fwd_adj(self._h, n, _ptrs(args), _ptrs(seeds), _ptrs(outs), _ptrs(adj), CXH_SOA, ld,
                               int(checkpoint_every), _stream_ptr(stream))
        L.check(rc, "cxh_map_fwd_adj")
        self._flagged(rc, n, "cxh_map_fwd_adj")
        return outs, adj

    # -- reduced maps (shared inputs, outputs summed over instances) ------------
    @staticmethod
    def _flags(flags, count, what):
        if flags is None:
            return [False] * count
        flags = [bool(f) for f in flags]
        if len(flags) != count:
            raise ValueError(f"{what} has {len(flags)} entries, expected {count}")
        return flags

    @staticmethod
    def _shared(t, nnz, what):
        """A block shared by all instances (or summed over them): nnz doubles."""
        torch = _torch()
        if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64:
            raise TypeError(f"{what} must be a float64 CUDA tensor")
        if t.numel() != nnz or (nnz > 1 and t.reshape(-1).stride(0) != 1):
            raise ValueError(f"{what} must hold the block's {nnz} nonzeros contiguously")
        return t.reshape(-1)

    def _red_args(self, tensors, nnz, flags, n, what):
        """Per-instance blocks checked as SoA, shared blocks as nnz vectors."""
        per = [None if (t is None or flags[i]) else t for i, t in enumerate(tensors)]
        ld = _check_soa(per, nnz, n, what)
        out = [None if t is None else (self._shared(t, nnz[i], f"{what}[{i}]") if flags[i] else t)
               for i, t in enumerate(tensors)]
        return out, ld

    def _red_alloc(self, nnz, flags, n, ld, like):
        torch = _torch()
        return [None if not k else
                torch.empty(k, dtype=torch.float64, device=like.device) if flags[i] else
                torch.empty((k, ld), dtype=torch.float64, device=like.device)[:, :n]
                for i, k in enumerate(nnz)]

    def eval_sum(self, args, n, reduce_in=None, reduce_out=None, stream=None):
        """Reduced map: inputs flagged in reduce_in are shared by all n
        instances (a vector of the block's nonzeros), outputs flagged in
        reduce_out are summed over the instances (a vector); the other blocks
        are (nnz, n) as in eval()."""
        ri = self._flags(reduce_in, len(self.nnz_in), "reduce_in")
        ro = self._flags(reduce_out, len(self.nnz_out), "reduce_out")
        like = self._like(args)
        args, ld = self._red_args(args, self.nnz_in, ri, n, "arg")
        out = self._red_alloc(self.nnz_out, ro, n, ld, like)
        L.check(L.cxh_map_eval_sum(self._h, n, _ptrs(args), _ptrs(out), (L.I32 * len(ri))(*ri),
                                   (L.I32 * len(ro))(*ro), ld, _stream_ptr(stream)), "cxh_map_eval_sum")
        return out

    def fwd_adj_sum(self, args, seeds, n, reduce_in=None, reduce_out=None, checkpoint_every=0, res=True,
                    stream=None):
        """Fused forward + adjoint of the reduced map: the seed of a reduced
        output (a vector) applies to every instance, the sensitivity of a
        reduced input is summed over the instances (a vector)."""
        ri = self._flags(reduce_in, len(self.nnz_in), "reduce_in")
        ro = self._flags(reduce_out, len(self.nnz_out), "reduce_out")
        like = self._like(args)
        args, ld = self._red_args(args, self.nnz_in, ri, n, "arg")
        seeds, ld_s = self._red_args(seeds, self.nnz_out, ro, n, "adj_out")
        per_in = any(k > 1 and not f for k, f in zip(self.nnz_in, ri))
        per_out = any(k > 1 and not f and s is not None for k, f, s in zip(self.nnz_out, ro, seeds))
        if per_in and per_out and ld_s != ld:
            raise ValueError("inputs and seeds must share the leading dimension")
        ld = ld if per_in else ld_s
        outs = self._red_alloc(self.nnz_out, ro, n, ld, like) if res else [None] * len(self.nnz_out)
        adj = self._red_alloc(self.nnz_in, ri, n, ld, like)
        L.check(L.cxh_map_fwd_adj_sum(self._h, n, _ptrs(args), _ptrs(seeds), _ptrs(outs), _ptrs(adj),
                                      (L.I32 * len(ri))(*ri), (L.I32 * len(ro))(*ro), ld, int(checkpoint_every),
                                      _stream_ptr(stream)), "cxh_map_fwd_adj_sum")
        return outs, adj

    # -- dependency (sparsity) propagation -------------------------------------
    @staticmethod
    def _check_words(tensors, nnz, n_words, what):
        """int64 CUDA tensors (nnz, >= n_words) holding 64-bit words; returns ld."""
        torch = _torch()
        ld = None
        for i, t in enumerate(tensors):
            if t is None:
                continue
            if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.int64:
                raise TypeError(f"{what}[{i}] must be an int64 CUDA tensor of 64-bit words")
            if t.dim() != 2 or t.shape[0] != nnz[i] or t.shape[1] < n_words:
                raise ValueError(f"{what}[{i}] has shape {tuple(t.shape)}, expected ({nnz[i]}, {n_words})")
            if nnz[i] > 1 and t.shape[1] > 1 and t.stride(1) != 1:
                raise ValueError(f"{what}[{i}] must have unit stride along words")
            s = t.stride(0) if nnz[i] > 1 else n_words
            if ld is None:
                ld = s
            elif nnz[i] > 1 and s != ld:
                raise ValueError(f"{what}[{i}]: every block must share the leading dimension")
        return ld if ld is not None else n_words

    def sp_forward(self, args, n_words, stream=None):
        """SXFunction::sp_forward over n_words 64-bit seed words per input
        nonzero (int64 tensors (nnz_in(i), n_words), None = no dependency):
        the dependency words of every output nonzero, (nnz_out(j), n_words)."""
        torch = _torch()
        ld = self._check_words(args, self.nnz_in, n_words, "arg")
        res = [torch.zeros((k, ld), dtype=torch.int64, device="cuda")[:, :n_words] if k else None
               for k in self.nnz_out]
        L.check(L.cxh_sp_forward(self._h, n_words, _ptrs(args), _ptrs(res), ld, _stream_ptr(stream)),
                "cxh_sp_forward")
        return res

    def sp_reverse(self, seeds, n_words, stream=None):
        """SXFunction::sp_reverse: seed words per output nonzero (left as they
        are; the sweep runs on a copy) to the dependency words of every input
        nonzero, (nnz_in(i), n_words)."""
        torch = _torch()
        self._check_words(seeds, self.nnz_out, n_words, "seed")
        work = [None if t is None else
                torch.empty((t.shape[0], n_words), dtype=torch.int64, device=t.device).copy_(t[:, :n_words])
                for t in seeds]
        ld = self._check_words(work, self.nnz_out, n_words, "seed")
        arg = [torch.zeros((k, ld), dtype=torch.int64, device="cuda")[:, :n_words] if k else None
               for k in self.nnz_in]
        L.check(L.cxh_sp_reverse(self._h, n_words, _ptrs(arg), _ptrs(work), ld, _stream_ptr(stream)),
                "cxh_sp_reverse")
        return arg

    def jac_sparsity(self, oind=0, iind=0):
        """Compact Jacobian pattern of output oind w.r.t. input iind
        (nnz_out(oind) x nnz_in(iind)) from GPU sweeps, one word per 64
        columns (forward) or rows (reverse), whichever needs fewer words --
        FunctionInternal::get_jac_sparsity's choice
        (casadi/core/function_internal.cpp:1928-1976)."""
        import numpy as np
        from .sx import Sparsity
        torch = _torch()
        n_in, n_out = self.nnz_in[iind], self.nnz_out[oind]
        w_fwd, w_adj = -(-n_in // 64), -(-n_out // 64)
        if n_in == 0 or n_out == 0:
            return Sparsity(n_out, n_in)

        def unit_words(n, nw):
            """(n, nw) words with bit r % 64 of word r // 64 set in row r, built on the device."""
            u = torch.zeros((n, nw), dtype=torch.int64, device="cuda")
            r = torch.arange(n, device="cuda")
            u[r, r // 64] = torch.bitwise_left_shift(torch.ones_like(r), r % 64)   # bit 63: the sign bit
            return u

        def set_bits(words):
            """(rows, nw) device words -> (row, 64 s + b) of every set bit b of
            word s; only the nonzero words leave the device."""
            idx = torch.nonzero(words)
            vals = words[idx[:, 0], idx[:, 1]].cpu().numpy()
            idx = idx.cpu().numpy()
            b = np.unpackbits(np.ascontiguousarray(vals).view(np.uint8).reshape(-1, 8), axis=1, bitorder="little")
            i, bit = np.nonzero(b)
            return idx[i, 0], idx[i, 1] * 64 + bit
        if w_fwd <= w_adj:
            seeds = [None] * len(self.nnz_in)
            seeds[iind] = unit_words(n_in, w_fwd)
            r, c = set_bits(self.sp_forward(seeds, w_fwd)[oind])      # output nonzero, input nonzero
        else:
            seeds = [None] * len(self.nnz_out)
            seeds[oind] = unit_words(n_out, w_adj)
            c, r = set_bits(self.sp_reverse(seeds, w_adj)[iind])      # input nonzero, output nonzero
        return Sparsity.triplet(n_out, n_in, r, c)

    def ckpt_size(self, n, checkpoint_every):
        v = L.I64()
        L.check(L.cxh_ckpt_size(self._h, n, int(checkpoint_every), -1, C.byref(v)))
        return v.value

    def fwd_ckpt(self, args, n, checkpoint_every, stream=None):
        torch = _torch()
        like = self._like(args)
        ld = _check_soa(args, self.nnz_in, n, "arg")
        outs = self._alloc(self.nnz_out, n, like)
        ckpt = torch.empty(max(self.ckpt_size(n, checkpoint_every), 1), dtype=torch.float64, device=like.device)
        rc = L.cxh_map_fwd_ckpt(self._h, n, _ptrs(args), _ptrs(outs), CXH_SOA, ld, int(checkpoint_every),
                                ckpt.data_ptr(), _stream_ptr(stream))
        L.check(rc, "cxh_map_fwd_ckpt")
        self._flagged(rc, n, "cxh_map_fwd_ckpt")
        return outs, ckpt

    def rev_ckpt(self, args, seeds, n, checkpoint_every, ckpt, stream=None):
        like = self._like(args)
        ld = _check_soa(args, self.nnz_in, n, "arg")
        _check_soa(seeds, self.nnz_out, n, "adj_out")
        adj = self._alloc(self.nnz_in, n, like)
        rc = L.cxh_map_rev_ckpt(self._h, n, _ptrs(args), _ptrs(seeds), _ptrs(adj), CXH_SOA, ld,
                                int(checkpoint_every), ckpt.data_ptr(), _stream_ptr(stream))
        L.check(rc, "cxh_map_rev_ckpt")
        self._flagged(rc, n, "cxh_map_rev_ckpt")
        return adj


class HipJacobian:
    """Fused primal + coloured forward seed sweeps + scatter into the nonzeros of
    the Jacobian (XFunction::jac evaluated numerically): of all outputs w.r.t.
    all inputs flattened, or with ``first_block`` of output 0 w.r.t. input 0
    (the other inputs are data).  ``symmetric`` selects star colouring
    (Hessians: the Jacobian of a gradient)."""

    def __init__(self, f, symmetric=False, tape=None, first_block=False, engine="interp"):
        self.f = f
        self.tape = tape if tape is not None else HipTape(f, engine=engine)
        self.plan = f.jac_plan(symmetric, first_block)
        p = self.plan
        if not p["forward"]:
            raise NotImplementedError("adjoint-mode colourings run through HipTape.fwd_adj per direction")
        self.jsp = p["jsp"]
        self.nnz_jac = self.jsp.nnz()
        h = L.P()
        sp, si, sd, sn = (np.ascontiguousarray(p[k], np.int64) for k in ("seed_ptr", "seed_idx", "src_dir", "src_nz"))
        L.check(L.cxh_jac_create(self.tape._h, 1, p["ndir"], sp.ctypes.data_as(L.PI64), si.ctypes.data_as(L.PI64),
                                 self.nnz_jac, sd.ctypes.data_as(L.PI64), sn.ctypes.data_as(L.PI64), C.byref(h)),
                "cxh_jac_create")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and L.cxh_jac_destroy is not None:
            L.cxh_jac_destroy(h)
            self._h = None

    def jit_source(self, segment=0):
        """Generated source of the specialised tangent kernel of one segment
        (all direction groups), with the number of groups and segments."""
        n, ng, ns = L.I64(), L.I32(), L.I32()
        L.check(L.cxh_jac_jit_source(self._h, int(segment), None, 0, C.byref(n), C.byref(ng), C.byref(ns)),
                "cxh_jac_jit_source")
        buf = C.create_string_buffer(n.value + 1)
        L.check(L.cxh_jac_jit_source(self._h, int(segment), buf, n.value + 1, C.byref(n), C.byref(ng),
                                     C.byref(ns)), "cxh_jac_jit_source")
        return buf.value.decode(), ng.value, ns.value

    def __call__(self, args, n=None, res=True, stream=None):
        torch = _torch()
        like = self.tape._like(args)
        n = n if n is not None else like.shape[1]
        ld = _check_soa(args, self.tape.nnz_in, n, "arg")
        outs = self.tape._alloc(self.tape.nnz_out, n, like) if res else [None] * len(self.tape.nnz_out)
        jac = torch.empty((max(self.nnz_jac, 1), ld), dtype=torch.float64, device=like.device)
        L.check(L.cxh_map_jac(self._h, n, _ptrs(args), _ptrs(outs), jac.data_ptr(), CXH_SOA, ld,
                              _stream_ptr(stream)), "cxh_map_jac")
        return outs, jac[: self.nnz_jac, :n]


class HipBatch:
    """Heterogeneous batch: forward + adjoint of several Functions over their
    own instance counts in ONE persistent kernel launch, replayed from a
    captured hipGraph (cxh_batch).  ``tune()`` runs the autotuner over block
    size, persistent grid and tile order."""

    def __init__(self, functions, counts, tapes=None):
        if len(functions) != len(counts):
            raise ValueError("one count per Function")
        self.functions = list(functions)
        self.counts = [int(c) for c in counts]
        self.tapes = tapes if tapes is not None else [HipTape(f) for f in self.functions]
        h = L.P()
        arr = L.ptr_array([t._h.value for t in self.tapes])
        cnt = np.array(self.counts, np.int64)
        L.check(L.cxh_batch_create(len(self.tapes), arr, cnt.ctypes.data_as(L.PI64), C.byref(h)),
                "cxh_batch_create")
        self._h = h
        self._keep = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and L.cxh_batch_destroy is not None:
            L.cxh_batch_destroy(h)
            self._h = None

    def bind(self, fun, args, seeds, res=True, adj=True):
        """Device buffers of Function ``fun``; allocates and returns
        (outputs, adjoints) as lists of (nnz, count) tensors."""
        t = self.tapes[fun]
        n = self.counts[fun]
        like = t._like(args)
        ld = _check_soa(args, t.nnz_in, n, "arg")
        _check_soa(seeds, t.nnz_out, n, "adj_out")
        outs = t._alloc(t.nnz_out, n, like) if res else [None] * len(t.nnz_out)
        adjs = t._alloc(t.nnz_in, n, like) if adj else [None] * len(t.nnz_in)
        for blk in outs + adjs:
            if blk is not None and blk.shape[0] > 1 and blk.stride(0) != ld:
                raise ValueError("outputs must share the inputs' leading dimension")
        L.check(L.cxh_batch_bind(self._h, fun, _ptrs(args), _ptrs(seeds), _ptrs(outs), _ptrs(adjs), CXH_SOA, ld),
                "cxh_batch_bind")
        self._keep[fun] = (args, seeds, outs, adjs)
        return outs, adjs

    def config(self, block=-1, grid_per_cu=-1, order=-1):
        L.check(L.cxh_batch_config(self._h, int(block), int(grid_per_cu), int(order)), "cxh_batch_config")

    def info(self):
        cfg = (L.I32 * 3)()
        nt = L.I64()
        L.check(L.cxh_batch_info(self._h, cfg, C.byref(nt)))
        return dict(block=cfg[0], grid_per_cu=cfg[1], order=cfg[2], n_tiles=nt.value)

    def tune(self, reps=3, stream=None):
        """Autotune the launch shape, then graph replay against direct launch;
        returns (config dict, ms per launch)."""
        best = (L.I32 * 3)()
        ms = L.D()
        L.check(L.cxh_batch_tune(self._h, _stream_ptr(stream), int(reps), best, C.byref(ms)), "cxh_batch_tune")
        return dict(block=best[0], grid_per_cu=best[1], order=best[2],
                    graph=bool(L.cxh_batch_prefers_graph(self._h) == 1)), ms.value

    def launch(self, stream=None, graph=True):
        """graph=True: replay the captured graph; False: launch directly;
        None: whichever the tuner found faster."""
        fn = L.cxh_batch_launch_best if graph is None else L.cxh_batch_launch if graph else L.cxh_batch_launch_direct
        L.check(fn(self._h, _stream_ptr(stream)), "cxh_batch_launch")


class HipMapAccum:
    """Batched mapaccum (cxh_mapaccum): every instance runs the N-step chain of
    ``step`` in one lane, looping over one specialised step body.  Blocks follow
    the mapaccum Function: inputs i < n_accum are (nnz_in(i), n), the others
    and all outputs (N*nnz, n) with step t's nonzeros after step t-1's."""

    def __init__(self, step, N, n_accum=1, tape=None):
        self.step = step
        self.N, self.n_accum = int(N), int(n_accum)
        self.tape = tape if tape is not None else HipTape(step)
        self.nnz_in = [step.nnz_in(i) if i < n_accum else self.N * step.nnz_in(i) for i in range(step.n_in())]
        self.nnz_out = [self.N * step.nnz_out(i) for i in range(step.n_out())]
        h = L.P()
        L.check(L.cxh_mapaccum_create(self.tape._h, self.N, self.n_accum, C.byref(h)), "cxh_mapaccum_create")
        self._h = h

    @classmethod
    def from_function(cls, F):
        """From a Function built by Function.mapaccum."""
        base, N, n_accum = F.macc_base
        return cls(base, N, n_accum)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and L.cxh_mapaccum_destroy is not None:
            L.cxh_mapaccum_destroy(h)
            self._h = None

    def source(self, mode=0, checkpoint_steps=0):
        """Generated kernel source: mode 0 forward chain, 1 fused forward +
        adjoint, 2 forward storing checkpoints, 3 adjoint from them."""
        n = L.I64()
        L.check(L.cxh_mapaccum_source(self._h, int(mode), int(checkpoint_steps), None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        L.check(L.cxh_mapaccum_source(self._h, int(mode), int(checkpoint_steps), buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    def ckpt_size(self, n, checkpoint_steps):
        v = L.I64()
        L.check(L.cxh_mapaccum_ckpt_size(self._h, n, int(checkpoint_steps), C.byref(v)))
        return v.value

    def fwd_ckpt(self, args, n, checkpoint_steps, stream=None):
        """Outputs, and the states every checkpoint_steps steps (for rev_ckpt)."""
        torch = _torch()
        like = next(a for a in args if a is not None)
        ld = _check_soa(args, self.nnz_in, n, "arg")
        outs = self._alloc(self.nnz_out, n, like)
        ckpt = torch.empty(max(self.ckpt_size(n, checkpoint_steps), 1), dtype=torch.float64, device=like.device)
        L.check(L.cxh_mapaccum_fwd_ckpt(self._h, n, _ptrs(args), _ptrs(outs), CXH_SOA, ld, int(checkpoint_steps),
                                        ckpt.data_ptr(), _stream_ptr(stream)), "cxh_mapaccum_fwd_ckpt")
        return outs, ckpt

    def rev_ckpt(self, args, seeds, n, checkpoint_steps, ckpt, stream=None):
        like = next(a for a in args if a is not None)
        ld = _check_soa(args, self.nnz_in, n, "arg")
        _check_soa(seeds, self.nnz_out, n, "adj_out")
        adj = self._alloc(self.nnz_in, n, like)
        L.check(L.cxh_mapaccum_rev_ckpt(self._h, n, _ptrs(args), _ptrs(seeds), _ptrs(adj), CXH_SOA, ld,
                                        int(checkpoint_steps), ckpt.data_ptr(), _stream_ptr(stream)),
                "cxh_mapaccum_rev_ckpt")
        return adj

    def _alloc(self, nnz, n, like):
        torch = _torch()
        return [torch.empty((k, n), dtype=torch.float64, device=like.device) if k else None for k in nnz]

    def eval(self, args, n=None, stream=None):
        like = next(a for a in args if a is not None)
        n = n if n is not None else like.shape[1]
        ld = _check_soa(args, self.nnz_in, n, "arg")
        out = self._alloc(self.nnz_out, n, like)
        L.check(L.cxh_mapaccum_eval(self._h, n, _ptrs(args), _ptrs(out), CXH_SOA, ld, _stream_ptr(stream)),
                "cxh_mapaccum_eval")
        return out

    def fwd_adj(self, args, seeds, n=None, checkpoint_steps=0, res=True, stream=None):
        like = next(a for a in args if a is not None)
        n = n if n is not None else like.shape[1]
        ld = _check_soa(args, self.nnz_in, n, "arg")
        _check_soa(seeds, self.nnz_out, n, "adj_out")
        outs = self._alloc(self.nnz_out, n, like) if res else [None] * len(self.nnz_out)
        adj = self._alloc(self.nnz_in, n, like)
        L.check(L.cxh_mapaccum_fwd_adj(self._h, n, _ptrs(args), _ptrs(seeds), _ptrs(outs), _ptrs(adj), CXH_SOA, ld,
                                       int(checkpoint_steps), _stream_ptr(stream)), "cxh_mapaccum_fwd_adj")
        return outs, adj


def batch_source(tapes):
    """Generated source of the batch kernel over these HipTapes."""
    arr = L.ptr_array([t._h.value for t in tapes])
    n = L.I64()
    L.check(L.cxh_batch_source(len(tapes), arr, None, 0, C.byref(n)), "cxh_batch_source")
    buf = C.create_string_buffer(n.value + 1)
    L.check(L.cxh_batch_source(len(tapes), arr, buf, n.value + 1, C.byref(n)), "cxh_batch_source")
    return buf.value.decode()


def jit_compile_check(src, arch="gfx950"):
    """Compile generated source through hiprtc without a GPU (raises on error)."""
    L.check(L.cxh_jit_compile_check(src.encode(), arch.encode()), "cxh_jit_compile_check")


def device_info():
    ncu, lds = L.I32(), L.I32()
    name = C.create_string_buffer(64)
    L.check(L.cxh_device_info(C.byref(ncu), C.byref(lds), name, 64))
    return dict(n_cu=ncu.value, lds_per_cu_kb=lds.value, arch=name.value.decode())
