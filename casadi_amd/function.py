"""Function and Map: CasADi's callable-graph API, Python side.

``Function(name, [x, u], [expr], ...)`` compiles the SX graph into a tape
(SXFunction::init, casadi/core/sx_function.cpp:476-841, restated natively in
casadi_amd/csrc/sx_function.cpp).  ``f.map(n, parallelization)`` batches it
(Map::create, casadi/core/map.cpp:39-51) with ``"serial"``, ``"openmp"`` or
the new ``"hip"`` parallelization, which runs on the GPU through the C ABI.
"""
import ctypes as C
import numbers

import numpy as np

from . import _lib as L
from .sx import SX, Sparsity

INSTR_DTYPE = np.dtype([("op", "<i4"), ("i0", "<i4"), ("i1", "<i4"), ("i2", "<i4")])

PARALLELIZATIONS = ("serial", "openmp", "hip", "unroll", "inline", "thread")


class Function:
    """An SX Function (casadi::Function backed by an SXFunction)."""

    def __init__(self, name, ex_in=None, ex_out=None, name_in=None, name_out=None, opts=None,
                 _handle=None):
        if _handle is not None:
            self._h = _handle
        else:
            opts = dict(opts or {})
            ex_in = [x if isinstance(x, SX) else SX.from_value(x) for x in ex_in]
            ex_out = [x if isinstance(x, SX) else SX.from_value(x) for x in ex_out]
            sps_in = [x.sparsity.compressed() for x in ex_in]
            sps_out = [x.sparsity.compressed() for x in ex_out]
            nzs_in = [x.nz for x in ex_in]
            nzs_out = [x.nz for x in ex_out]
            keep = [sps_in, sps_out, nzs_in, nzs_out]
            p_sp_in = L.ptr_array([a.ctypes.data for a in sps_in])
            p_sp_out = L.ptr_array([a.ctypes.data for a in sps_out])
            p_nz_in = L.ptr_array([a.ctypes.data if a.size else None for a in nzs_in])
            p_nz_out = L.ptr_array([a.ctypes.data if a.size else None for a in nzs_out])
            names_in = [n.encode() for n in (name_in or [f"i{i}" for i in range(len(ex_in))])]
            names_out = [n.encode() for n in (name_out or [f"o{i}" for i in range(len(ex_out))])]
            c_in = (C.c_char_p * max(len(names_in), 1))(*names_in)
            c_out = (C.c_char_p * max(len(names_out), 1))(*names_out)
            h = L.P()
            L.check(L.cxs_function_new(name.encode(), len(ex_in), p_sp_in, p_nz_in,
                                       C.cast(c_in, L.PP), len(ex_out), p_sp_out, p_nz_out,
                                       C.cast(c_out, L.PP), int(bool(opts.get("allow_free", False))),
                                       C.byref(h)), f"Function {name}")
            del keep
            self._h = h
        self._cache = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and L is not None and L.cxs_function_free is not None:
            L.cxs_function_free(h)
            self._h = None

    # -- introspection -------------------------------------------------------
    def _info(self):
        if "info" not in self._cache:
            ni, no, nin, ws, nf = L.I32(), L.I32(), L.I64(), L.I64(), L.I64()
            L.check(L.cxs_function_info(self._h, C.byref(ni), C.byref(no), C.byref(nin), C.byref(ws),
                                        C.byref(nf)))
            self._cache["info"] = (ni.value, no.value, nin.value, ws.value, nf.value)
        return self._cache["info"]

    def name(self):
        return L.cxs_function_name(self._h).decode()

    def n_in(self):
        return self._info()[0]

    def n_out(self):
        return self._info()[1]

    def n_instructions(self):
        return self._info()[2]

    def sz_w(self):
        return self._info()[3]

    def has_free(self):
        return self._info()[4] > 0

    def name_in(self, i=None):
        if i is None:
            return [self.name_in(k) for k in range(self.n_in())]
        return L.cxs_function_name_in(self._h, i).decode()

    def name_out(self, i=None):
        if i is None:
            return [self.name_out(k) for k in range(self.n_out())]
        return L.cxs_function_name_out(self._h, i).decode()

    def index_in(self, name):
        return self.name_in().index(name)

    def index_out(self, name):
        return self.name_out().index(name)

    def sparsity_in(self, i):
        i = self.index_in(i) if isinstance(i, str) else i
        key = ("spi", i)
        if key not in self._cache:
            h = L.P()
            L.check(L.cxs_function_sparsity_in(self._h, i, C.byref(h)))
            self._cache[key] = Sparsity.from_compressed(L.take_buf(h))
        return self._cache[key]

    def sparsity_out(self, i):
        i = self.index_out(i) if isinstance(i, str) else i
        key = ("spo", i)
        if key not in self._cache:
            h = L.P()
            L.check(L.cxs_function_sparsity_out(self._h, i, C.byref(h)))
            self._cache[key] = Sparsity.from_compressed(L.take_buf(h))
        return self._cache[key]

    def nnz_in(self, i=None):
        if i is None:
            return sum(self.nnz_in(k) for k in range(self.n_in()))
        return self.sparsity_in(i).nnz()

    def nnz_out(self, i=None):
        if i is None:
            return sum(self.nnz_out(k) for k in range(self.n_out()))
        return self.sparsity_out(i).nnz()

    def size_in(self, i):
        return self.sparsity_in(i).shape

    def size_out(self, i):
        return self.sparsity_out(i).shape

    def sx_out(self, i):
        h = L.P()
        L.check(L.cxs_function_nz_out(self._h, i, C.byref(h)))
        return SX(self.sparsity_out(i), L.take_buf(h, "uint64"))

    def sx_in(self, i):
        h = L.P()
        L.check(L.cxs_function_nz_in(self._h, i, C.byref(h)))
        return SX(self.sparsity_in(i), L.take_buf(h, "uint64"))

    def instructions(self):
        """The compiled tape as a structured array (op, i0, i1, i2); constants of
        OP_CONST records are the float64 overlaid on (i1, i2)."""
        if "tape" not in self._cache:
            a = np.zeros(self.n_instructions(), INSTR_DTYPE)
            if a.size:
                L.check(L.cxs_function_tape(self._h, a.ctypes.data))
            self._cache["tape"] = a
        return self._cache["tape"]

    def instruction_constants(self):
        a = self.instructions()
        return a.view(np.uint8).reshape(-1, 16)[:, 8:].copy().view("<f8").reshape(-1)

    # -- evaluation -----------------------------------------------------------
    def _npar(self, args):
        """Number of parallel evaluations implied by the argument shapes
        (FunctionInternal::check_mat, casadi/core/function_internal.cpp:3649-3670):
        an argument with n_rows == size1_in(i) and P*size2_in(i) columns asks
        for P evaluations; all such arguments must agree."""
        npar = 1
        for i, a in enumerate(args):
            if a is None:
                continue
            v = np.asarray(a, dtype=float)
            nr, nc = self.sparsity_in(i).shape
            if v.ndim != 2 or v.size <= 1 or v.shape == (nr, nc):
                continue
            if v.shape[0] == nr and v.shape[1] > nc and nc > 0 and v.shape[1] % (npar * nc) == 0:
                npar *= v.shape[1] // (npar * nc)
        return npar

    def _numeric_args(self, args):
        out = []
        for i, a in enumerate(args):
            sp = self.sparsity_in(i)
            if a is None:
                out.append(None)
                continue
            v = np.asarray(a, dtype=float)
            nr, nc = sp.shape
            if v.ndim == 0 or v.size == 1:
                dense = np.full(sp.shape, float(v.reshape(-1)[0]) if v.size else 0.0)
            elif v.size == 0:
                dense = np.zeros(sp.shape)
            else:
                if v.ndim == 1:
                    v = v.reshape(-1, 1) if nc == 1 else v.reshape(1, -1)
                if v.shape != sp.shape and min(v.shape) == 1 and v.shape[::-1] == sp.shape:
                    v = v.T                          # a transposed vector
                if v.shape != sp.shape and v.shape[0] == nr and 0 < v.shape[1] < nc and nc % v.shape[1] == 0:
                    v = np.tile(v, (1, nc // v.shape[1]))   # repeated horizontally
                if v.shape != sp.shape:
                    raise ValueError(f"{self.name()}: input {i} has shape {v.shape}, expected {sp.shape}")
                dense = v
            out.append(np.ascontiguousarray(dense[sp.row, sp.get_col()]))
        return out

    def _dense_out(self, i, nz):
        sp = self.sparsity_out(i)
        d = np.zeros(sp.shape)
        d[sp.row, sp.get_col()] = nz
        return d

    def __call__(self, *args, **kwargs):
        if kwargs:
            if args:
                raise ValueError("mix of positional and keyword arguments")
            res = self.call({k: v for k, v in kwargs.items()})
            return res
        res = self.call(list(args))
        return res[0] if len(res) == 1 else tuple(res)

    def call(self, args):
        named = isinstance(args, dict)
        if named:
            lst = [None] * self.n_in()
            for k, v in args.items():
                lst[self.index_in(k)] = v
            args = lst
        if len(args) != self.n_in():
            raise ValueError(f"{self.name()}: expected {self.n_in()} inputs, got {len(args)}")
        symbolic = any(isinstance(a, SX) for a in args)
        res = self._call_sx(args) if symbolic else self._call_num(args)
        if named:
            return {self.name_out(i): r for i, r in enumerate(res)}
        return res

    def _call_num(self, args):
        npar = self._npar(args)
        if npar > 1:   # evaluation with multiple arguments: a serial map
            from .map import MapFunction
            return MapFunction(self, npar, "serial").call(list(args))
        arg = self._numeric_args(args)
        res = [np.zeros(self.nnz_out(i)) for i in range(self.n_out())]
        pa = L.ptr_array([a.ctypes.data if a is not None and a.size else None for a in arg])
        pr = L.ptr_array([r.ctypes.data if r.size else None for r in res])
        L.check(L.cxs_function_eval(self._h, pa, pr), self.name())
        return [self._dense_out(i, r) for i, r in enumerate(res)]

    def _call_sx(self, args):
        arrs = []
        for i, a in enumerate(args):
            sp = self.sparsity_in(i)
            if a is None:
                x = SX.zeros(sp)
            else:
                x = a if isinstance(a, SX) else SX.from_value(a)
                if x.is_scalar() and sp.shape != (1, 1):
                    x = SX.from_dense_nodes(np.full(sp.shape, int(x.nz[0]), np.uint64))
                if x.shape != sp.shape:
                    if x.numel() == sp.numel():
                        x = SX.from_dense_nodes(x.dense_nodes().T.reshape(sp.shape[::-1]).T)
                    else:
                        raise ValueError(f"{self.name()}: input {i} has shape {x.shape}, expected {sp.shape}")
                d = x.dense_nodes()
                x = SX(sp, d[sp.row, sp.get_col()])
            arrs.append(np.ascontiguousarray(x.nz))
        res = [np.zeros(self.nnz_out(i), np.uint64) for i in range(self.n_out())]
        pa = L.ptr_array([a.ctypes.data if a.size else None for a in arrs])
        pr = L.ptr_array([r.ctypes.data if r.size else None for r in res])
        L.check(L.cxs_function_call(self._h, pa, pr), self.name())
        return [SX(self.sparsity_out(i), r) for i, r in enumerate(res)]

    # -- derivatives (casadi/core/function_internal.cpp:2220-2440) -------------
    def _derived(self, fn, *extra):
        h = L.P()
        L.check(fn(self._h, *extra, C.byref(h)), self.name())
        return Function(None, _handle=h)

    def jacobian(self):
        if "jac" not in self._cache:
            self._cache["jac"] = self._derived(L.cxs_function_jacobian)
        return self._cache["jac"]

    def forward(self, nfwd):
        key = ("fwd", nfwd)
        if key not in self._cache:
            self._cache[key] = self._derived(L.cxs_function_forward, nfwd)
        return self._cache[key]

    def reverse(self, nadj):
        key = ("adj", nadj)
        if key not in self._cache:
            self._cache[key] = self._derived(L.cxs_function_reverse, nadj)
        return self._cache[key]

    def jac_sparsity(self, oind=0, iind=0, compact=True, symmetric=False):
        h = L.P()
        L.check(L.cxs_function_jac_sparsity(self._h, oind, iind, int(compact), int(symmetric), C.byref(h)))
        return Sparsity.from_compressed(L.take_buf(h))

    def jac_plan(self, symmetric=False, first_block=False):
        """Colouring + scatter plan of the compressed Jacobian (XFunction::jac
        bookkeeping) of all outputs w.r.t. all inputs flattened, or with
        ``first_block`` of output 0 w.r.t. input 0 only: dict with jsp,
        forward, seeds (CSR), src_dir, src_nz."""
        key = ("plan", bool(symmetric), bool(first_block))
        if key not in self._cache:
            fwd = L.I32()
            hs = [L.P() for _ in range(5)]
            L.check(L.cxs_function_jac_plan(self._h, int(symmetric), int(first_block), C.byref(fwd),
                                            *[C.byref(h) for h in hs]))
            jsp, sptr, sidx, sdir, snz = [L.take_buf(h) for h in hs]
            self._cache[key] = dict(jsp=Sparsity.from_compressed(jsp), forward=bool(fwd.value),
                                    seed_ptr=sptr, seed_idx=sidx, src_dir=sdir, src_nz=snz,
                                    ndir=len(sptr) - 1)
        return self._cache[key]

    # -- chains -----------------------------------------------------------------
    def mapaccum(self, name, N=None, n_accum=1):
        """Function::mapaccum (casadi/core/function.cpp:669-754): N chained
        calls, the first n_accum outputs fed back as the first n_accum inputs;
        the other inputs and all outputs stacked horizontally.  Built here as
        one SX Function (the chain unrolled); ``macc_base`` keeps the step
        Function so that batched GPU evaluation can loop over one compiled step
        (casadi_amd.hip.HipMapAccum).  ``mapaccum(N)`` names it mapaccum_<f>."""
        if N is None:
            name, N = "mapaccum_" + self.name(), name
        N = int(N)
        if N <= 0:
            raise ValueError("mapaccum: N must be positive")
        if n_accum > min(self.n_in(), self.n_out()):
            raise ValueError("mapaccum: too many accumulators")
        for i in range(n_accum):
            if self.sparsity_in(i).shape != self.sparsity_out(i).shape:
                raise ValueError(f"mapaccum: input {i} and output {i} differ in shape")
        ins = [SX.sym(self.name_in(i), *self.sparsity_in(i).shape) if i < n_accum else
               SX.sym(self.name_in(i), self.sparsity_in(i).size1(), N * self.sparsity_in(i).size2())
               for i in range(self.n_in())]
        state = ins[:n_accum]
        outs = [[] for _ in range(self.n_out())]
        for k in range(N):
            args = list(state)
            for i in range(n_accum, self.n_in()):
                c = self.sparsity_in(i).size2()
                args.append(ins[i][:, k * c:(k + 1) * c])
            res = self.call(args)
            for o in range(self.n_out()):
                outs[o].append(res[o])
            state = res[:n_accum]
        from .sx import horzcat
        F = Function(name, ins, [horzcat(*r) for r in outs], self.name_in(), self.name_out())
        F.macc_base = (self, N, n_accum)
        return F

    def fold(self, N):
        """Function::fold (casadi/core/function.cpp:661-667): mapaccum with only
        the last step's block of output 0."""
        base = self.mapaccum("mapaccum_" + self.name(), N)
        ins = [SX.sym(base.name_in(i), *base.sparsity_in(i).shape) for i in range(base.n_in())]
        out = base.call(ins)
        c = self.sparsity_out(0).size2()
        out[0] = out[0][:, (N - 1) * c:N * c]
        return Function("fold_" + self.name(), ins, out, base.name_in(), base.name_out())

    # -- batching ---------------------------------------------------------------
    def to_torch(self, checkpoint_every=0, engine="jit"):
        """Batched GPU evaluation as a differentiable torch op whose backward
        is the checkpointed reverse sweep (casadi_amd.autograd.TorchFunction).
        checkpoint_every counts tape instructions, or steps for a Function
        built by mapaccum (which then runs as loop kernels)."""
        from .autograd import TorchFunction
        if getattr(self, "macc_base", None) is not None and engine == "jit":
            from .hip import HipMapAccum
            return TorchFunction(self, checkpoint_every, engine, tape=HipMapAccum.from_function(self))
        return TorchFunction(self, checkpoint_every, engine)

    def map(self, *args, **kwargs):
        """Function::map overloads (casadi/core/function.cpp:789-896):

        * ``map(n, parallelization="serial", max_num_threads=None, opts=None)``:
          Map::create (casadi/core/map.cpp:39-51);
        * ``map(n, reduce_in, reduce_out, opts=None)`` with lists of bools: the
          serial MapSum (casadi/core/function.cpp:789-794);
        * ``map(name, parallelization, n, reduce_in, reduce_out, opts=None)``
          with lists of input / output indices or names: a map whose listed
          inputs are shared and listed outputs summed
          (casadi/core/function.cpp:796-826).
        """
        from .map import MapFunction, MapSumFunction
        if args and isinstance(args[0], str):
            name, parallelization, n = args[0], args[1], int(args[2])
            rin = list(args[3]) if len(args) > 3 else list(kwargs.pop("reduce_in", []))
            rout = list(args[4]) if len(args) > 4 else list(kwargs.pop("reduce_out", []))
            opts = args[5] if len(args) > 5 else kwargs.pop("opts", None)
            self._check_par(parallelization)
            ri = [False] * self.n_in()
            for i in rin:
                ri[self.index_in(i) if isinstance(i, str) else int(i)] = True
            ro = [False] * self.n_out()
            for j in rout:
                ro[self.index_out(j) if isinstance(j, str) else int(j)] = True
            return MapSumFunction(self, n, parallelization, ri, ro, name=name, opts=opts or {})
        if len(args) >= 3 and isinstance(args[1], (list, tuple)):
            opts = args[3] if len(args) > 3 else kwargs.pop("opts", None)
            return MapSumFunction(self, int(args[0]), "serial", args[1], args[2], opts=opts or {})
        n = int(args[0])
        parallelization = args[1] if len(args) > 1 else kwargs.pop("parallelization", "serial")
        max_num_threads = args[2] if len(args) > 2 else kwargs.pop("max_num_threads", None)
        opts = args[3] if len(args) > 3 else kwargs.pop("opts", None)
        self._check_par(parallelization)
        return MapFunction(self, n, parallelization, max_num_threads, opts or {})

    @staticmethod
    def _check_par(parallelization):
        if parallelization not in PARALLELIZATIONS:
            raise ValueError(f"Unknown parallelization '{parallelization}' for map: "
                             f"choose one of {PARALLELIZATIONS}")

    def __repr__(self):
        ins = ", ".join(f"{self.name_in(i)}[{self.size_in(i)[0]}x{self.size_in(i)[1]}]" for i in range(self.n_in()))
        outs = ", ".join(f"{self.name_out(i)}[{self.size_out(i)[0]}x{self.size_out(i)[1]}]" for i in range(self.n_out()))
        return f"{self.name()}:({ins})->({outs}) SXFunction"
