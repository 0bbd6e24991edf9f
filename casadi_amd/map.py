"""Map: one Function evaluated over n instances (casadi/core/map.cpp).

``serial`` / ``openmp`` / ``thread`` run the host VM (Map::eval_gen,
casadi/core/map.cpp:141-157 and OmpMap::eval, casadi/core/map.cpp:340-386);
``hip`` runs the gfx950 kernels: by default the straight-line kernels
specialised from the tape (opts {"engine": "jit"}), or the tape interpreter
(opts {"engine": "interp"}), which needs no compilation; with opts
{"jit_async": True} the kernels compile on a worker thread and the calls made
meanwhile run on the interpreter.  Inputs follow Map's signature:
input i is ``repmat(sparsity_in(i), 1, n)`` -- for a column-vector input an
``(nrow, n)`` array -- and a non-repeated input is broadcast to all instances.
"""
import os

import numpy as np

from . import _lib as L


class MapFunction:
    def __init__(self, f, n, parallelization, max_num_threads=None, opts=None):
        self.f = f
        self.n = n
        self.parallelization = parallelization
        self.opts = dict(opts or {})
        if max_num_threads is None:
            max_num_threads = os.cpu_count() or 1
        self.n_threads = 1 if parallelization in ("serial", "unroll", "inline") else int(max_num_threads)
        self._tape = None

    def name(self):
        return f"map{self.n}_{self.f.name()}"

    def n_in(self):
        return self.f.n_in()

    def n_out(self):
        return self.f.n_out()

    @property
    def tape(self):
        if self._tape is None:
            from .hip import HipTape
            self._tape = HipTape(self.f, block=self.opts.get("block", 64),
                                 w_in_lds=self.opts.get("w_in_lds", True),
                                 engine=self.opts.get("engine", "jit"))
            if self.opts.get("regularity_check", False):
                # CasADi's option: NaN / Inf in an output raises; here checked
                # per instance on the GPU (cxh_tape_regularity_check)
                self._tape.regularity_check(self.n)
            if self.opts.get("jit_async", False):
                # calls before the specialised kernels are built run on the
                # interpreter (cxh_tape_jit_async)
                self._tape.jit_async(True)
        return self._tape

    def _aos(self, i, a):
        """AoS nonzeros (instance-major) of map input i."""
        sp = self.f.sparsity_in(i)
        nr, nc = sp.shape
        if a is None:
            return None
        v = np.asarray(a, dtype=float)
        if v.ndim <= 1:
            v = v.reshape(nr, -1) if v.size != 1 else np.full((nr, nc), float(v.reshape(-1)[0]))
        if v.shape == (nr, nc):
            v = np.tile(v, (1, self.n))
        if v.shape != (nr, nc * self.n):
            raise ValueError(f"{self.name()}: input {i} has shape {v.shape}, expected {(nr, nc * self.n)}")
        blocks = v.reshape(nr, self.n, nc).transpose(1, 2, 0)  # [k][col][row]
        rows, cols = sp.row, sp.get_col()
        return np.ascontiguousarray(blocks[:, cols, rows])     # [k][nz]

    def _dense_out(self, i, aos):
        sp = self.f.sparsity_out(i)
        nr, nc = sp.shape
        out = np.zeros((self.n, nc, nr))
        out[:, sp.get_col(), sp.row] = aos.reshape(self.n, -1)
        return out.transpose(2, 0, 1).reshape(nr, self.n * nc)

    def __call__(self, *args):
        res = self.call(list(args))
        return res[0] if len(res) == 1 else tuple(res)

    # -- derivatives: the map of the derivative (Map::get_forward / get_reverse,
    # casadi/core/map.cpp:219-325); a map's forward/reverse Function has the
    # base Function's derivative signature with every argument stacked n times
    def forward(self, nfwd):
        return MapFunction(self.f.forward(nfwd), self.n, self.parallelization, self.n_threads, self.opts)

    def reverse(self, nadj):
        return MapFunction(self.f.reverse(nadj), self.n, self.parallelization, self.n_threads, self.opts)

    def jacobian_blocks(self):
        """Per-instance Jacobian of the base Function, mapped (the block-diagonal
        nonzeros of the map's Jacobian)."""
        return MapFunction(self.f.jacobian(), self.n, self.parallelization, self.n_threads, self.opts)

    def call(self, args):
        if len(args) != self.f.n_in():
            raise ValueError(f"{self.name()}: expected {self.f.n_in()} inputs")
        try:
            import torch
            if any(isinstance(a, torch.Tensor) and a.is_cuda for a in args):
                if self.parallelization != "hip":
                    raise TypeError("device tensors require parallelization 'hip'")
                return self.tape.eval(list(args), self.n)
        except ImportError:
            pass
        aos = [self._aos(i, a) for i, a in enumerate(args)]
        if self.parallelization == "hip":
            res = self.tape.eval_host([a.reshape(-1) if a is not None else None for a in aos], self.n)
        else:
            res = [np.zeros(self.n * self.f.nnz_out(i)) for i in range(self.f.n_out())]
            pa = L.ptr_array([a.ctypes.data if a is not None and a.size else None for a in aos])
            pr = L.ptr_array([r.ctypes.data if r.size else None for r in res])
            L.check(L.cxs_function_map_eval(self.f._h, self.n, pa, pr, self.n_threads), self.name())
        return [self._dense_out(i, r) for i, r in enumerate(res)]


class MapSumFunction:
    """A map whose reduced inputs are shared by all n instances and whose
    reduced outputs are summed over them.

    ``Function.map(n, reduce_in, reduce_out)`` is CasADi's serial MapSum
    (MapSum::eval_gen, casadi/core/mapsum.cpp:155-186);
    ``Function.map(name, parallelization, n, reduce_in, reduce_out)`` is the
    reduced form of any map (casadi/core/function.cpp:796-817: repmat of the
    reduced inputs, repsum of the reduced outputs).  With ``hip`` the shared
    inputs are read in place by every instance and the sums run on the GPU in a
    fixed order (``cxh_map_eval_sum``); the host parallelizations add instance
    by instance, as MapSum does.  A reduced input takes the base Function's
    input (its nonzeros as a vector on the device), a reduced output is the base
    Function's output."""

    def __init__(self, f, n, parallelization, reduce_in, reduce_out, name=None, max_num_threads=None,
                 opts=None):
        self.f = f
        self.n = int(n)
        self.parallelization = parallelization
        self.reduce_in = [bool(x) for x in reduce_in]
        self.reduce_out = [bool(x) for x in reduce_out]
        if len(self.reduce_in) != f.n_in() or len(self.reduce_out) != f.n_out():
            raise ValueError("Dimension mismatch")  # casadi/core/mapsum.cpp:38-39
        self._name = name or f"mapsum_{self.n}_{f.name()}"
        self.map = MapFunction(f, self.n, parallelization, max_num_threads, opts)

    def name(self):
        return self._name

    def n_in(self):
        return self.f.n_in()

    def n_out(self):
        return self.f.n_out()

    @property
    def tape(self):
        return self.map.tape

    def __call__(self, *args):
        res = self.call(list(args))
        return res[0] if len(res) == 1 else tuple(res)

    def _device(self, args):
        try:
            import torch
        except ImportError:
            return False
        return any(isinstance(a, torch.Tensor) and a.is_cuda for a in args)

    def _nz(self, sp, a):
        """Nonzeros of one (shared) block given densely, as a vector."""
        v = np.asarray(a, dtype=float)
        nr, nc = sp.shape
        if v.size == 1 and sp.nnz() != 1:
            v = np.full((nr, nc), float(v.reshape(-1)[0]))
        v = v.reshape(nr, nc) if v.size == nr * nc else v
        if v.shape != (nr, nc):
            raise ValueError(f"{self.name()}: shared input has shape {v.shape}, expected {(nr, nc)}")
        return np.ascontiguousarray(v[sp.row, sp.get_col()])

    def _dense(self, sp, nz):
        nr, nc = sp.shape
        out = np.zeros((nr, nc))
        out[sp.row, sp.get_col()] = nz
        return out

    def call(self, args):
        if len(args) != self.f.n_in():
            raise ValueError(f"{self.name()}: expected {self.f.n_in()} inputs")
        if self._device(args):
            if self.parallelization != "hip":
                raise TypeError("device tensors require parallelization 'hip'")
            return self.tape.eval_sum(list(args), self.n, self.reduce_in, self.reduce_out)
        if self.parallelization == "hip":
            import torch
            dev = [None if a is None else
                   torch.tensor(self._nz(self.f.sparsity_in(i), a), device="cuda") if self.reduce_in[i] else
                   torch.tensor(self.map._aos(i, a).T.copy(), device="cuda")
                   for i, a in enumerate(args)]
            out = self.tape.eval_sum(dev, self.n, self.reduce_in, self.reduce_out)
            return [self._dense(self.f.sparsity_out(j), o.cpu().numpy()) if self.reduce_out[j] else
                    self.map._dense_out(j, o.cpu().numpy().T.reshape(-1))
                    for j, o in enumerate(out)]
        # host: every instance reads the shared blocks; sums in instance order
        full = [None if a is None else
                np.tile(self._dense(self.f.sparsity_in(i), self._nz(self.f.sparsity_in(i), a)), (1, self.n))
                if self.reduce_in[i] else a for i, a in enumerate(args)]
        res = self.map.call(full)
        out = []
        for j, r in enumerate(res):
            if not self.reduce_out[j]:
                out.append(r)
                continue
            nr, nc = self.f.sparsity_out(j).shape
            blocks = r.reshape(nr, self.n, nc)
            acc = np.zeros((nr, nc))
            for k in range(self.n):
                acc = acc + blocks[:, k, :]
            out.append(acc)
        return out

    def to_torch(self, checkpoint_every=0):
        """The reduced map as a differentiable torch op on device tensors
        (casadi_amd.autograd.TorchMapSum): gradients of shared inputs come
        back summed over the instances."""
        from .autograd import TorchMapSum
        return TorchMapSum(self, checkpoint_every)

    def fwd_adj(self, args, seeds, checkpoint_every=0, res=True):
        """Device tensors: outputs and the reverse sweep of the reduced map in
        one pass (the MapSum's reverse, casadi/core/mapsum.cpp:354-415): the
        seed of a reduced output applies to every instance, the sensitivity of
        a reduced input is summed over the instances."""
        if self.parallelization != "hip":
            raise TypeError("fwd_adj runs on the GPU: parallelization 'hip'")
        return self.tape.fwd_adj_sum(list(args), list(seeds), self.n, self.reduce_in, self.reduce_out,
                                     checkpoint_every=checkpoint_every, res=res)
