"""Map: one Function evaluated over n instances (casadi/core/map.cpp).

``serial`` / ``openmp`` / ``thread`` run the host VM (Map::eval_gen,
casadi/core/map.cpp:141-157 and OmpMap::eval, casadi/core/map.cpp:340-386);
``hip`` runs the gfx950 kernels: by default the straight-line kernels
specialised from the tape (opts {"engine": "jit"}), or the tape interpreter
(opts {"engine": "interp"}), which needs no compilation.  Inputs follow Map's signature:
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
