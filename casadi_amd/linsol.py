"""Linsol: CasADi's linear-solver handle with the "ldl" and "qr" plugins,
batched over instances on the GPU.

``Linsol(name, solver, sp)`` (casadi/core/linsol.cpp:35-39) computes the
symbolic factorisation once -- ``Sparsity::ldl`` with AMD for "ldl"
(casadi/solvers/linsol_ldl.cpp:67-100), ``Sparsity::qr_sparse`` with AMD for
"qr" (casadi/solvers/linsol_qr.cpp:67-83) -- then ``nfact`` / ``solve`` run the
batched kernels (LinsolLdl::nfact/solve, casadi/solvers/linsol_ldl.cpp:119-132;
LinsolQr::nfact/solve, casadi/solvers/linsol_qr.cpp:126-178).

The numeric state follows CasADi's contract: ``nfact`` must precede ``solve``
(casadi/core/linsol.cpp:211-218).  Factor and solve run fused in one launch
per solve, on the values handed to the last ``nfact``.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .sx import Sparsity

PLUGINS = ("ldl", "qr")


class Linsol:
    def __init__(self, name, solver, sp, opts=None):
        if solver not in PLUGINS:
            raise ValueError(f"Plugin '{solver}' is not found (available: {PLUGINS})")
        opts = dict(opts or {})
        self._name = name
        self.solver = solver
        self.sp = sp
        self.n = sp.size1()
        if sp.size1() != sp.size2():
            raise ValueError("Linsol requires a square matrix")
        self._h = None
        self._A = None
        if solver == "ldl":
            if not sp.is_symmetric():
                raise ValueError("LDL factorization requires a symmetric matrix")
            self.sp_lt, self.p = sp.ldl(amd=opts.get("amd", True))
            spa, splt, p = (np.ascontiguousarray(v, np.int64) for v in (sp.compressed(), self.sp_lt.compressed(), self.p))
            h = L.P()
            L.check(L.cxh_ldl_create(spa.ctypes.data_as(L.PI64), splt.ctypes.data_as(L.PI64),
                                     p.ctypes.data_as(L.PI64), C.byref(h)), "cxh_ldl_create")
            engine = opts.get("engine", "jit")
            if engine not in ("jit", "interp"):
                raise ValueError(f"unknown engine '{engine}' (jit, interp)")
            L.check(L.cxh_ldl_engine(h, 1 if engine == "jit" else 0), "cxh_ldl_engine")
        else:
            self.eps = float(opts.get("eps", 1e-12))
            self.sp_v, self.sp_r, self.prinv, self.pc = sp.qr_sparse(amd=opts.get("amd", True))
            arrs = [np.ascontiguousarray(v, np.int64) for v in (sp.compressed(), self.sp_v.compressed(),
                                                                 self.sp_r.compressed(), self.prinv, self.pc)]
            h = L.P()
            L.check(L.cxh_qr_create(*[a.ctypes.data_as(L.PI64) for a in arrs], self.eps, C.byref(h)),
                    "cxh_qr_create")
            engine = opts.get("engine", "jit")
            if engine not in ("jit", "interp"):
                raise ValueError(f"unknown engine '{engine}' (jit, interp)")
            L.check(L.cxh_qr_engine(h, 1 if engine == "jit" else 0), "cxh_qr_engine")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is None:
            return
        fn = L.cxh_ldl_destroy if self.solver == "ldl" else L.cxh_qr_destroy
        if fn is not None:
            fn(h)
        self._h = None

    def jit_source(self, nrhs=1, tr=False, segment=0):
        """Generated source of the specialised factor + solve kernel (LDL: one
        kernel; QR: segment `segment`, with the number of segments)."""
        n = L.I64()
        if self.solver == "ldl":
            L.check(L.cxh_ldl_jit_source(self._h, int(nrhs), None, 0, C.byref(n)), "cxh_ldl_jit_source")
            buf = C.create_string_buffer(n.value + 1)
            L.check(L.cxh_ldl_jit_source(self._h, int(nrhs), buf, n.value + 1, C.byref(n)), "cxh_ldl_jit_source")
            return buf.value.decode()
        ns = L.I32()
        L.check(L.cxh_qr_jit_source(self._h, int(nrhs), int(tr), int(segment), None, 0, C.byref(n), C.byref(ns)),
                "cxh_qr_jit_source")
        buf = C.create_string_buffer(n.value + 1)
        L.check(L.cxh_qr_jit_source(self._h, int(nrhs), int(tr), int(segment), buf, n.value + 1, C.byref(n),
                                    C.byref(ns)), "cxh_qr_jit_source")
        return buf.value.decode(), ns.value

    def name(self):
        return self._name

    def plugin_name(self):
        return self.solver

    # -- CasADi's sfact / nfact / solve contract -------------------------------
    def sfact(self, A=None):
        """Symbolic factorisation: already done at construction (fixed pattern)."""
        return 0

    def nfact(self, A):
        """Store the numeric values (nnz,) or (nnz, n_inst) for the next solve."""
        self._A = A
        return 0

    def solve(self, A, B=None, tr=False):
        """solve(A, B) or, after nfact(A), solve(B): X = A^-1 B.

        Host arrays: A (nnz,) + B (n,) / (n, nrhs) for one system, or
        A (nnz, n_inst) + B (n, n_inst) for a batch.  CUDA tensors in the same
        [nonzero][instance] layout stay on the device."""
        if B is None:
            if self._A is None:
                raise RuntimeError("Linsol::solve: 'nfact' must be called before 'solve'")
            A, B = self._A, A
        return self.solve_batch(A, B, tr=tr)[0]

    def solve_batch(self, A, B, tr=False, nrhs=1):
        """Factor + solve a batch; returns (X, status) with status[k] the
        number of zero pivots (ldl) or |R_ii| < eps entries (qr) of instance k."""
        import torch
        dev_in = isinstance(A, torch.Tensor) and A.is_cuda
        At = A if dev_in else torch.as_tensor(np.asarray(A, dtype=float), dtype=torch.float64, device="cuda")
        Bt = B if isinstance(B, torch.Tensor) and B.is_cuda else \
            torch.as_tensor(np.asarray(B, dtype=float), dtype=torch.float64, device="cuda")
        single = At.dim() == 1
        if single:
            At = At.reshape(-1, 1)
            Bt = Bt.reshape(self.n, -1).T.reshape(-1, 1) if Bt.dim() == 2 else Bt.reshape(-1, 1)
            nrhs = Bt.shape[0] // self.n
        n_inst = At.shape[1]
        if At.shape[0] != self.sp.nnz():
            raise ValueError(f"A has {At.shape[0]} nonzeros, the pattern has {self.sp.nnz()}")
        if Bt.shape[0] != self.n * nrhs or Bt.shape[1] != n_inst:
            raise ValueError(f"B must be ({self.n * nrhs}, {n_inst})")
        At = At.contiguous()
        X = Bt.clone().contiguous()
        status = torch.zeros(n_inst, dtype=torch.int32, device=At.device)
        stream = torch.cuda.current_stream().cuda_stream
        if self.solver == "ldl":
            L.check(L.cxh_ldl_factor_solve(self._h, n_inst, At.data_ptr(), X.data_ptr(), nrhs,
                                           status.data_ptr(), 0, stream), "cxh_ldl_factor_solve")
        else:
            L.check(L.cxh_qr_factor_solve(self._h, n_inst, At.data_ptr(), X.data_ptr(), nrhs, int(tr),
                                          status.data_ptr(), 0, stream), "cxh_qr_factor_solve")
        if dev_in:
            return X, status
        Xh = X.cpu().numpy()
        st = status.cpu().numpy()
        if single:
            return Xh.reshape(nrhs, self.n).T.reshape(np.shape(B)), st
        return Xh, st

    def rank(self, A):
        if self.solver != "qr":
            raise NotImplementedError("rank is provided by the qr plugin")
        _, st = self.solve_batch(A, np.zeros(self.n))
        return self.n - int(np.asarray(st).reshape(-1)[0])
