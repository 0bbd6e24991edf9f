"""Linsol: CasADi's linear-solver handle with the "ldl" and "qr" plugins,
batched over instances on the GPU.

``Linsol(name, solver, sp)`` (casadi/core/linsol.cpp:35-39) computes the
symbolic factorisation once -- ``Sparsity::ldl`` with AMD for "ldl", or with
the "incomplete" option the pattern of the permuted matrix itself (no fill-in;
casadi/solvers/linsol_ldl.cpp:67-100), ``Sparsity::qr_sparse`` with AMD for
"qr" (casadi/solvers/linsol_qr.cpp:67-83).

The numeric state follows CasADi's contract (casadi/core/linsol.cpp:104-218):
``nfact`` factors at once and the factors stay on the device
(``cxh_ldl_nfact`` / ``cxh_qr_nfact``), so later changes to A do not reach
them; ``solve`` needs a preceding ``nfact`` and may be called any number of
times on its factors (``cxh_ldl_solve`` / ``cxh_qr_solve``).  ``neig`` /
``rank`` (ldl) and ``det`` (qr) read the factors of the last ``nfact``.
``solve_batch`` runs factor and solve fused in one launch and keeps nothing.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .sx import Sparsity

PLUGINS = ("ldl", "qr")


def _incomplete_ldl_pattern(sp, amd):
    """LinsolLdl's symbolic step with "incomplete": p = AMD (or the identity)
    and the strictly upper triangle of A(p, p), no fill-in
    (casadi/solvers/linsol_ldl.cpp:85-95)."""
    n = sp.size1()
    p = np.asarray(sp.amd(), np.int64) if amd else np.arange(n, dtype=np.int64)
    pinv = np.empty(n, np.int64)
    pinv[p] = np.arange(n)
    rows, cols = np.asarray(sp.row, np.int64), np.asarray(sp.get_col(), np.int64)
    pr, pc = pinv[rows], pinv[cols]
    keep = pr < pc
    return Sparsity.triplet(n, n, pr[keep], pc[keep]), p


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
        self._n_fact = None      # instances factored by the last nfact
        self._status = None
        if solver == "ldl":
            if not sp.is_symmetric():
                raise ValueError("LDL factorization requires a symmetric matrix")
            if opts.get("incomplete", False):
                self.sp_lt, self.p = _incomplete_ldl_pattern(sp, opts.get("amd", True))
            else:
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
            L.check(L.cxh_qr_cache(h, int(opts.get("cache", 0))), "cxh_qr_cache")
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
        """Symbolic factorisation: done at construction (the pattern is fixed);
        a new sfact invalidates the numeric factors (casadi/core/linsol.cpp:112-127)."""
        self._n_fact = None
        return 0

    @staticmethod
    def _device(a):
        import torch
        if isinstance(a, torch.Tensor) and a.is_cuda:
            return a
        return torch.as_tensor(np.asarray(a, dtype=float), dtype=torch.float64, device="cuda")

    def nfact(self, A):
        """Numeric factorisation of one system (A: (nnz,)) or a batch (A:
        (nnz, n_inst), SoA), factored now: the factors stay on the device
        until the next nfact.  For "qr" a singular instance (|R_ii| < eps)
        makes it fail, as LinsolQr::nfact does; "ldl" only counts zero pivots
        (status())."""
        import torch
        At = self._device(A)
        single = At.dim() == 1
        At = (At.reshape(-1, 1) if single else At).contiguous()
        if At.shape[0] != self.sp.nnz():
            raise ValueError(f"A has {At.shape[0]} nonzeros, the pattern has {self.sp.nnz()}")
        n_inst = At.shape[1]
        status = torch.zeros(n_inst, dtype=torch.int32, device=At.device)
        stream = torch.cuda.current_stream().cuda_stream
        fn = L.cxh_ldl_nfact if self.solver == "ldl" else L.cxh_qr_nfact
        self._n_fact = None
        L.check(fn(self._h, n_inst, At.data_ptr(), status.data_ptr(), 0, stream), fn.__name__)
        self._n_fact, self._single, self._status = n_inst, single, status
        if self.solver == "qr" and int(status.max()) > 0:
            self._n_fact = None
            raise RuntimeError(f"'nfact' failed: {int((status > 0).sum())} of {n_inst} instances are singular")
        return 0

    def status(self):
        """Per instance of the last nfact: zero or non-finite pivots (ldl),
        |R_ii| < eps entries (qr)."""
        if self._status is None:
            raise RuntimeError("Linear system has not been factorized")
        return self._status.cpu().numpy()

    def _require_nfact(self):
        if self._n_fact is None:
            raise RuntimeError("Linsol::solve: 'nfact' must be called before 'solve'")

    def solve(self, A, B=None, tr=False):
        """solve(A, B, tr): sfact, nfact and solve (Linsol::solve(DM, DM, bool),
        casadi/core/linsol.cpp:73-99); solve(B, tr) after nfact: B (n,) or
        (n, nrhs) for one system, (n, n_inst) or (n*nrhs, n_inst) for the
        batch of the last nfact.  Device tensors stay on the device."""
        import torch
        if B is None or isinstance(B, bool):
            tr = bool(B) if isinstance(B, bool) else tr
            B = A
        else:
            self.sfact(A)
            self.nfact(A)
        self._require_nfact()
        dev_in = isinstance(B, torch.Tensor) and B.is_cuda
        Bt = self._device(B)
        n_inst = self._n_fact
        if self._single:
            shape = tuple(Bt.shape)
            X = (Bt.reshape(self.n, -1).T.reshape(-1, 1) if Bt.dim() == 2 else Bt.reshape(-1, 1)).clone()
            nrhs = X.shape[0] // self.n
        else:
            X = Bt.clone().contiguous()
            nrhs = X.shape[0] // self.n
            if X.shape[1] != n_inst:
                raise ValueError(f"B has {X.shape[1]} instances, the last nfact factored {n_inst}")
            if X.shape[0] != self.n * nrhs:
                raise ValueError(f"B must have {self.n} * nrhs rows")
        X = X.contiguous()
        stream = torch.cuda.current_stream().cuda_stream
        fn = L.cxh_ldl_solve if self.solver == "ldl" else L.cxh_qr_solve
        L.check(fn(self._h, n_inst, X.data_ptr(), nrhs, int(bool(tr)), 0, stream), fn.__name__)
        if self._single:
            X = X.reshape(nrhs, self.n).T.reshape(shape)
        return X if dev_in else X.cpu().numpy()

    def _counts(self, fn, what):
        import torch
        self._require_nfact()
        out = torch.zeros(max(self._n_fact, 1), dtype=torch.int32, device="cuda")
        L.check(fn(self._h, out.data_ptr(), torch.cuda.current_stream().cuda_stream), what)
        v = out[: self._n_fact].cpu().numpy()
        return int(v[0]) if self._single else v

    def neig(self, A=None):
        """Number of negative eigenvalues (negative D entries) of the last
        nfact's factorisation, per instance (LinsolLdl::neig)."""
        if self.solver != "ldl":
            raise NotImplementedError(f"'neig' not defined for {self.solver}")
        return self._counts(L.cxh_ldl_neig, "cxh_ldl_neig")

    def rank(self, A=None):
        """Number of nonzero D entries of the last nfact's factorisation, per
        instance (LinsolLdl::rank)."""
        if self.solver != "ldl":
            raise NotImplementedError(f"'rank' not defined for {self.solver}")
        return self._counts(L.cxh_ldl_rank, "cxh_ldl_rank")

    def det(self, A):
        """Determinant: sfact, nfact, then the determinant from the QR factors
        (Linsol::det(DM), casadi/core/linsol.cpp:195-201; LinsolQr::det)."""
        import torch
        if self.solver != "qr":
            raise NotImplementedError(f"'det' not defined for {self.solver}")
        self.sfact(A)
        self.nfact(A)
        out = torch.zeros(max(self._n_fact, 1), dtype=torch.float64, device="cuda")
        L.check(L.cxh_qr_det(self._h, out.data_ptr(), torch.cuda.current_stream().cuda_stream), "cxh_qr_det")
        v = out[: self._n_fact].cpu().numpy()
        return float(v[0]) if self._single else v

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
