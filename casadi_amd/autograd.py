"""A Function evaluated over a batch on the GPU as a differentiable torch op.

``TorchFunction(f, checkpoint_every)`` maps ``f`` over the instances of a batch
(``f.map(n, "hip")``) and registers its adjoint with torch.autograd:

* forward: the forward sweep of every instance, storing the tape state at the
  start of every checkpoint segment into a checkpoint buffer kept for backward
  (cxh_map_fwd_ckpt);
* backward: the hand-written reverse sweep, restarting each segment from its
  checkpoint, seeded with the incoming output cotangents (cxh_map_rev_ckpt).

The gradient is ``J(x)^T g`` per instance, the quantity CasADi's
``f.reverse(1)`` Function computes (XFunction::get_reverse,
casadi/core/x_function.hpp:892-935), evaluated numerically instead of through
a symbolic reverse tape.

Tensors are float64 CUDA tensors laid out ``(nnz_in(i), n)`` -- one row per
nonzero of the input, one column per instance, as ``repmat(x, 1, n)`` stored
row-major (see ``casadi_amd.hip``).
"""
from .hip import HipTape


def _torch():
    import torch
    return torch


class TorchFunction:
    """Differentiable batched evaluation of ``f`` (one lane per instance)."""

    def __init__(self, f, checkpoint_every=0, engine="jit", tape=None):
        self.f = f
        self.tape = tape if tape is not None else HipTape(f, engine=engine)
        self.checkpoint_every = int(checkpoint_every)
        self._fn = _make_autograd_function()

    def __call__(self, *args):
        if len(args) != len(self.tape.nnz_in):
            raise ValueError(f"{self.f.name()} takes {len(self.tape.nnz_in)} inputs, got {len(args)}")
        outs = self._fn.apply(self, *args)
        return outs[0] if len(outs) == 1 else outs


def _make_autograd_function():
    torch = _torch()

    class _TapeAdjoint(torch.autograd.Function):
        @staticmethod
        def forward(ctx, owner, *args):
            args = [a.contiguous() if a is not None else None for a in args]
            n = next(a for a in args if a is not None).shape[1]
            outs, ckpt = owner.tape.fwd_ckpt(args, n, owner.checkpoint_every)
            ctx.owner, ctx.n, ctx.ckpt = owner, n, ckpt
            ctx.save_for_backward(*[a if a is not None else torch.empty(0) for a in args])
            ctx.present = [a is not None for a in args]
            like = args[0] if args[0] is not None else next(a for a in args if a is not None)
            return tuple(o if o is not None else like.new_zeros((0, n)) for o in outs)

        @staticmethod
        def backward(ctx, *grads):
            owner, n = ctx.owner, ctx.n
            saved = ctx.saved_tensors
            args = [a if p else None for a, p in zip(saved, ctx.present)]
            seeds = []
            for g, k in zip(grads, owner.tape.nnz_out):
                if g is None or k == 0:
                    seeds.append(None)
                else:
                    seeds.append(g.to(torch.float64).contiguous())
            adj = owner.tape.rev_ckpt(args, seeds, n, owner.checkpoint_every, ctx.ckpt)
            res = [None]
            for a, need in zip(adj, ctx.needs_input_grad[1:]):
                res.append(a if need else None)
            return tuple(res)

    return _TapeAdjoint


class TorchMapSum:
    """Differentiable reduced map: inputs flagged in ``reduce_in`` are one block
    shared by every instance (a vector of its nonzeros), outputs flagged in
    ``reduce_out`` are summed over the instances (a vector); the other blocks
    are ``(nnz, n)`` as for TorchFunction.  Forward is cxh_map_eval_sum;
    backward is the fused forward + adjoint of the reduced map
    (cxh_map_fwd_adj_sum, the reverse of MapSum, casadi/core/mapsum.cpp:354-415):
    the cotangent of a summed output seeds every instance, the gradient of a
    shared input is summed over the instances.  Backward re-runs the forward
    sweep inside the same kernels instead of keeping checkpoints, so nothing
    but the inputs is held between forward and backward."""

    def __init__(self, mapsum, checkpoint_every=0):
        if mapsum.parallelization != "hip":
            raise TypeError("to_torch needs parallelization 'hip'")
        self.mapsum = mapsum
        self.tape = mapsum.tape
        self.checkpoint_every = int(checkpoint_every)
        self._fn = _make_mapsum_function()

    def __call__(self, *args):
        if len(args) != len(self.tape.nnz_in):
            raise ValueError(f"{self.mapsum.name()} takes {len(self.tape.nnz_in)} inputs, got {len(args)}")
        outs = self._fn.apply(self, *args)
        return outs[0] if len(outs) == 1 else outs


def _make_mapsum_function():
    torch = _torch()

    class _MapSumAdjoint(torch.autograd.Function):
        @staticmethod
        def forward(ctx, owner, *args):
            ms = owner.mapsum
            args = [a.contiguous() if a is not None else None for a in args]
            outs = owner.tape.eval_sum(args, ms.n, ms.reduce_in, ms.reduce_out)
            ctx.owner = owner
            ctx.save_for_backward(*[a if a is not None else torch.empty(0) for a in args])
            ctx.present = [a is not None for a in args]
            like = next(a for a in args if a is not None)
            return tuple(o if o is not None else
                         like.new_zeros((0,) if ms.reduce_out[j] else (0, ms.n)) for j, o in enumerate(outs))

        @staticmethod
        def backward(ctx, *grads):
            owner = ctx.owner
            ms = owner.mapsum
            args = [a if p else None for a, p in zip(ctx.saved_tensors, ctx.present)]
            seeds = [None if g is None or k == 0 else g.to(torch.float64).contiguous()
                     for g, k in zip(grads, owner.tape.nnz_out)]
            _, adj = owner.tape.fwd_adj_sum(args, seeds, ms.n, ms.reduce_in, ms.reduce_out,
                                            checkpoint_every=owner.checkpoint_every, res=False)
            return (None,) + tuple(a if need else None for a, need in zip(adj, ctx.needs_input_grad[1:]))

    return _MapSumAdjoint
