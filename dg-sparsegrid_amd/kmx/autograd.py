"""Differentiation through the operator apply.

`ScaledOperator` holds a plan and groups of its terms with one scalar scale
per group (e.g. the terms acting in dimension d): A(alpha) = sum_g alpha_g A_g.
`apply(op, x, alpha)` is a torch.autograd.Function whose backward uses the
adjoint kernel (kmx_apply_transpose) for dL/dx = A(alpha)^T gy and one masked
forward apply per group for dL/dalpha_g = gy . (A_g x).

`rk3_trajectory` runs Kutta RK3 (time_advance.m:151-204) through that
function with checkpointed recomputation every `every` steps, so memory is
one state per segment instead of every stage.
"""
import numpy as np

from .timestep import A21, A31, A32, B1, B2, B3, C2, C3


def _torch():
    import torch

    return torch


class ScaledOperator:
    def __init__(self, plan, groups, prec="fp64"):
        """groups: list of lists of term indices (each term in at most one group;
        terms in no group keep scale 1)."""
        self.plan, self.groups, self.prec = plan, [list(g) for g in groups], prec
        self.grouped = {t for g in self.groups for t in g}
        self._cur = None

    def set_scales(self, alpha):
        a = [float(v) for v in alpha]
        if a == self._cur:
            return
        for g, terms in enumerate(self.groups):
            for t in terms:
                self.plan.set_term_scale(t, a[g])
        self._cur = a

    def group_apply(self, x, g):
        """A_g x (unscaled)."""
        keep = self._cur
        for t in range(self.plan.T):
            self.plan.set_term_scale(t, 1.0 if t in self.groups[g] else 0.0)
        y = self.plan.apply(x.contiguous(), prec=self.prec)
        self._cur = None
        if keep is not None:
            self._reset(keep)
        return y

    def _reset(self, alpha):
        for t in range(self.plan.T):
            if t not in self.grouped:
                self.plan.set_term_scale(t, 1.0)
        self._cur = None
        self.set_scales(alpha)


def _make_fn():
    torch = _torch()

    class _Apply(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, alpha, op):
            op.set_scales(alpha.detach().cpu().tolist())
            y = op.plan.apply(x.detach().contiguous(), prec=op.prec)
            ctx.op = op
            ctx.save_for_backward(x, alpha)
            return y

        @staticmethod
        def backward(ctx, gy):
            x, alpha = ctx.saved_tensors
            op = ctx.op
            a = alpha.detach().cpu().tolist()
            op.set_scales(a)
            gy = gy.contiguous()
            gx = op.plan.apply_transpose(gy, prec=op.prec) if ctx.needs_input_grad[0] else None
            ga = None
            if ctx.needs_input_grad[1]:
                ga = torch.stack([torch.dot(gy, op.group_apply(x.detach(), g)) for g in range(len(op.groups))])
                op._reset(a)
            return gx, ga, None

    return _Apply


_FN = None


def apply(op, x, alpha):
    """y = A(alpha) x, differentiable in x and alpha."""
    global _FN
    if _FN is None:
        _FN = _make_fn()
    return _FN.apply(x, alpha, op)


def rk3_trajectory(op, f0, alpha, dt, nsteps, every=5, source=None, t0=0.0):
    """f_T after nsteps RK3 steps of f' = A(alpha) f + s(t), differentiable in
    f0 and alpha; segments of `every` steps are recomputed in the backward pass
    (torch.utils.checkpoint)."""
    torch = _torch()
    from torch.utils.checkpoint import checkpoint

    def step(f, t):
        def rhs(tt):
            return 0.0 if source is None else source(tt)

        k1 = apply(op, f, alpha) + rhs(t)
        y1 = f + dt * A21 * k1
        k2 = apply(op, y1, alpha) + rhs(t + C2 * dt)
        y2 = f + dt * (A31 * k1 + A32 * k2)
        k3 = apply(op, y2, alpha) + rhs(t + C3 * dt)
        return f + dt * (B1 * k1 + B2 * k2 + B3 * k3)

    def segment(f, n, t):
        for i in range(n):
            f = step(f, t + i * dt)
        return f

    f = f0
    done = 0
    while done < nsteps:
        n = min(every, nsteps - done)
        t = t0 + done * dt
        if every and (f.requires_grad or alpha.requires_grad):
            f = checkpoint(segment, f, n, t, use_reentrant=False)
        else:
            f = segment(f, n, t)
        done += n
    del torch
    return f


def misfit(fT, target):
    """Final-time L2 misfit 0.5 ||f_T - target||^2."""
    d = fT - target
    return 0.5 * (d * d).sum()
