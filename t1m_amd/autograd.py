"""torch.autograd harness over contract / contract_vjp.

Forward: C = A . B (a fresh zero C accumulated by the library's contraction
kernels).  Backward: the library's vector-Jacobian product,
grad_A = grad_C . conj(B) and grad_B = conj(A) . grad_C, which is torch's
convention for complex inputs (grad_output of a real loss L is
dL/dRe(C) + i dL/dIm(C)).
"""
from __future__ import annotations

import torch

from . import contract, contract_vjp


def _out_shape(a, la, b, lb, lc):
    ext = {}
    for t, labels in ((a, la), (b, lb)):
        for ch, d in zip(labels, t.shape):
            ext[ch] = int(d)
    return [ext[ch] for ch in lc]


class ContractFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, b, la: str, lb: str, lc: str, precision: str = "native"):
        c = torch.zeros(_out_shape(a, la, b, lb, lc), dtype=a.dtype, device=a.device)
        contract(a, la, b, lb, c, lc, precision=precision)
        ctx.save_for_backward(a, b)
        ctx.labels = (la, lb, lc)
        ctx.precision = precision
        return c

    @staticmethod
    def backward(ctx, grad_c):
        a, b = ctx.saved_tensors
        la, lb, lc = ctx.labels
        grad_a = torch.zeros_like(a) if ctx.needs_input_grad[0] else None
        grad_b = torch.zeros_like(b) if ctx.needs_input_grad[1] else None
        if grad_a is not None or grad_b is not None:
            contract_vjp(a, la, b, lb, grad_c, lc, grad_a, grad_b,
                         precision=ctx.precision)
        return grad_a, grad_b, None, None, None, None


def contract_autograd(a, la: str, b, lb: str, lc: str, precision: str = "native"):
    """Differentiable C = A . B over labels."""
    return ContractFunction.apply(a, b, la, lb, lc, precision)
