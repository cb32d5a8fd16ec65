"""torch.autograd binding of the SP2 density solver and its hand-written adjoint.

    D = SP2Density.apply(H, bndfil, idemtol, ckpt)

H is an N x N symmetric float64 CUDA tensor.  Forward runs prg_sp2_alg2 on the
GPU (FP64 MFMA path) and records the branch sequence and the Gershgorin
bounds; backward runs the reverse sweep of the same recursion with those
frozen (prg_gpu_sp2_adjoint_device), recomputing the X_k from checkpoints
stored every `ckpt` steps.  The normalisation bounds are not differentiated
(the Gershgorin map is piecewise linear in H); gradients are those of the
purification polynomial at fixed bounds and sequence, which is what central
finite differences with the sequence and bounds frozen measure.
"""
from __future__ import annotations

import torch

from . import PREC_FP64, padded_ld, sp2_adjoint_device, sp2_alg2_device, gershgorin_device


def _pad(x, n, ld):
    out = torch.zeros((ld, ld), dtype=torch.float64, device=x.device)
    out[:n, :n] = x
    return out


class SP2Density(torch.autograd.Function):
    @staticmethod
    def forward(ctx, h, bndfil=0.5, idemtol=1e-10, ckpt=8):
        if h.dtype != torch.float64 or not h.is_cuda or h.dim() != 2 or h.shape[0] != h.shape[1]:
            raise ValueError("SP2Density expects an N x N float64 CUDA tensor")
        n = h.shape[0]
        ld = padded_ld(n)
        hp = _pad(h.detach(), n, ld)
        rho = torch.zeros_like(hp)
        lo, hi = gershgorin_device(hp, n)
        it, pp = sp2_alg2_device(hp, rho, n, bndfil, 1, 500, "Abs", idemtol, precision=PREC_FP64)
        ctx.save_for_backward(hp)
        ctx.meta = (n, ld, pp, lo, hi, ckpt)
        return rho[:n, :n].clone()

    @staticmethod
    def backward(ctx, grad_d):
        (hp,) = ctx.saved_tensors
        n, ld, pp, lo, hi, ckpt = ctx.meta
        gp = _pad(grad_d.contiguous().to(torch.float64), n, ld)
        out = torch.zeros_like(gp)
        sp2_adjoint_device(hp, gp, out, n, pp, lo, hi, ckpt)
        return out[:n, :n].clone(), None, None, None


def sp2_density(h, bndfil=0.5, idemtol=1e-10, ckpt=8):
    return SP2Density.apply(h, bndfil, idemtol, ckpt)
