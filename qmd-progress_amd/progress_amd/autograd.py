"""torch.autograd bindings of the SP2 and implicit-Fermi density solvers and
their hand-written adjoints.

    D = SP2Density.apply(H, bndfil, idemtol, ckpt)

H is an N x N symmetric float64 or float32 CUDA tensor (float32 input, as the
reference's mixed-precision entry takes, src/sp2_tensorcore/src/
tcore_hp_emulator.cu:81-130, is promoted exactly to float64; D and dL/dH are
returned in the input's dtype).  Forward runs prg_sp2_alg2 on the
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

from . import (PREC_FP64, gershgorin_device, implicit_fermi_adjoint_device,
               implicit_fermi_device, padded_ld, sp2_adjoint_device, sp2_alg2_device)


_DTYPES = (torch.float64, torch.float32)


def _check_input(h, who):
    if h.dtype not in _DTYPES or not h.is_cuda or h.dim() != 2 or h.shape[0] != h.shape[1]:
        raise ValueError(f"{who} expects an N x N float64 or float32 CUDA tensor")


def _pad(x, n, ld):
    out = torch.zeros((ld, ld), dtype=torch.float64, device=x.device)
    out[:n, :n] = x
    return out


class SP2Density(torch.autograd.Function):
    @staticmethod
    def forward(ctx, h, bndfil=0.5, idemtol=1e-10, ckpt=8):
        _check_input(h, "SP2Density")
        n = h.shape[0]
        ld = padded_ld(n)
        hp = _pad(h.detach().to(torch.float64), n, ld)
        rho = torch.zeros_like(hp)
        lo, hi = gershgorin_device(hp, n)
        it, pp = sp2_alg2_device(hp, rho, n, bndfil, 1, 500, "Abs", idemtol, precision=PREC_FP64)
        ctx.save_for_backward(hp)
        ctx.meta = (n, ld, pp, lo, hi, ckpt, h.dtype)
        return rho[:n, :n].to(h.dtype).clone()

    @staticmethod
    def backward(ctx, grad_d):
        (hp,) = ctx.saved_tensors
        n, ld, pp, lo, hi, ckpt, dtype = ctx.meta
        gp = _pad(grad_d.contiguous().to(torch.float64), n, ld)
        out = torch.zeros_like(gp)
        sp2_adjoint_device(hp, gp, out, n, pp, lo, hi, ckpt)
        return out[:n, :n].to(dtype).clone(), None, None, None


def sp2_density(h, bndfil=0.5, idemtol=1e-10, ckpt=8):
    return SP2Density.apply(h, bndfil, idemtol, ckpt)


class ImplicitFermiDensity(torch.autograd.Function):
    """P = prg_implicit_fermi(H) at a fixed chemical potential (method 1, k = 2,
    `nsteps` levels, inverse temperature `beta`, Newton-Schulz tolerance
    `tol`).  Backward: the reverse sweep of the level recursion
    (prg_gpu_implicit_fermi_adjoint_device), recomputing the levels from H;
    `canonical` gives the derivative at fixed occupation instead (the
    reference's first-order-response convention)."""

    @staticmethod
    def forward(ctx, h, nsteps, mu, beta, tol=1e-10, canonical=False):
        _check_input(h, "ImplicitFermiDensity")
        n = h.shape[0]
        ld = padded_ld(n)
        hp = _pad(h.detach().to(torch.float64), n, ld)
        p = torch.zeros_like(hp)
        implicit_fermi_device(hp, p, n, nsteps, mu, beta, tol)
        ctx.save_for_backward(hp)
        ctx.meta = (n, ld, nsteps, mu, beta, tol, canonical, h.dtype)
        return p[:n, :n].to(h.dtype).clone()

    @staticmethod
    def backward(ctx, grad_p):
        (hp,) = ctx.saved_tensors
        n, ld, nsteps, mu, beta, tol, canonical, dtype = ctx.meta
        g = grad_p.contiguous().to(torch.float64)
        ap = _pad(0.5 * (g + g.T), n, ld)
        out = torch.zeros_like(ap)
        implicit_fermi_adjoint_device(hp, ap, out, n, nsteps, mu, beta, tol, canonical)
        return out[:n, :n].to(dtype).clone(), None, None, None, None, None


def implicit_fermi_density(h, nsteps, mu, beta, tol=1e-10, canonical=False):
    return ImplicitFermiDensity.apply(h, nsteps, mu, beta, tol, canonical)
