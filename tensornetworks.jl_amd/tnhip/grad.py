"""Energy <psi|H(c)|psi> of an MPO built from an operator list and its gradient with respect
to the term coefficients c, on the GPU.

The device pass (tnh_mpo_expect_grad) returns dE/dW_i for every MPO tensor from a forward
environment sweep with checkpoints every `block` sites and a backward sweep that recomputes
one block at a time. A coefficient c_t enters the automaton MPO at exactly one tensor entry
(mpo_tensors' slots), so dE/dc_t = sum_{s,t'} dE/dW_site[row, s, t', col] op[s, t'].
`energy_function` wraps this as a torch.autograd.Function of the coefficient vector, the
Python counterpart of a ChainRulesCore rrule. The reference has no differentiation; its
energy is inner(psi, H, psi), src/structures/mps/mpo.jl:181-217.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .core import Context, Tensor
from .models import OpList, Sitetypes, mpo_tensors
from .tebd import DeviceMPS


def mpo_expectation(psi: DeviceMPS, W, block: int = 8, grad: bool = False):
    """<psi|H|psi> for the MPO tensors W (numpy arrays or device Tensors), and with
    grad=True the list of dE/dW_i as numpy arrays of W_i's shape."""
    ctx = psi.ctx
    N = len(psi)
    if len(W) != N:
        raise ValueError("MPO and MPS have different lengths.")
    Wd = [w if isinstance(w, Tensor) else Tensor.from_numpy(np.asarray(w), ctx=ctx) for w in W]
    ps = (C.c_void_p * N)(*[t.h.value for t in psi.sites])
    ws = (C.c_void_p * N)(*[t.h.value for t in Wd])
    er, ei = C.c_double(), C.c_double()
    out = (C.c_void_p * N)() if grad else None
    ctx.call("tnh_mpo_expect_grad", N, ps, ws, int(block), C.byref(er), C.byref(ei), out)
    E = complex(er.value, ei.value)
    if not grad:
        return E
    G = [Tensor(ctx, C.c_void_p(out[i])) for i in range(N)]
    return E, [g.numpy() for g in G]


def coupling_gradient(st: Sitetypes, psi: DeviceMPS, H: OpList, block: int = 8):
    """(E, dE/dc) for H = sum_t c_t O_t: E = <psi|H|psi> (no normalisation) and
    dE/dc_t = <psi|O_t|psi> obtained by the backward pass through the MPO."""
    slots: list = []
    W = mpo_tensors(st, H, slots)
    E, G = mpo_expectation(psi, W, block=block, grad=True)
    g = np.array([np.sum(G[site - 1][row, :, :, col] * op) for site, row, col, op in slots])
    return E, g


def energy_function(st: Sitetypes, psi: DeviceMPS, H: OpList, block: int = 8):
    """A callable c -> E(c) on a torch float64 vector of term coefficients (same order as
    H's terms) whose backward is the device gradient: the torch.autograd counterpart of a
    ChainRulesCore rrule for the coupling dependence of <psi|H(c)|psi>."""
    import torch

    template = H

    class _Energy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, c):
            ol = OpList(template.length)
            for ops, sites, ci in zip(template.ops, template.sites, c.detach().cpu().numpy().tolist()):
                ol.add(ops, sites, ci)
            E, g = coupling_gradient(st, psi, ol, block=block)
            ctx.save_for_backward(torch.as_tensor(g.real, dtype=c.dtype))
            return torch.as_tensor(E.real, dtype=c.dtype)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g

    return _Energy.apply
