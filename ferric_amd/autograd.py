"""torch.autograd.Function over the RI-MP2 energy and its hand-written adjoint.

``ri_mp2_energy(b_ov, eps, nocc, nvir, first_occ)`` returns the differentiable
pair (E_os, E_ss). The forward runs the fused energy kernel; the backward runs
the adjoint kernel with the incoming gradient as the spin weights
(E = g_os E_os + g_ss E_ss), so any scalar built from the two components --
plain MP2, SCS-MP2 (crates/ferric-mp2/src/scs.rs:93-114), SOS -- gets
dE/dB (4 x_ov) and dE/deps (2 diag P) without storing amplitudes.
"""
from __future__ import annotations

import torch

from ._lib import Precision
from .rimp2 import energy_adjoint, spin_components


def _adjoint_precision(p: Precision) -> Precision:
    # The adjoint runs in FP64, BF16x3 or INT8x4: those stay, everything else is FP64.
    return p if p in (Precision.BF16X3, Precision.INT8X4) else Precision.FP64


class RiMp2Function(torch.autograd.Function):
    """forward(b_ov, eps_occ, eps_vir, precision) -> tensor([E_os, E_ss])."""

    @staticmethod
    def forward(ctx, b_ov, eps_occ, eps_vir, precision=int(Precision.FP64)):
        nocc, nvir = eps_occ.numel(), eps_vir.numel()
        eps = torch.cat([eps_occ.detach(), eps_vir.detach()]).to("cpu", torch.float64).numpy()
        b = b_ov.detach().contiguous()
        sc = spin_components(b, eps, nocc, nvir, 0, nocc, None, Precision(precision))
        ctx.save_for_backward(b, eps_occ, eps_vir)
        ctx.precision = Precision(precision)
        return torch.tensor([sc.e_os, sc.e_ss], dtype=torch.float64, device=b_ov.device)

    @staticmethod
    def backward(ctx, grad):
        b, eps_occ, eps_vir = ctx.saved_tensors
        nocc, nvir = eps_occ.numel(), eps_vir.numel()
        g = grad.detach().to("cpu", torch.float64)
        eps = torch.cat([eps_occ.detach(), eps_vir.detach()]).to("cpu", torch.float64).numpy()
        _, _, gamma, deps = energy_adjoint(b, eps, nocc, nvir, 0, nocc, float(g[0]), float(g[1]),
                                           _adjoint_precision(ctx.precision))
        if not torch.is_tensor(gamma):  # host entry (CPU b_ov): a numpy array
            gamma = torch.from_numpy(gamma)
        gb = gamma.to(b.device, b.dtype) if ctx.needs_input_grad[0] else None
        d = torch.from_numpy(deps)
        go = d[:nocc].to(eps_occ.device, eps_occ.dtype) if ctx.needs_input_grad[1] else None
        gv = d[nocc:].to(eps_vir.device, eps_vir.dtype) if ctx.needs_input_grad[2] else None
        return gb, go, gv, None


def ri_mp2_energy(b_ov: torch.Tensor, eps: torch.Tensor, nocc: int, nvir: int,
                  first_occ: int = 0, nocc_total: int | None = None,
                  precision: Precision = Precision.FP64):
    """Differentiable (E_os, E_ss) over the active pairs; ``eps`` is the full
    orbital-energy vector (frozen-core entries get zero gradient)."""
    if nocc_total is None:
        nocc_total = first_occ + nocc
    e_occ = eps[first_occ:first_occ + nocc]
    e_vir = eps[nocc_total:nocc_total + nvir]
    out = RiMp2Function.apply(b_ov, e_occ, e_vir, int(precision))
    return out[0], out[1]
