"""Differentiating through a structure-factor potential with torch.autograd.

``structure_factor_loss(pot, x, S0, V)`` returns phi(x; S0, V) as a 0-d
float64 tensor whose backward gives d phi / d x, d phi / d S0 and d phi / d V.

The coordinate gradient is the potential's own fused energy+gradient pass
(spread -> FFT -> loss/pullback -> one inverse FFT -> kernel-derivative
interpolation in the adjoint precisions), run once in forward when x needs a
gradient and scaled from the reference's convention (d phi / d(2 pi p / L),
nuwu.hpp:294-314) to d phi / d x for the tensor passed in.  The target
gradients are closed forms over the modes evaluated on the device from the
same forward's rho (fresco_pot_backward).  Nothing of size N x modes is ever
stored: the spread grid is rebuilt on every evaluation, so memory is
O(N + fine grid) whatever the mode count.
"""
from __future__ import annotations

import torch

from ._device import as_device_f64

__all__ = ["StructureFactorLoss", "structure_factor_loss"]


class StructureFactorLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, S0, V, pot):
        need_x, need_s, need_v = ctx.needs_input_grad[0], ctx.needs_input_grad[1], ctx.needs_input_grad[2]
        pot.set_targets(S0.detach(), V.detach())
        xd, _ = as_device_f64(x, pot.ndof)
        if need_x:
            E, g = pot.get_energy_gradient(xd)
            scale = torch.as_tensor(pot.x_grad_scale(), dtype=torch.float64, device=g.device)
            gx = (g * scale).reshape(x.shape).to(device=x.device, dtype=x.dtype)
        else:
            E = pot.get_energy(xd)
            gx = None
        gS0 = gV = None
        if need_s or need_v:
            dS0, dV = pot.target_grads()
            dS0, dV = pot._grads_user(dS0, dV, S0)
            gS0 = dS0.to(device=S0.device, dtype=S0.dtype) if need_s else None
            gV = dV.to(device=V.device, dtype=V.dtype) if need_v else None
        ctx.grads = (gx, gS0, gV)
        return torch.tensor(E, dtype=torch.float64, device=x.device)

    @staticmethod
    def backward(ctx, gE):
        gx, gS0, gV = ctx.grads
        out = []
        for g in (gx, gS0, gV):
            out.append(None if g is None else g * gE.to(device=g.device, dtype=g.dtype))
        return out[0], out[1], out[2], None


def structure_factor_loss(pot, x, S0=None, V=None):
    """phi of potential ``pot`` at coordinates (or grid values) ``x`` with
    targets ``S0`` and weights ``V`` (same layout as the constructor's Sk / V;
    defaults: the constructor's, as constants)."""
    if S0 is None or V is None:
        s_user, v_user = pot.targets()
        S0 = torch.as_tensor(s_user) if S0 is None else S0
        V = torch.as_tensor(v_user) if V is None else V
    if not isinstance(x, torch.Tensor):
        x = torch.as_tensor(x, dtype=torch.float64)
    return StructureFactorLoss.apply(x, S0, V, pot)
