"""Differentiating through a structure-factor potential with torch.autograd.

``structure_factor_loss(pot, x, S0, V)`` returns phi(x; S0, V) as a 0-d
float64 tensor whose backward gives d phi / d x, d phi / d S0 and d phi / d V.

The potential's fused energy+gradient pass (spread -> transform ->
loss/pullback -> one inverse transform -> kernel-derivative interpolation in
the adjoint precisions) is split where the two directions meet: forward
runs the pass up to the loss and its pullback spectrum (the energy), and
backward runs only the rest (the inverse transform and the interpolation of
the gradient at the points the forward sorted), so nothing is computed for a
gradient that backward never asks for and nothing is computed twice when it
does.  If another evaluation of the same potential, or a target update, comes
between the two, backward evaluates again at the forward's coordinates and
targets.  Potentials without the split (UwU, UwNU, NUwNU, batched NUwU)
compute their gradient in forward, as one pass.

The coordinate gradient is scaled from the reference's convention
(d phi / d(2 pi p / L), nuwu.hpp:294-314) to d phi / d x for the tensor passed
in.  The target gradients are closed forms over the modes evaluated on the
device from the evaluation's rho (fresco_pot_backward).  Targets are sent to
the library only when they changed since the last call (tensor identity and
version); when only S0 changed the update needs no host synchronisation.
Nothing of size N x modes is ever stored: memory is O(N + fine grid)
whatever the mode count.
"""
from __future__ import annotations

import numpy as np
import torch

from ._device import as_device_f64

__all__ = ["StructureFactorLoss", "structure_factor_loss"]


def _update_targets(pot, S0, V, given):
    """set_targets for the targets the caller gave (given = (S0 given, V
    given)) unless they are the tensors, at the same version, that this
    function last sent to ``pot``; V unchanged goes as None (no host
    synchronisation)."""
    s_given, v_given = given
    if not (s_given or v_given):
        return
    rec = getattr(pot, "_ag_targets", None) or (None, None, None, None)
    same_s = not s_given or (rec[0] is S0 and rec[1] == S0._version)
    same_v = not v_given or (rec[2] is V and rec[3] == V._version)
    if same_s and same_v:
        return
    pot.set_targets(S0.detach(), None if same_v else V.detach())
    pot._ag_targets = (S0 if s_given else None, S0._version, V if v_given else None, V._version)


def _x_scale(pot, dev):
    """pot.x_grad_scale() as a float64 tensor on dev, cached on the potential."""
    cache = getattr(pot, "_ag_xscale", None)
    if cache is None:
        cache = pot._ag_xscale = {}
    key = str(dev)
    if key not in cache:
        cache[key] = torch.as_tensor(np.asarray(pot.x_grad_scale()), dtype=torch.float64, device=dev)
    return cache[key]


def _target_grads(pot, S0, V, need_s, need_v):
    dS0, dV = pot.target_grads()
    dS0, dV = pot._grads_user(dS0, dV, S0)
    gS0 = dS0.to(device=S0.device, dtype=S0.dtype) if need_s else None
    gV = dV.to(device=V.device, dtype=V.dtype) if need_v else None
    return gS0, gV


def _x_grad(pot, g, x_shape, x_device, x_dtype):
    return (g * _x_scale(pot, g.device)).reshape(x_shape).to(device=x_device, dtype=x_dtype)


class StructureFactorLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, S0, V, pot, given=(True, True)):
        need_x, need_s, need_v = ctx.needs_input_grad[0], ctx.needs_input_grad[1], ctx.needs_input_grad[2]
        _update_targets(pot, S0, V, given)
        xd, _ = as_device_f64(x, pot.ndof)
        ctx.pot = pot
        ctx.kept = False
        ctx.grads = (None, None, None)
        if not (need_x or need_s or need_v):
            E = pot.get_energy(xd)
        else:
            token, E = pot.energy_keep(xd)
            if token:
                # backward finishes the pass (or evaluates again)
                ctx.kept = token
                ctx.xd = xd
                ctx.S0, ctx.V = S0.detach(), V.detach()
                ctx.x_meta = (x.shape, x.device, x.dtype)
            else:
                gx = None
                if need_x:
                    E, g = pot.get_energy_gradient(xd)
                    gx = _x_grad(pot, g, x.shape, x.device, x.dtype)
                else:
                    E = pot.get_energy(xd)
                gS0, gV = (None, None)
                if need_s or need_v:
                    gS0, gV = _target_grads(pot, S0, V, need_s, need_v)
                ctx.grads = (gx, gS0, gV)
        return torch.tensor(E, dtype=torch.float64, device=x.device)

    @staticmethod
    def backward(ctx, gE):
        if ctx.kept:
            gx, gS0, gV = StructureFactorLoss._finish(ctx)
        else:
            gx, gS0, gV = ctx.grads
        out = []
        for g in (gx, gS0, gV):
            out.append(None if g is None else g * gE.to(device=g.device, dtype=g.dtype))
        return out[0], out[1], out[2], None, None

    @staticmethod
    def _finish(ctx):
        pot, xd = ctx.pot, ctx.xd
        need_x, need_s, need_v = ctx.needs_input_grad[0], ctx.needs_input_grad[1], ctx.needs_input_grad[2]
        g = torch.empty_like(xd) if need_x else None
        restore = None
        if not pot.gradient_kept(ctx.kept, g):
            # another evaluation or target update came in between: evaluate
            # again at the forward's coordinates and targets, then put the
            # potential's current targets back
            restore = getattr(pot, "_user_targets", None), getattr(pot, "_ag_targets", None)
            pot.set_targets(ctx.S0, ctx.V)
            if need_x:
                _, g = pot.get_energy_gradient(xd)
            else:
                pot.get_energy(xd)
        gx = _x_grad(pot, g, *ctx.x_meta) if need_x else None
        gS0, gV = (None, None)
        if need_s or need_v:
            gS0, gV = _target_grads(pot, ctx.S0, ctx.V, need_s, need_v)
        if restore is not None and restore[0] is not None:
            pot.set_targets(*restore[0])
            pot._ag_targets = restore[1]
        return gx, gS0, gV


def structure_factor_loss(pot, x, S0=None, V=None):
    """phi of potential ``pot`` at coordinates (or grid values) ``x`` with
    targets ``S0`` and weights ``V`` (same layout as the constructor's Sk / V;
    an omitted one: the potential's current one, as a constant, and it is
    not sent to the library)."""
    given = (S0 is not None, V is not None)
    if S0 is None or V is None:
        s_user, v_user = pot.targets()
        S0 = torch.as_tensor(s_user) if S0 is None else S0
        V = torch.as_tensor(v_user) if V is None else V
    if not isinstance(x, torch.Tensor):
        x = torch.as_tensor(x, dtype=torch.float64)
    return StructureFactorLoss.apply(x, S0, V, pot, given)
