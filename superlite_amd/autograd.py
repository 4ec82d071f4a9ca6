"""torch.autograd.Function over the opacity build and its hand-written backward.

Forward: physical_opacity + opacity_planckmean + opacity_rossmean (GAS/gas_update.f:95-100)
through slgpu_opacity_build.  Backward: slgpu_opacity_vjp, which recomputes the line terms
on the device.  Differentiable inputs are the build's own: temperature, natom, n_e, the
mass fractions (30, nc) and the ion populations oc (nc, nion); vol, mass and ginv are
constants.  Outputs are the FP64 group opacity (nc, ng), the Planck mean (nc) and the
Rosseland mean (nc).  Tensors are host (CPU) float64 tensors; the arithmetic runs on the
GPU behind the library.  The backward re-runs the forward on the saved inputs first, so
it never depends on which build the context ran last.
"""
from __future__ import annotations

import numpy as np
import torch


class OpacityFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, slctx, consts, temp, natom, nelec, natomfr, oc):
        args = [t.detach().cpu().numpy().astype(np.float64) for t in (temp, natom, nelec, natomfr, oc)]
        out = OpacityFunction._build(slctx, consts, *args)
        ctx.slctx, ctx.consts, ctx.args = slctx, consts, args
        return (torch.from_numpy(out["cap64"]), torch.from_numpy(out["capgrey"]),
                torch.from_numpy(out["capross"]))

    @staticmethod
    def _build(slctx, consts, temp, natom, nelec, natomfr, oc):
        return slctx.opacity_build(temp, natom, consts["vol"], consts["mass"], nelec, natomfr, oc,
                                   consts["ginv"], flags=consts.get("flags", 0),
                                   want=("cap64", "capgrey", "capross"))

    @staticmethod
    def backward(ctx, g_cap, g_grey, g_ross):
        OpacityFunction._build(ctx.slctx, ctx.consts, *ctx.args)
        g = ctx.slctx.opacity_vjp(
            None if g_cap is None else g_cap.detach().cpu().numpy(),
            None if g_grey is None else g_grey.detach().cpu().numpy(),
            None if g_ross is None else g_ross.detach().cpu().numpy())
        return (None, None, torch.from_numpy(g["temp"]), torch.from_numpy(g["natom"]),
                torch.from_numpy(g["nelec"]), torch.from_numpy(g["natomfr"]),
                torch.from_numpy(g["oc"]))


def opacity(slctx, consts, temp, natom, nelec, natomfr, oc):
    """(cap64, capgrey, capross) = opacity build, differentiable in its tensor arguments."""
    return OpacityFunction.apply(slctx, consts, temp, natom, nelec, natomfr, oc)
