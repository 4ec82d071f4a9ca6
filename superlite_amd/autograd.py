"""torch.autograd.Functions over the opacity build and its hand-written backward.

OpacityFunction: physical_opacity + opacity_planckmean + opacity_rossmean
(GAS/gas_update.f:95-100) through slgpu_opacity_build, differentiable in the build's own
inputs (temperature, natom, n_e, mass fractions (30, nc), ion populations oc (nc, nion));
backward: slgpu_opacity_vjp, which recomputes the line terms on the device.

GasOpacityFunction: the same outputs as functions of the physical state, temperature,
density and element mass fractions (30, nc): slgpu_eos (massfr2natomfr + the LTE EOS,
gas_setup.f:101-160, eos_update.f:23-45) feeds slgpu_opacity_build on the device; backward:
slgpu_opacity_vjp chained into slgpu_eos_vjp (the implicit adjoint of the converged electron
fraction), so dL/dT gathers the opacity's direct terms and the ionisation balance's.

Outputs are the FP64 group opacity (nc, ng), the Planck mean (nc) and the Rosseland mean
(nc).  Tensors are host (CPU) float64 tensors; the arithmetic runs on the GPU behind the
library.  A context keeps the state of its last build for the backward; the backward
re-runs the forward only when the context has built something else since.
"""
from __future__ import annotations

import numpy as np
import torch


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64)


class OpacityFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, slctx, consts, temp, natom, nelec, natomfr, oc):
        args = [_np(t) for t in (temp, natom, nelec, natomfr, oc)]
        out = OpacityFunction._build(slctx, consts, *args)
        ctx.slctx, ctx.consts, ctx.args, ctx.token = slctx, consts, args, slctx.builds
        return (torch.from_numpy(out["cap64"]), torch.from_numpy(out["capgrey"]),
                torch.from_numpy(out["capross"]))

    @staticmethod
    def _build(slctx, consts, temp, natom, nelec, natomfr, oc):
        return slctx.opacity_build(temp, natom, consts["vol"], consts["mass"], nelec, natomfr, oc,
                                   consts["ginv"], flags=consts.get("flags", 0),
                                   want=("cap64", "capgrey", "capross"))

    @staticmethod
    def backward(ctx, g_cap, g_grey, g_ross):
        if ctx.slctx.builds != ctx.token:
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


class GasOpacityFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, slctx, consts, temp, rho, massfr):
        args = [_np(t) for t in (temp, rho, massfr)]
        out = GasOpacityFunction._build(slctx, consts, *args)
        ctx.slctx, ctx.consts, ctx.args, ctx.token = slctx, consts, args, slctx.builds
        return (torch.from_numpy(out["cap64"]), torch.from_numpy(out["capgrey"]),
                torch.from_numpy(out["capross"]))

    @staticmethod
    def _build(slctx, consts, temp, rho, massfr):
        vol = consts["vol"]
        mass = rho * vol
        e = slctx.eos(temp, mass, massfr, nelec=consts.get("nelec0"), acc=consts.get("acc", 0.0))
        return slctx.opacity_build(temp, e["natom"], vol, mass, e["nelec"], e["natomfr"], e["oc"],
                                   e["ginv"], flags=consts.get("flags", 0),
                                   want=("cap64", "capgrey", "capross"))

    @staticmethod
    def backward(ctx, g_cap, g_grey, g_ross):
        if ctx.slctx.builds != ctx.token:
            GasOpacityFunction._build(ctx.slctx, ctx.consts, *ctx.args)
        g = ctx.slctx.opacity_vjp(
            None if g_cap is None else g_cap.detach().cpu().numpy(),
            None if g_grey is None else g_grey.detach().cpu().numpy(),
            None if g_ross is None else g_ross.detach().cpu().numpy())
        e = ctx.slctx.eos_vjp(g["natom"], g["natomfr"], g["nelec"], g["oc"])
        vol = ctx.consts["vol"]
        return (None, None, torch.from_numpy(g["temp"] + e["temp"]),
                torch.from_numpy(e["mass"] * vol), torch.from_numpy(e["massfr"]))


def gas_opacity(slctx, consts, temp, rho, massfr):
    """(cap64, capgrey, capross) as functions of temperature, density and mass fractions
    (30, nc), through the device EOS; consts: vol (nc), optional nelec0 (start values of the
    EOS solve), acc (its accuracy; the reference's 1e-8 by default) and flags."""
    return GasOpacityFunction.apply(slctx, consts, temp, rho, massfr)
