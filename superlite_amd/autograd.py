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
(nc).  The arithmetic runs on the GPU behind the library.  With host (CPU) input tensors the
outputs and gradients are host tensors.  With input tensors on the context's GPU, the
outputs are device tensors filled on the device (slgpu_opacity_outputs_dev), and the
backward reads device output gradients in place (slgpu_opacity_vjp_dev), so the nc x ng
arrays never cross PCIe; the small per-cell inputs and input gradients still do.  A context
keeps the state of its last build for the backward; the backward re-runs the forward only
when the context has built something else since.
"""
from __future__ import annotations

import numpy as np
import torch


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64)


def _device_outputs(slctx, dev):
    """the last build's (cap64, capgrey, capross) as float64 tensors on `dev`"""
    cap = torch.empty((slctx.nx, slctx.ng), dtype=torch.float64, device=dev)
    grey = torch.empty(slctx.nx, dtype=torch.float64, device=dev)
    ross = torch.empty(slctx.nx, dtype=torch.float64, device=dev)
    torch.cuda.current_stream(dev).synchronize()         # the library writes on its own stream
    slctx.opacity_outputs_dev(cap.data_ptr(), grey.data_ptr(), ross.data_ptr())
    return cap, grey, ross


def _vjp(slctx, g_cap, g_grey, g_ross):
    """slgpu_opacity_vjp, reading device output gradients in place"""
    gs = [g for g in (g_cap, g_grey, g_ross) if g is not None]
    if gs and gs[0].is_cuda:
        dev = gs[0].device
        keep = [None if g is None else g.detach().to(device=dev, dtype=torch.float64).contiguous()
                for g in (g_cap, g_grey, g_ross)]
        torch.cuda.current_stream(dev).synchronize()
        return slctx.opacity_vjp_dev(*[0 if k is None else k.data_ptr() for k in keep]), dev
    return slctx.opacity_vjp(*[None if g is None else g.detach().cpu().numpy() for g in
                               (g_cap, g_grey, g_ross)]), None


def _t(a, dev):
    t = torch.from_numpy(a)
    return t if dev is None else t.to(dev)


class OpacityFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, slctx, consts, temp, natom, nelec, natomfr, oc):
        args = [_np(t) for t in (temp, natom, nelec, natomfr, oc)]
        dev = temp.device if temp.is_cuda else None
        out = OpacityFunction._build(slctx, consts, *args, host=dev is None)
        ctx.slctx, ctx.consts, ctx.args, ctx.token = slctx, consts, args, slctx.builds
        ctx.dev = dev
        if dev is not None:
            return _device_outputs(slctx, dev)
        return (torch.from_numpy(out["cap64"]), torch.from_numpy(out["capgrey"]),
                torch.from_numpy(out["capross"]))

    @staticmethod
    def _build(slctx, consts, temp, natom, nelec, natomfr, oc, host=True):
        return slctx.opacity_build(temp, natom, consts["vol"], consts["mass"], nelec, natomfr, oc,
                                   consts["ginv"], flags=consts.get("flags", 0),
                                   want=("cap64", "capgrey", "capross") if host else ())

    @staticmethod
    def backward(ctx, g_cap, g_grey, g_ross):
        if ctx.slctx.builds != ctx.token:
            OpacityFunction._build(ctx.slctx, ctx.consts, *ctx.args, host=False)
        g, dev = _vjp(ctx.slctx, g_cap, g_grey, g_ross)
        dev = ctx.dev if ctx.dev is not None else dev
        return (None, None, _t(g["temp"], dev), _t(g["natom"], dev), _t(g["nelec"], dev),
                _t(g["natomfr"], dev), _t(g["oc"], dev))


def opacity(slctx, consts, temp, natom, nelec, natomfr, oc):
    """(cap64, capgrey, capross) = opacity build, differentiable in its tensor arguments."""
    return OpacityFunction.apply(slctx, consts, temp, natom, nelec, natomfr, oc)


class GasOpacityFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, slctx, consts, temp, rho, massfr):
        args = [_np(t) for t in (temp, rho, massfr)]
        dev = temp.device if temp.is_cuda else None
        out = GasOpacityFunction._build(slctx, consts, *args, host=dev is None)
        ctx.slctx, ctx.consts, ctx.args, ctx.token = slctx, consts, args, slctx.builds
        ctx.dev = dev
        if dev is not None:
            return _device_outputs(slctx, dev)
        return (torch.from_numpy(out["cap64"]), torch.from_numpy(out["capgrey"]),
                torch.from_numpy(out["capross"]))

    @staticmethod
    def _build(slctx, consts, temp, rho, massfr, host=True):
        vol = consts["vol"]
        mass = rho * vol
        e = slctx.eos(temp, mass, massfr, nelec=consts.get("nelec0"), acc=consts.get("acc", 0.0))
        return slctx.opacity_build(temp, e["natom"], vol, mass, e["nelec"], e["natomfr"], e["oc"],
                                   e["ginv"], flags=consts.get("flags", 0),
                                   want=("cap64", "capgrey", "capross") if host else ())

    @staticmethod
    def backward(ctx, g_cap, g_grey, g_ross):
        if ctx.slctx.builds != ctx.token:
            GasOpacityFunction._build(ctx.slctx, ctx.consts, *ctx.args, host=False)
        g, dev = _vjp(ctx.slctx, g_cap, g_grey, g_ross)
        dev = ctx.dev if ctx.dev is not None else dev
        e = ctx.slctx.eos_vjp(g["natom"], g["natomfr"], g["nelec"], g["oc"])
        vol = ctx.consts["vol"]
        return (None, None, _t(g["temp"] + e["temp"], dev), _t(e["mass"] * vol, dev),
                _t(e["massfr"], dev))


def gas_opacity(slctx, consts, temp, rho, massfr):
    """(cap64, capgrey, capross) as functions of temperature, density and mass fractions
    (30, nc), through the device EOS; consts: vol (nc), optional nelec0 (start values of the
    EOS solve), acc (its accuracy; the reference's 1e-8 by default) and flags."""
    return GasOpacityFunction.apply(slctx, consts, temp, rho, massfr)
