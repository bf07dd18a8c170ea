"""torch.autograd.Function over the solver: TOA (rho_I, rho_Q, rho_U) of a
batch of explicitly layered instances, differentiable with respect to every
per-layer optical property and the wind speed through the hand-written
adjoint (ocrt_plan_adjoint).

    rho = toa_stokes(layers, wind, base)        # layers [B, 4, nt] float64, wind [B]
    rho.backward(g)                             # fills layers.grad, wind.grad

`layers` rows: Rayleigh tau, aerosol tau, aerosol single-scattering albedo,
absorption tau, top down.  `base` is a list of B keyword dicts for
ocrt_amd.abi.make_instance (geometry, surface, Greek coefficients, streams,
modes, ...); their layers and wind are replaced by the tensors.  The solve
runs the FP64 form with IPSS off (the adjoint's domain).
"""
import numpy as np
import torch

from . import Plan
from .abi import make_instance
from .adjoint import layer_gradients


def _descriptors(layers, wind, base):
    descs = []
    for b, kw in enumerate(base):
        kw = dict(kw)
        kw["layers"] = layers[b]
        kw["wind"] = float(wind[b])
        kw["ipss"] = 0
        descs.append(make_instance(**kw))
    return descs


class TOAStokes(torch.autograd.Function):
    @staticmethod
    def forward(ctx, layers, wind, base):
        lay = layers.detach().to("cpu", torch.float64).numpy()
        wnd = wind.detach().to("cpu", torch.float64).numpy()
        descs = _descriptors(lay, wnd, base)
        plan = Plan(descs, precision="fp64", adjoint=True)
        plan.execute()
        res = plan.results()
        if not (res["status"] == 0).all():
            plan.close()
            raise RuntimeError("forward solve failed for instances %s" % np.nonzero(res["status"])[0].tolist())
        ctx.plan = plan
        ctx.layers = lay
        ctx.aerosol = [d.L_max > 0 and d.greek for d in descs]
        ctx.keep = descs
        out = np.stack([res["rho_I"], res["rho_Q"], res["rho_U"]], axis=1)
        return torch.as_tensor(out, dtype=torch.float64, device=layers.device)

    @staticmethod
    def backward(ctx, grad_out):
        g = grad_out.detach().to("cpu", torch.float64).numpy()
        adj = ctx.plan.adjoint(g)
        ctx.plan.close()
        gl = np.stack([layer_gradients(ctx.layers[b], adj["grad_h"][b], adj["grad_x"][b], adj["grad_y"][b],
                                       aerosol=bool(ctx.aerosol[b])) for b in range(len(ctx.layers))])
        dev = grad_out.device
        return (torch.as_tensor(gl, dtype=torch.float64, device=dev),
                torch.as_tensor(adj["grad_wind"], dtype=torch.float64, device=dev), None)


def toa_stokes(layers, wind, base):
    return TOAStokes.apply(layers, wind, base)
