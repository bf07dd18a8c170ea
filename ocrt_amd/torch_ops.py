"""torch.autograd.Function over the solver: TOA (rho_I, rho_Q, rho_U) of a
batch of explicitly layered instances, differentiable with respect to every
per-layer optical property and the wind speed through the hand-written
adjoint (ocrt_plan_adjoint).

    rho = toa_stokes(layers, wind, base)        # layers [B, 4, nt] float64, wind [B]
    rho.backward(g)                             # fills layers.grad, wind.grad
    rho = toa_stokes(layers, wind, base, precision="mixed")

`layers` rows: Rayleigh tau, aerosol tau, aerosol single-scattering albedo,
absorption tau, top down.  `base` is a list of B keyword dicts for
ocrt_amd.abi.make_instance (geometry, surface, Greek coefficients, streams,
modes, ...); their layers and wind are replaced by the tensors.  The solve
runs with IPSS off (the adjoint's domain), in the FP64 form (the default) or,
with precision="mixed", in the mixed adjoint form: the forward's fields kept
as bf16 pairs, fp32 adjoint fields and bf16 MFMA operator products (a few
1e-6 of each gradient's scale from the FP64 form's).

Two ways through the boundary:

* device_io (the default for CUDA tensors): the tensors stay on the device.
  A plan per shape signature (base descriptors, batch, layers) is made once,
  from the first call's values, and kept; every call hands the plan the
  layers' and winds' data_ptr()s (ocrt_plan_set_inputs), runs it on
  torch.cuda.current_stream(), and has the results and the gradients written
  straight into device tensors (ocrt_plan_rho, ocrt_plan_adjoint_layers).
  Nothing is copied to the host after the plan exists and nothing
  synchronises.  The plan keeps the structure it was made with: an instance
  whose aerosol column appears or vanishes, or whose sea turns calm or rough,
  comes back as NaN (status OCRT_ERR_ARGS) -- make a plan for that structure
  by calling with it first, or use the host path.
* host path (device_io=False): descriptors from host copies and a fresh plan
  per call, results and gradients through host memory.

A cached plan serves every forward of its signature; a backward whose
forward is no longer the plan's last one reruns that forward from its saved
inputs before the adjoint, so interleaved forwards and repeated backwards
(retain_graph) give the gradients of their own inputs.
"""
from collections import OrderedDict

import numpy as np
import torch

from . import Plan
from .abi import make_instance
from .adjoint import layer_gradients

_CACHE = OrderedDict()
_CACHE_MAX = 4


def _check_precision(precision):
    if precision not in ("fp64", "mixed"):
        raise ValueError("precision must be 'fp64' or 'mixed', not %r" % (precision,))
    return precision


def _descriptors(layers, wind, base):
    descs = []
    for b, kw in enumerate(base):
        kw = dict(kw)
        kw["layers"] = layers[b]
        kw["wind"] = float(wind[b])
        kw["ipss"] = 0
        descs.append(make_instance(**kw))
    return descs


def _hashable(v):
    if isinstance(v, np.ndarray):
        return ("nd", v.dtype.str, v.shape, v.tobytes())
    if isinstance(v, (list, tuple)):
        return tuple(_hashable(x) for x in v)
    if isinstance(v, dict):
        return tuple(sorted((k, _hashable(x)) for k, x in v.items()))
    return v


def _signature(layers, base, precision):
    kws = tuple(_hashable({k: v for k, v in kw.items() if k not in ("layers", "wind")}) for kw in base)
    return (tuple(layers.shape), str(layers.device), precision, kws)


class _CachedPlan:
    """An adjoint plan with device inputs, and the forward it last ran."""

    def __init__(self, layers, wind, base, precision):
        descs = _descriptors(layers.detach().cpu().numpy(), wind.detach().cpu().numpy(), base)
        self.plan = Plan(descs, precision=precision, adjoint=True, device_inputs=True)
        self.keep = descs
        self.last = 0      # token of the forward whose state the plan holds
        self.issued = 0

    def run(self, layers, wind, stream):
        self.plan.set_inputs(layers.data_ptr(), wind.data_ptr(), stream)
        self.plan.execute(stream)
        self.issued += 1
        self.last = self.issued
        return self.last


def _cached_plan(layers, wind, base, precision):
    key = _signature(layers, base, precision)
    entry = _CACHE.get(key)
    if entry is None:
        entry = _CachedPlan(layers, wind, base, precision)
        _CACHE[key] = entry
        while len(_CACHE) > _CACHE_MAX:
            _CACHE.popitem(last=False)   # the plan lives on while a graph still holds it
    else:
        _CACHE.move_to_end(key)
    return entry


class TOAStokes(torch.autograd.Function):
    @staticmethod
    def forward(ctx, layers, wind, base, device_io, precision):
        if device_io:
            lay = layers.detach().to(torch.float64).contiguous()
            wnd = wind.detach().to(device=lay.device, dtype=torch.float64).contiguous()
            entry = _cached_plan(lay, wnd, base, precision)
            stream = torch.cuda.current_stream(lay.device)
            token = entry.run(lay, wnd, stream)
            out = torch.empty((lay.shape[0], 3), dtype=torch.float64, device=lay.device)
            entry.plan.rho_device(out.data_ptr(), 0, stream)
            ctx.entry, ctx.token, ctx.device_io = entry, token, True
            ctx.save_for_backward(lay, wnd)
            return out
        lay = layers.detach().to("cpu", torch.float64).numpy()
        wnd = wind.detach().to("cpu", torch.float64).numpy()
        descs = _descriptors(lay, wnd, base)
        plan = Plan(descs, precision=precision, adjoint=True)
        plan.execute()
        res = plan.results()
        if not (res["status"] == 0).all():
            plan.close()
            raise RuntimeError("forward solve failed for instances %s" % np.nonzero(res["status"])[0].tolist())
        ctx.plan = plan
        ctx.layers = lay
        ctx.aerosol = [d.L_max > 0 and d.greek for d in descs]
        ctx.keep = descs
        ctx.device_io = False
        out = np.stack([res["rho_I"], res["rho_Q"], res["rho_U"]], axis=1)
        return torch.as_tensor(out, dtype=torch.float64, device=layers.device)

    @staticmethod
    def backward(ctx, grad_out):
        if ctx.device_io:
            lay, wnd = ctx.saved_tensors
            entry = ctx.entry
            stream = torch.cuda.current_stream(lay.device)
            if entry.last != ctx.token:
                # another forward ran on this plan since: redo this one's
                ctx.token = entry.run(lay, wnd, stream)
            g = grad_out.detach().to(device=lay.device, dtype=torch.float64).contiguous()
            gl = torch.empty_like(lay)
            gw = torch.empty_like(wnd)
            entry.plan.adjoint_layers(g.data_ptr(), gl.data_ptr(), gw.data_ptr(), stream)
            return gl, gw, None, None, None
        g = grad_out.detach().to("cpu", torch.float64).numpy()
        adj = ctx.plan.adjoint(g)
        gl = np.stack([layer_gradients(ctx.layers[b], adj["grad_h"][b], adj["grad_x"][b], adj["grad_y"][b],
                                       aerosol=bool(ctx.aerosol[b])) for b in range(len(ctx.layers))])
        dev = grad_out.device
        return (torch.as_tensor(gl, dtype=torch.float64, device=dev),
                torch.as_tensor(adj["grad_wind"], dtype=torch.float64, device=dev), None, None, None)


def toa_stokes(layers, wind, base, device_io=None, precision="fp64"):
    """TOA (rho_I, rho_Q, rho_U) [B, 3]; device_io defaults to layers.is_cuda;
    precision "fp64" or "mixed" (the adjoint form the forward and the
    gradients run in)."""
    if device_io is None:
        device_io = bool(layers.is_cuda)
    return TOAStokes.apply(layers, wind, base, bool(device_io), _check_precision(precision))
