"""Jacobians of the TOA Stokes reflectance through the hand-written adjoint.

The library's adjoint returns, per instance, the gradient of
g . (rho_I, rho_Q, rho_U) with respect to the per-level medium (level optical
depth h[k], scattering fractions xdel[k], ydel[k]) and the wind speed.  This
module maps those onto the per-layer inputs of an explicitly layered
descriptor (Rayleigh tau, aerosol tau, aerosol single-scattering albedo,
absorption tau; `layers` of include/ocrt_amd.h) by the level convention of
the medium (rt_atm.c:343-360 and :521-538):

    h[k]     = sum_{l<k} dtau_l,   dtau_l = tau_R,l + tau_a,l + tau_abs,l
    xdel[l+1] = ssa_l tau_a,l / dtau_l,   ydel[l+1] = tau_R,l / dtau_l
    level 0: (0, 1) with Rayleigh present, else (ssa_0, 0).
"""
import numpy as np


def layer_gradients(layers, g_h, g_x, g_y, aerosol=True):
    """Per-layer gradients [4, nt] (Rayleigh tau, aerosol tau, aerosol ssa,
    absorption tau) from per-level gradients [nt+1]."""
    R, A, W, G = (np.asarray(v, dtype=np.float64) for v in layers)
    g_h, g_x, g_y = (np.asarray(v, dtype=np.float64) for v in (g_h, g_x, g_y))
    dt = R + A + G
    # h[k] depends on every layer above level k
    g_dt = np.cumsum(g_h[::-1])[::-1][1:]
    gx1, gy1 = (g_x[1:], g_y[1:]) if aerosol else (np.zeros_like(dt), g_y[1:])
    x1 = W * A / dt if aerosol else np.zeros_like(dt)
    y1 = R / dt
    common = g_dt - (gx1 * x1 + gy1 * y1) / dt
    gR = common + gy1 / dt
    gA = common + (gx1 * W / dt if aerosol else 0.0)
    gW = gx1 * A / dt if aerosol else np.zeros_like(dt)
    gG = common.copy()
    if aerosol and not R.sum() > 0.0:
        gW[0] += g_x[0]
    return np.stack([gR, gA, gW, gG])


def aod_scale_gradient(layers, g_layers):
    """d/ds at s = 1 of the solve with every aerosol optical depth scaled by s."""
    return float(np.dot(np.asarray(layers)[1], np.asarray(g_layers)[1]))
