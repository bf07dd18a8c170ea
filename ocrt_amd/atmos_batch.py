"""The reference Python package's batched atmosphere entry point on the GPU
solver: `solve_atm_batch` with the arguments and result keys of
code/OCRT_Python/ocrt_py/atmos_batch.py:885-1013.

Each `atms[b]` is a medium object as the reference's builders make it
(build_atm_rayleigh / build_atm_aerosol, ocrt_py/atmos.py:33-79, 314-376):
  rm, gb   ring cosines and weights, [2 n_mu + 1], solar slot at index n_mu,
           the positive nodes (Gauss-Legendre plus the zero-weight view node)
           at n_mu + 1 ..
  h        level optical depths [nt + 1] (total extinction, top down)
  xdel, ydel  aerosol and Rayleigh scattering fractions per level
  beta2    Rayleigh anisotropy (gives the depolarisation)
  betal, gammal, alphal, zetal  aerosol Greek coefficients (aerosol modes)
`vjs[b]` is the positive ring index of the view node (1-based), `mu_suns[b]`
the solar cosine, `dphis[b]` the relative azimuth in radians (the Fourier
reconstruction uses dphi + pi, rt_fourier.c:9-33), `wind_list[b]` the wind
speed of the Cox-Munk ocean the reference's batch path always uses.

Every case becomes one descriptor with explicit layering: per layer k the
Rayleigh scattering depth ydel[k+1] dtau_k, the aerosol scattering depth
xdel[k+1] dtau_k (single-scattering albedo 1) and the rest of dtau_k as
absorption, so the solver sees the same optical depths and scattering
fractions; level 0 keeps the case's own fractions.  The cases run as one
uneven batch on the GPU (FP64 by default, like the reference's float64
arrays).  With use_value_kernel and aer_p11 (per case a pair of ascending
scattering angles in degrees and normalised P11 values) the aerosol I -> I
phase table is the azimuth mean of that table, as the reference's batched
value kernel (code/OCRT_Python/ocrt_py/atmos_batch.py:929-947,
aerosol.py:530-549).  Not served (ValueError): the coupled-ocean bottom
source, azimuth rules other than 1024 points, and the degenerate wind <= 0
with the Cox-Munk sigma law.
"""
import math

import numpy as np

from . import Plan
from .abi import make_instance

_DEPOL_DEFAULT = 0.0279


def _depolarisation(beta2):
    # beta2 = ron / 2 with ron = 2 (1 - d) / (2 + d) (rt_molecular_phase.c:5-34)
    ron = 2.0 * beta2
    return (2.0 - 2.0 * ron) / (ron + 2.0)


def _layers(atm, nt):
    h = np.asarray(atm.h, dtype=np.float64)
    xd = np.asarray(atm.xdel, dtype=np.float64)
    yd = np.asarray(atm.ydel, dtype=np.float64)
    if h.shape != (nt + 1,) or xd.shape != (nt + 1,) or yd.shape != (nt + 1,):
        raise ValueError("medium arrays must have nt + 1 levels")
    dt = np.diff(h)
    if np.any(dt <= 0.0):
        raise ValueError("level optical depths must increase downwards")
    ray = yd[1:] * dt
    aer = xd[1:] * dt
    absorb = np.maximum(dt - ray - aer, 0.0)
    lay = np.zeros((4, nt))
    lay[0] = ray
    lay[1] = aer
    lay[2] = np.where(aer > 0.0, 1.0, 0.0)
    lay[3] = absorb
    return lay, xd[0], yd[0]


def solve_atm_batch(atms, vjs, mu_suns, dphis, wind_list, n_mu, nt, m_max, max_iterations, tolerance,
                    mode='black_fresnel_ocean', L_max=2, sigma_type=1, n_water=1.34, q_convention=1,
                    n_phi_quad=1024, bottom_source=None, use_value_kernel=False, aer_p11=None,
                    precision="fp64"):
    """Batched atmosphere SOS over the GPU solver; returns the reference's
    dict: rho_I/Q/U, I/Q/U_TOA [B]; boa_I/Q/U [B, m_max + 1, n_mu] (BOA
    downward field at the positive nodes, ascending mu); mu_quad, w_quad
    [B, n_mu]; T_diff_dn_hemi [B]; n_orders [B]."""
    if mode == 'coxmunk':
        mode = 'black_fresnel_ocean'
    elif mode == 'aerosol_coxmunk':
        mode = 'aerosol_black_fresnel_ocean'
    if mode not in ('black_fresnel_ocean', 'aerosol_black_fresnel_ocean'):
        raise ValueError(f"mode {mode!r} is not a batched atmosphere mode")
    if bottom_source is not None:
        raise ValueError("the coupled-ocean bottom source is not served")
    if n_phi_quad != 1024:
        raise ValueError("the Cox-Munk azimuth rule has 1024 points")
    greek_mode = mode == 'aerosol_black_fresnel_ocean'
    B = len(atms)
    value_kernel = greek_mode and use_value_kernel and aer_p11 is not None
    if value_kernel and len(aer_p11) != B:
        raise ValueError("one (theta, P11) table per case")
    if not (len(vjs) == len(mu_suns) == len(dphis) == len(wind_list) == B):
        raise ValueError("one view index, solar cosine, azimuth and wind per case")
    descs, mu_quad, w_quad = [], np.zeros((B, n_mu)), np.zeros((B, n_mu))
    for b, atm in enumerate(atms):
        rm = np.asarray(atm.rm, dtype=np.float64)
        gb = np.asarray(atm.gb, dtype=np.float64)
        if rm.shape != (2 * n_mu + 1,):
            raise ValueError("ring arrays must have 2 n_mu + 1 directions")
        mu_pos, w_pos = rm[n_mu + 1:], gb[n_mu + 1:]
        mu_quad[b], w_quad[b] = mu_pos, w_pos
        vj = int(vjs[b])
        n_gl = int(np.count_nonzero(w_pos > 0.0))
        mu_view = float(mu_pos[vj - 1])
        lay, x0, y0 = _layers(atm, nt)
        if not greek_mode:
            # without Greek coefficients the reference's batch path gives the
            # aerosol no phase function: its scattering never enters
            lay[3] += lay[1]
            lay[1] = 0.0
            lay[2] = 0.0
            x0 = 0.0
        ws = float(wind_list[b])
        if ws <= 0.0:
            if sigma_type != 1:
                raise ValueError("wind <= 0 with the Cox-Munk sigma law gives a degenerate slope variance")
            ws = 0.01   # the OCRT law floors the wind at 0.01 m/s
        greek = None
        if greek_mode:
            L = min(int(L_max), len(atm.betal) - 1)
            greek = np.stack([np.asarray(getattr(atm, k), dtype=np.float64)[:L + 1]
                              for k in ("betal", "gammal", "alphal", "zetal")])
        # level 0 carries (0, 1) when Rayleigh is present and the aerosol
        # albedo with no Rayleigh (rt_atm.c:343-360)
        tau_R = float(lay[0].sum()) if y0 > 0.0 else 0.0
        if y0 == 0.0 and x0 > 0.0:
            lay[2, 0] = x0   # level 0 reads the top layer's albedo
            lay[3, 0] = max(lay[3, 0] - lay[1, 0] * (1.0 / x0 - 1.0), 0.0)
            lay[1, 0] = lay[1, 0] / x0
        d = make_instance(sza=math.degrees(math.acos(float(mu_suns[b]))), vza=math.degrees(math.acos(mu_view)),
                          raa=math.degrees(float(dphis[b])), wavelength=555.0, surface="black_fresnel_ocean",
                          n_water=n_water, wind=ws, sigma_type=sigma_type, q_convention=q_convention,
                          decouple_sunglint=1, tau_R=tau_R,
                          depol=_depolarisation(atm.beta2) if hasattr(atm, "beta2") else _DEPOL_DEFAULT,
                          tau_a=float(lay[1].sum()), ssa_a=1.0, greek=greek, n_mu=n_gl, m_max=m_max,
                          max_orders=max_iterations, tolerance=tolerance, ipss=0, layers=lay,
                          phase_table=(aer_p11[b][0], aer_p11[b][1]) if value_kernel else None)
        descs.append(d)
    plan = Plan(descs, precision=precision)
    plan.export_nodes()
    plan.execute()
    res = plan.results()
    _, boa = plan.node_fields(n_mu)
    plan.close()
    boa = boa[:, : m_max + 1]
    mus = np.asarray(mu_suns, dtype=np.float64)
    # T_diff_dn_hemi = 2 pi sum mu w boa_I[m = 0] / (pi mu_sun), view node w = 0
    F_dn = np.sum(mu_quad * w_quad * boa[:, 0, 0, :], axis=1) * 2.0 * math.pi
    return dict(rho_I=res["rho_I"], rho_Q=res["rho_Q"], rho_U=res["rho_U"],
                I_TOA=res["I_TOA"], Q_TOA=res["Q_TOA"], U_TOA=res["U_TOA"],
                boa_I=boa[:, :, 0, :].copy(), boa_Q=boa[:, :, 1, :].copy(), boa_U=boa[:, :, 2, :].copy(),
                mu_quad=mu_quad, w_quad=w_quad,
                T_diff_dn_hemi=F_dn / (math.pi * mus), n_orders=res["n_orders_used"].astype(int))
