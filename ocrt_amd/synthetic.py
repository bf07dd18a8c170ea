"""Synthetic inputs of the benchmark configurations.

Aerosol phase matrices enter the solver as the four coefficient families the
reference carries (rt_aerosol_input_t, code/OCRT_C/src/rt_solver.h:94-114):
  beta  : ordinary Legendre moments of P11, beta_0 = 1
  gamma : moments of P12 on the m = 0 spin-2 functions S_l (rt_phase_moments.c:5-20)
  alpha, zeta : spin-2 completion of the P22 = P11 and P33 moments
                (rt_phase_moments.c:22-64, called as in rt_aerosol.c:754-758).

The synthetic phase matrix is defined in angle space (mu = cos Theta):
  P11 = Henyey-Greenstein truncated at order L: beta_l = (2l+1) g^l
  P22 = P11
  P12 = -p * (1 - mu^2) * P11          (single-scatter DoLP = p sin^2 Theta, p = 0.2)
  P33 = mu * P11
and projected with a Gauss-Legendre rule that is exact for these polynomials.
"""
import math

import numpy as np

from .abi import make_instance

RAYLEIGH_TAU_555 = 0.0941
DEPOL = 0.0279
BANDS_NM = (412.0, 443.0, 490.0, 555.0, 660.0, 680.0, 745.0, 865.0)


def legendre_table(L, mu):
    """P_l(mu) for l = 0..L; shape [L+1, len(mu)]."""
    P = np.zeros((L + 1, mu.size))
    P[0] = 1.0
    if L >= 1:
        P[1] = mu
    for l in range(1, L):
        P[l + 1] = ((2 * l + 1) * mu * P[l] - l * P[l - 1]) / (l + 1)
    return P


def spin2_m0_table(L, mu):
    """S_l(mu) of the reference (rt_phase_spin2_m0): S_2 = 3(1-mu^2)/(2 sqrt 6)."""
    S = np.zeros((L + 1, mu.size))
    if L >= 2:
        S[2] = 3.0 * (1.0 - mu * mu) / (2.0 * math.sqrt(6.0))
    for l in range(2, L):
        scale = (2.0 * l + 1.0) / math.sqrt((l - 1.0) * (l + 3.0))
        back = math.sqrt((l - 2.0) * (l + 2.0)) / (2.0 * l + 1.0)
        S[l + 1] = scale * (mu * S[l] - back * S[l - 1])
    return S


def complete_spherical(beta11, beta22, delta33):
    """Spin-2 completion of the diagonal families (rt_phase_moments.c:22-64)."""
    L = beta11.size - 1
    alpha = np.zeros(L + 1)
    zeta = np.zeros(L + 1)
    for l in range(2, L + 1):
        e = float(l)
        lower = e * (e - 1.0) / ((e + 1.0) * (e + 2.0))
        cpl = 4.0 * (2.0 * e + 1.0) / (e * (e - 1.0) * (e + 1.0) * (e + 2.0))
        base = (e - 1.0) ** 2
        o = np.arange(l - 1, -1, -1)
        gap = l - o
        w = base - 1.5 * (gap - 1) * (l + o)
        even = (gap % 2) == 0
        sb = np.sum(w[even] * beta11[o[even]])
        sd = np.sum(w[even] * delta33[o[even]])
        ob = np.sum(w[~even] * beta22[o[~even]])
        od = np.sum(w[~even] * delta33[o[~even]])
        zeta[l] = lower * delta33[l] - cpl * (sd - ob)
        alpha[l] = lower * beta11[l] - cpl * (sb - od)
    return alpha, zeta


def hg_greek(g, L, dolp=0.2):
    """Greek arrays [4, L+1] of the synthetic aerosol (see module doc)."""
    l = np.arange(L + 1, dtype=np.float64)
    beta = (2.0 * l + 1.0) * np.power(g, l)
    x, w = np.polynomial.legendre.leggauss(L + 16)
    P = legendre_table(L, x)
    p11 = beta @ P
    p12 = -dolp * (1.0 - x * x) * p11
    p33 = x * p11
    S = spin2_m0_table(L, x)
    gamma = (2.0 * l + 1.0) * 0.5 * (S * (p12 * w)).sum(axis=1)
    d33 = (2.0 * l + 1.0) * 0.5 * (P * (p33 * w)).sum(axis=1)
    alpha, zeta = complete_spherical(beta, beta, d33)
    n = beta[0]
    return np.ascontiguousarray(np.stack([beta / n, gamma / n, alpha / n, zeta / n]))


def config0():
    """Rayleigh only, 50 layers, flat Fresnel n=1.34, 24 streams, m 0..2."""
    return [make_instance(sza=40.0, vza=55.0, raa=90.0, wavelength=555.0, surface="flat",
                          tau_R=RAYLEIGH_TAU_555, n_mu=24, n_layers=50, m_max=2)], []


def config1(m_max=95, n_layers=400, n_mu=48, L=180, ipss=1):
    """Rayleigh + HG aerosol (g 0.7, L 180, ssa 0.95, AOD 0.1, H 2 km), rough ocean
    wind 5, 48 streams, m 0..95, IPSS on."""
    gk = hg_greek(0.70, L)
    d = make_instance(sza=40.0, vza=55.0, raa=90.0, wavelength=555.0,
                      surface="black_fresnel_ocean", wind=5.0, tau_R=RAYLEIGH_TAU_555,
                      tau_a=0.1, ssa_a=0.95, aer_h_km=2.0, greek=gk, n_mu=n_mu,
                      n_layers=n_layers, m_max=m_max, ipss=ipss)
    return [d], [gk]


def batch_config(n_geom=1024, bands=BANDS_NM, seed=2024, m_max=16, L=80,
                 layers=(50, 400), streams=(16, 64), fixed_layers=None, fixed_streams=None,
                 ipss=0):
    """Config 2 (and config 3 with fixed layers/streams): bands x geometries of
    uneven instances drawn from one seed."""
    rng = np.random.default_rng(seed)
    descs, keep = [], []
    for _ in range(n_geom):
        sza = rng.uniform(0.0, 80.0)
        vza = rng.uniform(0.0, 70.0)
        raa = rng.uniform(0.0, 180.0)
        aod865 = math.exp(rng.uniform(math.log(0.01), math.log(0.5)))
        ang = rng.uniform(0.0, 2.0)
        g = rng.uniform(0.6, 0.8)
        ssa = rng.uniform(0.85, 1.0)
        wind = rng.uniform(1.0, 15.0)
        nt = fixed_layers or int(rng.integers(layers[0], layers[1] + 1))
        nmu = fixed_streams or int(rng.integers(streams[0], streams[1] + 1))
        gk = hg_greek(g, L)
        keep.append(gk)
        for lam in bands:
            tau_r = RAYLEIGH_TAU_555 * (lam / 555.0) ** -4.08
            aod = aod865 * (lam / 865.0) ** -ang
            descs.append(make_instance(sza=sza, vza=vza, raa=raa, wavelength=lam,
                                       surface="black_fresnel_ocean", wind=wind, tau_R=tau_r,
                                       tau_a=aod, ssa_a=ssa, aer_h_km=2.0, greek=gk,
                                       n_mu=nmu, n_layers=nt, m_max=m_max, ipss=ipss))
    return descs, keep


def config2(n_geom=1024, **kw):
    return batch_config(n_geom=n_geom, **kw)


def config3(n_inst=256, seed=3, **kw):
    return batch_config(n_geom=n_inst, bands=(555.0,), seed=seed, fixed_layers=200,
                        fixed_streams=32, **kw)
