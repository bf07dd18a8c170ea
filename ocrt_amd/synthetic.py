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


# ---------------------------------------------------------------------------
# Coupled atmosphere-ocean workload: 100 wavelengths over 330-1100 nm, water
# bodies with chlorophyll-driven optical properties and a strongly
# forward-peaked 1000-term hydrosol.

def pure_water_absorption(wl):
    """Smooth synthetic pure-water absorption (m^-1): low in the blue, rising
    through the red to strong near-infrared absorption."""
    wl = np.asarray(wl, dtype=np.float64)
    return (0.005 + 0.05 * np.exp((wl - 550.0) / 80.0) + 2.0 / (1.0 + np.exp(-(wl - 740.0) / 25.0))
            + 20.0 / (1.0 + np.exp(-(wl - 1000.0) / 40.0)))


def pure_water_scattering(wl):
    """Pure-water scattering coefficient (m^-1), lambda^-4.32 law."""
    return 0.0038 * (400.0 / np.asarray(wl, dtype=np.float64)) ** 4.32


def phyto_shape(wl):
    """Smooth synthetic phytoplankton absorption shape: blue and red peaks."""
    wl = np.asarray(wl, dtype=np.float64)
    return np.exp(-((wl - 440.0) / 40.0) ** 2) + 0.5 * np.exp(-((wl - 675.0) / 15.0) ** 2) + 0.1


def water_iops(wl, chl):
    """(a, b_w, b_p) in m^-1: absorption a = a_w + 0.06 Chl^0.65 shape(wl),
    pure-water scattering b_w, particle scattering b_p = 0.3 Chl^0.62 (550/wl)."""
    a = pure_water_absorption(wl) + 0.06 * chl ** 0.65 * phyto_shape(wl)
    return a, pure_water_scattering(wl), 0.3 * chl ** 0.62 * (550.0 / wl)


def hydrosol_greek(L=1000, g1=0.95, g2=-0.3, w=0.98, dolp=0.3):
    """Greek set of a two-term HG hydrosol (weight w on g1), with Rayleigh-like
    polarisation of amplitude dolp (P12 = -dolp (1 - mu^2) P11)."""
    return np.ascontiguousarray(w * hg_greek(g1, L, dolp=dolp) + (1.0 - w) * hg_greek(g2, L, dolp=dolp))


def config4(n_inst=512, seed=4, n_mu=64, n_layers=300, water_layers=100, m_max=127, depth_m=200.0,
            tau_cap=30.0, L_hyd=1000, sza=40.0, vza=30.0, raa=90.0, wind=5.0, chl=None, aod_scale=1.0):
    """Coupled atmosphere-ocean instances: wavelength i mod 100 on a uniform
    330-1100 nm grid, Chl ~ logU(0.01, 10) mg m^-3 per instance (or the
    given array), atmosphere as the batched workload (Rayleigh tau lambda^-4.08,
    AOD(865) ~ logU(0.01, 0.5) with an Angstrom law, HG g 0.7 aerosol with 80
    Greek terms, SSA ~ U(0.85, 1)) over a rough sea (wind 5 m/s, n 1.34) in
    300 layers; water 200 m deep in 100 layers (optical depth capped at
    tau_cap: the black bottom lies below), 64 streams both sides, modes 0..127.
    Returns (atmosphere descriptors, Water list, per-instance dict)."""
    from .ocean import Water
    rng = np.random.default_rng(seed)
    wls = np.linspace(330.0, 1100.0, 100)
    hyd = hydrosol_greek(L_hyd)
    gk = hg_greek(0.70, 80)
    descs, waters = [], []
    meta = {"wavelength": [], "chl": [], "aod865": [], "angstrom": []}
    for i in range(n_inst):
        wl = float(wls[i % 100])
        c_chl = float(chl[i]) if chl is not None else math.exp(rng.uniform(math.log(0.01), math.log(10.0)))
        aod865 = math.exp(rng.uniform(math.log(0.01), math.log(0.5)))
        ang = rng.uniform(0.0, 2.0)
        ssa = rng.uniform(0.85, 1.0)
        tau_r = RAYLEIGH_TAU_555 * (wl / 555.0) ** -4.08
        aod = aod_scale * aod865 * (wl / 865.0) ** -ang
        descs.append(make_instance(sza=sza, vza=vza, raa=raa, wavelength=wl, surface="black_fresnel_ocean",
                                   wind=wind, tau_R=tau_r, tau_a=aod, ssa_a=ssa, aer_h_km=2.0, greek=gk,
                                   n_mu=n_mu, n_layers=n_layers, m_max=m_max))
        a, bw, bp = water_iops(wl, c_chl)
        c = a + bw + bp
        waters.append(Water(min(c * depth_m, tau_cap), bw / c, bp / c, greek=hyd, n_layers=water_layers,
                            n_mu=n_mu))
        meta["wavelength"].append(wl)
        meta["chl"].append(c_chl)
        meta["aod865"].append(aod865)
        meta["angstrom"].append(ang)
    return descs, waters, {k: np.array(v) for k, v in meta.items()}
