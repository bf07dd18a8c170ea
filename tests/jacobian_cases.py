"""Shared by the Jacobian pins: the adjoint-workload instances of
tests/golden/ref_jacobians.json (reference finite differences of wind and AOD
scale), their explicitly layered equivalents and the layer Jacobian of the
AOD scale.

A profile instance (Rayleigh and aerosol columns on the level grid uniform in
total optical depth, rt_atm.c:177-400) is the layered instance whose layers
are its medium's per-layer Rayleigh and aerosol depths; scaling the aerosol
column by s moves every level, so d rho / ds = sum over layers of the
adjoint's per-layer gradient times d(layer)/ds.  d(layer)/ds is taken by
central differences of the oracle's medium construction (no solve)."""
import ctypes as C
import json
import os

import numpy as np

from ocrt_amd.abi import make_instance

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "ref_jacobians.json")


def load():
    with open(FIXTURE) as fh:
        return json.load(fh)


def profile(dd, scale=1.0):
    return make_instance(sza=dd["sza_deg"], vza=dd["vza_deg"], raa=dd["raa_deg"], wavelength=dd["wavelength_nm"],
                         surface="black_fresnel_ocean", wind=dd["wind_speed"], tau_R=dd["tau_R"],
                         tau_a=dd["tau_a"] * scale, ssa_a=dd["ssa_a"], aer_h_km=dd["aer_h_km"],
                         greek=np.array(dd["greek"]), n_mu=dd["n_mu"], n_layers=dd["n_layers"], m_max=dd["m_max"],
                         ipss=0, max_orders=dd["max_orders"], tolerance=dd["tolerance"])


def medium_layers(oracle, d):
    """[4, nt] layers reproducing the medium of a profile instance."""
    L = d.n_layers + 1
    h, ch, x, y, z = (np.zeros(L) for _ in range(5))
    assert oracle.oracle_medium(C.byref(d), *(a.ctypes.data_as(C.c_void_p) for a in (h, ch, x, y, z))) == 0
    dt = np.diff(h)
    aer = x[1:] * dt / d.ssa_a
    ray = y[1:] * dt
    return np.stack([ray, aer, np.full_like(dt, d.ssa_a), np.clip(dt - ray - aer, 0.0, None)])


def layered(oracle, dd):
    """(layered descriptor, its layers, d(layers)/d(AOD scale))."""
    lay = medium_layers(oracle, profile(dd))
    h = 1e-5
    dlay = (medium_layers(oracle, profile(dd, 1 + h)) - medium_layers(oracle, profile(dd, 1 - h))) / (2 * h)
    d = make_instance(sza=dd["sza_deg"], vza=dd["vza_deg"], raa=dd["raa_deg"], wavelength=dd["wavelength_nm"],
                      surface="black_fresnel_ocean", wind=dd["wind_speed"], greek=np.array(dd["greek"]),
                      n_mu=dd["n_mu"], m_max=dd["m_max"], layers=lay, max_orders=dd["max_orders"],
                      tolerance=dd["tolerance"])
    return d, lay, dlay
