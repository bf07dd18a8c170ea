"""Write golden vectors of the atmosphere solve driven by an external bottom
source (run in the build container only): the reference solver, compiled
from its own sources into oracle/_ref/libocrt_ref.so, solves seeded cases with
rt_options_t.ext_bottom_per_m_I/Q/U and bottom_source_only set, the coupled
ocean's pass 2 (code/OCRT_C/src/rt_types.h:722-747; rt_solver.c:949-977;
rt_surface_boundary.c:495-567).  The sources are smooth synthetic upward
Fourier fields at the ring nodes.

    python tests/golden/make_bottom_source_golden.py  ->  ref_bottom_source.json
"""
import ctypes as C
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ocrt_amd.abi import Result, RESULT_FIELDS, make_instance  # noqa: E402
from ocrt_amd import synthetic as S  # noqa: E402
from make_golden import desc_to_dict  # noqa: E402

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libocrt_ref.so")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
OUT = os.path.join(HERE, "ref_bottom_source.json")


def ring(n_mu, vza):
    lib = C.CDLL(ORACLE_LIB)
    mu, w, vi = np.zeros(n_mu + 1), np.zeros(n_mu + 1), C.c_int()
    n = lib.oracle_ring(n_mu, C.c_double(math.cos(math.radians(vza))), mu.ctypes.data_as(C.c_void_p),
                        w.ctypes.data_as(C.c_void_p), C.byref(vi))
    return mu[:n]


def source(mu, m_src, scale=0.02):
    """[3, m_src + 1, n] upward I, Q, U: smooth in mu, decaying in m."""
    m = np.arange(m_src + 1)[:, None]
    x = mu[None, :]
    I = scale * x * np.exp(-m / 2.0) * (1.0 + 0.3 * x)
    Q = -0.2 * scale * (1 - x * x) * np.exp(-m / 3.0)
    U = 0.15 * scale * x * (1 - x) * np.exp(-m / 2.5) * (m > 0)
    return np.stack([I, Q, U])


def cases():
    gk = S.hg_greek(0.7, 40)
    out = []
    out.append(("bs_rough_only", dict(surface="black_fresnel_ocean", wind=5.0, tau_a=0.2, ssa_a=0.95, greek=gk,
                                      n_mu=16, n_layers=50, m_max=6, sza=50.0, vza=30.0, raa=60.0,
                                      decouple_sunglint=1), 6, 1))
    out.append(("bs_flat_only_fewer_modes", dict(surface="flat", tau_a=0.1, ssa_a=0.9, greek=gk, n_mu=12,
                                                 n_layers=40, m_max=6, sza=35.0, vza=48.0, raa=120.0), 3, 1))
    out.append(("bs_rough_plus_sun", dict(surface="black_fresnel_ocean", wind=7.0, tau_a=0.15, ssa_a=0.93,
                                          greek=gk, n_mu=16, n_layers=60, m_max=8, sza=42.0, vza=22.0, raa=100.0,
                                          decouple_sunglint=1), 8, 0))
    out.append(("bs_rayleigh_rough_glint", dict(surface="black_fresnel_ocean", wind=3.0, n_mu=24, n_layers=40,
                                                m_max=4, sza=30.0, vza=30.0, raa=180.0), 4, 1))
    out.append(("bs_black_ignored", dict(surface="black", tau_a=0.2, ssa_a=0.95, greek=gk, n_mu=12, n_layers=30,
                                         m_max=3, sza=20.0, vza=10.0, raa=30.0), 3, 1))
    out.append(("bs_c3_shape_ipss", dict(surface="black_fresnel_ocean", wind=6.0, tau_a=0.12, ssa_a=0.93,
                                         greek=S.hg_greek(0.75, 80), n_mu=32, n_layers=200, m_max=16, sza=55.0,
                                         vza=33.0, raa=95.0, ipss=1, decouple_sunglint=1), 16, 1))
    return out


def main():
    lib = C.CDLL(REF_LIB)
    recs = []
    for name, kw, m_src, only in cases():
        mu = ring(kw["n_mu"], kw["vza"])
        src = source(mu, m_src)
        d = make_instance(bottom_source=src, bottom_source_only=only, **kw)
        mm = d.m_max
        r = Result()
        modes = (C.c_double * (3 * (mm + 1)))()
        orders = (C.c_int * (mm + 1))()
        rc = lib.ref_solve(C.byref(d), C.byref(r), modes, orders, None)
        dd = desc_to_dict(d, kw.get("greek"))
        for k in ("n_ang_phase", "theta_phase", "p11_phase", "layers", "gas", "bottom_source", "bottom_n_mu",
                  "bottom_m_max", "bottom_source_only"):
            dd.pop(k, None)
        recs.append({"name": name, "desc": dd, "bottom_source": src.tolist(), "bottom_source_only": only,
                     "rc": rc, "result": {f: getattr(r, f) for f in RESULT_FIELDS},
                     "toa_modes": list(modes), "orders_per_m": list(orders)})
        print(name, rc, r.rho_I, r.rho_Q, r.rho_U, list(orders))
    with open(OUT, "w") as f:
        json.dump(recs, f)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
