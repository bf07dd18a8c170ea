"""Write golden vectors of the aerosol value kernel (run in the build container
only): the reference solver, compiled from its own sources into
oracle/_ref/libocrt_ref.so, solves seeded cases whose aerosol carries an
angle-space P11 table next to its Greek coefficients
(rt_aerosol_input_t.use_value_kernel, code/OCRT_C/src/rt_solver.h:102-113).
The tables are synthetic (no Mie files are needed): Henyey-Greenstein
functions, a two-term HG and an HG with a backscatter glory, on 0.5 and 1
degree grids, each normalised so that (1/2) int P11 dmu = 1.

    python tests/golden/make_value_kernel_golden.py  ->  ref_value_kernel.json
"""
import ctypes as C
import json
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
OUT = os.path.join(HERE, "ref_value_kernel.json")


def table(kind, step):
    """(theta_deg, P11) of a synthetic phase function, normalised."""
    th = np.arange(0.0, 180.0 + 1e-9, step)
    mu = np.cos(np.radians(th))

    def hg(g):
        return (1 - g * g) / (1 + g * g - 2 * g * mu) ** 1.5
    if kind == "hg70":
        p = hg(0.70)
    elif kind == "tthg":
        p = 0.9 * hg(0.8) + 0.1 * hg(-0.4)
    else:   # HG plus a backscatter glory lobe
        p = hg(0.65) + 0.8 * np.exp(-((180.0 - th) / 4.0) ** 2)
    # normalise with the trapezoid rule in mu on a fine resampling
    x = np.linspace(-1.0, 1.0, 200001)
    pf = np.interp(np.degrees(np.arccos(x))[::-1], th, p)[::-1]
    norm = 0.5 * np.trapezoid(pf, x)
    return th, p / norm


def cases():
    out = []
    gk60 = S.hg_greek(0.70, 60)
    gk80 = S.hg_greek(0.75, 80)
    out.append(("vk_hg_rough_ipss", dict(surface="black_fresnel_ocean", wind=5.0, tau_a=0.2, ssa_a=0.95, greek=gk60,
                                         n_mu=16, n_layers=60, m_max=8, ipss=1, sza=50.0, vza=30.0, raa=60.0),
                ("hg70", 0.5)))
    out.append(("vk_hg_flat", dict(surface="flat", tau_a=0.3, ssa_a=0.9, greek=gk60, n_mu=12, n_layers=40, m_max=6,
                                   sza=35.0, vza=48.0, raa=120.0), ("hg70", 0.5)))
    out.append(("vk_glory_black", dict(surface="black", tau_R=0.0, tau_a=0.25, ssa_a=0.97, greek=gk80, n_mu=16,
                                       n_layers=50, m_max=10, sza=20.0, vza=25.0, raa=175.0), ("glory", 0.5)))
    out.append(("vk_tthg_rough", dict(surface="black_fresnel_ocean", wind=9.0, tau_a=0.15, ssa_a=0.92, greek=gk80,
                                      n_mu=20, n_layers=80, m_max=12, sza=60.0, vza=10.0, raa=30.0), ("tthg", 1.0)))
    out.append(("vk_glory_ipss_sza70", dict(surface="black_fresnel_ocean", wind=3.0, tau_a=0.1, ssa_a=0.99,
                                            greek=gk80, n_mu=16, n_layers=100, m_max=8, ipss=1, sza=70.0, vza=40.0,
                                            raa=0.0), ("glory", 1.0)))
    out.append(("vk_c3_shape", dict(surface="black_fresnel_ocean", wind=6.0, tau_a=0.12, ssa_a=0.93, greek=gk80,
                                    n_mu=32, n_layers=200, m_max=16, sza=42.0, vza=33.0, raa=95.0), ("hg70", 0.5)))
    return out


def main():
    lib = C.CDLL(REF_LIB)
    recs = []
    for name, kw, (tk, step) in cases():
        th, p11 = table(tk, step)
        d = make_instance(phase_table=(th, p11), **kw)
        mm = d.m_max
        r = Result()
        modes = (C.c_double * (3 * (mm + 1)))()
        orders = (C.c_int * (mm + 1))()
        rc = lib.ref_solve(C.byref(d), C.byref(r), modes, orders, None)
        greek = kw.get("greek")
        dd = desc_to_dict(d, greek)
        for k in ("n_ang_phase", "theta_phase", "p11_phase", "layers", "gas"):
            dd.pop(k, None)
        recs.append({"name": name, "desc": dd, "table": {"theta": th.tolist(), "p11": p11.tolist()},
                     "rc": rc, "result": {f: getattr(r, f) for f in RESULT_FIELDS},
                     "toa_modes": list(modes), "orders_per_m": list(orders)})
        print(name, rc, r.rho_I, r.rho_Q, r.rho_U, list(orders))
    with open(OUT, "w") as f:
        json.dump(recs, f)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
