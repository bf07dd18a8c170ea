"""Write golden vectors of the in-water SOS (run in the build container only):
the reference's own order loops for the water body -- rt_solver_sos_pol,
rt_solver_sos_pol_intrefl (flat water-side reflection at the top) and
rt_solver_sos_pol_intrefl_rough (Cox-Munk water-side kernels) -- driven
through oracle/ref_harness.c:ref_water_solve on in-water columns filled as the
reference fills its water medium (rt_water_rt.c:2104-2236): pure water and
water with a forward-peaked hydrosol, refracted sun, black bottom.

    python tests/golden/make_water_golden.py  ->  ref_water_sos.json
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

from ocrt_amd.abi import make_instance  # noqa: E402
from ocrt_amd import synthetic as S  # noqa: E402
from make_golden import desc_to_dict  # noqa: E402

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libocrt_ref.so")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
OUT = os.path.join(HERE, "ref_water_sos.json")


def gl_nodes(n):
    lib = C.CDLL(ORACLE_LIB)
    mu, w = np.zeros(n), np.zeros(n)
    lib.oracle_gauss_legendre_pos(n, mu.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p))
    return mu


def two_term_hg(g1, g2, w, L):
    """Greek set of a two-term HG hydrosol (weight w on g1) with Rayleigh-like
    polarisation (P12 = -0.3 (1 - mu^2) P11, P33 = mu P11 projected)."""
    a = S.hg_greek(g1, L, dolp=0.3)
    b = S.hg_greek(g2, L, dolp=0.3)
    return np.ascontiguousarray(w * a + (1.0 - w) * b)


def cases():
    hyd = two_term_hg(0.95, -0.3, 0.98, 120)
    out = []
    out.append(("w_pure_black_top", dict(water=(20.0, 0.28, 0.0), top_reflect=0, n_mu=16, n_layers=60, m_max=4,
                                         sza=30.0), 10))
    out.append(("w_pure_flat_top", dict(water=(20.0, 0.28, 0.0), top_reflect=1, n_mu=16, n_layers=60, m_max=4,
                                        sza=30.0), 10))
    out.append(("w_hydrosol_flat", dict(water=(25.0, 0.05, 0.6), top_reflect=1, n_mu=24, n_layers=100, m_max=8,
                                        sza=20.0, greek=hyd), 14))
    out.append(("w_hydrosol_rough", dict(water=(25.0, 0.05, 0.6), top_reflect=2, wind=6.0, n_mu=24, n_layers=100,
                                         m_max=8, sza=35.0, greek=hyd), 14))
    out.append(("w_pure_rough_sigma0", dict(water=(15.0, 0.4, 0.0), top_reflect=2, wind=11.0, sigma_type=0,
                                            n_mu=16, n_layers=50, m_max=3, sza=40.0), 6))
    return out


def main():
    lib = C.CDLL(REF_LIB)
    recs = []
    for name, kw, vnode in cases():
        n, mm = kw["n_mu"], kw["m_max"]
        mu = gl_nodes(n)
        # the view sits on a Gauss node, so the ring is the Gauss rule itself
        vza = math.degrees(math.acos(mu[vnode]))
        d = make_instance(surface="black", vza=vza, raa=0.0, depol=0.039, n_water=1.34, max_orders=80,
                          tolerance=1e-7, **kw)
        top = np.zeros((mm + 1) * 3 * n)
        bot = np.zeros((mm + 1) * 3 * n)
        orders = np.zeros(mm + 1, dtype=np.int32)
        rc = lib.ref_water_solve(C.byref(d), top.ctypes.data_as(C.c_void_p), bot.ctypes.data_as(C.c_void_p),
                                 orders.ctypes.data_as(C.c_void_p))
        dd = desc_to_dict(d, kw.get("greek"))
        for k in ("n_ang_phase", "theta_phase", "p11_phase", "layers", "gas", "bottom_source", "top_source"):
            dd.pop(k, None)
        recs.append({"name": name, "desc": dd, "rc": rc, "n": n, "top_up": top.tolist(), "bot_dn": bot.tolist(),
                     "orders_per_m": orders.tolist()})
        print(name, rc, top[:2], orders.tolist())
    with open(OUT, "w") as f:
        json.dump(recs, f)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
