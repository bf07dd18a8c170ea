"""Write tests/golden/ref_atmos_batch.json from the reference Python package's
own batch entry point (code/OCRT_Python/ocrt_py/atmos_batch.py:885,
solve_atm_batch) on media made by its own builders (ocrt_py/atmos.py:33,
build_atm_rayleigh; :314, build_atm_aerosol): an aerosol + Rayleigh batch, an
aerosol-only batch (tau_R = 0) and a Rayleigh-only batch.  Run in the build
container (the reference tree is not on the GPU box); the fixture stores
every medium array the GPU mirror reads, the call's arguments and the
reference's outputs.

    python tests/golden/make_atmos_batch_golden.py
"""
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, "/root/reference/code/OCRT_Python")
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from ocrt_py import atmos as RA            # noqa: E402
from ocrt_py import atmos_batch as RB      # noqa: E402
from ocrt_amd import synthetic as S        # noqa: E402

MEDIUM_KEYS = ("rm", "gb", "h", "xdel", "ydel", "beta2")
GREEK_KEYS = ("betal", "gammal", "alphal", "zetal")


def arr(x):
    return np.asarray(x, dtype=np.float64).tolist()


def batch(name, mode, tau_R, tau_a, cases, nt=30, n_gl=12, m_max=5, L_max=40):
    gk = S.hg_greek(0.72, L_max)
    atms, vjs, mus, dphis, winds = [], [], [], [], []
    for sza, vza, raa, ws, ssa in cases:
        mu_sun, mu_view = math.cos(math.radians(sza)), math.cos(math.radians(vza))
        if mode == "aerosol_black_fresnel_ocean":
            atm, vj, _ = RA.build_atm_aerosol(nt, tau_R, tau_a, ssa, mu_sun, mu_view, n_gl, gk[0].copy(),
                                              gk[1].copy(), gk[2].copy(), gk[3].copy(), L_max)
        else:
            atm, vj, _ = RA.build_atm_rayleigh(nt, tau_R, mu_sun, mu_view, n_gl)
        atms.append(atm)
        vjs.append(int(vj))
        mus.append(mu_sun)
        dphis.append(math.radians(raa))
        winds.append(ws)
    n_mu = len(np.asarray(atms[0].rm)) // 2
    args = dict(n_mu=n_mu, nt=nt, m_max=m_max, max_iterations=30, tolerance=1e-7, mode=mode, L_max=L_max)
    out = RB.solve_atm_batch(atms, vjs, mus, dphis, winds, n_mu, nt, m_max, 30, 1e-7, mode=mode, L_max=L_max)
    media = []
    for atm in atms:
        m = {k: arr(getattr(atm, k)) for k in MEDIUM_KEYS}
        if mode == "aerosol_black_fresnel_ocean":
            m.update({k: arr(getattr(atm, k)) for k in GREEK_KEYS})
        media.append(m)
    keys = ("rho_I", "rho_Q", "rho_U", "I_TOA", "Q_TOA", "U_TOA", "boa_I", "boa_Q", "boa_U", "T_diff_dn_hemi",
            "n_orders")
    return dict(name=name, args=args, vjs=vjs, mu_suns=mus, dphis=dphis, winds=winds, media=media,
                out={k: arr(out[k]) for k in keys if k in out})


def main():
    geo = [(35.0, 20.0, 60.0, 4.0, 0.9), (62.0, 48.0, 150.0, 9.0, 0.95), (15.0, 5.0, 10.0, 1.5, 0.85)]
    rec = [batch("aerosol_rayleigh", "aerosol_black_fresnel_ocean", 0.1, 0.2, geo),
           batch("aerosol_only", "aerosol_black_fresnel_ocean", 0.0, 0.2, geo),
           batch("rayleigh_only", "black_fresnel_ocean", 0.1, 0.0, geo)]
    path = os.path.join(HERE, "ref_atmos_batch.json")
    with open(path, "w") as f:
        json.dump(rec, f)
    for r in rec:
        print(r["name"], "rho_I", r["out"]["rho_I"], "n_orders", r["out"].get("n_orders"))
    print("wrote", path)


if __name__ == "__main__":
    main()
