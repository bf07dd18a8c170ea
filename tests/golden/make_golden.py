"""Write the golden vectors of the test suite (run in the build container only).

Two sources, both data only:
  1. the reference OCRT solver itself, compiled from its own sources into
     oracle/_ref/libocrt_ref.so by `make -C oracle ref`, run on seeded inputs
     -> ref_cases.json (single-view solves) and ref_lut.json (full-grid solves);
  2. validation records the reference repository holds, sub-sampled
     -> ref_fixtures/*.csv (copied rows, unchanged values).

    python tests/golden/make_golden.py
"""
import csv
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ocrt_amd.abi import Instance, Result, RESULT_FIELDS, make_instance  # noqa: E402
from ocrt_amd import synthetic as S  # noqa: E402

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libocrt_ref.so")
REF_TREE = "/root/reference/code/OCRT_C"
REF_DOCS = "/root/reference/docs"

DESC_FIELDS = [f for f, _ in Instance._fields_ if f != "greek"]


def desc_to_dict(d, greek):
    out = {f: getattr(d, f) for f in DESC_FIELDS}
    out["greek"] = None if greek is None else np.asarray(greek).tolist()
    return out


def ref_solve(lib, d):
    r = Result()
    mm = d.m_max
    modes = (C.c_double * (3 * (mm + 1)))()
    orders = (C.c_int * (mm + 1))()
    resid = (C.c_double * (mm + 1))()
    rc = lib.ref_solve(C.byref(d), C.byref(r), modes, orders, resid)
    return rc, r, list(modes), list(orders), list(resid)


def cases():
    """(name, descriptor, greek) list covering configs 0-3 and the surface,
    IPSS and option variants the reference exercises."""
    out = []
    d, _ = S.config0()
    out.append(("c0_flat", d[0], None))
    out.append(("c0_black", make_instance(surface="black"), None))
    out.append(("c0_lambert", make_instance(surface="lambert"), None))
    out.append(("c0_bfo_w0", make_instance(surface="black_fresnel_ocean", wind=0.0, sza=55.0, vza=55.0,
                                           raa=180.0), None))
    out.append(("c0_bfo_w7_decoupled", make_instance(surface="black_fresnel_ocean", wind=7.0,
                                                     decouple_sunglint=1), None))
    out.append(("c0_bfo_w5_sigma0_q0", make_instance(surface="black_fresnel_ocean", wind=5.0, sigma_type=0,
                                                     q_convention=0, raa=150.0), None))
    out.append(("c0_ipss_sza75", make_instance(surface="black_fresnel_ocean", wind=5.0, sza=75.0, ipss=1,
                                               n_layers=40), None))
    out.append(("c0_max_orders_2", make_instance(surface="flat", max_orders=2), None))
    out.append(("c0_max_orders_1", make_instance(surface="black_fresnel_ocean", wind=4.0, max_orders=1), None))
    out.append(("c0_nadir_view", make_instance(surface="flat", vza=0.0, raa=0.0), None))
    ds, gk = S.config1(m_max=32)
    out.append(("c1_m32_ipss", ds[0], gk[0]))
    ds, gk = S.config1(m_max=8, ipss=0, n_layers=120, n_mu=24)
    out.append(("c1_m8_nt120", ds[0], gk[0]))
    gk = S.hg_greek(0.75, 60)
    out.append(("aer_flat_ipss", make_instance(surface="flat", tau_a=0.3, ssa_a=0.9, greek=gk, n_layers=60,
                                               m_max=6, n_mu=16, ipss=1, sza=62.0, vza=30.0, raa=20.0), gk))
    out.append(("aer_only_black", make_instance(surface="black", tau_R=0.0, tau_a=0.2, ssa_a=0.97, greek=gk,
                                                n_layers=50, m_max=6, n_mu=16), gk))
    ds3, gk3 = S.config3(n_inst=8)
    for i, d in enumerate(ds3):
        out.append((f"c3_{i}", d, gk3[i]))
    ds2, gk2 = S.config2(n_geom=2)
    for i, d in enumerate(ds2):
        out.append((f"c2_{i}", d, gk2[i // 8]))
    return out


def lut_cases():
    vza = [0.0, 10.0, 25.0, 40.0, 55.0, 70.0]
    raa = [0.0, 30.0, 90.0, 150.0, 180.0, 270.0]
    gk = S.hg_greek(0.7, 40)
    return [
        ("lut_black_ray", make_instance(surface="black", vza=0.0, raa=0.0, sza=60.0, tau_R=0.25, n_layers=40),
         None, vza, raa),
        ("lut_black_ray_ipss", make_instance(surface="black", vza=0.0, raa=0.0, sza=70.0, tau_R=0.25,
                                             n_layers=40, ipss=1), None, vza, raa),
        ("lut_bfo_w3", make_instance(surface="black_fresnel_ocean", wind=3.0, vza=0.0, raa=0.0, sza=40.0,
                                     tau_R=0.0935485118, n_layers=40), None, vza, raa),
        ("lut_aer_flat", make_instance(surface="flat", vza=0.0, raa=0.0, sza=35.0, tau_a=0.15, ssa_a=0.9,
                                       greek=gk, n_layers=30, m_max=8, n_mu=16), gk, vza, raa),
    ]


def copy_reference_fixtures():
    dst = os.path.join(HERE, "ref_fixtures")
    os.makedirs(dst, exist_ok=True)
    # IPSS geometry oracle emitted by the reference package (whole file, data)
    src = os.path.join(REF_TREE, "tests", "fixtures", "ipss_oracle_geometry.csv")
    with open(src) as f, open(os.path.join(dst, "ipss_oracle_geometry.csv"), "w") as g:
        g.write(f.read())
    # the non-IPSS gate of the reference (gas absorption on, whole file, gzip)
    import gzip
    src = os.path.join(REF_TREE, "tests", "fixtures", "ipss_nonpssa_reference.csv")
    with open(src, "rb") as f, gzip.open(os.path.join(dst, "ipss_nonpssa_reference.csv.gz"), "wb", 9) as g:
        g.write(f.read())
    # IPSS paper cases: sub-sample VZA {0,20,40,60,70} x RAA {0,45,90,135,180}
    keep_v = {0.0, 20.0, 40.0, 60.0, 70.0}
    keep_r = {0.0, 45.0, 90.0, 135.0, 180.0}
    cdir = os.path.join(REF_TREE, "validation", "ipss_2026-09-05", "cases")
    for tau in ("t025", "t100"):
        for s in ("s0p0", "s70p47", "s84p26"):
            for mode in ("pp", "ipss"):
                name = f"case_{tau}_{s}_{mode}.csv"
                with open(os.path.join(cdir, name)) as f:
                    rows = list(csv.reader(f))
                head, body = rows[0], rows[1:]
                sel = [r for r in body if float(r[2]) in keep_v and float(r[3]) in keep_r]
                with open(os.path.join(dst, name), "w", newline="") as g:
                    w = csv.writer(g)
                    w.writerow(head)
                    w.writerows(sel)
    src = os.path.join(REF_DOCS, "validation", "cross_check_fr631_20260823", "atm_rhoIQU_bfo.csv")
    with open(src) as f, open(os.path.join(dst, "atm_rhoIQU_bfo.csv"), "w") as g:
        g.write(f.read())


def main():
    lib = C.CDLL(REF_LIB)
    recs = []
    for name, d, greek in cases():
        rc, r, modes, orders, resid = ref_solve(lib, d)
        recs.append({"name": name, "desc": desc_to_dict(d, greek), "rc": rc,
                     "result": {f: getattr(r, f) for f in RESULT_FIELDS},
                     "toa_modes": modes, "orders_per_m": orders, "residual_per_m": resid})
        print(name, rc, r.rho_I, r.n_orders_used, flush=True)
    with open(os.path.join(HERE, "ref_cases.json"), "w") as f:
        json.dump(recs, f, indent=0)
    luts = []
    for name, d, greek, vza, raa in lut_cases():
        nv, nr = len(vza), len(raa)
        arrs = {k: (C.c_double * (nv * nr))() for k in ("rho_I", "rho_Q", "rho_U", "T_diff_dn_dir",
                                                        "T_sg_up_dir", "T_total_up_dir")}
        tdd, tdh = C.c_double(), C.c_double()
        rc = lib.ref_solve_lut(C.byref(d), nv, (C.c_double * nv)(*vza), nr, (C.c_double * nr)(*raa),
                               arrs["rho_I"], arrs["rho_Q"], arrs["rho_U"], arrs["T_diff_dn_dir"],
                               arrs["T_sg_up_dir"], arrs["T_total_up_dir"], C.byref(tdd), C.byref(tdh))
        luts.append({"name": name, "desc": desc_to_dict(d, greek), "rc": rc, "vza": vza, "raa": raa,
                     "out": {k: list(v) for k, v in arrs.items()},
                     "T_dir_dn": tdd.value, "T_diff_dn_hemi": tdh.value})
        print(name, rc, flush=True)
    with open(os.path.join(HERE, "ref_lut.json"), "w") as f:
        json.dump(luts, f, indent=0)
    copy_reference_fixtures()


if __name__ == "__main__":
    main()
