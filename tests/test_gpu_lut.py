"""LUT / full-grid mode on the GPU (rt_solve_case_pol_lut): every direction of
the grid against the reference's own outputs.

- the golden LUT vectors written by the reference solver (FP64, 1e-11);
- the reference's validation cases (Rayleigh tau 0.25 / 1.0, black surface,
  400 layers, IPSS off and on; rough-ocean cross-check at wind 3), at the
  precision their CSV files carry;
- the mixed form within its error budget of the FP64 form.
"""
import math
import os

import numpy as np
import pytest

from conftest import desc_from_dict
from ocrt_amd.abi import make_instance

pytestmark = pytest.mark.gpu
FIX = os.path.join(os.path.dirname(__file__), "golden", "ref_fixtures")


@pytest.fixture(scope="module")
def ocrt():
    import ocrt_amd
    ocrt_amd.library()
    return ocrt_amd


def lut_solve(ocrt, descs, vza, raa, precision="fp64"):
    p = ocrt.Plan(descs, precision=precision)
    p.set_lut(vza, raa)
    p.execute()
    out = p.lut()
    res = p.results()
    p.close()
    return out, res


def test_lut_matches_reference_golden(ocrt, golden_luts):
    for rec in golden_luts:
        d, keep = desc_from_dict(rec["desc"])
        print(f"[gpu] LUT {rec['name']}", flush=True)
        out, _ = lut_solve(ocrt, [d], rec["vza"], rec["raa"])
        nv, nr = len(rec["vza"]), len(rec["raa"])
        scale = np.max(np.abs(rec["out"]["rho_I"]))
        for k in rec["out"]:
            ref = np.array(rec["out"][k]).reshape(nv, nr)
            assert np.max(np.abs(out[k][0] - ref)) <= 1e-11 * max(scale, np.max(np.abs(ref))), (rec["name"], k)
        assert abs(out["T_dir_dn"][0] - rec["T_dir_dn"]) <= 1e-13 * rec["T_dir_dn"]
        assert abs(out["T_diff_dn_hemi"][0] - rec["T_diff_dn_hemi"]) <= 1e-11 * rec["T_diff_dn_hemi"]


def _read(path):
    import csv
    with open(path) as f:
        rows = list(csv.reader(f))
    return [[float(x) for x in r] for r in rows[1:]]


@pytest.mark.parametrize("mode", ["pp", "ipss"])
def test_lut_reference_validation_cases(ocrt, mode):
    """validation/ipss_2026-09-05/cases (subsampled copies): all six
    (tau, SZA) cases of one mode in one batched plan."""
    cases = []
    for tau in ("t025", "t100"):
        tau_R = -math.log(_read(os.path.join(FIX, f"case_{tau}_s0p0_pp.csv"))[0][7])
        for sza in ("s0p0", "s70p47", "s84p26"):
            cases.append((tau_R, _read(os.path.join(FIX, f"case_{tau}_{sza}_{mode}.csv"))))
    vza = sorted({r[2] for _, rows in cases for r in rows})
    raa = sorted({r[3] for _, rows in cases for r in rows})
    descs = [make_instance(surface="black", sza=rows[0][0], vza=0.0, raa=0.0, tau_R=tau_R, n_layers=400,
                           ipss=1 if mode == "ipss" else 0) for tau_R, rows in cases]
    out, _ = lut_solve(ocrt, descs, vza, raa)
    worst = 0.0
    for i, (tau_R, rows) in enumerate(cases):
        for r in rows:
            a, b = vza.index(r[2]), raa.index(r[3])
            worst = max(worst, abs(out["rho_I"][i, a, b] - r[4]) / abs(r[4]),
                        abs(out["rho_Q"][i, a, b] - r[5]) / abs(r[4]),
                        abs(out["rho_U"][i, a, b] - r[6]) / abs(r[4]),
                        abs(out["T_diff_dn_dir"][i, a, b] - r[10]) / abs(r[10]))
            assert abs(out["T_dir_dn"][i] - r[7]) <= 1e-9 * r[7]
            assert abs(out["T_diff_dn_hemi"][i] - r[9]) <= 1e-8 * r[9]
    print(f"[gpu] validation cases ({mode}) worst rel {worst:.2e}", flush=True)
    assert worst < 2e-9


def test_lut_bfo_cross_check(ocrt):
    rows = _read(os.path.join(FIX, "atm_rhoIQU_bfo.csv"))
    vza = sorted({abs(r[2]) for r in rows})
    raa = sorted({r[3] for r in rows})
    d = make_instance(surface="black_fresnel_ocean", wind=3.0, decouple_sunglint=1, sza=40.0, vza=0.0, raa=0.0,
                      tau_R=0.0935485118, n_layers=40)
    out, _ = lut_solve(ocrt, [d], vza, raa)
    for r in rows:
        a, b = vza.index(abs(r[2])), raa.index(r[3])
        assert abs(out["rho_I"][0, a, b] - r[4]) < 1e-8 * abs(r[4])
        assert abs(out["rho_Q"][0, a, b] - r[5]) < 1e-8 * abs(r[4])
        assert abs(out["rho_U"][0, a, b] - r[6]) < 1e-8 * abs(r[4])


def test_lut_mixed_within_budget(ocrt, golden_luts):
    from ocrt_amd import synthetic as S
    gk = S.hg_greek(0.7, 40)
    descs = [make_instance(surface="black_fresnel_ocean", wind=6.0, tau_a=0.2, ssa_a=0.93, greek=gk, n_mu=16,
                           n_layers=80, m_max=8, sza=45.0, vza=20.0, raa=30.0),
             make_instance(surface="flat", tau_a=0.1, ssa_a=0.97, greek=gk, n_mu=16, n_layers=60, m_max=6,
                           sza=30.0, vza=40.0, raa=100.0, ipss=1)]
    vza = np.arange(0.0, 80.0, 10.0)
    raa = np.arange(0.0, 360.0, 30.0)
    a, _ = lut_solve(ocrt, descs, vza, raa, "fp64")
    b, _ = lut_solve(ocrt, descs, vza, raa, "mixed")
    ei = np.max(np.abs(b["rho_I"] - a["rho_I"]) / np.abs(a["rho_I"]))
    eq = max(np.max(np.abs(b["rho_Q"] - a["rho_Q"])), np.max(np.abs(b["rho_U"] - a["rho_U"])))
    print(f"[gpu] LUT mixed vs fp64: rel I {ei:.2e}, abs Q/U {eq:.2e}", flush=True)
    assert ei < 1e-4 and eq < 1e-4
