"""The oracle (C restatement) against the reference: golden vectors produced by
the reference solver itself and the validation records the reference holds."""
import csv
import ctypes as C
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, desc_from_dict, oracle_solve
from ocrt_amd.abi import make_instance

FIX = os.path.join(GOLDEN, "ref_fixtures")
with open(os.path.join(GOLDEN, "ref_cases.json")) as _f:
    N_CASES = len(__import__("json").load(_f))


def rel(a, b, floor=1e-300):
    return abs(a - b) / max(abs(b), floor)


@pytest.mark.parametrize("idx", range(N_CASES))
def test_oracle_matches_reference_single_view(oracle, golden_cases, idx):
    rec = golden_cases[idx]
    d, keep = desc_from_dict(rec["desc"])
    rc, r, modes, orders = oracle_solve(oracle, d)
    assert rc == rec["rc"] == 0
    ref = rec["result"]
    scale = max(abs(ref["rho_I"]), 1e-30)
    for f in ("rho_I", "rho_Q", "rho_U"):
        assert abs(getattr(r, f) - ref[f]) <= 1e-12 * scale, (rec["name"], f, getattr(r, f), ref[f])
    for f in ("T_dir_dn", "T_diff_dn_hemi", "T_total_dn_hemi", "T_diff_dn_dir", "T_dir_up_view",
              "T_total_up_dir"):
        assert rel(getattr(r, f), ref[f]) <= 1e-12, (rec["name"], f)
    for f in ("T_sg_up_dir", "rho_glint_direct_I", "rho_glint_direct_Q", "rho_glint_direct_U"):
        assert abs(getattr(r, f) - ref[f]) <= 1e-12 * max(abs(ref[f]), 1e-12), (rec["name"], f)
    assert r.n_orders_used == ref["n_orders_used"] and r.converged == ref["converged"]
    rm = np.array(rec["toa_modes"]).reshape(-1, 3)
    mscale = np.max(np.abs(rm[:, 0])) + 1e-30
    assert np.max(np.abs(modes - rm)) <= 1e-12 * mscale
    assert list(orders) == rec["orders_per_m"]


def test_oracle_matches_reference_lut(oracle, golden_luts):
    for rec in golden_luts:
        d, keep = desc_from_dict(rec["desc"])
        vza, raa = rec["vza"], rec["raa"]
        nv, nr = len(vza), len(raa)
        outs = {k: np.zeros(nv * nr) for k in rec["out"]}
        tdd, tdh = C.c_double(), C.c_double()
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        rc = oracle.oracle_solve_lut(C.byref(d), nv, (C.c_double * nv)(*vza), nr, (C.c_double * nr)(*raa),
                                     p(outs["rho_I"]), p(outs["rho_Q"]), p(outs["rho_U"]),
                                     p(outs["T_diff_dn_dir"]), p(outs["T_sg_up_dir"]),
                                     p(outs["T_total_up_dir"]), C.byref(tdd), C.byref(tdh))
        assert rc == 0
        scale = np.max(np.abs(rec["out"]["rho_I"]))
        for k, v in outs.items():
            ref = np.array(rec["out"][k])
            assert np.max(np.abs(v - ref)) <= 1e-12 * max(scale, np.max(np.abs(ref))), (rec["name"], k)
        assert rel(tdd.value, rec["T_dir_dn"]) < 1e-13 and rel(tdh.value, rec["T_diff_dn_hemi"]) < 1e-12


def _read(path):
    with open(path) as f:
        rows = list(csv.reader(f))
    return rows[0], [[float(x) for x in r] for r in rows[1:]]


@pytest.mark.parametrize("tau", ["t025", "t100"])
@pytest.mark.parametrize("sza", ["s0p0", "s70p47", "s84p26"])
@pytest.mark.parametrize("mode", ["pp", "ipss"])
def test_oracle_vs_reference_ipss_paper_cases(oracle, tau, sza, mode):
    """validation/ipss_2026-09-05/cases: Rayleigh only, black surface, gas off,
    400 layers, 24 streams, full-grid output, IPSS off (pp) or on.  The column
    optical depth is recovered from the SZA-0 direct transmittance column."""
    head, rows = _read(os.path.join(FIX, f"case_{tau}_s0p0_pp.csv"))
    tau_R = -math.log(rows[0][7])
    head, rows = _read(os.path.join(FIX, f"case_{tau}_{sza}_{mode}.csv"))
    s = rows[0][0]
    vza = sorted({r[2] for r in rows})
    raa = sorted({r[3] for r in rows})
    d = make_instance(surface="black", sza=s, vza=0.0, raa=0.0, tau_R=tau_R, n_layers=400,
                      ipss=1 if mode == "ipss" else 0)
    nv, nr = len(vza), len(raa)
    out = {k: np.zeros(nv * nr) for k in ("I", "Q", "U", "dd", "sg", "tu")}
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    tdd, tdh = C.c_double(), C.c_double()
    rc = oracle.oracle_solve_lut(C.byref(d), nv, (C.c_double * nv)(*vza), nr, (C.c_double * nr)(*raa),
                                 p(out["I"]), p(out["Q"]), p(out["U"]), p(out["dd"]), p(out["sg"]),
                                 p(out["tu"]), C.byref(tdd), C.byref(tdh))
    assert rc == 0
    worst = 0.0
    for r in rows:
        ix = vza.index(r[2]) * nr + raa.index(r[3])
        worst = max(worst, rel(out["I"][ix], r[4]), abs(out["Q"][ix] - r[5]) / abs(r[4]),
                    abs(out["U"][ix] - r[6]) / abs(r[4]), rel(out["dd"][ix], r[10]))
        assert rel(tdd.value, r[7]) < 1e-9 and rel(tdh.value, r[9]) < 1e-8
    # the fixture carries 11 significant digits
    assert worst < 2e-9, worst


def test_oracle_vs_reference_bfo_cross_check(oracle):
    """docs/validation/cross_check_fr631_20260823/atm_rhoIQU_bfo.csv: Rayleigh
    tau 0.0935485118, rough ocean wind 3, direct glint decoupled, 40 layers,
    SZA 40, full-grid sampling (9 significant digits)."""
    head, rows = _read(os.path.join(FIX, "atm_rhoIQU_bfo.csv"))
    vza = sorted({abs(r[2]) for r in rows})
    raa = sorted({r[3] for r in rows})
    d = make_instance(surface="black_fresnel_ocean", wind=3.0, decouple_sunglint=1, sza=40.0, vza=0.0,
                      raa=0.0, tau_R=0.0935485118, n_layers=40)
    nv, nr = len(vza), len(raa)
    out = {k: np.zeros(nv * nr) for k in ("I", "Q", "U")}
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = oracle.oracle_solve_lut(C.byref(d), nv, (C.c_double * nv)(*vza), nr, (C.c_double * nr)(*raa),
                                 p(out["I"]), p(out["Q"]), p(out["U"]), None, None, None, None, None)
    assert rc == 0
    for r in rows:
        ix = vza.index(abs(r[2])) * nr + raa.index(r[3])
        assert rel(out["I"][ix], r[4]) < 1e-8
        assert abs(out["Q"][ix] - r[5]) < 1e-8 * abs(r[4])
        assert abs(out["U"][ix] - r[6]) < 1e-8 * abs(r[4])


def test_oracle_ipss_geometry_vs_reference_fixture(oracle):
    """tests/fixtures/ipss_oracle_geometry.csv of the reference (gate G-I0a/c:
    rel 1e-9 and identical shadow classification) on the homogeneous
    tau 0.25, 100-layer test atmosphere."""
    with open(os.path.join(FIX, "ipss_oracle_geometry.csv")) as f:
        rows = list(csv.reader(f))[1:]
    nv = ns = 0
    for kind, a, b, c, dd, val in rows:
        a, b, c, dd, val = map(float, (a, b, c, dd, val))
        if kind == "tau_view":
            t = oracle.oracle_ipss_homog_tau_view(C.c_double(0.25), 100, C.c_double(a))
            assert rel(t, val, 1.0) < 1e-9
            nv += 1
        elif kind == "tau_sun":
            rb = (C.c_double * 3)(a, b, c)
            t = oracle.oracle_ipss_homog_tau_sun(C.c_double(0.25), 100, C.c_double(dd), rb)
            assert (val < 0.0) == (t < 0.0)
            if val >= 0.0:
                assert abs(t - val) / (val if val > 1e-12 else 1.0) < 1e-9
            ns += 1
    assert nv >= 9 and ns > 100
