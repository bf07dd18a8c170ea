"""Reference-compatible boundary (include/ocrt_rt.h) without a GPU.

- Every field of rt_case_t, rt_options_t, rt_result_t, rt_aerosol_input_t,
  rt_lut_grid_out_t and rt_solver_sos_options_t sits at the offset the
  reference's headers give it (tests/golden/rt_layout.json, written from
  code/OCRT_C/src/rt_solver.h by tests/golden/make_rt_layout.py; re-derived
  from the reference here when its tree is present), and the ctypes view
  read from the header agrees.
- libocrt_amd.so exports rt_solve_case_pol, rt_solve_case_pol_aerosol,
  rt_solve_case_pol_lut and every other function the header declares.
- The Rayleigh optical depth of Bodhaine et al. (1999) equals the reference's
  rt_rayleigh_tau_model_full (compiled from its sources into oracle/_ref) and
  follows the closed-form scalings tau ~ P and tau ~ 1/g.
- Requests the solver does not serve are refused with -1 before any device
  work.
"""
import ctypes as C
import json
import math
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_rt_layout  # noqa: E402

from ocrt_amd import rt_abi  # noqa: E402

REF_SRC = "/root/reference/code/OCRT_C/src"
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libocrt_ref.so")
FIXTURE = os.path.join(ROOT, "tests", "golden", "rt_layout.json")


def test_header_layout_equals_reference_fixture():
    want = json.load(open(FIXTURE))
    got = make_rt_layout.layout(os.path.join(ROOT, "include"), "ocrt_rt.h")
    for s in want:
        assert got[s]["size"] == want[s]["size"], s
        assert got[s]["fields"] == want[s]["fields"], s
    assert sum(len(v["fields"]) for v in want.values()) >= 250


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference headers not present")
def test_fixture_equals_reference_headers():
    want = json.load(open(FIXTURE))
    assert make_rt_layout.layout(REF_SRC, make_rt_layout.REF_HEADERS) == want


def test_ctypes_view_matches_fixture():
    want = json.load(open(FIXTURE))
    classes = rt_abi.structs()
    for s, v in want.items():
        cls = classes[s]
        assert C.sizeof(cls) == v["size"], s
        for f, off in v["fields"].items():
            assert getattr(cls, f).offset == off, (s, f)


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    # static inline helpers live in the header, not the library
    text = re.sub(r"static inline[^{]*\{.*?\n\}", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:rt|ocrt|surface)_[a-zA-Z_0-9]+)\s*\(", text)))


def test_library_exports_reference_entry_points():
    import ocrt_amd
    out = subprocess.run(["nm", "-D", "--defined-only", ocrt_amd.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    names = declared("ocrt_rt.h")
    assert {"rt_solve_case_pol", "rt_solve_case_pol_aerosol", "rt_solve_case_pol_lut"} <= set(names)
    assert set(names) <= exported, set(names) - exported


def _ref_tau():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref/libocrt_ref.so not built")
    ref = C.CDLL(REF_LIB)
    f = ref.rt_rayleigh_tau_model_full
    f.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]
    f.restype = C.c_double
    return f


def test_rayleigh_tau_equals_reference():
    ref = _ref_tau()
    lib = rt_abi.library()
    for wl in (330.0, 412.0, 443.0, 490.0, 555.0, 660.0, 865.0, 1100.0):
        for p in (1013.25, 1019.2233, 800.0, 2707.819666):
            for lat, alt in ((45.0, 0.0), (10.0, 1500.0)):
                a = lib.ocrt_rayleigh_tau_bodhaine1999(wl, p, lat, alt)
                b = ref(wl, p, lat, alt, 0)
                # the reference forms n^2 - 1 from n^2 (n - 1 ~ 3e-4), which
                # costs it ~1e-13 relative; 1e-12 is its own libm gate
                assert abs(a - b) <= 1e-12 * abs(b), (wl, p, lat, alt, a, b)


def test_rayleigh_tau_closed_form_scalings():
    lib = rt_abi.library()
    t = lib.ocrt_rayleigh_tau_bodhaine1999
    # linear in pressure
    assert abs(t(555.0, 2026.5, 45.0, 0.0) / t(555.0, 1013.25, 45.0, 0.0) - 2.0) < 1e-14
    # inverse in gravity at the column altitude z_c = 0.73737 z + 5517.56 m
    def g(lat, z):
        s1, s2 = math.sin(math.radians(lat)), math.sin(math.radians(2 * lat))
        return 9.780327 * (1 + 0.0053024 * s1 * s1 - 0.0000058 * s2 * s2) - 3.086e-6 * z
    r = t(555.0, 1013.25, 0.0, 0.0) / t(555.0, 1013.25, 90.0, 0.0)
    assert abs(r - g(90.0, 5517.56) / g(0.0, 5517.56)) < 1e-14
    # lambda^-4 dominated: tau(412) / tau(865) between 19 and 22 (dispersion and King factor)
    assert 19.0 < t(412.0, 1013.25, 45.0, 0.0) / t(865.0, 1013.25, 45.0, 0.0) < 22.0
    # published order of magnitude at 555 nm, sea level: 0.09 .. 0.1
    assert 0.09 < t(555.0, 1013.25, 45.0, 0.0) < 0.1


def _case(**kw):
    cs = rt_abi.Case()
    cs.sza_deg, cs.vza_deg, cs.raa_deg, cs.wavelength_nm = 40.0, 30.0, 90.0, 555.0
    cs.surface = rt_abi.SURFACES["black"]
    cs.rayleigh_on, cs.pressure_hpa, cs.n_water = 1, 1013.25, 1.34
    for k, v in kw.items():
        setattr(cs, k, v)
    return cs


@pytest.mark.parametrize("change", [
    dict(case=dict(wavelength_nm=320.0)), dict(case=dict(sza_deg=90.0)), dict(case=dict(vza_deg=-1.0)),
    dict(opts=dict(fourier_m_max=33)), dict(opts=dict(n_mu=1)), dict(opts=dict(n_layers=1)),
    dict(case=dict(surface=4)), dict(opts=dict(integration_method=1)), dict(opts=dict(view_as_node=0)),
    dict(case=dict(tau_R_input=0.0), opts=dict(tau_r_source=1)),
])
def test_unserved_or_invalid_requests_refused(change):
    lib = rt_abi.library()
    cs = _case(**change.get("case", {}))
    opts = rt_abi.options_default()
    for k, v in change.get("opts", {}).items():
        setattr(opts, k, v)
    out = rt_abi.Result()
    out.rho_I = 123.0
    assert lib.rt_solve_case_pol(C.byref(cs), C.byref(opts), C.byref(out)) == -1
    assert out.rho_I == 123.0   # refused before the output is touched


def test_aerosol_input_checks():
    lib = rt_abi.library()
    beta = (C.c_double * 3)(1.0, 0.5, 0.2)
    for bad in (dict(tau_a=-0.1), dict(ssa_a=1.5), dict(L_max=1), dict(betal_aer=None)):
        aer = rt_abi.Aerosol()
        aer.tau_a, aer.ssa_a, aer.L_max = 0.1, 0.9, 2
        aer.betal_aer = aer.gammal_aer = aer.alphal_aer = aer.zetal_aer = beta
        for k, v in bad.items():
            setattr(aer, k, v)
        out = rt_abi.Result()
        assert lib.rt_solve_case_pol_aerosol(C.byref(_case()), None, C.byref(aer), C.byref(out)) == -1
