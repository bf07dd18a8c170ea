"""The aerosol value kernel on the GPU (setup.hip:value_phase, ipss.hip) through
the C ABI:
  - FP64 against the reference build's golden vectors
    (tests/golden/ref_value_kernel.json), 1e-11 relative with equal order
    counts; the mixed form within its budget (1e-4 relative in I, 1e-4
    absolute in Q and U);
  - instances sharing one table and one operator (two bands of a geometry)
    equal the instances solved alone, bit for bit;
  - the reference-compatible rt_solve_case_pol_aerosol with use_value_kernel
    against the reference library on the same structs (1e-11);
  - the Python batch mirror with use_value_kernel against the oracle."""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, oracle_solve
from test_value_kernel import CASES, vk_desc

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11
MIXED_TOL = 1e-4


@pytest.fixture(scope="module")
def ocrt():
    import ocrt_amd
    ocrt_amd.library()
    return ocrt_amd


@pytest.fixture(scope="module")
def batch():
    keep, descs = [], []
    for rec in CASES:
        d, k = vk_desc(rec)
        descs.append(d)
        keep.append(k)
    return descs, keep


def test_value_kernel_fp64_matches_reference(ocrt, batch):
    descs, keep = batch
    print(f"[gpu] value kernel: {len(descs)} golden cases, fp64", flush=True)
    out = ocrt.solve(descs, precision="fp64", modes=True)
    for i, rec in enumerate(CASES):
        ref = rec["result"]
        assert out["status"][i] == 0
        for f in ("rho_I", "rho_Q", "rho_U"):
            assert abs(out[f][i] - ref[f]) <= FP64_TOL * abs(ref["rho_I"]), (rec["name"], f, out[f][i], ref[f])
        mm = len(rec["orders_per_m"])
        rm = np.array(rec["toa_modes"]).reshape(-1, 3)
        assert np.max(np.abs(out["toa_modes"][i, :mm] - rm)) <= FP64_TOL * np.max(np.abs(rm[:, 0])), rec["name"]
        assert list(out["orders_per_m"][i, :mm]) == rec["orders_per_m"], rec["name"]


def test_value_kernel_mixed_within_budget(ocrt, batch):
    descs, keep = batch
    print("[gpu] value kernel: mixed form", flush=True)
    out = ocrt.solve(descs, precision="mixed")
    for i, rec in enumerate(CASES):
        ref = rec["result"]
        assert out["status"][i] == 0
        assert abs(out["rho_I"][i] - ref["rho_I"]) <= MIXED_TOL * abs(ref["rho_I"]), rec["name"]
        assert abs(out["rho_Q"][i] - ref["rho_Q"]) <= MIXED_TOL, rec["name"]
        assert abs(out["rho_U"][i] - ref["rho_U"]) <= MIXED_TOL, rec["name"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_value_kernel_shared_operator_equals_alone(ocrt, precision):
    from ocrt_amd import synthetic as S
    from ocrt_amd.abi import make_instance
    rec = CASES[0]
    th, p11 = np.array(rec["table"]["theta"]), np.array(rec["table"]["p11"])
    gk = S.hg_greek(0.7, 60)
    descs = [make_instance(surface="black_fresnel_ocean", wind=6.0, tau_R=0.0941 * (wl / 555.0) ** -4.08,
                           tau_a=0.1 * (wl / 555.0) ** -1.2, ssa_a=0.94, greek=gk, n_mu=14, n_layers=45, m_max=7,
                           sza=38.0, vza=22.0, raa=80.0, wavelength=wl, phase_table=(th, p11))
             for wl in (443.0, 555.0, 865.0)]
    batch_out = ocrt.solve(descs, precision=precision, modes=True)
    for i, d in enumerate(descs):
        one = ocrt.solve([d], precision=precision, modes=True)
        for f in ("rho_I", "rho_Q", "rho_U", "n_orders_used"):
            assert np.array_equal(batch_out[f][i], one[f][0]), (precision, i, f)


def test_rt_compat_value_kernel_equals_reference(ref_lib):
    from ocrt_amd import rt_abi
    from test_gpu_rt_compat import Aer, case, compare, options, run
    mine = rt_abi.library()
    rt_abi.bind(ref_lib)
    rec = CASES[0]
    th = np.ascontiguousarray(rec["table"]["theta"], dtype=np.float64)
    p11 = np.ascontiguousarray(rec["table"]["p11"], dtype=np.float64)
    for ckw, okw in ((dict(surface="black_fresnel_ocean", wind=6.0, sza=55.0, vza=25.0, raa=40.0), dict(m_max=6, pssa=1)),
                     (dict(surface="flat", sza=30.0, vza=50.0, raa=160.0), dict(m_max=5))):
        aer = Aer(tau_a=0.2, ssa=0.95, g=0.7, L=40)
        aer.s.use_value_kernel = 1
        aer.s.n_ang_phase = th.size
        aer.s.theta_phase = th.ctypes.data_as(C.POINTER(C.c_double))
        aer.s.P11_phase = p11.ctypes.data_as(C.POINTER(C.c_double))
        cs, o = case(**ckw), options(**okw)
        print(f"[gpu] rt_solve_case_pol_aerosol value kernel: {ckw} {okw}", flush=True)
        ra, a = run(mine, cs, o, aer)
        rb, b = run(ref_lib, cs, o, aer)
        assert ra == rb == 0
        compare(a, b, o.fourier_m_max)
