"""The aerosol value kernel (rt_aerosol_input_t.use_value_kernel,
code/OCRT_C/src/rt_solver.h:102-113; rt_solver.c:190-240 and 922-925;
rt_ipss.c:81-146) in the oracle:
  - against golden vectors written by the reference build
    (tests/golden/ref_value_kernel.json: HG, two-term HG and glory tables, flat
    and rough surfaces, IPSS on and off), 1e-12 relative, equal order counts;
  - against the mathematics: a P11 table sampled from the Legendre sum of the
    case's own beta coefficients must reproduce the moment-kernel solve (the
    addition theorem: the azimuth mean of P_l(cos Theta) cos(m phi) is
    P_l^m(mu_j) P_l^m(mu_k)), up to the table's linear interpolation error."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, desc_from_dict, oracle_solve
from ocrt_amd import synthetic as S
from ocrt_amd.abi import make_instance

with open(os.path.join(GOLDEN, "ref_value_kernel.json")) as _f:
    CASES = json.load(_f)


def vk_desc(rec):
    d, greek = desc_from_dict(rec["desc"])
    th = np.ascontiguousarray(rec["table"]["theta"], dtype=np.float64)
    p11 = np.ascontiguousarray(rec["table"]["p11"], dtype=np.float64)
    d.n_ang_phase = th.size
    d.theta_phase = th.ctypes.data_as(C.POINTER(C.c_double))
    d.p11_phase = p11.ctypes.data_as(C.POINTER(C.c_double))
    return d, (greek, th, p11)


@pytest.mark.parametrize("idx", range(len(CASES)))
def test_oracle_value_kernel_matches_reference(oracle, idx):
    rec = CASES[idx]
    d, keep = vk_desc(rec)
    rc, r, modes, orders = oracle_solve(oracle, d)
    assert rc == rec["rc"] == 0
    ref = rec["result"]
    scale = abs(ref["rho_I"])
    for f in ("rho_I", "rho_Q", "rho_U"):
        assert abs(getattr(r, f) - ref[f]) <= 1e-12 * scale, (rec["name"], f, getattr(r, f), ref[f])
    rm = np.array(rec["toa_modes"]).reshape(-1, 3)
    assert np.max(np.abs(modes - rm)) <= 1e-12 * np.max(np.abs(rm[:, 0]))
    assert list(orders) == rec["orders_per_m"]


def test_value_kernel_of_the_legendre_sum_is_the_moment_kernel(oracle):
    gk = S.hg_greek(0.6, 30)
    th = np.linspace(0.0, 180.0, 7201)
    mu = np.cos(np.radians(th))
    p11 = gk[0] @ S.legendre_table(30, mu)
    kw = dict(surface="black_fresnel_ocean", wind=4.0, tau_a=0.3, ssa_a=0.95, greek=gk, n_mu=12, n_layers=30,
              m_max=6, sza=45.0, vza=35.0, raa=70.0, ipss=1)
    rc0, r0, m0, o0 = oracle_solve(oracle, make_instance(**kw))
    rc1, r1, m1, o1 = oracle_solve(oracle, make_instance(phase_table=(th, p11), **kw))
    assert rc0 == rc1 == 0
    # linear interpolation on a 0.025-degree grid: ~1e-7 of the phase function
    assert abs(r1.rho_I - r0.rho_I) <= 2e-6 * abs(r0.rho_I)
    assert abs(r1.rho_Q - r0.rho_Q) <= 2e-6 * abs(r0.rho_I)
    assert np.max(np.abs(m1 - m0)) <= 2e-6 * np.max(np.abs(m0[:, 0]))
    # and a coarse table is visibly different: the check above has teeth
    thc = np.linspace(0.0, 180.0, 19)
    rc2, r2, _, _ = oracle_solve(oracle, make_instance(phase_table=(thc, gk[0] @ S.legendre_table(30, np.cos(np.radians(thc)))), **kw))
    assert abs(r2.rho_I - r0.rho_I) > 1e-4 * abs(r0.rho_I)
