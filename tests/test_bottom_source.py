"""The atmosphere solve driven by an external bottom source (the coupled
ocean's pass 2: rt_options_t.ext_bottom_per_m_* and bottom_source_only,
code/OCRT_C/src/rt_types.h:722-747; rt_solver.c:687-706, 949-977;
rt_surface_boundary.c:495-567) in the oracle, against golden vectors written
by the reference build (tests/golden/ref_bottom_source.json): rough and flat
surfaces, a source with fewer modes than the solve, the source with and
without the sun, the glint branch, the black surface (source ignored), and a
200-layer 32-stream case with IPSS.  Also pinned by linearity: the solve with
sun and source equals the solve with the sun alone plus the source alone."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, desc_from_dict, oracle_solve
from ocrt_amd import synthetic as S
from ocrt_amd.abi import make_instance

with open(os.path.join(GOLDEN, "ref_bottom_source.json")) as _f:
    CASES = json.load(_f)


def bs_desc(rec):
    d, greek = desc_from_dict(rec["desc"])
    src = np.ascontiguousarray(rec["bottom_source"], dtype=np.float64)
    d.bottom_source = src.ctypes.data_as(C.POINTER(C.c_double))
    d.bottom_m_max = src.shape[1] - 1
    d.bottom_n_mu = src.shape[2]
    d.bottom_source_only = rec["bottom_source_only"]
    return d, (greek, src)


@pytest.mark.parametrize("idx", range(len(CASES)))
def test_oracle_bottom_source_matches_reference(oracle, idx):
    rec = CASES[idx]
    d, keep = bs_desc(rec)
    rc, r, modes, orders = oracle_solve(oracle, d)
    assert rc == rec["rc"] == 0
    ref = rec["result"]
    scale = max(abs(ref["rho_I"]), 1e-300)
    for f in ("rho_I", "rho_Q", "rho_U"):
        assert abs(getattr(r, f) - ref[f]) <= 1e-12 * scale, (rec["name"], f, getattr(r, f), ref[f])
    rm = np.array(rec["toa_modes"]).reshape(-1, 3)
    assert np.max(np.abs(modes - rm)) <= 1e-12 * max(np.max(np.abs(rm[:, 0])), 1e-300)
    assert list(orders) == rec["orders_per_m"]


def test_bottom_source_is_linear(oracle):
    """sun + source = sun alone + source alone, for the TOA modes (the solve
    is linear in its sources; the order counts may differ, so a tight
    tolerance and many orders are used)."""
    rec = [c for c in CASES if c["name"] == "bs_rough_plus_sun"][0]
    d, keep = bs_desc(rec)
    d.tolerance, d.max_orders = 1e-14, 60
    _, r_both, m_both, _ = oracle_solve(oracle, d)
    d.bottom_source_only = 1
    _, r_src, m_src, _ = oracle_solve(oracle, d)
    d2, keep2 = desc_from_dict(rec["desc"])
    d2.tolerance, d2.max_orders = 1e-14, 60
    _, r_sun, m_sun, _ = oracle_solve(oracle, d2)
    assert np.max(np.abs(m_both - (m_src + m_sun))) <= 1e-12 * np.max(np.abs(m_both[:, 0]))


def test_bottom_source_shape_refused(oracle):
    gk = S.hg_greek(0.7, 20)
    d = make_instance(surface="flat", tau_a=0.1, greek=gk, n_mu=8, vza=30.0,
                      bottom_source=np.zeros((3, 3, 5)), bottom_source_only=1)   # ring has 9 nodes
    rc, _, _, _ = oracle_solve(oracle, d)
    assert rc == -7
