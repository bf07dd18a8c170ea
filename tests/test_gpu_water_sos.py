"""The in-water SOS on the GPU (in-water columns of a plan: uniform levels,
water-side reflection at the top fed back every order, diffuse top source)
through the C ABI: FP64 per-node fields against the reference's own water
order loops (tests/golden/ref_water_sos.json), 1e-11 of the mode-0 field with
equal order counts; the mixed form within 1e-4 of it; a diffuse top source
and a scaled beam against the oracle; instances sharing one set of rough
water-side kernels equal to the instances alone."""
import ctypes as C

import numpy as np
import pytest

from conftest import desc_from_dict
from test_water_sos import CASES, _top_source, nodes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ocrt():
    import ocrt_amd
    ocrt_amd.library()
    return ocrt_amd


def gpu_nodes(ocrt, descs, precision):
    plan = ocrt.Plan(descs, precision=precision)
    plan.export_nodes()
    plan.execute()
    res = plan.results(modes=True)
    pitch = max(d.n_mu + 1 for d in descs)
    up, dn = plan.node_fields(pitch)
    plan.close()
    return res, up, dn


def test_water_sos_matches_reference(ocrt):
    keep, descs = [], []
    for rec in CASES:
        d, g = desc_from_dict(rec["desc"])
        descs.append(d)
        keep.append(g)
    print(f"[gpu] in-water SOS: {len(descs)} golden cases, fp64 and mixed", flush=True)
    a, up, dn = gpu_nodes(ocrt, descs, "fp64")
    b, upm, dnm = gpu_nodes(ocrt, descs, "mixed")
    for i, rec in enumerate(CASES):
        n, mm = rec["n"], descs[i].m_max
        assert a["status"][i] == 0 and b["status"][i] == 0 and a["n_ring"][i] == n
        rt = np.array(rec["top_up"]).reshape(mm + 1, 3, n)
        rb = np.array(rec["bot_dn"]).reshape(mm + 1, 3, n)
        scale = np.max(np.abs(rt[0, 0]))
        assert np.max(np.abs(up[i, : mm + 1, :, :n] - rt)) <= 1e-11 * scale, rec["name"]
        assert np.max(np.abs(dn[i, : mm + 1, :, :n] - rb)) <= 1e-11 * max(np.max(np.abs(rb[0, 0])), 1e-300) + \
            1e-13 * scale, rec["name"]
        assert list(a["orders_per_m"][i, : mm + 1]) == rec["orders_per_m"], rec["name"]
        err = np.max(np.abs(upm[i, : mm + 1, :, :n] - rt)) / scale
        print(f"[gpu] {rec['name']}: mixed top field within {err:.2e} of the mode-0 maximum", flush=True)
        assert err <= 1e-4, rec["name"]


def test_water_top_source_and_beam_scale_match_oracle(ocrt, oracle):
    keep, descs = [], []
    for idx in (1, 3):
        d, g = desc_from_dict(CASES[idx]["desc"])
        n = CASES[idx]["n"]
        src = _top_source(n, d.m_max)
        d.top_source = src.ctypes.data_as(C.POINTER(C.c_double))
        d.top_m_max, d.top_n_mu = d.m_max, n
        d.beam_scale = 0.6
        descs.append(d)
        keep.append((g, src))
    print("[gpu] in-water SOS with a diffuse top source and beam weight 0.6", flush=True)
    a, up, dn = gpu_nodes(ocrt, descs, "fp64")
    for i, d in enumerate(descs):
        n, top, bot, orders = nodes(oracle, d)
        mm = d.m_max
        scale = np.max(np.abs(top[0, 0]))
        assert np.max(np.abs(up[i, : mm + 1, :, :n] - top)) <= 1e-11 * scale
        assert np.max(np.abs(dn[i, : mm + 1, :, :n] - bot)) <= 1e-11 * scale
        assert list(a["orders_per_m"][i, : mm + 1]) == list(orders)


def test_water_rough_kernels_shared(ocrt):
    rec = CASES[3]
    d1, g1 = desc_from_dict(rec["desc"])
    d2, g2 = desc_from_dict(rec["desc"])
    d2.water_omega_p = 0.4          # another particle load, same surface and ring: shares the kernels
    d3, g3 = desc_from_dict(rec["desc"])
    d3.wind_speed = 3.0             # another surface: its own kernels
    batch, up, dn = gpu_nodes(ocrt, [d1, d2, d3], "fp64")
    for i, d in enumerate((d1, d2, d3)):
        one, u1, _ = gpu_nodes(ocrt, [d], "fp64")
        assert np.array_equal(up[i, : d.m_max + 1], u1[0, : d.m_max + 1]), i
        assert np.array_equal(batch["orders_per_m"][i], one["orders_per_m"][0]), i
