"""Gas absorption tables without a GPU (csrc/gas.hip).

- rt_absorption_init of this library and of the reference (compiled from its
  sources into oracle/_ref) read the same files (tests/golden/gas_inputs: the
  US76 AFGL profile and the cross-section tables reduced to the test
  wavelengths) into equal states, and their column optical depths per gas
  agree to 1e-13.
- Known answers: the US76 ozone column is 346.9 DU (the value the AFGL file
  header states), the column optical depth is linear in an overridden total
  column and zero for a zero column.
- With the reference tree present, the full tables give the same optical
  depths as the reference at wavelengths across 330..1100 nm.
"""
import ctypes as C
import os

import numpy as np
import pytest

from ocrt_amd import rt_abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "tests", "golden", "gas_inputs")
AFGL, XSEC = os.path.join(DATA, "afgl_atm"), os.path.join(DATA, "xsec")
REF_INPUTS = "/root/reference/code/OCRT_C/inputs"
TEST_WL = [412.0, 443.0, 490.0, 555.0, 660.0, 865.0, 554.0, 556.0]


@pytest.fixture(scope="module")
def libs(ref_lib):
    return rt_abi.library(), rt_abi.bind(ref_lib)


def arr(p, n):
    return np.ctypeslib.as_array(p, shape=(n,)).copy()


def test_states_equal_reference(libs):
    mine, ref = libs
    a = rt_abi.absorption(mine, AFGL, XSEC)
    b = rt_abi.absorption(ref, AFGL, XSEC)
    try:
        n = b.atm.n_levels
        assert a.atm.n_levels == n == 40
        for f in ("z_km", "P_mbar", "T_K", "n_total_cm3"):
            assert np.array_equal(arr(getattr(a.atm, f), n), arr(getattr(b.atm, f), n)), f
        for g in range(6):
            assert np.array_equal(arr(a.atm.mr[g], n), arr(b.atm.mr[g], n))
            assert a.atm.column_default[g] == pytest.approx(b.atm.column_default[g], rel=1e-14)
            xa, xb = a.xsec[g], b.xsec[g]
            assert (xa.available, xa.n_wl, xa.n_layers) == (xb.available, xb.n_wl, xb.n_layers)
            assert np.array_equal(arr(xa.wl_nm, xa.n_wl), arr(xb.wl_nm, xb.n_wl))
            assert np.array_equal(arr(xa.sigma_cm2, xa.n_wl * xa.n_layers), arr(xb.sigma_cm2, xb.n_wl * xb.n_layers))
            assert a.column_eff[g] == pytest.approx(b.column_eff[g], rel=1e-14)
        for wl in TEST_WL:
            for g in range(6):
                ta = mine.rt_absorption_tau_per_gas(C.byref(a), g, wl)
                tb = ref.rt_absorption_tau_per_gas(C.byref(b), g, wl)
                assert ta == pytest.approx(tb, rel=1e-13, abs=1e-300), (wl, rt_abi.GASES[g], ta, tb)
            assert mine.rt_absorption_tau_total(C.byref(a), wl) == pytest.approx(
                ref.rt_absorption_tau_total(C.byref(b), wl), rel=1e-13)
    finally:
        mine.rt_absorption_free(C.byref(a))
        ref.rt_absorption_free(C.byref(b))


def test_known_answers():
    import ocrt_amd
    lib = rt_abi.library()
    st = rt_abi.absorption(lib, AFGL, XSEC)
    o3_du = st.atm.column_default[1] / 2.6868e16
    lib.rt_absorption_free(C.byref(st))
    assert abs(o3_du - 346.9) < 0.5, o3_du
    base = ocrt_amd.GasTable(AFGL, XSEC)
    col = base.tau(555.0, 1)
    assert col > 0.0
    for du in (150.0, 300.0, 450.0):
        t = ocrt_amd.GasTable(AFGL, XSEC, columns={"o3": du * 2.6868e16})
        assert t.tau(555.0, 1) == pytest.approx(col * du / o3_du, rel=1e-12)
    zero = ocrt_amd.GasTable(AFGL, XSEC, columns={g: 0.0 for g in rt_abi.GASES})
    assert zero.tau(555.0) == 0.0
    # the handle equals the state's per-gas depths
    st = rt_abi.absorption(lib, AFGL, XSEC)
    for g in range(6):
        assert base.tau(443.0, g) == lib.rt_absorption_tau_per_gas(C.byref(st), g, 443.0)
    lib.rt_absorption_free(C.byref(st))


@pytest.mark.skipif(not os.path.isdir(REF_INPUTS), reason="reference data tree not present")
def test_full_tables_equal_reference(libs):
    mine, ref = libs
    afgl, xsec = os.path.join(REF_INPUTS, "afgl_atm"), os.path.join(REF_INPUTS, "xsec")
    rng = np.random.default_rng(1)
    for prof in (0, 1):
        a = rt_abi.absorption(mine, afgl, xsec, prof)
        b = rt_abi.absorption(ref, afgl, xsec, prof)
        for wl in np.concatenate([[330.0, 1100.0, 760.5], rng.uniform(330, 1100, 12)]):
            for g in range(6):
                ta = mine.rt_absorption_tau_per_gas(C.byref(a), g, wl)
                tb = ref.rt_absorption_tau_per_gas(C.byref(b), g, wl)
                assert ta == pytest.approx(tb, rel=1e-13, abs=1e-300), (prof, wl, g)
        mine.rt_absorption_free(C.byref(a))
        ref.rt_absorption_free(C.byref(b))
