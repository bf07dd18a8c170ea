"""The external bottom source on the GPU (sos.hip:order_one) through the C ABI:
FP64 against the reference build's golden vectors
(tests/golden/ref_bottom_source.json), 1e-11 relative with equal order
counts; the mixed form within its budget; and the reference-compatible
rt_solve_case_pol with ext_bottom_per_m_* and bottom_source_only against the
reference library on the same structs."""
import ctypes as C
import math

import numpy as np
import pytest

from test_bottom_source import CASES, bs_desc

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11


@pytest.fixture(scope="module")
def ocrt():
    import ocrt_amd
    ocrt_amd.library()
    return ocrt_amd


def test_bottom_source_fp64_and_mixed_match_reference(ocrt):
    keep, descs = [], []
    for rec in CASES:
        d, k = bs_desc(rec)
        descs.append(d)
        keep.append(k)
    print(f"[gpu] bottom source: {len(descs)} golden cases, fp64 and mixed", flush=True)
    a = ocrt.solve(descs, precision="fp64", modes=True)
    b = ocrt.solve(descs, precision="mixed")
    for i, rec in enumerate(CASES):
        ref = rec["result"]
        assert a["status"][i] == 0 and b["status"][i] == 0
        scale = max(abs(ref["rho_I"]), 1e-300)
        for f in ("rho_I", "rho_Q", "rho_U"):
            assert abs(a[f][i] - ref[f]) <= FP64_TOL * scale, (rec["name"], f, a[f][i], ref[f])
        mm = len(rec["orders_per_m"])
        rm = np.array(rec["toa_modes"]).reshape(-1, 3)
        assert np.max(np.abs(a["toa_modes"][i, :mm] - rm)) <= FP64_TOL * max(np.max(np.abs(rm[:, 0])), 1e-300)
        assert list(a["orders_per_m"][i, :mm]) == rec["orders_per_m"], rec["name"]
        assert abs(b["rho_I"][i] - ref["rho_I"]) <= 1e-4 * scale + 1e-300, rec["name"]
        assert abs(b["rho_Q"][i] - ref["rho_Q"]) <= 1e-4 and abs(b["rho_U"][i] - ref["rho_U"]) <= 1e-4


def test_rt_compat_bottom_source_equals_reference(ref_lib):
    from ocrt_amd import rt_abi
    from test_gpu_rt_compat import Aer, case, compare, options, run
    mine = rt_abi.library()
    rt_abi.bind(ref_lib)
    import ocrt_amd
    lib = ocrt_amd.library()
    for ckw, okw, only in ((dict(surface="black_fresnel_ocean", wind=6.0, sza=55.0, vza=25.0, raa=40.0,
                                 decouple=1), dict(m_max=5, n_mu=16), 1),
                           (dict(surface="flat", sza=30.0, vza=50.0, raa=160.0), dict(m_max=4, n_mu=12), 0)):
        o = options(**okw)
        mu = np.zeros(o.n_mu + 1)
        n = lib.ocrt_ring(o.n_mu, C.c_double(ckw["vza"]), mu.ctypes.data_as(C.c_void_p), None)
        mb = okw["m_max"] - 1   # one mode fewer than the solve: the top mode carries no source
        x = mu[None, :n]
        m = np.arange(mb + 1)[:, None]
        src = [np.ascontiguousarray(v) for v in (0.02 * x * np.exp(-m / 2.0) + 0 * m,
                                                 -0.004 * (1 - x * x) * np.exp(-m / 3.0),
                                                 0.003 * x * (1 - x) * (m > 0) + 0 * x)]
        P = C.POINTER(C.c_double)
        o.ext_bottom_per_m_I, o.ext_bottom_per_m_Q, o.ext_bottom_per_m_U = [v.ctypes.data_as(P) for v in src]
        o.ext_bottom_n_mu, o.ext_bottom_m_max, o.bottom_source_only = n, mb, only
        cs = case(**ckw)
        aer = Aer(tau_a=0.2, ssa=0.95, g=0.7, L=40)
        print(f"[gpu] rt_solve_case_pol_aerosol bottom source: {ckw} {okw} only={only}", flush=True)
        ra, a = run(mine, cs, o, aer)
        rb, b = run(ref_lib, cs, o, aer)
        assert ra == rb == 0
        compare(a, b, o.fourier_m_max)
        # a source on the wrong ring is refused like the reference (-7)
        o.ext_bottom_n_mu = n + 1
        ra, _ = run(mine, cs, o, aer)
        rb, _ = run(ref_lib, cs, o, aer)
        assert ra == rb == -7
