"""The oracle's reverse mode against the reference's own perturbation runs:
d(g . rho)/d wind and d(g . rho)/d(AOD scale) of four of the twelve
adjoint-workload instances (the GPU test takes all twelve) (555 nm, 200 layers, 32 streams, modes 0..16) against central
differences (h = 1e-4 relative) of the reference solver
(tests/golden/ref_jacobians.json).  The AOD scale goes through the profile's
layer Jacobian (tests/jacobian_cases.py)."""
import ctypes as C

import numpy as np

from jacobian_cases import layered, load
from ocrt_amd.adjoint import layer_gradients


def test_oracle_adjoint_matches_reference_finite_differences(oracle):
    fx = load()
    g = np.ascontiguousarray(fx["g"], dtype=np.float64)
    worst_w = worst_a = 0.0
    for c in fx["cases"][:4]:
        d, lay, dlay = layered(oracle, c["desc"])
        L = d.n_layers + 1
        gh, gx, gy = np.zeros(L), np.zeros(L), np.zeros(L)
        gw = C.c_double()
        assert oracle.oracle_adjoint(C.byref(d), g.ctypes.data_as(C.c_void_p), gh.ctypes.data_as(C.c_void_p),
                                     gx.ctypes.data_as(C.c_void_p), gy.ctypes.data_as(C.c_void_p),
                                     C.byref(gw)) == 0
        gl = layer_gradients(lay, gh, gx, gy)
        ga = float(np.sum(gl * dlay))
        worst_w = max(worst_w, abs(gw.value - c["fd_wind"]) / abs(c["fd_wind"]))
        worst_a = max(worst_a, abs(ga - c["fd_aod"]) / abs(c["fd_aod"]))
    print(f"oracle adjoint vs reference FD: wind {worst_w:.1e}, AOD scale {worst_a:.1e}")
    assert worst_w < 1e-6 and worst_a < 1e-6
