"""Reference finite-difference Jacobians of the adjoint workload (run in the
build container only): instances drawn as the adjoint configuration (555 nm,
200 layers, 32 streams, modes 0..16, HG aerosol L = 80, rough ocean), each
solved by the reference OCRT solver (oracle/_ref/libocrt_ref.so, its
rt_solve_case_pol_aerosol through oracle/ref_harness.c) at the wind speed and
the aerosol optical depth scaled by 1 +- 1e-4, giving central differences of
g . (rho_I, rho_Q, rho_U) for g = (1, 0.5, -0.5):

    fd_wind = (f(wind (1 + h)) - f(wind (1 - h))) / (2 h wind)
    fd_aod  = (f(tau_a (1 + h)) - f(tau_a (1 - h))) / (2 h)      (d/ds at s = 1)

-> ref_jacobians.json.

    python tests/golden/make_jacobian_golden.py
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from ocrt_amd.abi import Instance, Result  # noqa: E402
from ocrt_amd import synthetic as S  # noqa: E402
from make_golden import REF_LIB  # noqa: E402

H = 1e-4
G = np.array([1.0, 0.5, -0.5])
N_INST = 12


def desc_to_dict(d, greek):
    """The descriptor's scalar fields and its Greek coefficients."""
    out = {}
    for f, t in Instance._fields_:
        v = getattr(d, f)
        if isinstance(v, (int, float)):
            out[f] = v
    out["greek"] = np.asarray(greek).tolist()
    return out


def solve_all(lib, descs):
    arr = (Instance * len(descs))(*descs)
    out = (Result * len(descs))()
    threads = lib.ref_max_threads()
    assert lib.ref_solve_batch(arr, len(descs), out, threads) == 0
    for r in out:
        assert r.status == 0
    return np.array([[r.rho_I, r.rho_Q, r.rho_U] for r in out])


def main():
    lib = C.CDLL(REF_LIB)
    lib.ref_solve_batch.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    gen, keep = S.config3(n_inst=N_INST, seed=3)
    runs = []
    for d in gen:
        base = Instance.from_buffer_copy(d)
        wp, wm = Instance.from_buffer_copy(d), Instance.from_buffer_copy(d)
        wp.wind_speed = d.wind_speed * (1 + H)
        wm.wind_speed = d.wind_speed * (1 - H)
        ap, am = Instance.from_buffer_copy(d), Instance.from_buffer_copy(d)
        ap.tau_a = d.tau_a * (1 + H)
        am.tau_a = d.tau_a * (1 - H)
        runs += [base, wp, wm, ap, am]
    rho = solve_all(lib, runs).reshape(len(gen), 5, 3)
    f = rho @ G
    cases = []
    for i, d in enumerate(gen):
        cases.append({"desc": desc_to_dict(d, keep[i]), "rho": rho[i, 0].tolist(),
                      "fd_wind": (f[i, 1] - f[i, 2]) / (2 * H * d.wind_speed),
                      "fd_aod": (f[i, 3] - f[i, 4]) / (2 * H)})
    with open(os.path.join(HERE, "ref_jacobians.json"), "w") as fh:
        json.dump({"g": G.tolist(), "h": H, "cases": cases}, fh, indent=1)
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main()
