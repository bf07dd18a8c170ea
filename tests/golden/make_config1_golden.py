"""Write the full-size Rayleigh + aerosol single-instance vectors (48 streams,
400 layers, 180-term HG Greek set, rough ocean at 5 m/s, IPSS on, Fourier
modes 0..95) from the oracle, the C restatement of the reference path.

The reference solver caps the atmosphere at m <= 32 (code/OCRT_C/src/rt_solver.c:630),
so it cannot write this case.  Modes 0..32 of this record are checked
against the reference build's m <= 32 record of the same case
(ref_cases.json 'c1_m32_ipss') by tests/test_config1_fixture.py; the modes
above 32 run through the same oracle code as the ones below.

    python tests/golden/make_config1_golden.py      (about a minute on one core)
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ocrt_amd.abi import Result, RESULT_FIELDS  # noqa: E402
from ocrt_amd import synthetic as S  # noqa: E402

OUT = os.path.join(HERE, "oracle_config1_m95.json")


def main():
    lib = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    ds, gk = S.config1(m_max=95)
    d = ds[0]
    mm = d.m_max
    r = Result()
    modes = np.zeros(3 * (mm + 1))
    orders = np.zeros(mm + 1, dtype=np.int32)
    rc = lib.oracle_solve(C.byref(d), C.byref(r), modes.ctypes.data_as(C.c_void_p),
                          orders.ctypes.data_as(C.c_void_p), None)
    assert rc == 0, rc
    rec = {"config": "config1: n 48, nt 400, L 180, HG g 0.70, AOD 0.1, ssa 0.95, wind 5, IPSS, m 0..95",
           "m_max": mm, "result": {f: getattr(r, f) for f in RESULT_FIELDS},
           "toa_modes": modes.tolist(), "orders_per_m": orders.tolist()}
    with open(OUT, "w") as f:
        json.dump(rec, f, indent=0)
    print(f"wrote {OUT}: rho_I {r.rho_I:.15e} orders {r.n_orders_used}")


if __name__ == "__main__":
    main()
