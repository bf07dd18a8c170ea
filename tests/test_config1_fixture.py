"""The full-size single-instance Rayleigh + aerosol record (48 streams, 400
layers, 180-term Greek set, IPSS, Fourier modes 0..95), written by the oracle
(tests/golden/make_config1_golden.py) because the reference caps the
atmosphere at m <= 32 (code/OCRT_C/src/rt_solver.c:630).

Its modes 0..32 are pinned to the reference build's m <= 32 solve of the same
case (ref_cases.json 'c1_m32_ipss'): the Fourier modes are independent, so
the per-mode TOA values and order counts must agree, and the modes above 32
must not move rho beyond what they weigh."""
import json
import os

import numpy as np

from conftest import GOLDEN


def _load():
    with open(os.path.join(GOLDEN, "oracle_config1_m95.json")) as f:
        rec = json.load(f)
    with open(os.path.join(GOLDEN, "ref_cases.json")) as f:
        ref = [c for c in json.load(f) if c["name"] == "c1_m32_ipss"][0]
    return rec, ref


def test_config1_low_modes_match_reference():
    rec, ref = _load()
    modes = np.array(rec["toa_modes"]).reshape(-1, 3)
    rm = np.array(ref["toa_modes"]).reshape(-1, 3)
    assert modes.shape == (96, 3) and rm.shape == (33, 3)
    scale = np.max(np.abs(rm[:, 0]))
    assert np.max(np.abs(modes[:33] - rm)) <= 1e-12 * scale
    assert rec["orders_per_m"][:33] == ref["orders_per_m"]


def test_config1_high_modes_weigh_what_rho_moves():
    rec, ref = _load()
    modes = np.array(rec["toa_modes"]).reshape(-1, 3)
    mu_s = np.cos(np.radians(40.0))
    weight = 2.0 * np.abs(modes[33:]).sum(axis=0) / mu_s * rec["result"]["kappa"]
    for c, f in enumerate(("rho_I", "rho_Q", "rho_U")):
        d = abs(rec["result"][f] - ref["result"][f])
        assert d <= weight[c] + 1e-13 * abs(ref["result"]["rho_I"]), (f, d, weight[c])
    # the modes decay: nothing above 32 exceeds 1e-12 of mode 0
    assert np.max(np.abs(modes[33:])) <= 1e-12 * abs(modes[0, 0])
