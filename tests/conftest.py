"""Shared fixtures.  Tests marked `gpu` need a ROCm device and the built HIP
library; everything else runs on the CPU."""
import ctypes as C
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from ocrt_amd.abi import Instance, Result, make_instance  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libocrt_ref.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


def _build_oracle():
    if not os.path.exists(ORACLE_LIB):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       stdout=subprocess.DEVNULL)


@pytest.fixture(scope="session")
def oracle():
    """The C restatement (test infrastructure only)."""
    _build_oracle()
    lib = C.CDLL(ORACLE_LIB)
    lib.oracle_ipss_kappa.restype = C.c_double
    lib.oracle_ipss_tau_dir_dn.restype = C.c_double
    lib.oracle_reconstruct_cos.restype = C.c_double
    lib.oracle_reconstruct_sin.restype = C.c_double
    lib.oracle_us62_grid_fraction_above.restype = C.c_double
    lib.oracle_ipss_homog_tau_view.restype = C.c_double
    lib.oracle_ipss_homog_tau_sun.restype = C.c_double
    return lib


@pytest.fixture(scope="session")
def ref_lib():
    if not os.path.exists(REF_LIB):
        pytest.skip("reference build oracle/_ref/libocrt_ref.so not present")
    return C.CDLL(REF_LIB)


def desc_from_dict(dd):
    """Golden record -> (descriptor, greek array kept alive)."""
    greek = None if dd["greek"] is None else np.ascontiguousarray(dd["greek"], dtype=np.float64)
    d = make_instance()
    for k, v in dd.items():
        if k == "greek":
            continue
        setattr(d, k, v)
    if greek is not None:
        d.greek = greek.ctypes.data_as(C.POINTER(C.c_double))
        d.L_max = greek.shape[1] - 1
    else:
        d.greek = None
    return d, greek


@pytest.fixture(scope="session")
def golden_cases():
    with open(os.path.join(GOLDEN, "ref_cases.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def golden_luts():
    with open(os.path.join(GOLDEN, "ref_lut.json")) as f:
        return json.load(f)


def oracle_solve(lib, d):
    mm = d.m_max
    r = Result()
    modes = np.zeros(3 * (mm + 1))
    orders = np.zeros(mm + 1, dtype=np.int32)
    rc = lib.oracle_solve(C.byref(d), C.byref(r), modes.ctypes.data_as(C.c_void_p),
                          orders.ctypes.data_as(C.c_void_p), None)
    return rc, r, modes.reshape(-1, 3), orders


@pytest.fixture(scope="session")
def orc(oracle):
    """The oracle with the argument types of its interface-kernel functions."""
    D = C.POINTER(C.c_double)
    oracle.oracle_interface_fourier.argtypes = [C.c_int, D, C.c_int, D, C.c_int, C.c_int, C.c_int, C.c_double,
                                                C.c_int, C.c_double, C.c_int, D]
    oracle.oracle_interface_trig.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                             C.c_int, C.c_double, C.c_int, D]
    oracle.oracle_interface_trig.restype = None
    oracle.oracle_interface_direct.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_double]
    oracle.oracle_interface_direct.restype = C.c_double
    return oracle
