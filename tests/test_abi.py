"""The C ABI boundary without a GPU: the library loads, exports every entry
point include/ocrt_amd.h declares, the ctypes mirrors match the header's
struct layout, and invalid descriptors are rejected before any device work."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import ocrt_amd
from ocrt_amd.abi import Instance, Result, make_instance

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ocrt_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(ocrt_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ocrt_amd.library()
    names = declared_functions()
    assert len(names) >= 11
    for name in names:
        assert hasattr(lib, name), name
    assert lib.ocrt_version().decode().startswith("ocrt_amd")


def test_no_declared_symbol_missing_from_dynamic_table():
    out = subprocess.run(["nm", "-D", "--defined-only", ocrt_amd.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(declared_functions()) <= exported


def _c_layout(tmp_path):
    src = tmp_path / "layout.c"
    exe = tmp_path / "layout"
    fields_i = [f for f, _ in Instance._fields_]
    fields_r = [f for f, _ in Result._fields_]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ocrt_amd.h"', "int main(void){",
             'printf("inst %zu\\n", sizeof(ocrt_instance_t));',
             'printf("res %zu\\n", sizeof(ocrt_result_t));']
    for f in fields_i:
        lines.append(f'printf("i.{f} %zu\\n", offsetof(ocrt_instance_t, {f}));')
    for f in fields_r:
        lines.append(f'printf("r.{f} %zu\\n", offsetof(ocrt_result_t, {f}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return dict((k, int(v)) for k, v in (line.split() for line in out.splitlines()))


def test_ctypes_mirrors_match_header_layout(tmp_path):
    lay = _c_layout(tmp_path)
    assert lay["inst"] == C.sizeof(Instance)
    assert lay["res"] == C.sizeof(Result)
    for f, _ in Instance._fields_:
        assert lay[f"i.{f}"] == getattr(Instance, f).offset, f
    for f, _ in Result._fields_:
        assert lay[f"r.{f}"] == getattr(Result, f).offset, f


@pytest.mark.parametrize("bad", [
    dict(sza=90.0), dict(sza=-1.0), dict(vza=95.0), dict(wavelength=200.0), dict(wavelength=1200.0),
    dict(n_mu=1), dict(n_layers=1), dict(m_max=-1), dict(max_orders=0), dict(tolerance=0.0),
    dict(tau_R=-0.1), dict(tau_R=0.0), dict(surface=7), dict(n_water=0.0),
])
def test_invalid_descriptor_is_rejected(bad):
    lib = ocrt_amd.library()
    kw = dict(bad)
    surface = kw.pop("surface", None)
    d = make_instance(**kw)
    if surface is not None:
        d.surface = surface
    plan = C.c_void_p()
    rc = lib.ocrt_plan_create(C.byref(d), 1, ocrt_amd.PREC_FP64, 0, C.byref(plan))
    assert rc == -1
    assert not plan.value


def test_bad_arguments_rejected():
    lib = ocrt_amd.library()
    d = make_instance()
    plan = C.c_void_p()
    assert lib.ocrt_plan_create(None, 1, 0, 0, C.byref(plan)) == -1
    assert lib.ocrt_plan_create(C.byref(d), 0, 0, 0, C.byref(plan)) == -1
    assert lib.ocrt_plan_create(C.byref(d), 1, 5, 0, C.byref(plan)) == -1
    assert lib.ocrt_plan_create(C.byref(d), 1, 0, 0, None) == -1


def test_aerosol_ssa_out_of_range_rejected():
    from ocrt_amd import synthetic as S
    lib = ocrt_amd.library()
    d = make_instance(tau_a=0.1, ssa_a=1.2, greek=S.hg_greek(0.7, 20))
    plan = C.c_void_p()
    assert lib.ocrt_plan_create(C.byref(d), 1, 0, 0, C.byref(plan)) == -1


@pytest.mark.parametrize("table", [
    (np.array([0.0, 90.0, 45.0]), np.ones(3)),          # not ascending
    (np.array([0.0, 200.0]), np.ones(2)),               # past 180 degrees
    (np.array([0.0, 180.0]), np.array([1.0, np.nan])),  # non-finite value
])
def test_bad_value_kernel_table_rejected(table):
    from ocrt_amd import synthetic as S
    lib = ocrt_amd.library()
    d = make_instance(tau_a=0.1, greek=S.hg_greek(0.7, 20), phase_table=table)
    plan = C.c_void_p()
    assert lib.ocrt_plan_create(C.byref(d), 1, 0, 0, C.byref(plan)) == -1
    assert not plan.value


def test_bottom_source_shape_rejected():
    from ocrt_amd import synthetic as S
    lib = ocrt_amd.library()
    mu = np.zeros(10)
    n = lib.ocrt_ring(8, C.c_double(30.0), mu.ctypes.data_as(C.c_void_p), None)
    assert n == 9 and np.all(np.diff(mu[:n]) > 0)
    plan = C.c_void_p()
    for shape in ((3, 2, n - 1), (3, 2, n + 1)):
        d = make_instance(surface="flat", tau_a=0.1, greek=S.hg_greek(0.7, 20), n_mu=8, vza=30.0,
                          bottom_source=np.zeros(shape), bottom_source_only=1)
        assert lib.ocrt_plan_create(C.byref(d), 1, 0, 0, C.byref(plan)) == -7
        assert not plan.value
    # the adjoint does not take a bottom source
    d = make_instance(surface="flat", tau_a=0.1, greek=S.hg_greek(0.7, 20), n_mu=8, vza=30.0,
                      bottom_source=np.zeros((3, 2, n)))
    assert lib.ocrt_plan_create_ex(C.byref(d), 1, 0, 2, C.byref(plan)) == -1
