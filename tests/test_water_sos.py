"""The in-water SOS (the coupled ocean's water body) in the oracle:
  - against golden vectors from the reference's own order loops for the water
    body (rt_solver_sos_pol, rt_solver_sos_pol_intrefl and
    rt_solver_sos_pol_intrefl_rough, code/OCRT_C/src/rt_solver.c:1987-2922,
    driven as rt_water_rt.c:5966-6034 does through oracle/ref_harness.c):
    per-mode upward field at the top and downward field at the bottom at
    every node, 1e-12 of the mode-0 field, equal order counts;
  - the diffuse top source (the transmitted skylight) by linearity: a solve
    with the sun and the source equals the sun alone plus the source alone
    (beam_scale 0);
  - the beam weight scales the solution linearly."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, desc_from_dict

with open(os.path.join(GOLDEN, "ref_water_sos.json")) as _f:
    CASES = json.load(_f)


def water_desc(rec):
    d, greek = desc_from_dict(rec["desc"])
    return d, greek


def nodes(lib, d):
    mm = d.m_max
    n = d.n_mu + 1
    top = np.zeros((mm + 1) * 3 * n)
    bot = np.zeros((mm + 1) * 3 * n)
    orders = np.zeros(mm + 1, dtype=np.int32)
    n = lib.oracle_solve_nodes(C.byref(d), top.ctypes.data_as(C.c_void_p), bot.ctypes.data_as(C.c_void_p),
                               orders.ctypes.data_as(C.c_void_p))
    assert n > 0, n
    k = (mm + 1) * 3 * n
    return n, top[:k].reshape(mm + 1, 3, n), bot[:k].reshape(mm + 1, 3, n), orders


@pytest.mark.parametrize("idx", range(len(CASES)))
def test_oracle_water_sos_matches_reference(oracle, idx):
    rec = CASES[idx]
    d, keep = water_desc(rec)
    n, top, bot, orders = nodes(oracle, d)
    assert n == rec["n"]   # the view sits on a Gauss node: no extra ring node
    rt = np.array(rec["top_up"]).reshape(top.shape)
    rb = np.array(rec["bot_dn"]).reshape(bot.shape)
    assert np.max(np.abs(top - rt)) <= 1e-12 * np.max(np.abs(rt[0, 0])), rec["name"]
    assert np.max(np.abs(bot - rb)) <= 1e-12 * np.max(np.abs(rb[0, 0])), rec["name"]
    assert list(orders) == rec["orders_per_m"], rec["name"]


def _top_source(n, mm, seed=3):
    rng = np.random.default_rng(seed)
    x = np.linspace(0.1, 1.0, n)[None, :]
    m = np.arange(mm + 1)[:, None]
    I = 0.05 * (0.5 + x) * np.exp(-m / 2.0) * (1 + 0.1 * rng.random((mm + 1, n)))
    Q = -0.01 * (1 - x * x) * np.exp(-m / 3.0)
    U = 0.005 * x * (1 - x) * np.exp(-m / 2.5) * (m > 0)
    return np.stack([I, Q, U])


@pytest.mark.parametrize("idx", [1, 3])
def test_top_source_is_linear(oracle, idx):
    rec = CASES[idx]
    d, keep = water_desc(rec)
    d.tolerance, d.max_orders = 1e-14, 200
    n = rec["n"]
    src = _top_source(n, d.m_max)
    d.top_source = src.ctypes.data_as(C.POINTER(C.c_double))
    d.top_m_max, d.top_n_mu = d.m_max, n
    _, t_both, b_both, _ = nodes(oracle, d)
    d.beam_scale = 0.0
    _, t_src, b_src, _ = nodes(oracle, d)
    d.beam_scale = 1.0
    d.top_source = None
    d.top_m_max, d.top_n_mu = -1, 0
    _, t_sun, b_sun, _ = nodes(oracle, d)
    scale = np.max(np.abs(t_both[0, 0]))
    assert np.max(np.abs(t_both - (t_src + t_sun))) <= 1e-11 * scale
    assert np.max(np.abs(b_both - (b_src + b_sun))) <= 1e-11 * scale   # deep bottom: tiny field
    # the source is real: it changes the field
    assert np.max(np.abs(t_src)) > 1e-3 * scale


def test_beam_scale_is_linear(oracle):
    rec = CASES[2]
    d, keep = water_desc(rec)
    _, t1, b1, o1 = nodes(oracle, d)
    d.beam_scale = 0.37
    _, t2, b2, o2 = nodes(oracle, d)
    assert np.max(np.abs(t2 - 0.37 * t1)) <= 1e-13 * np.max(np.abs(t1))
    assert list(o1) == list(o2)


def test_water_column_arguments_refused(oracle):
    from ocrt_amd.abi import make_instance
    for kw in (dict(surface="flat"), dict(ipss=1), dict(top_reflect=3), dict(top_reflect=2, wind=0.0)):
        args = dict(surface="black", water=(10.0, 0.3, 0.0), n_mu=8, n_layers=10, m_max=2, depol=0.039)
        args.update(kw)
        d = make_instance(**args)
        assert oracle.oracle_solve_nodes(C.byref(d), None, None, None) < 0, kw
