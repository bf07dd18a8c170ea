"""The coupled atmosphere-ocean chain on the GPU (ocrt_amd.ocean.coupled_solve)
against the same chain run on the CPU from pinned parts: the oracle's solves
(atmosphere pass 1 with its bottom-of-atmosphere field, the in-water column
with the transmitted skylight as its top source, atmosphere pass 2 with the
water-leaving bottom source; each pinned to the reference's own solvers) and
the reference's own coupling functions and direct transmittance
(rt_air_water_couple_atm_to_water / _water_to_atm,
surface_T_aw_coxmunk_direct, from oracle/_ref).  FP64 within 1e-10 of rho_I;
the mixed form within the error budget.  Also: a water body that does not
scatter sends nothing back (the water-leaving part is zero and the total is
pass 1)."""
import ctypes as C
import math

import numpy as np
import pytest

from test_gpu_coupling import BoaExport, Coupled, _bind, _bind_up

pytestmark = pytest.mark.gpu
D = C.POINTER(C.c_double)


def _p(a):
    return None if a is None else a.ctypes.data_as(D)


def _nodes(oracle, d):
    mm = d.m_max
    top = np.zeros((mm + 1) * 3 * 700)
    bot = np.zeros((mm + 1) * 3 * 700)
    orders = np.zeros(mm + 1, dtype=np.int32)
    n = oracle.oracle_solve_nodes(C.byref(d), _p(top), _p(bot), orders.ctypes.data_as(C.c_void_p))
    assert n > 0, n
    k = (mm + 1) * 3 * n
    return n, top[:k].reshape(mm + 1, 3, n), bot[:k].reshape(mm + 1, 3, n)


def cpu_chain(oracle, ref_lib, d, water, ring):
    """The coupled chain on the CPU from the oracle and the reference's couplings."""
    from ocrt_amd import ocean
    from ocrt_amd.abi import Result, make_instance
    down, up = _bind(ref_lib), _bind_up(ref_lib)
    ref_lib.surface_T_aw_coxmunk_direct.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
    ref_lib.surface_T_aw_coxmunk_direct.restype = C.c_double
    mm = d.m_max
    r1 = Result()
    assert oracle.oracle_solve(C.byref(d), C.byref(r1), None, None, None) == 0
    n_a, _, bot = _nodes(oracle, d)
    mu_a, w_a = ring(d.n_mu, d.vza_deg)
    assert mu_a.size == n_a
    mus, muv = math.cos(math.radians(d.sza_deg)), math.cos(math.radians(d.vza_deg))
    mu_sw, mu_vw = ocean.snell_water(mus, d.n_water), ocean.snell_water(muv, d.n_water)
    vza_w = math.degrees(math.acos(mu_vw))
    mu_w, w_wr = ring(water.n_mu, vza_w)
    n_w = mu_w.size
    wind = d.wind_speed if d.surface == 2 else 0.0
    # air -> water: the reference's coupling of the oracle's BOA field
    I, Q, U = (np.ascontiguousarray(bot[:, s, :]).reshape(-1) for s in range(3))
    b = BoaExport(mm, n_a, _p(mu_a), _p(I), _p(Q), _p(U), 0.0, mus, 0)
    sky = [np.zeros((mm + 1) * n_w) for _ in range(3)]
    mw = np.zeros(n_w)
    c = Coupled(mm, n_w, _p(mw), _p(sky[0]), _p(sky[1]), _p(sky[2]))
    assert down(C.byref(b), d.n_water, d.q_convention, _p(mu_w), n_w, _p(w_a), wind, d.sigma_type, C.byref(c)) == 0
    sky = np.stack([s.reshape(mm + 1, n_w) for s in sky])
    T = (ref_lib.surface_T_aw_coxmunk_direct(mus, d.n_water, wind, d.sigma_type) if wind > 0.0
         else ocean.flat_transmittance(mus, d.n_water))
    beam = r1.T_dir_dn * mus * T / mu_sw
    wd = make_instance(sza=math.degrees(math.acos(mu_sw)), vza=vza_w, raa=d.raa_deg, wavelength=d.wavelength_nm,
                       surface="black", n_water=d.n_water, wind=wind, sigma_type=d.sigma_type,
                       q_convention=d.q_convention, depol=water.depol, greek=water.greek, n_mu=water.n_mu,
                       n_layers=water.n_layers, m_max=mm, max_orders=150, tolerance=1e-7,
                       water=(water.tau, water.omega_w, water.omega_p), top_reflect=2 if wind > 0.0 else 1,
                       top_source=sky, beam_scale=beam)
    _, top_w, _ = _nodes(oracle, wd)
    # water -> air: the reference's coupling of the oracle's upwelling field
    Iw, Qw, Uw = (np.ascontiguousarray(top_w[:, s, :]).reshape(-1) for s in range(3))
    ww = np.ascontiguousarray(w_wr)
    wl = [np.zeros((mm + 1) * n_a) for _ in range(3)]
    assert up(_p(Iw), _p(Qw), _p(Uw), _p(mu_w), _p(ww), n_w, mm, d.n_water, d.q_convention, wind, d.sigma_type,
              _p(mu_a), n_a, _p(w_a), _p(wl[0]), _p(wl[1]), _p(wl[2])) == 0
    wl = np.stack([x.reshape(mm + 1, n_a) for x in wl])
    d2 = type(d).from_buffer_copy(d)
    d2.decouple_sunglint = 1
    d2.bottom_source = wl.ctypes.data_as(D)
    d2.bottom_m_max, d2.bottom_n_mu, d2.bottom_source_only = mm, n_a, 1
    r2 = Result()
    assert oracle.oracle_solve(C.byref(d2), C.byref(r2), None, None, None) == 0
    return {"rho_I": r1.rho_I + r2.rho_I, "rho_Q": r1.rho_Q + r2.rho_Q, "rho_U": r1.rho_U + r2.rho_U,
            "rho_wl_I": r2.rho_I, "beam": beam}


def _cases():
    from ocrt_amd import ocean, synthetic as S
    from ocrt_amd.abi import make_instance
    gk = S.hg_greek(0.7, 30)
    hyd = S.hg_greek(0.9, 60, dolp=0.3)
    atm = [make_instance(surface="black_fresnel_ocean", wind=5.0, tau_a=0.15, ssa_a=0.95, greek=gk, n_mu=12,
                         n_layers=30, m_max=4, sza=40.0, vza=30.0, raa=60.0, wavelength=490.0,
                         tau_R=0.0941 * (490 / 555.0) ** -4.08),
           make_instance(surface="flat", tau_a=0.1, ssa_a=0.9, greek=gk, n_mu=10, n_layers=25, m_max=3,
                         sza=30.0, vza=20.0, raa=120.0, wavelength=555.0),
           make_instance(surface="black_fresnel_ocean", wind=9.0, n_mu=12, n_layers=20, m_max=4, sza=55.0,
                         vza=45.0, raa=150.0, wavelength=443.0, tau_R=0.0941 * (443 / 555.0) ** -4.08)]
    water = [ocean.Water(12.0, 0.25, 0.5, greek=hyd, n_layers=40, n_mu=16),
             ocean.Water(20.0, 0.35, 0.0, n_layers=50, n_mu=12),
             ocean.Water(8.0, 0.1, 0.6, greek=hyd, n_layers=30, n_mu=14)]
    return atm, water


def test_coupled_chain_equals_cpu_chain(oracle, ref_lib):
    from ocrt_amd import ocean
    atm, water = _cases()
    print("[gpu] coupled atmosphere-ocean chain: 3 instances, fp64 and mixed", flush=True)
    a = ocean.coupled_solve(atm, water, precision="fp64", max_orders_water=150)
    b = ocean.coupled_solve(atm, water, precision="mixed", max_orders_water=150)
    assert (a["status"] == 0).all() and (b["status"] == 0).all()
    for i, (d, w) in enumerate(zip(atm, water)):
        ref = cpu_chain(oracle, ref_lib, d, w, ocean.ring)
        scale = abs(ref["rho_I"])
        assert abs(a["beam_scale"][i] - ref["beam"]) <= 1e-12 * ref["beam"], i
        assert ref["rho_wl_I"] > 1e-4 * scale, i   # the water body contributes
        for f in ("rho_I", "rho_Q", "rho_U", "rho_wl_I"):
            assert abs(a[f][i] - ref[f]) <= 1e-10 * scale, (i, f, a[f][i], ref[f])
        ei = abs(b["rho_I"][i] - ref["rho_I"]) / scale
        eq = max(abs(b["rho_Q"][i] - ref["rho_Q"]), abs(b["rho_U"][i] - ref["rho_U"]))
        print(f"[gpu] coupled instance {i}: rho_I {ref['rho_I']:.6f} (water-leaving {ref['rho_wl_I']:.2e}); "
              f"mixed rel I {ei:.1e}, abs Q/U {eq:.1e}", flush=True)
        assert ei <= 1e-4 and eq <= 1e-4


def test_non_scattering_water_leaves_nothing():
    from ocrt_amd import ocean
    atm, water = _cases()
    dark = [ocean.Water(w.tau, 0.0, 0.0, n_layers=w.n_layers, n_mu=w.n_mu) for w in water]
    a = ocean.coupled_solve(atm, dark, precision="fp64")
    assert np.all(a["rho_wl_I"] == 0.0) and np.all(a["rho_I"] == a["rho_atm_I"])


def test_config4_full_batch_mixed_vs_fp64():
    """The full coupled workload: 512 instances over 100 wavelengths
    (330-1100 nm), Chl-driven water bodies with the 1000-term hydrosol,
    300 + 100 layers, 64 streams both sides, modes 0..127: every instance of
    the mixed form within the error budget of the FP64 form."""
    from ocrt_amd import ocean, synthetic as S
    d, w, meta = S.config4()
    print(f"[gpu] coupled workload: {len(d)} instances, mixed then fp64", flush=True)
    st = {}
    m = ocean.coupled_solve(d, w, precision="mixed", stage_times=st)
    print(f"[gpu] mixed stages {st}", flush=True)
    f = ocean.coupled_solve(d, w, precision="fp64")
    assert (m["status"] == 0).all() and (f["status"] == 0).all()
    ei = np.abs(m["rho_I"] - f["rho_I"]) / np.abs(f["rho_I"])
    eq = np.maximum(np.abs(m["rho_Q"] - f["rho_Q"]), np.abs(m["rho_U"] - f["rho_U"]))
    print(f"[gpu] coupled workload: mixed vs fp64 worst rel I {ei.max():.2e}, abs Q/U {eq.max():.2e}; "
          f"water-leaving share of rho_I {np.min(f['rho_wl_I'] / f['rho_I']):.3f}..{np.max(f['rho_wl_I'] / f['rho_I']):.3f}",
          flush=True)
    assert ei.max() <= 1e-4 and eq.max() <= 1e-4
    assert np.all(f["rho_wl_I"] >= 0.0)


def test_config4_jacobians_converge_in_the_step():
    """Jacobians of the TOA Stokes vector with respect to ln Chl and ln AOD by
    central differences (FP64): steps 1e-3 and 1e-4 agree within 1e-3 of the
    Jacobian's scale on a subset of the coupled workload."""
    from ocrt_amd import ocean, synthetic as S
    _, _, meta = S.config4(n_inst=6, seed=11, n_mu=24, n_layers=60, water_layers=40, m_max=16, L_hyd=400)
    chl = meta["chl"]

    def make(c, a):
        d, w, _ = S.config4(n_inst=6, seed=11, n_mu=24, n_layers=60, water_layers=40, m_max=16, L_hyd=400, chl=c,
                            aod_scale=a)
        return d, w
    print("[gpu] coupled Jacobians (ln Chl, ln AOD), 6 instances, two steps", flush=True)
    v1, c1, a1 = ocean.coupled_jacobian(make, chl, rel=1e-3)
    v2, c2, a2 = ocean.coupled_jacobian(make, chl, rel=1e-4)
    assert np.array_equal(v1, v2)
    for J1, J2 in ((c1, c2), (a1, a2)):
        scale = np.max(np.abs(J1[:, 0]))
        assert scale > 0.0
        assert np.max(np.abs(J1 - J2)) <= 1e-3 * scale, (np.max(np.abs(J1 - J2)), scale)
