"""The reference Python package's batch entry point (ocrt_amd.atmos_batch,
mirroring code/OCRT_Python/ocrt_py/atmos_batch.py:885-1013).

CPU: the medium objects map onto explicit layering that reproduces the
levels' optical depths and scattering fractions; unserved options are
refused before any device work.
GPU: solve_atm_batch over medium objects built like the reference's builders
equals the oracle solving the same atmosphere (FP64, 1e-11 of rho_I), and
its BOA fields give the oracle's hemispheric diffuse transmittance.
"""
import ctypes as C
import math
from types import SimpleNamespace

import numpy as np
import pytest

from conftest import oracle_solve

RON = 2.0 * (1.0 - 0.0279) / (2.0 + 0.0279)


def ring(oracle, n_gl, mu_view):
    mu = np.zeros(n_gl + 1)
    w = np.zeros(n_gl + 1)
    vi = C.c_int()
    n = oracle.oracle_ring(n_gl, C.c_double(mu_view), mu.ctypes.data_as(C.c_void_p),
                           w.ctypes.data_as(C.c_void_p), C.byref(vi))
    return mu[:n], w[:n], vi.value


def medium(mu_ring, w_ring, mu_sun, ray, aer_ext, ssa, greek=None):
    """A medium object laid out like the reference's build_atm_aerosol
    (ocrt_py/atmos.py:33-79, 314-376): ring with the solar slot, levels top
    down, level 0 Rayleigh-only, level k carrying layer k - 1's fractions."""
    n = mu_ring.size
    nt = ray.size
    dt = ray + aer_ext
    h = np.concatenate([[0.0], np.cumsum(dt)])
    xdel = np.concatenate([[0.0], ssa * aer_ext / dt])
    ydel = np.concatenate([[1.0], ray / dt])
    rm = np.zeros(2 * n + 1)
    gb = np.zeros(2 * n + 1)
    rm[n] = -mu_sun
    rm[n + 1:] = mu_ring
    rm[:n] = -mu_ring[::-1]
    gb[n + 1:] = w_ring
    gb[:n] = w_ring[::-1]
    a = SimpleNamespace(h=h, ch=0.5 * np.exp(-h / mu_sun), xdel=xdel, ydel=ydel, rm=rm, gb=gb, mu_sun=mu_sun,
                        beta0=1.0, beta2=0.5 * RON, gamma2=-RON * math.sqrt(1.5), alpha2=3.0 * RON)
    if greek is not None:
        a.betal, a.gammal, a.alphal, a.zetal = (greek[k] for k in range(4))
    assert nt == h.size - 1
    return a


def profile(nt, seed):
    rng = np.random.default_rng(seed)
    ray = 0.11 / nt * (1.0 + 0.3 * np.sin(np.arange(nt)))
    aer = 0.2 / nt * np.exp(-np.arange(nt)[::-1] / (0.3 * nt)) * rng.uniform(0.8, 1.2, nt)
    return ray, aer


def test_layers_reproduce_the_medium():
    from ocrt_amd.atmos_batch import _layers
    nt = 17
    ray, aer = profile(nt, 1)
    a = medium(np.array([0.2, 0.6, 0.9, 1.0]), np.array([0.3, 0.4, 0.3, 0.0]), 0.7, ray, aer, 0.88)
    lay, x0, y0 = _layers(a, nt)
    dt = lay[0] + lay[1] + lay[3]
    h = np.concatenate([[0.0], np.cumsum(dt)])
    assert np.allclose(h, a.h, rtol=1e-15, atol=1e-17)
    assert np.allclose(lay[2] * lay[1] / dt, a.xdel[1:], rtol=1e-14)
    assert np.allclose(lay[0] / dt, a.ydel[1:], rtol=1e-14)
    assert (x0, y0) == (0.0, 1.0)


@pytest.mark.parametrize("kw", [dict(bottom_source=(0, 0, 0)),
                                dict(use_value_kernel=True, aer_p11=[], mode="aerosol_coxmunk"),
                                dict(n_phi_quad=512), dict(mode="flat")])
def test_unserved_options_are_refused(kw):
    from ocrt_amd.atmos_batch import solve_atm_batch
    a = SimpleNamespace()
    with pytest.raises(ValueError):
        solve_atm_batch([a], [1], [0.8], [0.5], [5.0], 3, 4, 2, 20, 1e-7, **kw)


@pytest.mark.gpu
@pytest.mark.parametrize("value_kernel", [False, True])
def test_solve_atm_batch_equals_oracle(oracle, value_kernel):
    from ocrt_amd import synthetic as S
    from ocrt_amd.abi import make_instance
    from ocrt_amd.atmos_batch import solve_atm_batch
    nt, n_gl, m_max = 40, 16, 6
    gk = S.hg_greek(0.72, 40)
    # the value kernel's table: HG with a backscatter lobe on a 1-degree grid
    th = np.arange(0.0, 180.5, 1.0)
    cth = np.cos(np.radians(th))
    p11 = (1 - 0.72 ** 2) / (1 + 0.72 ** 2 - 2 * 0.72 * cth) ** 1.5 + 0.5 * np.exp(-((180.0 - th) / 5.0) ** 2)
    table = (th, p11) if value_kernel else None
    geo = [(35.0, 20.0, 60.0, 4.0, 0.9), (62.0, 48.0, 150.0, 9.0, 0.95), (15.0, 5.0, 10.0, 1.5, 0.85)]
    atms, vjs, mus, dphis, winds, descs = [], [], [], [], [], []
    for b, (sza, vza, raa, ws, ssa) in enumerate(geo):
        mu_ring, w_ring, vi = ring(oracle, n_gl, math.cos(math.radians(vza)))
        ray, aer = profile(nt, 10 + b)
        mu_sun = math.cos(math.radians(sza))
        atms.append(medium(mu_ring, w_ring, mu_sun, ray, aer, ssa, greek=gk))
        vjs.append(vi + 1)
        mus.append(mu_sun)
        dphis.append(math.radians(raa))
        winds.append(ws)
        descs.append(make_instance(surface="black_fresnel_ocean", wind=ws, decouple_sunglint=1, sza=sza, vza=vza,
                                   raa=raa, tau_R=float(ray.sum()), tau_a=float(aer.sum()), ssa_a=ssa, greek=gk,
                                   n_mu=n_gl, m_max=m_max, layers=np.stack([ray, aer, np.full(nt, ssa),
                                                                            np.zeros(nt)]), phase_table=table))
    n_mu = n_gl + 1
    print("[gpu] solve_atm_batch: 3 cases, 16 streams + view node, 40 layers, m 0..6", flush=True)
    out = solve_atm_batch(atms, vjs, mus, dphis, winds, n_mu, nt, m_max, 20, 1e-7,
                          mode="aerosol_black_fresnel_ocean", L_max=40, use_value_kernel=value_kernel,
                          aer_p11=[table] * 3 if value_kernel else None)
    assert out["boa_I"].shape == (3, m_max + 1, n_mu)
    for b, d in enumerate(descs):
        rc, r, modes, orders = oracle_solve(oracle, d)
        assert rc == 0
        for f in ("rho_I", "rho_Q", "rho_U", "I_TOA", "Q_TOA", "U_TOA"):
            assert abs(out[f][b] - getattr(r, f)) <= 1e-11 * abs(r.rho_I), (b, f, out[f][b], getattr(r, f))
        # the BOA export gives the hemispheric diffuse transmittance of the solve
        assert abs(out["T_diff_dn_hemi"][b] - r.T_diff_dn_hemi) <= 1e-11 * abs(r.T_diff_dn_hemi), b
        assert out["n_orders"][b] == max(orders)
        assert np.array_equal(out["mu_quad"][b], atms[b].rm[n_mu + 1:])


@pytest.mark.gpu
def test_node_fields_agree_with_the_view_modes(oracle):
    """The per-node TOA export at the view node is the per-mode TOA value the
    results report (rough-ocean seed removed in both), bit for bit, and every
    field is zero past the instance's ring and modes."""
    import ocrt_amd
    from ocrt_amd import synthetic as S
    from ocrt_amd.abi import make_instance
    gk = S.hg_greek(0.7, 30)
    descs = [make_instance(surface="black_fresnel_ocean", wind=7.0, tau_a=0.2, ssa_a=0.9, greek=gk, n_mu=12,
                           n_layers=30, m_max=5, sza=40.0, vza=25.0, raa=80.0),
             make_instance(surface="flat", tau_a=0.1, ssa_a=0.95, greek=gk, n_mu=20, n_layers=45, m_max=3,
                           sza=55.0, vza=62.0, raa=130.0)]
    plan = ocrt_amd.Plan(descs, precision="mixed")
    plan.export_nodes()
    plan.execute()
    res = plan.results(modes=True)
    pitch = 24
    up, dn = plan.node_fields(pitch)
    plan.close()
    for b, d in enumerate(descs):
        _, _, vi = ring(oracle, d.n_mu, math.cos(math.radians(d.vza_deg)))
        n_ring = int(res["n_ring"][b])
        for m in range(d.m_max + 1):
            assert np.array_equal(up[b, m, :, vi], res["toa_modes"][b, m]), (b, m)
        assert not up[b, d.m_max + 1:].any() and not dn[b, d.m_max + 1:].any()
        assert not up[b, :, :, n_ring:].any() and not dn[b, :, :, n_ring:].any()
        assert np.all(dn[b, 0, 0, :n_ring] > 0.0)   # diffuse skylight reaches every node at m = 0


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["aerosol_rayleigh", "aerosol_only", "rayleigh_only"])
def test_solve_atm_batch_equals_reference_package(name):
    """The mirror against the reference package's own solve_atm_batch on media
    from its own builders (tests/golden/ref_atmos_batch.json, written by
    make_atmos_batch_golden.py in the build container): an aerosol + Rayleigh
    batch, aerosol only (tau_R = 0) and Rayleigh only.  TOA values, the per-mode
    BOA fields and the hemispheric diffuse transmittance within 1e-11, equal
    order counts."""
    import json
    import os
    from conftest import GOLDEN
    from ocrt_amd.atmos_batch import solve_atm_batch
    with open(os.path.join(GOLDEN, "ref_atmos_batch.json")) as f:
        rec = {r["name"]: r for r in json.load(f)}[name]
    a = rec["args"]
    atms = [SimpleNamespace(**{k: np.asarray(v) for k, v in m.items()}) for m in rec["media"]]
    print(f"[gpu] solve_atm_batch vs the reference package: {name}, {len(atms)} cases", flush=True)
    out = solve_atm_batch(atms, rec["vjs"], rec["mu_suns"], rec["dphis"], rec["winds"], a["n_mu"], a["nt"],
                          a["m_max"], a["max_iterations"], a["tolerance"], mode=a["mode"], L_max=a["L_max"])
    ref = {k: np.asarray(v) for k, v in rec["out"].items()}
    scale = np.abs(ref["rho_I"])
    for f in ("rho_I", "rho_Q", "rho_U", "I_TOA", "Q_TOA", "U_TOA"):
        assert np.all(np.abs(out[f] - ref[f]) <= 1e-11 * scale), (f, out[f], ref[f])
    for f in ("boa_I", "boa_Q", "boa_U"):
        assert out[f].shape == ref[f].shape, f
        assert np.max(np.abs(out[f] - ref[f])) <= 1e-11 * np.max(np.abs(ref["boa_I"])), f
    assert np.all(np.abs(out["T_diff_dn_hemi"] - ref["T_diff_dn_hemi"]) <= 1e-11 * np.abs(ref["T_diff_dn_hemi"]))
    assert list(out["n_orders"]) == [int(v) for v in ref["n_orders"]]


def test_reference_built_media_map_onto_layers():
    """The media of the reference package's own builders (stored in
    ref_atmos_batch.json) go through the mirror's layering with their level
    optical depths and scattering fractions reproduced."""
    import json
    import os
    from conftest import GOLDEN
    from ocrt_amd.atmos_batch import _layers
    with open(os.path.join(GOLDEN, "ref_atmos_batch.json")) as f:
        recs = json.load(f)
    for rec in recs:
        nt = rec["args"]["nt"]
        for m in rec["media"]:
            a = SimpleNamespace(**{k: np.asarray(v) for k, v in m.items()})
            lay, x0, y0 = _layers(a, nt)
            dt = lay[0] + lay[1] + lay[3]
            h = np.concatenate([[0.0], np.cumsum(dt)])
            assert np.allclose(h, a.h, rtol=1e-14, atol=1e-16), rec["name"]
            assert np.allclose(lay[1] / dt, a.xdel[1:], rtol=1e-13, atol=1e-16), rec["name"]
            assert np.allclose(lay[0] / dt, a.ydel[1:], rtol=1e-13, atol=1e-16), rec["name"]
            assert (x0, y0) == (a.xdel[0], a.ydel[0])
