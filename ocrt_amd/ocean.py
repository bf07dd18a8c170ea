"""Coupled atmosphere-ocean solve on the GPU: the atmosphere over a rough
(or flat) sea, the water body beneath it, and the light that crosses the
surface both ways.

The chain follows the reference's coupled driver
(code/OCRT_C/src/rt_solver.c:3468-6663, rt_solve_case_ocean) in its pass
structure, on batched plans:

  1. atmosphere pass 1 (ocrt_plan, the sea surface as the lower boundary):
     TOA Stokes vector and the diffuse downward field at the bottom of the
     atmosphere at every ring node (rt_solve_case_pol_for_ocean, :1468);
  2. air -> water (ocrt_couple_atm_to_water_batch, rt_air_water_coupling.c:
     244-478): that field mapped onto the water ring, just below the surface,
     for every instance that shares rings and surface in one call;
  3. in-water SOS (in-water columns of an ocrt_plan: rt_water_rt.c:2104-2236,
     5966-6034): the refracted direct beam, weighted by the atmosphere's
     direct transmittance and the surface's direct transmission, plus the
     transmitted skylight as a diffuse top source, with the water-side
     reflection of the surface fed back every order;
  4. water -> air (ocrt_couple_water_to_atm_batch,
     rt_air_water_coupling.c:598-780):
     the upwelling field just below the surface mapped onto the atmosphere's
     ring just above it, the water-leaving radiance;
  5. atmosphere pass 2 (ocrt_plan with the bottom source, bottom_source_only,
     rt_solver.c:5131-5160): the water-leaving radiance carried to the TOA
     through the scattering atmosphere and the surface's reflection.

TOA result = pass 1 + pass 2.  The light that leaves the water and is sent
back into it by the atmosphere is not followed again (one crossing each way,
as in the reference's pass structure).

Normalisation: the reference's Convention B (F0 = pi, rho = I_TOA / mu_sun).
The direct beam enters the water with normal flux pi mu_s e^(-tau/mu_s)
T(mu_s) / mu_w (T the direct transmittance of the surface: flat Fresnel or
the slope-integrated Cox-Munk one, ocrt_interface_direct), i.e. the in-water
column's beam weight is mu_s e^(-tau/mu_s) T(mu_s) / mu_w.

Not reproduced: the reference's equivalent-beam injection of the skylight
(rt_water_rt.c:681-900) -- here the transmitted skylight enters as the
column's diffuse top boundary field -- and its many environment-switched
diagnostic variants.
"""
import ctypes as C
import math

import numpy as np

from . import Plan, library
from .abi import make_instance

VP = C.c_void_p
D = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(VP)


def ring(n_mu, vza_deg):
    """(mu, w) of the ring a plan builds for n_mu Gauss nodes and a view angle."""
    lib = library()
    mu, w = np.zeros(n_mu + 1), np.zeros(n_mu + 1)
    n = lib.ocrt_ring(n_mu, C.c_double(vza_deg), _p(mu), _p(w))
    if n < 2:
        raise ValueError("bad ring")
    return mu[:n].copy(), w[:n].copy()


def snell_water(mu_air, n_water):
    """In-water cosine of a direction of cosine mu_air in air."""
    s = math.sqrt(max(0.0, 1.0 - mu_air * mu_air)) / n_water
    return math.sqrt(max(0.0, 1.0 - s * s))


def flat_transmittance(mu_air, n_water):
    """Unpolarised flux transmittance of the flat surface, air -> water."""
    st = math.sqrt(max(0.0, 1.0 - mu_air * mu_air)) / n_water
    ct = math.sqrt(max(0.0, 1.0 - st * st))
    rs = (mu_air - n_water * ct) / (mu_air + n_water * ct)
    rp = (n_water * mu_air - ct) / (n_water * mu_air + ct)
    return 1.0 - 0.5 * (rs * rs + rp * rp)


def _groups(keys):
    """Indices grouped by equal key, in first-seen order."""
    g = {}
    for i, k in enumerate(keys):
        g.setdefault(k, []).append(i)
    return list(g.values())


def _couple_down(lib, m_max, mu_a, w_a, fields, n_water, q_conv, mu_w, wind, sigma_type):
    """ocrt_couple_atm_to_water_batch for instances sharing nodes and surface:
    fields [B][M][3][>= n_a] -> [B][3][M][n_w]."""
    B, M, n_a, n_w = len(fields), m_max + 1, mu_a.size, mu_w.size
    f = np.ascontiguousarray(np.stack([np.transpose(x[:, :, :n_a], (1, 0, 2)) for x in fields]))
    out = np.zeros((B, 3, M, n_w))
    rc = lib.ocrt_couple_atm_to_water_batch(B, m_max, n_a, _p(mu_a), _p(w_a), _p(f), C.c_double(n_water), q_conv,
                                            _p(mu_w), n_w, C.c_double(wind), sigma_type, _p(out), None)
    if rc != 0:
        raise RuntimeError(f"ocrt_couple_atm_to_water_batch failed with code {rc}")
    return out


def _couple_up(lib, m_max, mu_w, fields, n_water, q_conv, wind, sigma_type, mu_a):
    """ocrt_couple_water_to_atm_batch for instances sharing nodes and surface:
    fields [B][M][3][>= n_w] -> [B][3][M][n_a]."""
    B, M, n_w, n_a = len(fields), m_max + 1, mu_w.size, mu_a.size
    f = np.ascontiguousarray(np.stack([np.transpose(x[:, :, :n_w], (1, 0, 2)) for x in fields]))
    out = np.zeros((B, 3, M, n_a))
    rc = lib.ocrt_couple_water_to_atm_batch(B, m_max, n_w, _p(mu_w), _p(f), C.c_double(n_water), q_conv,
                                            C.c_double(wind), sigma_type, n_a, _p(mu_a), _p(out), None)
    if rc != 0:
        raise RuntimeError(f"ocrt_couple_water_to_atm_batch failed with code {rc}")
    return out


def direct_transmittance(lib, mu_s, wind, sigma_type, n_water):
    """Direct-beam transmittance of the surface at solar cosine mu_s."""
    if wind > 0.0:
        mu = np.array([mu_s])
        out = np.zeros(1)
        rc = lib.ocrt_interface_direct(3, _p(mu), 1, C.c_double(wind), sigma_type, C.c_double(n_water), _p(out),
                                       None)
        if rc != 0:
            raise RuntimeError(f"ocrt_interface_direct failed with code {rc}")
        return float(out[0])
    return flat_transmittance(mu_s, n_water)


class Water:
    """The water body under one atmosphere instance: optical depth, pure-water
    and particle scattering albedos, particle Greek coefficients [4, L+1]
    (None: pure water), layers and Gauss nodes per hemisphere."""

    def __init__(self, tau, omega_w, omega_p, greek=None, n_layers=100, n_mu=48, depol=0.039):
        self.tau, self.omega_w, self.omega_p = float(tau), float(omega_w), float(omega_p)
        self.greek = None if greek is None else np.ascontiguousarray(greek, dtype=np.float64)
        self.n_layers, self.n_mu, self.depol = int(n_layers), int(n_mu), float(depol)


def coupled_solve(atm_descs, waters, precision="mixed", max_orders_water=150, tolerance=1e-7, stage_times=None):
    """Coupled solve of instances (atmosphere descriptors over a sea surface,
    surface "black_fresnel_ocean" or "flat", with one Water each).  Returns a
    dict of arrays: rho_I/Q/U (TOA total), rho_atm_* (pass 1), rho_wl_*
    (pass 2: the water-leaving part at the TOA), Lu_I (upwelling radiance just
    below the surface at the refracted view, mode-summed at the view
    azimuth), beam_scale, orders_water."""
    import time
    lib = library()
    N = len(atm_descs)
    if len(waters) != N:
        raise ValueError("one Water per instance")
    t = time.perf_counter()
    mm = max(d.m_max for d in atm_descs)
    for d in atm_descs:
        if d.surface not in (2, 3):
            raise ValueError("the coupled ocean needs a sea surface (black_fresnel_ocean or flat)")
    # 1. atmosphere, pass 1
    pa = Plan(atm_descs, precision=precision)
    pa.export_nodes()
    pa.execute()
    ra = pa.results(modes=True)
    pitch_a = max(d.n_mu + 1 for d in atm_descs)
    _, boa = pa.node_fields(pitch_a)
    pa.close()
    t1 = time.perf_counter()
    # 2. air -> water, and the in-water columns; instances that share their
    # rings and surface are coupled in one batched call
    keep, rings_a, rings_w, winds, mus_w = [], [], [], [], []
    beams = np.zeros(N)
    for i, (d, w) in enumerate(zip(atm_descs, waters)):
        rings_a.append(ring(d.n_mu, d.vza_deg))
        mu_vw = snell_water(math.cos(math.radians(d.vza_deg)), d.n_water)
        rings_w.append(ring(w.n_mu, math.degrees(math.acos(mu_vw)))[0])
        winds.append(d.wind_speed if d.surface == 2 else 0.0)
    skies = [None] * N
    keys = [(d.m_max, rings_a[i][0].tobytes(), rings_a[i][1].tobytes(), rings_w[i].tobytes(), d.n_water,
             d.q_convention, winds[i], d.sigma_type) for i, d in enumerate(atm_descs)]
    for g in _groups(keys):
        d = atm_descs[g[0]]
        mu_a, w_a = rings_a[g[0]]
        sky = _couple_down(lib, d.m_max, mu_a, w_a, [boa[i, : d.m_max + 1] for i in g], d.n_water, d.q_convention,
                           rings_w[g[0]], winds[g[0]], d.sigma_type)
        for j, i in enumerate(g):
            skies[i] = np.ascontiguousarray(sky[j])
    wdescs = []
    for i, (d, w) in enumerate(zip(atm_descs, waters)):
        mus = math.cos(math.radians(d.sza_deg))
        mu_sw = snell_water(mus, d.n_water)
        vza_w = math.degrees(math.acos(snell_water(math.cos(math.radians(d.vza_deg)), d.n_water)))
        wind = winds[i]
        T_dir = direct_transmittance(lib, mus, wind, d.sigma_type, d.n_water)
        beams[i] = ra["T_dir_dn"][i] * mus * T_dir / mu_sw
        greek = w.greek if (w.greek is not None and w.omega_p > 0.0) else None
        wd = make_instance(sza=math.degrees(math.acos(mu_sw)), vza=vza_w, raa=d.raa_deg, wavelength=d.wavelength_nm,
                           surface="black", n_water=d.n_water, wind=max(wind, 0.0), sigma_type=d.sigma_type,
                           q_convention=d.q_convention, depol=w.depol, greek=greek, n_mu=w.n_mu,
                           n_layers=w.n_layers, m_max=d.m_max, max_orders=max_orders_water, tolerance=tolerance,
                           water=(w.tau, w.omega_w, w.omega_p), top_reflect=2 if wind > 0.0 else 1,
                           top_source=skies[i], beam_scale=beams[i])
        wdescs.append(wd)
        keep.append(skies[i])
    t2 = time.perf_counter()
    pw = Plan(wdescs, precision=precision)
    pw.export_nodes()
    pw.execute()
    rw = pw.results(modes=True)
    pitch_w = max(wd.n_mu + 1 for wd in wdescs)
    up_w, _ = pw.node_fields(pitch_w)
    pw.close()
    t3 = time.perf_counter()
    # 4. water -> air, batched like step 2, and 5. atmosphere pass 2 with the
    # water-leaving source
    wls = [None] * N
    keys = [(d.m_max, rings_w[i].tobytes(), d.n_water, d.q_convention, winds[i], d.sigma_type,
             rings_a[i][0].tobytes()) for i, d in enumerate(atm_descs)]
    for g in _groups(keys):
        d = atm_descs[g[0]]
        n_w = rings_w[g[0]].size
        wl = _couple_up(lib, d.m_max, rings_w[g[0]], [up_w[i, : d.m_max + 1, :, :n_w] for i in g], d.n_water,
                        d.q_convention, winds[g[0]], d.sigma_type, rings_a[g[0]][0])
        for j, i in enumerate(g):
            wls[i] = np.ascontiguousarray(wl[j])
    bdescs = []
    for i, d in enumerate(atm_descs):
        # the same atmosphere (every field, its aerosol tables included), with
        # the water-leaving field as its only source
        bd = type(d).from_buffer_copy(d)
        bd.decouple_sunglint = 1
        bd.bottom_source = wls[i].ctypes.data_as(D)
        bd.bottom_m_max, bd.bottom_n_mu = d.m_max, rings_a[i][0].size
        bd.bottom_source_only = 1
        bdescs.append(bd)
        keep.append(wls[i])
    t4 = time.perf_counter()
    pb = Plan(bdescs, precision=precision)
    pb.execute()
    rb = pb.results(modes=True)
    pb.close()
    t5 = time.perf_counter()
    if stage_times is not None:
        stage_times.update({"atm_pass1_s": t1 - t, "couple_down_s": t2 - t1, "water_s": t3 - t2,
                            "couple_up_s": t4 - t3, "atm_pass2_s": t5 - t4})
    out = {"status": np.maximum(np.abs(ra["status"]), np.maximum(np.abs(rw["status"]), np.abs(rb["status"])))}
    for f in ("rho_I", "rho_Q", "rho_U"):
        out[f] = ra[f] + rb[f]
        out[f.replace("rho", "rho_atm")] = ra[f]
        out[f.replace("rho", "rho_wl")] = rb[f]
    # upwelling just below the surface at the refracted view: the water
    # column's view-node modes reconstructed at the view azimuth
    phi = np.radians([d.raa_deg for d in atm_descs]) + math.pi
    m = np.arange(mm + 1)
    wts = np.where(m == 0, 1.0, 2.0 * np.cos(np.outer(phi, m)))
    out["Lu_I"] = np.sum(rw["toa_modes"][:, :, 0] * wts, axis=1)
    out["beam_scale"] = beams
    out["orders_water"] = rw["n_orders_used"]
    out["orders_atm"] = np.maximum(ra["n_orders_used"], rb["n_orders_used"])
    out["toa_modes"] = ra["toa_modes"] + rb["toa_modes"]
    return out


def coupled_jacobian(make, chl, aod_scale=1.0, rel=1e-3, precision="fp64", **solve_kw):
    """Jacobians of the TOA (rho_I, rho_Q, rho_U) with respect to ln Chl and
    ln AOD by central differences, every perturbed instance in one batched
    coupled solve with the unperturbed ones.  make(chl, aod_scale) returns
    (atmosphere descriptors, Water list) for the chlorophyll array and the
    AOD scale; rel is the relative step.  Returns (values [N, 3],
    d/d ln Chl [N, 3], d/d ln AOD [N, 3])."""
    chl = np.asarray(chl, dtype=np.float64)
    N = chl.size
    sets = [(chl, aod_scale), (chl * (1.0 + rel), aod_scale), (chl * (1.0 - rel), aod_scale),
            (chl, aod_scale * (1.0 + rel)), (chl, aod_scale * (1.0 - rel))]
    descs, waters = [], []
    for c, a in sets:
        d, w = make(c, a)
        descs += d
        waters += w
    out = coupled_solve(descs, waters, precision=precision, **solve_kw)
    v = np.stack([out["rho_I"], out["rho_Q"], out["rho_U"]], axis=1).reshape(5, N, 3)
    step = np.log(1.0 + rel) - np.log(1.0 - rel)   # the steps in ln Chl and ln AOD
    dchl = (v[1] - v[2]) / step
    daod = (v[3] - v[4]) / step
    return v[0], dchl, daod
