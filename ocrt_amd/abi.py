"""ctypes mirror of include/ocrt_amd.h (instance descriptor and result records).

The same records are understood by the HIP library (libocrt_amd.so), by the
CPU restatement used as the test oracle and by the reference adapter, so one
descriptor array drives all three.
"""
import ctypes as C

import numpy as np

SURFACE_BLACK = 0
SURFACE_LAMBERT = 1
SURFACE_BLACK_FRESNEL_OCEAN = 2
SURFACE_FLAT = 3

PREC_FP64 = 0
PREC_MIXED = 1

SURFACE_NAMES = {
    "black": SURFACE_BLACK,
    "lambert": SURFACE_LAMBERT,
    "black_fresnel_ocean": SURFACE_BLACK_FRESNEL_OCEAN,
    "coxmunk": SURFACE_BLACK_FRESNEL_OCEAN,
    "flat": SURFACE_FLAT,
}


class Instance(C.Structure):
    _fields_ = [
        ("sza_deg", C.c_double), ("vza_deg", C.c_double),
        ("raa_deg", C.c_double), ("wavelength_nm", C.c_double),
        ("surface", C.c_int), ("decouple_sunglint", C.c_int),
        ("n_water", C.c_double), ("wind_speed", C.c_double),
        ("sigma_type", C.c_int), ("q_convention", C.c_int),
        ("tau_R", C.c_double), ("depol", C.c_double),
        ("tau_a", C.c_double), ("ssa_a", C.c_double), ("aer_h_km", C.c_double),
        ("L_max", C.c_int), ("n_mu", C.c_int),
        ("greek", C.POINTER(C.c_double)),
        ("n_layers", C.c_int), ("m_max", C.c_int),
        ("max_orders", C.c_int), ("ipss", C.c_int),
        ("tolerance", C.c_double),
        ("layers", C.POINTER(C.c_double)),
        ("gas", C.c_void_p),
        ("n_ang_phase", C.c_int),
        ("theta_phase", C.POINTER(C.c_double)),
        ("p11_phase", C.POINTER(C.c_double)),
        ("bottom_source", C.POINTER(C.c_double)),
        ("bottom_n_mu", C.c_int), ("bottom_m_max", C.c_int),
        ("bottom_source_only", C.c_int),
        ("water_tau", C.c_double), ("water_omega_w", C.c_double),
        ("water_omega_p", C.c_double), ("beam_scale", C.c_double),
        ("top_reflect", C.c_int), ("top_n_mu", C.c_int), ("top_m_max", C.c_int), ("pad_water", C.c_int),
        ("top_source", C.POINTER(C.c_double)),
    ]


class Result(C.Structure):
    _fields_ = [
        ("rho_I", C.c_double), ("rho_Q", C.c_double), ("rho_U", C.c_double),
        ("I_TOA", C.c_double), ("Q_TOA", C.c_double), ("U_TOA", C.c_double),
        ("tau_R_total", C.c_double),
        ("T_dir_dn", C.c_double), ("T_diff_dn_hemi", C.c_double),
        ("T_total_dn_hemi", C.c_double), ("T_diff_dn_dir", C.c_double),
        ("T_dir_up_view", C.c_double),
        ("T_sg_up_dir", C.c_double), ("T_total_up_dir", C.c_double),
        ("rho_glint_direct_I", C.c_double), ("rho_glint_direct_Q", C.c_double),
        ("rho_glint_direct_U", C.c_double),
        ("kappa", C.c_double),
        ("n_orders_used", C.c_int), ("converged", C.c_int),
        ("status", C.c_int), ("n_ring", C.c_int),
    ]


RESULT_FIELDS = [f for f, _ in Result._fields_]


def make_instance(sza=40.0, vza=55.0, raa=90.0, wavelength=555.0,
                  surface="flat", n_water=1.34, wind=0.0, sigma_type=1,
                  q_convention=1, decouple_sunglint=0, tau_R=0.0941, depol=0.0279,
                  tau_a=0.0, ssa_a=1.0, aer_h_km=2.0, greek=None, n_mu=24,
                  n_layers=50, m_max=2, max_orders=20, tolerance=1e-7, ipss=0, layers=None, gas=None,
                  phase_table=None, bottom_source=None, bottom_source_only=0, water=None,
                  top_reflect=0, top_source=None, beam_scale=1.0):
    """Fill one descriptor.  `greek` is a float64 array [4, L+1] (beta, gamma,
    alpha, zeta) or None.  `layers` is None (column generator) or a float64
    array [4, n_layers] of per-layer Rayleigh tau, aerosol tau, aerosol SSA and
    absorption tau, top down; n_layers then follows its shape.  `gas` is an
    ocrt_gas handle (GasTable) or None.  `phase_table` is None (moment kernel)
    or a pair (theta_deg, P11) of equal-length arrays, ascending in theta: the
    angle-space aerosol phase function of the value kernel.  `bottom_source`
    is None or an array [3, m_src + 1, n_ring] of the upward I, Q, U Fourier
    field injected at the bottom (n_ring = ring_nodes(n_mu, vza)[0].size);
    `bottom_source_only` drops the solar source.  `water` = (tau, omega_w,
    omega_p) selects an in-water column (sza/vza then in-water angles, depol
    0.039 for water, surface "black"); `top_reflect` 1 flat or 2 rough
    water-side reflection at the top; `top_source` [3, m_src + 1, n_ring] the
    diffuse field entering at the top; `beam_scale` the direct beam's weight.
    Converted copies are kept alive by the descriptor."""
    d = Instance()
    d.sza_deg, d.vza_deg, d.raa_deg, d.wavelength_nm = sza, vza, raa, wavelength
    d.surface = SURFACE_NAMES[surface] if isinstance(surface, str) else int(surface)
    d.decouple_sunglint = int(decouple_sunglint)
    d.n_water, d.wind_speed = n_water, wind
    d.sigma_type, d.q_convention = sigma_type, q_convention
    d.tau_R, d.depol = tau_R, depol
    d.tau_a, d.ssa_a, d.aer_h_km = tau_a, ssa_a, aer_h_km
    if greek is not None:
        g = np.ascontiguousarray(greek, dtype=np.float64)
        if g.ndim != 2 or g.shape[0] != 4:
            raise ValueError("greek must have shape [4, L+1]")
        d.L_max = g.shape[1] - 1
        d.greek = g.ctypes.data_as(C.POINTER(C.c_double))
        d._greek_keepalive = g   # a converted copy lives as long as the descriptor
    else:
        d.L_max = 0
        d.greek = None
    d.n_mu, d.n_layers, d.m_max = n_mu, n_layers, m_max
    d.max_orders, d.tolerance, d.ipss = max_orders, tolerance, ipss
    if layers is not None:
        lay = np.ascontiguousarray(layers, dtype=np.float64)
        if lay.ndim != 2 or lay.shape[0] != 4:
            raise ValueError("layers must have shape [4, n_layers]")
        d.n_layers = lay.shape[1]
        d.layers = lay.ctypes.data_as(C.POINTER(C.c_double))
        d._layers_keepalive = lay
    else:
        d.layers = None
    d.gas = gas.handle if gas is not None else None
    d._gas_keepalive = gas
    if phase_table is not None:
        th = np.ascontiguousarray(phase_table[0], dtype=np.float64)
        p11 = np.ascontiguousarray(phase_table[1], dtype=np.float64)
        if th.ndim != 1 or th.shape != p11.shape or th.size < 2:
            raise ValueError("phase_table must be two equal 1-d arrays of at least 2 angles")
        d.n_ang_phase = th.size
        d.theta_phase = th.ctypes.data_as(C.POINTER(C.c_double))
        d.p11_phase = p11.ctypes.data_as(C.POINTER(C.c_double))
        d._phase_keepalive = (th, p11)
    else:
        d.n_ang_phase = 0
        d.theta_phase = None
        d.p11_phase = None
    if bottom_source is not None:
        b = np.ascontiguousarray(bottom_source, dtype=np.float64)
        if b.ndim != 3 or b.shape[0] != 3:
            raise ValueError("bottom_source must have shape [3, m_src + 1, n_ring]")
        d.bottom_source = b.ctypes.data_as(C.POINTER(C.c_double))
        d.bottom_m_max = b.shape[1] - 1
        d.bottom_n_mu = b.shape[2]
        d._bottom_keepalive = b
    else:
        d.bottom_source = None
        d.bottom_m_max = -1
        d.bottom_n_mu = 0
    d.bottom_source_only = int(bottom_source_only)
    if water is not None:
        d.water_tau, d.water_omega_w, d.water_omega_p = (float(v) for v in water)
    else:
        d.water_tau = d.water_omega_w = d.water_omega_p = 0.0
    d.beam_scale = beam_scale
    d.top_reflect = int(top_reflect)
    if top_source is not None:
        t = np.ascontiguousarray(top_source, dtype=np.float64)
        if t.ndim != 3 or t.shape[0] != 3:
            raise ValueError("top_source must have shape [3, m_src + 1, n_ring]")
        d.top_source = t.ctypes.data_as(C.POINTER(C.c_double))
        d.top_m_max = t.shape[1] - 1
        d.top_n_mu = t.shape[2]
        d._top_keepalive = t
    else:
        d.top_source = None
        d.top_m_max = -1
        d.top_n_mu = 0
    return d


def instance_array(descs):
    arr = (Instance * len(descs))()
    for i, d in enumerate(descs):
        arr[i] = d
    return arr


def results_to_dict(res):
    """Array of Result -> dict of numpy arrays."""
    out = {}
    for f in RESULT_FIELDS:
        out[f] = np.array([getattr(r, f) for r in res])
    return out
