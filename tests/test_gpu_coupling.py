"""Atmosphere -> water coupling of the coupled ocean on the GPU
(rt_air_water_couple_atm_to_water through this library's C ABI) against the
reference's own function (oracle/_ref, built from its sources), on the same
BOA fields: the rough Fourier-transmission contraction with air weights, the
flat Snell map with interpolation, and the rough m = 0 BTDF override without
weights; every output within 1e-12 of the largest value of its field."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
D = C.POINTER(C.c_double)


class BoaExport(C.Structure):
    _fields_ = [("m_max", C.c_int), ("n_mu", C.c_int), ("mu_quad_pos", D), ("I_per_m", D), ("Q_per_m", D),
                ("U_per_m", D), ("tau_atm_total", C.c_double), ("mu_sun_air", C.c_double), ("allocated", C.c_int)]


class Coupled(C.Structure):
    _fields_ = [("m_max", C.c_int), ("n_mu_water", C.c_int), ("mu_water_pos", D), ("I_inwater_per_m", D),
                ("Q_inwater_per_m", D), ("U_inwater_per_m", D)]


def _bind(lib):
    f = lib.rt_air_water_couple_atm_to_water
    f.argtypes = [C.POINTER(BoaExport), C.c_double, C.c_int, D, C.c_int, D, C.c_double, C.c_int, C.POINTER(Coupled)]
    f.restype = C.c_int
    return f


@pytest.fixture(scope="module")
def libs(ref_lib):
    import ocrt_amd
    return _bind(ocrt_amd.library()), _bind(ref_lib)


def _p(a):
    return None if a is None else a.ctypes.data_as(D)


def _run(fn, boa_arrays, mu_w, w_a, wind, sigma_type, m_max):
    mu_a, I, Q, U = boa_arrays
    b = BoaExport(m_max, mu_a.size, _p(mu_a), _p(I), _p(Q), _p(U), 0.3, 0.8, 0)
    n_w = mu_w.size
    out = [np.zeros((m_max + 1) * n_w) for _ in range(3)]
    mw_copy = np.zeros(n_w)
    c = Coupled(m_max, n_w, _p(mw_copy), _p(out[0]), _p(out[1]), _p(out[2]))
    rc = fn(C.byref(b), 1.34, 1, _p(mu_w), n_w, _p(w_a), wind, sigma_type, C.byref(c))
    return rc, [o.reshape(m_max + 1, n_w) for o in out], mw_copy


@pytest.mark.parametrize("case", ["rough_weights", "flat_weights", "rough_no_weights", "flat_law0"])
def test_coupling_equals_reference(libs, case):
    ours, ref = libs
    rng = np.random.default_rng(7)
    m_max = 4
    x, w = np.polynomial.legendre.leggauss(32)
    mu_gl, w_gl = x[16:], w[16:]
    mu_a = np.concatenate([mu_gl, [0.57]])            # a zero-weight view node after the Gauss nodes
    w_a = np.concatenate([w_gl, [0.0]])
    base = np.exp(-1.5 * (1 - mu_a))
    I = np.concatenate([base * (1 + 0.1 * m) * (0.5 ** m) for m in range(m_max + 1)]) * 0.05
    Q = I * rng.uniform(-0.3, 0.1, I.size)
    U = I * rng.uniform(-0.1, 0.1, I.size)
    xw, _ = np.polynomial.legendre.leggauss(24)
    mu_w = np.ascontiguousarray(xw[12:])
    wind, st, wa = {"rough_weights": (6.0, 1, w_a), "flat_weights": (0.0, 1, w_a),
                    "rough_no_weights": (5.0, 1, None), "flat_law0": (1e-12, 0, w_a)}[case]
    print(f"[gpu] coupling {case}", flush=True)
    rc_o, got, mw_o = _run(ours, (mu_a, I, Q, U), mu_w, wa, wind, st, m_max)
    rc_r, want, mw_r = _run(ref, (mu_a, I, Q, U), mu_w, wa, wind, st, m_max)
    assert rc_o == rc_r == 0
    assert np.array_equal(mw_o, mw_r)
    for s in range(3):
        scale = max(np.abs(want[s]).max(), 1e-300)
        assert np.abs(got[s] - want[s]).max() <= 1e-12 * scale, (case, s, np.abs(got[s] - want[s]).max(), scale)


def test_coupling_argument_checks(libs):
    ours, ref = libs
    mu_a = np.array([0.2, 0.5, 0.9])
    z = np.zeros(3)
    for fn in (ours, ref):
        b = BoaExport(0, 3, _p(mu_a), _p(z), _p(z), _p(z), 0.0, 1.0, 0)
        c = Coupled(1, 3, _p(np.zeros(3)), _p(np.zeros(6)), _p(np.zeros(6)), _p(np.zeros(6)))
        assert fn(C.byref(b), 1.34, 1, _p(mu_a), 3, None, 0.0, 1, C.byref(c)) == -1   # m_max mismatch
        assert fn(C.byref(b), 1.0, 1, _p(mu_a), 3, None, 0.0, 1, C.byref(c)) == -1    # n_water <= 1


def _bind_up(lib):
    f = lib.rt_air_water_couple_water_to_atm
    f.argtypes = [D, D, D, D, D, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, D, C.c_int, D, D, D, D]
    f.restype = C.c_int
    return f


@pytest.mark.parametrize("wind", [5.0, 0.0])
def test_reverse_coupling_equals_reference(ref_lib, wind):
    """Water -> atmosphere: the rough composite-grid transmission (wind 5) and
    the flat reverse-Snell map (wind 0) against the reference build."""
    import ocrt_amd
    ours, ref = _bind_up(ocrt_amd.library()), _bind_up(ref_lib)
    m_max = 3
    xw, ww = np.polynomial.legendre.leggauss(40)
    mu_w, w_w = np.ascontiguousarray(xw[20:]), np.ascontiguousarray(ww[20:])
    n_w = mu_w.size
    rng = np.random.default_rng(3)
    I = np.concatenate([np.exp(-2.0 * (1 - mu_w)) * 0.02 * 0.6 ** m for m in range(m_max + 1)])
    Q = I * rng.uniform(-0.2, 0.2, I.size)
    U = I * rng.uniform(-0.1, 0.1, I.size)
    xa, wa = np.polynomial.legendre.leggauss(24)
    mu_a = np.ascontiguousarray(np.concatenate([xa[12:], [0.43]]))
    w_a = np.ascontiguousarray(np.concatenate([wa[12:], [0.0]]))
    print(f"[gpu] reverse coupling wind {wind}", flush=True)
    res = []
    for fn in (ours, ref):
        out = [np.zeros((m_max + 1) * mu_a.size) for _ in range(3)]
        rc = fn(_p(I), _p(Q), _p(U), _p(mu_w), _p(w_w), n_w, m_max, 1.34, 1, wind, 1, _p(mu_a), mu_a.size, _p(w_a),
                _p(out[0]), _p(out[1]), _p(out[2]))
        assert rc == 0
        res.append(out)
    for s in range(3):
        scale = max(np.abs(res[1][s]).max(), 1e-300)
        assert np.abs(res[0][s] - res[1][s]).max() <= 1e-12 * scale, (wind, s)


def _fields(rng, m_max, mu, n_inst):
    """n_inst random Fourier fields [3][(m_max+1) * n] of a smooth shape."""
    base = np.exp(-1.5 * (1 - mu))
    out = []
    for _ in range(n_inst):
        I = np.concatenate([base * rng.uniform(0.5, 1.5) * 0.05 * 0.6 ** m for m in range(m_max + 1)])
        out.append(np.stack([I, I * rng.uniform(-0.3, 0.1, I.size), I * rng.uniform(-0.1, 0.1, I.size)]))
    return np.ascontiguousarray(np.stack(out))


@pytest.mark.parametrize("case", ["rough_weights", "flat", "rough_btdf"])
def test_batched_down_coupling_equals_single_calls(case):
    """ocrt_couple_atm_to_water_batch: every instance of a batch bit for bit
    what ocrt_couple_atm_to_water gives it alone."""
    import ocrt_amd
    lib = ocrt_amd.library()
    VP = C.c_void_p
    rng = np.random.default_rng(11)
    m_max, n_inst = 3, 5
    x, w = np.polynomial.legendre.leggauss(24)
    mu_a = np.ascontiguousarray(np.concatenate([x[12:], [0.61]]))
    w_a = np.ascontiguousarray(np.concatenate([w[12:], [0.0]]))
    mu_w = np.ascontiguousarray(np.polynomial.legendre.leggauss(20)[0][10:])
    wind, wa = {"rough_weights": (7.0, w_a), "flat": (0.0, w_a), "rough_btdf": (4.0, None)}[case]
    f = _fields(rng, m_max, mu_a, n_inst)
    n_a, n_w, M = mu_a.size, mu_w.size, m_max + 1
    print(f"[gpu] batched down coupling {case}", flush=True)
    out = np.zeros((n_inst, 3, M * n_w))
    q = lambda a: None if a is None else a.ctypes.data_as(VP)
    rc = lib.ocrt_couple_atm_to_water_batch(n_inst, m_max, n_a, q(mu_a), q(wa), q(f), C.c_double(1.34), 1, q(mu_w),
                                            n_w, C.c_double(wind), 1, q(out), None)
    assert rc == 0
    for i in range(n_inst):
        one = [np.zeros(M * n_w) for _ in range(3)]
        fi = [np.ascontiguousarray(f[i, s]) for s in range(3)]
        rc = lib.ocrt_couple_atm_to_water(m_max, n_a, q(mu_a), q(wa), q(fi[0]), q(fi[1]), q(fi[2]), C.c_double(1.34),
                                          1, q(mu_w), n_w, C.c_double(wind), 1, q(one[0]), q(one[1]), q(one[2]), None)
        assert rc == 0
        assert np.array_equal(out[i], np.stack(one)), (case, i)
    assert np.abs(out).max() > 0


@pytest.mark.parametrize("wind", [6.0, 0.0])
def test_batched_up_coupling_equals_single_calls(wind):
    """ocrt_couple_water_to_atm_batch: every instance of a batch bit for bit
    what ocrt_couple_water_to_atm gives it alone."""
    import ocrt_amd
    lib = ocrt_amd.library()
    VP = C.c_void_p
    rng = np.random.default_rng(12)
    m_max, n_inst = 2, 4
    mu_w = np.ascontiguousarray(np.polynomial.legendre.leggauss(32)[0][16:])
    x = np.polynomial.legendre.leggauss(16)[0]
    mu_a = np.ascontiguousarray(np.concatenate([x[8:], [0.38]]))
    f = _fields(rng, m_max, mu_w, n_inst)
    n_a, n_w, M = mu_a.size, mu_w.size, m_max + 1
    print(f"[gpu] batched up coupling wind {wind}", flush=True)
    out = np.zeros((n_inst, 3, M * n_a))
    q = lambda a: a.ctypes.data_as(VP)
    rc = lib.ocrt_couple_water_to_atm_batch(n_inst, m_max, n_w, q(mu_w), q(f), C.c_double(1.34), 1, C.c_double(wind),
                                            1, n_a, q(mu_a), q(out), None)
    assert rc == 0
    for i in range(n_inst):
        one = [np.zeros(M * n_a) for _ in range(3)]
        fi = [np.ascontiguousarray(f[i, s]) for s in range(3)]
        rc = lib.ocrt_couple_water_to_atm(m_max, n_w, q(mu_w), q(fi[0]), q(fi[1]), q(fi[2]), C.c_double(1.34), 1,
                                          C.c_double(wind), 1, n_a, q(mu_a), q(one[0]), q(one[1]), q(one[2]), None)
        assert rc == 0
        assert np.array_equal(out[i], np.stack(one)), (wind, i)
    assert np.abs(out).max() > 0
