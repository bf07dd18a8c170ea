"""The Greek-coefficient converter (reference: code/OCRT_C/src/rt_phase_moments.c:6-64,
rt_phase_spin2_m0 and rt_phase_complete_spherical) in its two restatements:
ocrt_amd.synthetic (which builds every synthetic aerosol the tests and the
benchmark use) and the oracle's oracle_complete_spherical.

Pinned three ways:
  - against the reference functions themselves, compiled from the reference
    sources into oracle/_ref/libocrt_ref.so, on random inputs (1e-14);
  - against the mathematics: S_l(mu) = sqrt((l-2)!/(l+2)!) P_l^2(mu) (the
    m = 0 spin-2 functions are normalised associated Legendre functions of
    order 2), and alpha_l +- zeta_l equal the projections of P22 +- P33 onto
    the Wigner functions d^l_{2,+-2}(mu) = ((1 +- mu)/2)^2 P^{(0,4) or (4,0)}_{l-2}(mu)
    (Jacobi polynomials); the Rayleigh matrix gives alpha_2 = 3, zeta_2 = 0;
  - hg_greek's families reconstruct the phase matrix they are defined from:
    sum beta_l P_l = truncated Henyey-Greenstein, sum gamma_l S_l = -0.2 (1 - mu^2) P11.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.special import eval_jacobi, lpmv

from ocrt_amd import synthetic as S

D = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(D)


def _random_families(L, seed):
    rng = np.random.default_rng(seed)
    b11 = rng.normal(size=L + 1) / (1.0 + np.arange(L + 1)) ** 0.5
    b11[0] = 1.0
    b22 = b11 * (1.0 + 0.1 * rng.normal(size=L + 1))
    d33 = rng.normal(size=L + 1) / (1.0 + np.arange(L + 1)) ** 0.5
    return b11, b22, d33


def _ref_complete(lib, b11, b22, d33):
    L = b11.size - 1
    al, ze = np.zeros(L + 1), np.zeros(L + 1)
    rc = lib.rt_phase_complete_spherical(L, _p(b11), _p(b22), _p(d33), _p(al), _p(ze))
    assert rc == 0
    return al, ze


def _close(a, b, tol=1e-14):
    scale = max(np.max(np.abs(b)), 1.0)
    return np.max(np.abs(a - b)) <= tol * scale


@pytest.mark.parametrize("L", [20, 80, 180])
def test_complete_spherical_matches_reference(ref_lib, oracle, L):
    b11, b22, d33 = _random_families(L, seed=L)
    al_r, ze_r = _ref_complete(ref_lib, b11, b22, d33)
    al_s, ze_s = S.complete_spherical(b11, b22, d33)
    assert _close(al_s, al_r) and _close(ze_s, ze_r)
    al_o, ze_o = np.zeros(L + 1), np.zeros(L + 1)
    assert oracle.oracle_complete_spherical(L, _p(b11), _p(b22), _p(d33), _p(al_o), _p(ze_o)) == 0
    # the reference build contracts to FMA (-ffp-contract=fast), the oracle does not
    assert _close(al_o, al_r) and _close(ze_o, ze_r)


@pytest.mark.parametrize("L", [20, 80, 180])
def test_spin2_m0_matches_reference(ref_lib, L):
    mu = np.linspace(-1.0, 1.0, 41)
    ours = S.spin2_m0_table(L, mu)
    ref = np.zeros((L + 1, mu.size))
    row = np.zeros(L + 1)
    for j, x in enumerate(mu):
        ref_lib.rt_phase_spin2_m0(L, C.c_double(x), _p(row))
        ref[:, j] = row
    assert _close(ours, ref)


def test_spin2_m0_is_normalised_associated_legendre():
    L = 40
    mu = np.linspace(-0.99, 0.99, 23)
    ours = S.spin2_m0_table(L, mu)
    assert np.all(ours[:2] == 0.0)
    for l in range(2, L + 1):
        norm = math.sqrt(math.factorial(l - 2) / math.factorial(l + 2))
        want = norm * lpmv(2, l, mu)
        assert np.max(np.abs(ours[l] - want)) <= 1e-12 * max(1.0, np.max(np.abs(want))), l


def test_complete_spherical_rayleigh_closed_form(oracle):
    # Rayleigh without depolarisation: P11 = P22 = 3/4 (1 + mu^2), P33 = 3/2 mu
    b = np.array([1.0, 0.0, 0.5])
    d33 = np.array([0.0, 1.5, 0.0])
    al, ze = S.complete_spherical(b, b, d33)
    assert abs(al[2] - 3.0) <= 1e-15 and abs(ze[2]) <= 1e-15
    al_o, ze_o = np.zeros(3), np.zeros(3)
    oracle.oracle_complete_spherical(2, _p(b), _p(b), _p(d33), _p(al_o), _p(ze_o))
    assert abs(al_o[2] - 3.0) <= 1e-15 and abs(ze_o[2]) <= 1e-15


@pytest.mark.parametrize("L", [12, 40])
def test_complete_spherical_is_wigner_projection(L):
    # With P22 = P11 (the pairing the solver uses, rt_aerosol.c:754-758):
    #   alpha_l + zeta_l = (2l+1)/2 int (P22 + P33) d^l_{2,2}  dmu
    #   alpha_l - zeta_l = (2l+1)/2 int (P22 - P33) d^l_{2,-2} dmu
    b11, _, d33 = _random_families(L, seed=7 + L)
    al, ze = S.complete_spherical(b11, b11, d33)
    x, w = np.polynomial.legendre.leggauss(2 * L + 8)
    P = S.legendre_table(L, x)
    f22, f33 = b11 @ P, d33 @ P
    for l in range(2, L + 1):
        dp = ((1.0 + x) / 2.0) ** 2 * eval_jacobi(l - 2, 0, 4, x)
        dm = ((1.0 - x) / 2.0) ** 2 * eval_jacobi(l - 2, 4, 0, x)
        cp = (2 * l + 1) / 2.0 * np.sum(w * (f22 + f33) * dp)
        cm = (2 * l + 1) / 2.0 * np.sum(w * (f22 - f33) * dm)
        assert abs(al[l] + ze[l] - cp) <= 1e-11 * max(1.0, abs(cp)), l
        assert abs(al[l] - ze[l] - cm) <= 1e-11 * max(1.0, abs(cm)), l


@pytest.mark.parametrize("g,L", [(0.7, 180), (0.6, 80), (0.8, 80), (0.7, 40)])
def test_hg_greek_reconstructs_its_phase_matrix(g, L):
    gk = S.hg_greek(g, L)
    beta, gamma, alpha, zeta = gk
    x, _ = np.polynomial.legendre.leggauss(L + 1)
    P = S.legendre_table(L, x)
    Sp = S.spin2_m0_table(L, x)
    # truncated HG, normalised to beta_0 = 1
    l = np.arange(L + 1)
    hg_trunc = ((2 * l + 1) * g ** l) @ P
    p11 = beta @ P
    assert beta[0] == 1.0
    assert np.max(np.abs(p11 - hg_trunc)) <= 1e-12 * np.max(np.abs(hg_trunc))
    # (1 - mu^2) P11 has degree L + 2 while sum_{l<=L} gamma_l S_l = (1 - mu^2) x (degree L - 2):
    # gamma must be the orthogonal projection (residual orthogonal to every S_l), and the
    # residual is bounded by the dropped degrees L-1, L of P11, (2L+1) g^(L-1) in size
    xq, wq = np.polynomial.legendre.leggauss(L + 16)
    Pq, Sq = S.legendre_table(L, xq), S.spin2_m0_table(L, xq)
    hq = ((2 * l + 1) * g ** l) @ Pq
    resid = -0.2 * (1.0 - xq * xq) * hq - gamma @ Sq
    ortho = Sq @ (wq * resid)
    assert np.max(np.abs(ortho)) <= 1e-12 * np.max(np.abs(hq))
    bound = max(1e-12, 2.0 * (2 * L + 1) * g ** (L - 1)) * np.max(np.abs(hq))
    assert np.max(np.abs(resid)) <= bound
    p12 = gamma @ Sp
    want = -0.2 * (1.0 - x * x) * hg_trunc
    assert np.max(np.abs(p12 - want)) <= bound
    # alpha, zeta are the completion of (beta, beta, Legendre moments of mu * P11)
    d33 = np.zeros(L + 1)
    xx, ww = np.polynomial.legendre.leggauss(L + 16)
    PP = S.legendre_table(L, xx)
    d33 = (2 * l + 1) * 0.5 * (PP * (xx * (((2 * l + 1) * g ** l) @ PP) * ww)).sum(axis=1)
    a2, z2 = S.complete_spherical(beta, beta, d33)
    assert np.max(np.abs(a2 - alpha)) <= 1e-12 * max(1.0, np.max(np.abs(alpha)))
    assert np.max(np.abs(z2 - zeta)) <= 1e-12 * max(1.0, np.max(np.abs(zeta)))
