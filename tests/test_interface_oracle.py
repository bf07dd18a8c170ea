"""The oracle's rough air-water interface kernels (shared/surface.c of the
reference), pinned three ways on the CPU:
  - against the reference library's own outputs (tests/golden/ref_interface.npz,
    written by tests/golden/make_interface_golden.py from oracle/_ref);
  - by the discrete Fourier inversion identity of the midpoint azimuth rule,
    which ties every mode's normalisation, cos/sin split and sign fold to the
    Mueller matrix evaluated directly at the quadrature azimuths;
  - by the air -> water flux identity: the m = 0 intensity BTDF integrated over
    the water hemisphere equals the slope-integrated direct transmittance, two
    independent formulations of the same Cox-Munk transmission.
"""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "ref_interface.npz")
D = C.POINTER(C.c_double)
NW = 1.34


def _p(a):
    return a.ctypes.data_as(D)


def fourier(orc, kind, mu_o, mu_i, m_max, n_phi, ws, st, qc):
    mo = np.ascontiguousarray(mu_o, dtype=np.float64)
    mi = np.ascontiguousarray(mu_i, dtype=np.float64)
    out = np.zeros((m_max + 1, mo.size, mi.size, 9))
    rc = orc.oracle_interface_fourier(kind, _p(mo), mo.size, _p(mi), mi.size, m_max, n_phi, ws, st, NW, qc, _p(out))
    return rc, out


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD))


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_fourier_matches_reference(orc, gold, kind):
    for c, (ws, st, qc, nphi) in enumerate(gold["cfg"]):
        modes = gold[f"k{kind}_c{c}_modes"]
        want = gold[f"k{kind}_c{c}"]
        rc, got = fourier(orc, kind, gold["mu_o"], gold["mu_i"], int(modes.max()), int(nphi), ws, int(st), int(qc))
        assert rc == 0
        # the azimuth sums round at the size of the kernel itself (its m = 0
        # slab), whatever the size of the mode's own coefficients
        scale = np.abs(want[0]).max()
        err = np.abs(got[modes] - want).max() / scale
        assert err <= 1e-12, (kind, c, err)


def test_flat_taw_has_no_rough_kernel(orc, gold):
    rc, _ = fourier(orc, 3, gold["mu_o"], gold["mu_i"], 1, 128, 0.0, 0, 1)
    assert rc == int(gold["flat_taw_rc"]) == -2


@pytest.mark.parametrize("kind,name", [(3, "direct_taw"), (1, "direct_rww")])
def test_direct_terms_match_reference(orc, gold, kind, name):
    for c, (ws, st) in enumerate(gold["direct_cfg"]):
        for j, mu in enumerate(gold["direct_mu"]):
            want = gold[name][c, j]
            got = orc.oracle_interface_direct(kind, mu, ws, int(st), NW)
            assert abs(got - want) <= 1e-13 * max(abs(want), 1e-300), (kind, ws, st, mu, got, want)


COS_ENTRY = np.array([1, 1, 0, 1, 1, 0, 0, 0, 1], dtype=bool)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_fourier_inversion_identity(orc, kind):
    """With n_phi = 2K midpoint azimuths, modes 0..K reconstruct the Mueller
    matrix at every quadrature azimuth: cos entries from a_0 + sum a_m cos(m phi)
    (the Nyquist cosine vanishes at midpoints), sin entries from sum b_m sin(m phi)
    with the Nyquist term halved.  IU and QU carry the fold sign for m > 0
    (not in R_ww), which the reconstruction undoes."""
    K = 24
    n_phi = 2 * K
    mu_o = np.array([0.2, 0.55, 0.9])
    mu_i = np.array([0.35, 0.75, 1.0])
    for ws, st, qc in ((7.0, 1, 1), (2.0, 0, 0)):
        rc, F = fourier(orc, kind, mu_o, mu_i, K, n_phi, ws, st, qc)
        assert rc == 0
        F = F.copy()
        if kind != 1:
            F[1:, :, :, 2] *= -1.0
            F[1:, :, :, 5] *= -1.0
        phi = (np.arange(n_phi) + 0.5) * 2.0 * np.pi / n_phi
        m = np.arange(K + 1)
        cw = np.cos(np.outer(phi, m))
        sw = np.sin(np.outer(phi, m))
        sw[:, K] *= 0.5
        for o, mo in enumerate(mu_o):
            for i, mi in enumerate(mu_i):
                direct = np.zeros((n_phi, 9))
                for q, ph in enumerate(phi):
                    M = np.zeros(9)
                    orc.oracle_interface_trig(kind, mo, mi, np.cos(ph), np.sin(ph), ws, st, NW, qc, _p(M))
                    direct[q] = M
                recon = np.where(COS_ENTRY, cw @ F[:, o, i, :], sw @ F[:, o, i, :])
                scale = max(np.abs(direct).max(), 1e-300)
                assert np.abs(recon - direct).max() <= 1e-12 * scale, (kind, ws, o, i)


def test_taw_flux_equals_direct_transmittance(orc):
    """2 pi sum_o w_o mu_o T^0_II(mu_o <- mu_i) over a Gauss rule on the water
    hemisphere equals the slope-integrated direct transmittance T_aw(mu_i)."""
    n = 48
    x, w = np.polynomial.legendre.leggauss(2 * n)
    mu, wt = x[n:], w[n:]
    mu_i = np.array([0.5, 0.7, 0.866, 0.95, 1.0])
    for ws in (5.0, 12.0):
        rc, T = fourier(orc, 3, mu, mu_i, 0, 1024, ws, 1, 1)
        assert rc == 0
        flux = 2.0 * np.pi * np.einsum("o,o,oi->i", wt, mu, T[0, :, :, 0])
        direct = np.array([orc.oracle_interface_direct(3, m, ws, 1, NW) for m in mu_i])
        assert np.abs(flux / direct - 1.0).max() <= 1e-6, (ws, flux / direct)
