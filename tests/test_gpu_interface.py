"""Rough air-water interface kernels on the GPU (interface.hip) through the C
ABI: every mode of every kind against the reference library's outputs
(tests/golden/ref_interface.npz) and against the oracle at full quadrature
sizes (48 x 48 nodes, modes up to 127), plus the reference-named per-mode
wrappers and the argument checks.  Tolerance: the azimuth sums round at the
size of the kernel, so |GPU - reference| <= 1e-12 max |m = 0 slab| (5e-12 with
the grazing nodes of the full-size rule)."""
import ctypes as C

import numpy as np
import pytest

from test_interface_oracle import GOLD, NW, fourier

pytestmark = pytest.mark.gpu
D = C.POINTER(C.c_double)
KINDS = ["R_aa", "R_ww", "T_wa", "T_aw"]
WRAPPERS = ["surface_coxmunk_fourier_kernel", "surface_R_ww_coxmunk_fourier_kernel",
            "surface_T_wa_coxmunk_fourier_kernel", "surface_T_aw_coxmunk_fourier_kernel"]


@pytest.fixture(scope="module")
def ocrt():
    import ocrt_amd
    lib = ocrt_amd.library()
    for f in WRAPPERS:
        getattr(lib, f).argtypes = [D, C.c_int, D, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                                    C.c_int, D]
    for f in ("surface_T_aw_coxmunk_direct", "surface_R_ww_coxmunk_direct"):
        getattr(lib, f).argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
        getattr(lib, f).restype = C.c_double
    lib.surface_fkc_set_build_mmax.argtypes = [C.c_int]
    lib.surface_fkc_set_build_mmax.restype = None
    return ocrt_amd


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD))


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_gpu_matches_reference(ocrt, gold, kind):
    print(f"[gpu] interface kind {kind} vs reference", flush=True)
    for c, (ws, st, qc, nphi) in enumerate(gold["cfg"]):
        modes = gold[f"k{kind}_c{c}_modes"]
        want = gold[f"k{kind}_c{c}"]
        got = ocrt.interface_fourier(KINDS[kind], gold["mu_o"], gold["mu_i"], int(modes.max()), int(nphi), ws,
                                     int(st), NW, int(qc)).reshape(-1, gold["mu_o"].size, gold["mu_i"].size, 9)
        err = np.abs(got[modes] - want).max() / np.abs(want[0]).max()
        assert err <= 1e-12, (kind, c, err)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_gpu_matches_oracle_full_size(ocrt, orc, kind):
    n = 48
    x, _ = np.polynomial.legendre.leggauss(2 * n)
    mu = np.ascontiguousarray(x[n:])
    print(f"[gpu] interface kind {kind}: {n}x{n} nodes, modes 0..127, 1024 azimuths", flush=True)
    got = ocrt.interface_fourier(KINDS[kind], mu, mu, 127, 1024, 6.5, 1, NW, 1).reshape(128, n, n, 9)
    rc, want = fourier(orc, kind, mu, mu, 127, 1024, 6.5, 1, 1)
    assert rc == 0
    err = np.abs(got - want).max() / np.abs(want[0]).max()
    print(f"[gpu] interface kind {kind}: worst |GPU - oracle| / max |m = 0| = {err:.2e}", flush=True)
    # grazing nodes (mu ~ 1.6e-3) give lobes far above their azimuth mean;
    # the GPU forms the rotations by double angles, so the sums round apart
    # by a few 1e-13 of the kernel there
    assert err <= 5e-12, (kind, err)


def test_gpu_direct_terms(ocrt, gold):
    mu = gold["direct_mu"]
    for c, (ws, st) in enumerate(gold["direct_cfg"]):
        for kind, name in (("T_aw", "direct_taw"), ("R_ww", "direct_rww")):
            got = ocrt.interface_direct(kind, mu, ws, int(st), NW)
            want = gold[name][c]
            assert np.all(np.abs(got - want) <= 1e-13 * np.maximum(np.abs(want), 1e-300)), (kind, ws, got, want)


def test_reference_named_wrappers(ocrt, gold):
    """The per-mode functions under the reference's names return each mode's
    slab of one all-mode launch, bit for bit, also past the build cap."""
    lib = ocrt.library()
    mo, mi = gold["mu_o"], gold["mu_i"]
    p = lambda a: a.ctypes.data_as(D)
    for kind, f in enumerate(WRAPPERS):
        allm = ocrt.interface_fourier(KINDS[kind], mo, mi, 70, 360, 4.0, 1, NW, 1).reshape(71, mo.size, mi.size, 9)
        lib.surface_fkc_set_build_mmax(8)
        for m in (0, 3, 8, 9, 70, 2):
            r = np.zeros((mo.size, mi.size, 9))
            assert getattr(lib, f)(p(mo), mo.size, p(mi), mi.size, m, 360, 4.0, 1, NW, 1, p(r)) == 0
            assert np.array_equal(r, allm[m]), (f, m)
        lib.surface_fkc_set_build_mmax(-1)
    r = np.zeros((mo.size, mi.size, 9))
    assert lib.surface_T_aw_coxmunk_fourier_kernel(p(mo), mo.size, p(mi), mi.size, 1, 128, 0.0, 0, NW, 1, p(r)) == -2
    assert lib.surface_R_ww_coxmunk_fourier_kernel(p(mo), 0, p(mi), mi.size, 1, 128, 5.0, 1, NW, 1, p(r)) == -1
    assert lib.surface_R_ww_coxmunk_fourier_kernel(p(mo), mo.size, p(mi), mi.size, 1, 3, 5.0, 1, NW, 1, p(r)) == -1
    for c, (ws, st) in enumerate(gold["direct_cfg"]):
        for j, m in enumerate(gold["direct_mu"]):
            a = lib.surface_T_aw_coxmunk_direct(m, NW, ws, int(st))
            b = lib.surface_R_ww_coxmunk_direct(m, NW, ws, int(st))
            assert abs(a - gold["direct_taw"][c, j]) <= 1e-13 * max(gold["direct_taw"][c, j], 1e-300)
            assert abs(b - gold["direct_rww"][c, j]) <= 1e-13 * gold["direct_rww"][c, j]


def test_argument_checks(ocrt):
    import ocrt_amd
    mu = np.array([0.3, 0.6])
    with pytest.raises(ocrt_amd.OcrtError, match="-1"):
        ocrt.interface_fourier("R_ww", mu, mu, 256, 1024)          # more than 256 modes
    with pytest.raises(ocrt_amd.OcrtError, match="-1"):
        ocrt.interface_fourier("R_ww", np.array([0.0, 0.5]), mu, 4, 1024)   # cosine outside (0, 1]
    with pytest.raises(ocrt_amd.OcrtError, match="-5"):
        ocrt.interface_fourier("T_aw", mu, mu, 4, 128, wind=0.0, sigma_type=0)
