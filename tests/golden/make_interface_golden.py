"""Write tests/golden/ref_interface.npz: the rough-interface surface kernels of
the reference OCRT library (shared/surface.c), compiled from its own sources
into oracle/_ref/libocrt_ref.so by `make -C oracle ref`, on fixed node sets.
Run in the build container only (data only; no reference code is stored).

    python tests/golden/make_interface_golden.py

Contents (m-major like the reference's per-mode output, [n_o][n_i][9] per mode):
  mu_o, mu_i                    the node sets
  k{kind}_c{c}_modes            modes written for config c of kind
  k{kind}_c{c}                  [len(modes), n_o, n_i, 9]
  cfg                           [c] -> (wind, sigma_type, q_convention, n_phi)
  direct_mu, direct_cfg         cosines and (wind, sigma_type) of the direct terms
  direct_taw, direct_rww        [cfg, mu]
  flat_taw_rc                   return code of T_aw on a flat surface
kind: 0 air-side reflection, 1 water-side reflection, 2 water -> air,
3 air -> water transmission; n_water 1.34 throughout.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libocrt_ref.so")

FOURIER = ["surface_coxmunk_fourier_kernel", "surface_R_ww_coxmunk_fourier_kernel",
           "surface_T_wa_coxmunk_fourier_kernel", "surface_T_aw_coxmunk_fourier_kernel"]
CFG = [(5.0, 1, 1, 1024), (0.7, 0, 0, 128), (13.0, 1, 1, 360)]
MODES = [[0, 1, 2, 3, 7, 16, 33, 64, 70], [0, 1, 2, 5, 12], [0, 1, 4, 40]]
MU_O = np.array([0.05, 0.31, 0.62, 0.87, 1.0])
MU_I = np.array([0.12, 0.5, 0.93, 1.0])
DIRECT_MU = np.array([1e-10, 0.02, 0.1, 0.3, 0.5, 0.66, 0.8, 0.95, 1.0])
DIRECT_CFG = [(5.0, 1), (0.0, 0), (12.0, 0), (1.0, 1)]
NW = 1.34


def main():
    ref = C.CDLL(REF_LIB)
    D = C.POINTER(C.c_double)
    for f in FOURIER:
        getattr(ref, f).argtypes = [D, C.c_int, D, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                                    C.c_int, D]
    for f in ("surface_T_aw_coxmunk_direct", "surface_R_ww_coxmunk_direct"):
        getattr(ref, f).argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
        getattr(ref, f).restype = C.c_double
    p = lambda a: a.ctypes.data_as(D)
    out = {"mu_o": MU_O, "mu_i": MU_I, "cfg": np.array(CFG, dtype=np.float64),
           "direct_mu": DIRECT_MU, "direct_cfg": np.array(DIRECT_CFG, dtype=np.float64)}
    for kind, f in enumerate(FOURIER):
        for c, ((ws, st, qc, nphi), modes) in enumerate(zip(CFG, MODES)):
            tab = np.zeros((len(modes), MU_O.size, MU_I.size, 9))
            for j, m in enumerate(modes):
                r = np.zeros((MU_O.size, MU_I.size, 9))
                rc = getattr(ref, f)(p(MU_O), MU_O.size, p(MU_I), MU_I.size, m, nphi, ws, st, NW, qc, p(r))
                assert rc == 0, (f, m, rc)
                tab[j] = r
            out[f"k{kind}_c{c}"] = tab
            out[f"k{kind}_c{c}_modes"] = np.array(modes)
    r = np.zeros((MU_O.size, MU_I.size, 9))
    out["flat_taw_rc"] = np.array(ref.surface_T_aw_coxmunk_fourier_kernel(p(MU_O), MU_O.size, p(MU_I), MU_I.size, 1,
                                                                          128, 0.0, 0, NW, 1, p(r)))
    out["direct_taw"] = np.array([[ref.surface_T_aw_coxmunk_direct(m, NW, ws, st) for m in DIRECT_MU]
                                  for ws, st in DIRECT_CFG])
    out["direct_rww"] = np.array([[ref.surface_R_ww_coxmunk_direct(m, NW, ws, st) for m in DIRECT_MU]
                                  for ws, st in DIRECT_CFG])
    np.savez_compressed(os.path.join(HERE, "ref_interface.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_interface.npz"))


if __name__ == "__main__":
    main()
