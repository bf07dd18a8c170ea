"""Small gas-absorption data set for the tests (tests/golden/gas_inputs).

Copies the US Standard 1976 AFGL profile of the reference
(code/OCRT_C/inputs/afgl_atm/afgl_usstd76.dat) and reduces every layered
cross-section table (code/OCRT_C/inputs/xsec/xsec_<gas>.dat) to the
wavelength columns the tests use (each test wavelength and its two
neighbours, plus the 330 and 1100 nm ends that the loaders require).  At those
wavelengths the interpolated cross sections are the same as from the full
tables, so the reference and this solver read identical absorption from it.

    python tests/golden/make_gas_inputs.py [/root/reference]
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "gas_inputs")
TEST_WL = [412, 443, 490, 555, 660, 865]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "code", "OCRT_C", "inputs")
    os.makedirs(os.path.join(OUT, "afgl_atm"), exist_ok=True)
    os.makedirs(os.path.join(OUT, "xsec"), exist_ok=True)
    shutil.copyfile(os.path.join(src, "afgl_atm", "afgl_usstd76.dat"),
                    os.path.join(OUT, "afgl_atm", "afgl_usstd76.dat"))
    keep_wl = sorted({330, 1100} | {w + d for w in TEST_WL for d in (-1, 0, 1)})
    for gas in ("h2o", "o3", "o2", "co2", "no2", "ch4"):
        lines = open(os.path.join(src, "xsec", f"xsec_{gas}.dat")).read().splitlines()
        head = [l for l in lines if l.startswith("#")]
        data = [l.split() for l in lines if l.strip() and not l.startswith("#")]
        wl = [float(v) for v in data[0]]
        cols = [i for i, w in enumerate(wl) if round(w) in keep_wl and abs(w - round(w)) < 1e-9]
        assert len(cols) == len(keep_wl), (gas, len(cols))
        with open(os.path.join(OUT, "xsec", f"xsec_{gas}.dat"), "w") as f:
            f.write("\n".join(head) + "\n")
            f.write("# reduced to the wavelength columns of the tests\n")
            for row in data:
                f.write(" ".join(row[i] for i in cols) + "\n")
    print("wrote", OUT)


if __name__ == "__main__":
    main()
