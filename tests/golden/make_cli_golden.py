"""Golden outputs of the reference CLI for tests/test_gpu_cli.py.

Builds the reference command-line program from its sources
(`make -C oracle ref-cli` -> oracle/_ref/ocrt_ref), runs it from
code/OCRT_C of the reference tree (its inputs/ are read relative to that
directory) on each case below, and records exit status, stdout and the files
it wrote in tests/golden/cli_cases.json.  Gas absorption (always on in the
CLI) reads tests/golden/gas_inputs, copied to ./inputs of each run.  Needs the
reference tree, so it runs only in the container that has it:

    python tests/golden/make_cli_golden.py [/root/reference]
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "cli_cases.json")
GAS_INPUTS = os.path.join(HERE, "gas_inputs")

NOGAS = ["--gas-column-h2o", "0", "--gas-column-o3", "0", "--gas-column-no2", "0", "--gas-column-o2", "0",
         "--gas-column-co2", "0", "--gas-column-ch4", "0"]

# (name, args, env, input files {name: text}, output files to capture)
BATCH5 = "case_id,sza_deg,vza_deg,raa_deg,wavelength_nm\n1,40,30,90,555\n2,60,10,150,443\n3,20,55,10,865\n"
BATCH6 = ("case_id,sza_deg,vza_deg,raa_deg,wavelength_nm,tau_R_ref\n"
          "7,40,30,90,555,0.0935\n8,65,20,170,412,0.31\n")
CASES = [
    ("single_black", ["--surface", "black", "--sza", "40", "--vza", "30", "--raa", "90", "--wavelength", "555"] + NOGAS,
     {}, {}, []),
    ("single_flat", ["--surface", "flat", "--sza", "52", "--vza", "23", "--raa", "140", "--wavelength", "443"] + NOGAS,
     {}, {}, []),
    ("single_bfo", ["--surface", "black_fresnel_ocean", "--wind-speed", "7", "--sza", "38", "--vza", "47", "--raa",
                    "65", "--wavelength", "555"] + NOGAS, {}, {}, []),
    ("single_bfo_decoupled", ["--surface", "black_fresnel_ocean", "--wind-speed", "4", "--decouple-sunglint",
                              "--sza", "30", "--vza", "35", "--raa", "170", "--wavelength", "865"] + NOGAS,
     {}, {}, []),
    ("single_black_ipss", ["--surface", "black", "--pssa", "--sza", "70", "--vza", "30", "--raa", "45",
                           "--wavelength", "412"] + NOGAS, {}, {}, []),
    ("single_flat_pressure", ["--surface", "flat", "--pressure", "800", "--sza", "25", "--vza", "60", "--raa", "300",
                              "--wavelength", "660"] + NOGAS, {}, {}, []),
    ("single_advanced", ["--surface", "black_fresnel_ocean", "--wind-speed", "10", "--sigma-model", "nakajima-tanaka",
                         "--n-mu", "16", "--n-layers", "60", "--m-max", "4", "--sza", "50", "--vza", "20", "--raa",
                         "200", "--wavelength", "490"] + NOGAS, {"OCRT_ADVANCED": "1"}, {}, []),
    ("lut_bfo", ["--surface", "black_fresnel_ocean", "--wind-speed", "3", "--decouple-sunglint", "--sza", "40",
                 "--wavelength", "555", "--lut-vza-step", "10", "--lut-raa-step", "30", "--output-full-grid",
                 "lut_bfo.csv"] + NOGAS, {}, {}, ["lut_bfo.csv"]),
    ("lut_black_ipss", ["--surface", "black", "--pssa", "--sza", "60", "--wavelength", "443", "--lut-vza-step", "10",
                        "--lut-raa-step", "30", "--lut-vza-max", "80", "--output-full-grid", "lut_ipss"] + NOGAS,
     {}, {}, ["lut_ipss.csv"]),
    ("batch5_black", ["--surface", "black", "--batch", "in.csv", "--output", "out.csv"], {}, {"in.csv": BATCH5},
     ["out.csv"]),
    ("batch6_flat_advanced", ["--surface", "flat", "--batch", "in.csv", "--output", "out.csv", "--output-advanced"],
     {}, {"in.csv": BATCH6}, ["out.csv"]),
    ("gas_single_black", ["--surface", "black", "--sza", "40", "--vza", "30", "--raa", "90", "--wavelength", "555"],
     {}, {}, []),
    ("gas_single_bfo_ipss", ["--surface", "black_fresnel_ocean", "--wind-speed", "6", "--pssa", "--sza", "55",
                             "--vza", "35", "--raa", "120", "--wavelength", "443"], {}, {}, []),
    ("gas_single_o3_pwv", ["--surface", "flat", "--gas-column-o3", "300", "--pwv-g-cm", "2.5", "--sza", "35",
                           "--vza", "20", "--raa", "60", "--wavelength", "660"], {}, {}, []),
    ("gas_lut_black", ["--surface", "black", "--sza", "50", "--wavelength", "865", "--lut-vza-step", "10",
                       "--lut-raa-step", "30", "--output-full-grid", "lut_gas.csv"], {}, {}, ["lut_gas.csv"]),
    ("err_no_surface", ["--sza", "40", "--vza", "30", "--raa", "90", "--wavelength", "555"] + NOGAS, {}, {}, []),
    ("err_vza_only", ["--surface", "black", "--sza", "40", "--vza", "30", "--wavelength", "555"] + NOGAS, {}, {}, []),
    ("err_n_mu_gate", ["--surface", "black", "--n-mu", "16", "--sza", "40", "--wavelength", "555"] + NOGAS, {}, {},
     []),
    ("err_m_max_gate", ["--surface", "black", "--m-max", "4", "--sza", "40", "--wavelength", "555"] + NOGAS, {}, {},
     []),
    ("err_lambert", ["--surface", "lambert", "--sza", "40", "--wavelength", "555"], {}, {}, []),
    ("err_bfo_no_wind", ["--surface", "black_fresnel_ocean", "--sza", "40", "--vza", "10", "--raa", "0",
                         "--wavelength", "555"] + NOGAS, {}, {}, []),
    ("err_wavelength", ["--surface", "black", "--sza", "40", "--vza", "10", "--raa", "0", "--wavelength", "300"]
     + NOGAS, {}, {}, []),
    ("err_single_and_grid", ["--surface", "black", "--sza", "40", "--vza", "10", "--raa", "0", "--wavelength", "555",
                             "--lut-vza-step", "5"] + NOGAS, {}, {}, []),
    ("err_aerosol_without_aod", ["--surface", "black", "--aer-h-km", "3", "--sza", "40", "--vza", "10", "--raa",
                                 "0", "--wavelength", "555"] + NOGAS, {}, {}, []),
]


def run_case(exe, ref_dir, args, env, inputs, outputs):
    work = tempfile.mkdtemp()
    try:
        for name, text in inputs.items():
            open(os.path.join(work, name), "w").write(text)
        # relative input/output names resolve in the work dir; the gas data
        # (inputs/afgl_atm, inputs/xsec) is the reduced test set
        shutil.copytree(GAS_INPUTS, os.path.join(work, "inputs"))
        e = dict(os.environ, OMP_NUM_THREADS="1")
        e.pop("OCRT_ADVANCED", None)
        e.pop("OCRT_DEBUG", None)
        e.update(env)
        p = subprocess.run([exe] + args, cwd=work, env=e, capture_output=True, text=True, timeout=600)
        files = {}
        for name in outputs:
            path = os.path.join(work, name)
            files[name] = open(path).read() if os.path.exists(path) else None
        return p.returncode, p.stdout, files
    finally:
        shutil.rmtree(work)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref-cli", f"REF_ROOT={ref}"], check=True,
                   capture_output=True)
    exe = os.path.join(ROOT, "oracle", "_ref", "ocrt_ref")
    ref_dir = os.path.join(ref, "code", "OCRT_C")
    records = []
    for name, args, env, inputs, outputs in CASES:
        rc, out, files = run_case(exe, ref_dir, args, env, inputs, outputs)
        records.append(dict(name=name, args=args, env=env, inputs=inputs, outputs=outputs, rc=rc, stdout=out,
                            files=files))
        print(name, rc)
    with open(OUT, "w") as f:
        json.dump(records, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
