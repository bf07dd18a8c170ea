"""The `ocrt` command-line program against the reference CLI.

tests/golden/cli_cases.json holds, per case, the arguments, environment,
input files, exit status, stdout and written files of the reference CLI
(built from its sources and run by tests/golden/make_cli_golden.py).  The same
command is run here with ocrt_amd/ocrt on the GPU, with the same gas data
(tests/golden/gas_inputs as ./inputs): exit statuses must match;
for successful runs every printed number must agree to 2e-10 relative (the
outputs are printed with 11 significant digits) and every text token
(keys, column names, nan) must be identical, in stdout and in every file.
"""
import json
import math
import os
import shutil
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "ocrt_amd", "ocrt")
CASES = json.load(open(os.path.join(ROOT, "tests", "golden", "cli_cases.json")))
GAS_INPUTS = os.path.join(ROOT, "tests", "golden", "gas_inputs")


def tokens(text):
    out = []
    for tok in text.replace(",", " ").split():
        key, _, val = tok.rpartition("=")
        out.append((key, val))
    return out


def num(s):
    try:
        return float(s)
    except ValueError:
        return None


def same_text(got, want, where):
    gl = [l for l in got.splitlines() if l.strip()]
    wl = [l for l in want.splitlines() if l.strip()]
    assert len(gl) == len(wl), (where, len(gl), len(wl))
    if wl and "walltime_ms" in wl[0]:
        # batch detail CSV: the per-row wall time is a measurement, not a result
        col = wl[0].split(",").index("walltime_ms")
        assert gl[0] == wl[0]
        strip = lambda l: ",".join(v for i, v in enumerate(l.split(",")) if i != col)
        gl, wl = [strip(l) for l in gl[1:]], [strip(l) for l in wl[1:]]
    for a, b in zip(gl, wl):
        ta, tb = tokens(a), tokens(b)
        assert len(ta) == len(tb), (where, a, b)
        for (ka, va), (kb, vb) in zip(ta, tb):
            assert ka == kb, (where, ka, kb)
            fa, fb = num(va), num(vb)
            if fb is None or math.isnan(fb):
                assert va == vb, (where, ka, va, vb)
            elif "output=" in b and ka == "output":
                continue
            else:
                assert fa is not None and abs(fa - fb) <= 2e-10 * abs(fb) + 1e-15, (where, ka, va, vb)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_cli_matches_reference(case, tmp_path):
    assert os.path.exists(CLI), "build the CLI first (python -m ocrt_amd.build)"
    shutil.copytree(GAS_INPUTS, tmp_path / "inputs")
    for name, text in case["inputs"].items():
        (tmp_path / name).write_text(text)
    env = dict(os.environ)
    env.pop("OCRT_ADVANCED", None)
    env.pop("OCRT_DEBUG", None)
    env.update(case["env"])
    print(f"[gpu] ocrt {' '.join(case['args'])}", flush=True)
    p = subprocess.run([CLI] + case["args"], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=300)
    assert p.returncode == case["rc"], (p.returncode, case["rc"], p.stderr[-2000:])
    if case["rc"] != 0:
        assert "error" in p.stderr
        return
    # the batch summary line names the output path; everything else must agree
    same_text(p.stdout, case["stdout"], "stdout")
    for name, want in case["files"].items():
        got = (tmp_path / name).read_text()
        same_text(got, want, name)


def test_reference_gate_fixture_with_gas(tmp_path):
    """The reference's own non-IPSS gate (tests/fixtures/ipss_nonpssa_reference.csv
    of the reference, stored gzip-compressed; command of
    scripts/run_ipss_gates.sh:30-33): Rayleigh with the default US76 gas
    absorption over the rough ocean, SZA 75, 555 nm, full 5-degree LUT grid.
    Every value within 1e-12 absolute (the reference gate's floor for numerical
    zeros, run_ipss_gates.sh:49-88) or within the printing precision of the
    CSV (%.10e: 11 significant digits, so 2e-10 relative covers a flipped last
    digit)."""
    import csv
    import gzip
    shutil.copytree(GAS_INPUTS, tmp_path / "inputs")
    args = ["--sza", "75", "--wavelength", "555", "--pressure", "1013.25", "--surface", "black_fresnel_ocean",
            "--wind-speed", "5", "--decouple-sunglint", "--lut-vza-step", "5", "--lut-vza-max", "85",
            "--lut-raa-step", "5", "--output-full-grid", "gate.csv"]
    print(f"[gpu] ocrt {' '.join(args)}", flush=True)
    env = dict(os.environ)
    env.pop("OCRT_ADVANCED", None)
    p = subprocess.run([CLI] + args, cwd=tmp_path, env=env, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    want = list(csv.reader(gzip.open(os.path.join(ROOT, "tests", "golden", "ref_fixtures",
                                                  "ipss_nonpssa_reference.csv.gz"), "rt")))
    got = list(csv.reader(open(tmp_path / "gate.csv")))
    assert got[0] == want[0] and len(got) == len(want) == 1297
    worst = 0.0
    for ra, rb in zip(got[1:], want[1:]):
        for a, b in zip(ra, rb):
            fa, fb = float(a), float(b)
            d = abs(fa - fb)
            if d > 1e-12:
                worst = max(worst, d / abs(fb))
    print(f"[gpu] reference gate fixture: worst relative difference beyond 1e-12 absolute: {worst:.2e}", flush=True)
    assert worst <= 2e-10
