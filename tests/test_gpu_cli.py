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
