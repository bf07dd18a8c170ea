"""Field offsets and sizes of the reference-compatible structs.

Compiles a small C program with gcc that prints offsetof() of every field of
rt_case_t, rt_options_t, rt_result_t, rt_aerosol_input_t, rt_lut_grid_out_t,
rt_solver_sos_options_t and the gas absorption state rt_afgl_atm_t,
rt_xsec_t, rt_absorption_t (field names taken from include/ocrt_rt.h) and
sizeof() of each struct, against either this repository's header or the
reference's own headers (code/OCRT_C/src/rt_solver.h, which includes
rt_types.h, and rt_absorption.h).  Run as a script in a container that holds the reference, it
writes tests/golden/rt_layout.json from the reference headers:

    python tests/golden/make_rt_layout.py [/root/reference]
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REF_HEADERS = ["rt_solver.h", "rt_absorption.h"]
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

STRUCTS = ["rt_solver_sos_options_t", "rt_case_t", "rt_result_t", "rt_options_t", "rt_aerosol_input_t",
           "rt_lut_grid_out_t", "rt_afgl_atm_t", "rt_xsec_t", "rt_absorption_t"]


def field_names():
    import importlib.util
    spec = importlib.util.spec_from_file_location("rt_abi_parse", os.path.join(ROOT, "ocrt_amd", "rt_abi.py"))
    # parse without importing the package (no torch, no library load)
    src = open(os.path.join(ROOT, "ocrt_amd", "rt_abi.py")).read()
    ns = {"__file__": os.path.join(ROOT, "ocrt_amd", "rt_abi.py"), "__name__": "rt_abi_parse"}
    exec(compile(src.split("_S = structs()")[0], "rt_abi.py", "exec"), ns)
    specs = ns["parse_structs"]()
    return {s: [f for f, _ in specs[s]] for s in STRUCTS}


def layout(include_dir, headers):
    """{struct: {"size": n, "fields": {name: offset}}} as compiled by gcc."""
    names = field_names()
    headers = [headers] if isinstance(headers, str) else list(headers)
    lines = ["#include <stdio.h>", "#include <stddef.h>"] + [f'#include "{h}"' for h in headers] + ["int main(void) {"]
    for s in STRUCTS:
        lines.append(f'    printf("%s size %zu\\n", "{s}", sizeof({s}));')
        for f in names[s]:
            lines.append(f'    printf("%s %s %zu\\n", "{s}", "{f}", offsetof({s}, {f}));')
    lines.append("    return 0;\n}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(c, "w").write("\n".join(lines) + "\n")
        subprocess.run(["gcc", "-std=c11", "-O0", "-I", include_dir, c, "-o", exe, "-lm"], check=True,
                       capture_output=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    res = {s: {"size": None, "fields": {}} for s in STRUCTS}
    for line in out.splitlines():
        s, f, v = line.split()
        if f == "size":
            res[s]["size"] = int(v)
        else:
            res[s]["fields"][f] = int(v)
    return res


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    lay = layout(os.path.join(ref, "code", "OCRT_C", "src"), REF_HEADERS)
    with open(os.path.join(HERE, "rt_layout.json"), "w") as f:
        json.dump(lay, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "rt_layout.json"))


if __name__ == "__main__":
    main()
