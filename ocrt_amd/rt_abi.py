"""ctypes view of the reference-compatible C ABI (include/ocrt_rt.h).

The struct definitions are read from the header itself, so Python, C callers
and the layout check in tests/ all see one declaration:

    from ocrt_amd import rt_abi
    cs = rt_abi.Case(); opts = rt_abi.options_default(); out = rt_abi.Result()
    rc = rt_abi.library().rt_solve_case_pol(ctypes.byref(cs), ctypes.byref(opts), ctypes.byref(out))

`rt_case_t`, `rt_options_t`, `rt_result_t`, `rt_aerosol_input_t`,
`rt_lut_grid_out_t`, `rt_solver_sos_options_t` and the gas absorption state
`rt_absorption_t` are built by `structs()`; `Case`, `Options`, `Result`,
`Aerosol`, `LutGrid`, `Absorption` are the usual aliases.
"""
import ctypes as C
import os
import re

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "ocrt_rt.h")

_SCALARS = {
    "double": C.c_double, "int": C.c_int,
    "rt_surface_kind_t": C.c_int, "rt_sos_acceleration_t": C.c_int, "rt_integration_method_t": C.c_int,
    "rt_rayleigh_model_t": C.c_int, "rt_tau_r_source_t": C.c_int,
}
_POINTERS = {"double": C.POINTER(C.c_double), "char": C.c_char_p, "void": C.c_void_p}


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def parse_structs(path=HEADER):
    """Ordered field lists [(name, ctype-spec)] of every typedef struct in the header.

    A spec is ("scalar", base), ("pointer", base), ("array", base, n) or
    ("struct", typename)."""
    raw = open(path).read()
    macros = dict(re.findall(r"#define\s+(\w+)\s+(\d+)\s*$", raw, flags=re.M))
    text = _strip_comments(raw)
    for k, v in macros.items():
        text = re.sub(r"\b%s\b" % k, v, text)
    out = {}
    for body, name in re.findall(r"typedef\s+struct\s*\{(.*?)\}\s*(\w+)\s*;", text, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            decl = decl.replace("const ", "")
            m = re.match(r"(\w+)\s+(.*)$", decl)
            base, rest = m.group(1), m.group(2)
            for item in rest.split(","):
                item = item.strip()
                ptr = item.startswith("*")
                item = item.lstrip("*").strip()
                arr = re.match(r"(\w+)\[(\d+)\]$", item)
                if ptr and arr:
                    fields.append((arr.group(1), ("pointer_array", base, int(arr.group(2)))))
                elif ptr:
                    fields.append((item, ("pointer", base)))
                elif arr:
                    fields.append((arr.group(1), ("array", base, int(arr.group(2)))))
                elif base in _SCALARS:
                    fields.append((item, ("scalar", base)))
                else:
                    fields.append((item, ("struct", base)))
        out[name] = fields
    return out


def structs(path=HEADER):
    """ctypes.Structure classes for every struct of the header."""
    specs = parse_structs(path)
    made = {}

    def build(name):
        if name in made:
            return made[name]
        fields = []
        for fname, spec in specs[name]:
            kind = spec[0]
            if kind == "scalar":
                t = _SCALARS[spec[1]]
            elif kind == "pointer":
                t = _POINTERS[spec[1]]
            elif kind == "array":
                t = (_SCALARS[spec[1]] if spec[1] in _SCALARS else build(spec[1])) * spec[2]
            elif kind == "pointer_array":
                t = _POINTERS[spec[1]] * spec[2]
            else:
                t = build(spec[1])
            fields.append((fname, t))
        cls = type(name, (C.Structure,), {"_fields_": fields})
        made[name] = cls
        return cls

    for name in specs:
        build(name)
    return made


_S = structs()
Case = _S["rt_case_t"]
Options = _S["rt_options_t"]
Result = _S["rt_result_t"]
Aerosol = _S["rt_aerosol_input_t"]
LutGrid = _S["rt_lut_grid_out_t"]
SosOptions = _S["rt_solver_sos_options_t"]
Absorption = _S["rt_absorption_t"]
GASES = ("h2o", "o3", "o2", "co2", "no2", "ch4")

SURFACES = {"black": 0, "lambert": 1, "black_fresnel_ocean": 2, "flat": 3, "ocean": 4}


def options_default():
    """rt_options_default() of the header (code/OCRT_C/src/rt_types.h:760-798)."""
    o = Options()
    o.n_mu, o.n_layers, o.max_orders, o.l_max = 24, 40, 20, -1
    o.sos.max_iterations, o.sos.tolerance = 20, 1.0e-7
    o.fourier_m_max = 2
    o.view_as_node, o.vector_mode = 1, 1
    o.aerosol_aod_ref_nm, o.aerosol_l_max = 555.0, 80
    o.lut_vza_step, o.lut_raa_step, o.lut_vza_max = 2.5, 5.0, 85.0
    o.aer_h_km = 2.0
    o.ext_bottom_m_max = -1
    o.conv_tol = 1.0e-6
    return o


def bind(lib):
    """Declare the reference-compatible entry points on a loaded library
    (libocrt_amd.so, or any library exporting the same names)."""
    P = C.POINTER
    lib.rt_solve_case_pol.argtypes = [P(Case), P(Options), P(Result)]
    lib.rt_solve_case_pol_aerosol.argtypes = [P(Case), P(Options), P(Aerosol), P(Result)]
    lib.rt_solve_case_pol_lut.argtypes = [P(Case), P(Options), P(Aerosol), P(Result), P(LutGrid)]
    for f in (lib.rt_solve_case_pol, lib.rt_solve_case_pol_aerosol, lib.rt_solve_case_pol_lut):
        f.restype = C.c_int
    lib.rt_absorption_init.argtypes = [C.c_char_p, C.c_char_p, C.c_int, P(C.c_double), P(Absorption)]
    lib.rt_absorption_init.restype = C.c_int
    lib.rt_absorption_free.argtypes = [P(Absorption)]
    lib.rt_absorption_free.restype = None
    lib.rt_absorption_tau_per_gas.argtypes = [P(Absorption), C.c_int, C.c_double]
    lib.rt_absorption_tau_per_gas.restype = C.c_double
    lib.rt_absorption_tau_total.argtypes = [P(Absorption), C.c_double]
    lib.rt_absorption_tau_total.restype = C.c_double
    return lib


def absorption(lib, afgl_dir, xsec_dir, profile=0, override=None):
    """rt_absorption_init through `lib`; returns the state (free with
    lib.rt_absorption_free)."""
    st = Absorption()
    ov = (C.c_double * 6)(*(override if override is not None else [-1.0] * 6))
    rc = lib.rt_absorption_init(afgl_dir.encode(), xsec_dir.encode(), profile, ov, C.byref(st))
    if rc != 0:
        raise RuntimeError(f"rt_absorption_init failed with {rc}")
    return st


def library():
    """libocrt_amd.so with the reference-compatible entry points declared."""
    from . import library as _lib
    lib = _lib()
    bind(lib)
    lib.ocrt_rt_solve_batch.argtypes = [C.POINTER(Case), C.POINTER(Options), C.POINTER(Aerosol), C.c_int,
                                        C.POINTER(Result), C.POINTER(C.c_int)]
    lib.ocrt_rt_solve_batch.restype = C.c_int
    lib.ocrt_rayleigh_tau_bodhaine1999.argtypes = [C.c_double] * 4
    lib.ocrt_rayleigh_tau_bodhaine1999.restype = C.c_double
    return lib
