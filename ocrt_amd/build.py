"""Build the HIP library ocrt_amd/libocrt_amd.so for gfx950 (CPU-only host is fine:
hipcc cross-compiles the device code).

    python -m ocrt_amd.build          # or ocrt_amd.build.build()
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libocrt_amd.so")
CLI = os.path.join(HERE, "ocrt")
SOURCES = ["plan.hip", "setup.hip", "sos.hip", "ipss.hip", "adjoint.hip", "lut.hip", "rt_compat.hip", "gas.hip", "interface.hip"]
ARCH = os.environ.get("OCRT_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops off: hipcc -O3 must not emit v_pk_{add,mul,fma}_f32 on this target.
# Multiply-adds fuse everywhere, as with -ffp-contract=fast, except in the
# blocks that say `#pragma clang fp contract(off)`.
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-fno-gpu-rdc", "-ffp-contract=fast-honor-pragmas", "-Wno-unused-result"]


def build(verbose=False, jobs=4):
    """Compile every translation unit to an object, then link the shared library."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [hipcc, *FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = LIB + ".tmp"
    cmd = [hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_cli(verbose)
    return LIB


def build_cli(verbose=False):
    """The `ocrt` command-line program (csrc/cli.cpp), linked against the
    library next to it."""
    cxx = os.environ.get("CXX", "g++")
    tmp = CLI + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "cli.cpp"),
           "-L", HERE, "-locrt_amd", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-o", tmp]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, CLI)
    return CLI


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    text = out.decode(errors="replace")
    # the host pass does not know the packed-fp32 feature; that note is harmless
    lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l]
    if p.returncode != 0:
        sys.stderr.write("\n".join(lines) + "\n")
        raise RuntimeError(f"hipcc failed on {src}")
    if lines:
        sys.stderr.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
