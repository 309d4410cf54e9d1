"""Build the native libraries in-tree (hipcc for gfx950, g++ for host code).

Outputs go to cuvhs_amd/lib/: libcuvhs.so (decoder, C ABI of include/cuvhs.h),
librfgen.so (synthetic capture generator) and the `cuvhs` command-line tool.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")

HIP_SOURCES = ["k1_filters", "k1_mfma", "guard", "sync", "k2b", "k3", "k45", "k6_chroma", "k7_dropout", "context"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose=False, jobs=8):
    os.makedirs(LIB, exist_ok=True)
    objs = []
    cmds = []
    for s in HIP_SOURCES:
        o = os.path.join(LIB, s + ".o")
        objs.append(o)
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s + ".hip"), "-o", o])
    design_o = os.path.join(LIB, "design.o")
    cmds.append(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fopenmp", "-c",
                 os.path.join(CSRC, "design.cpp"), "-o", design_o])
    objs.append(design_o)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    so = os.path.join(LIB, "libcuvhs.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", so + ".tmp"] + objs + ["-lgomp"], verbose)
    os.replace(so + ".tmp", so)
    gen = os.path.join(LIB, "librfgen.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-shared", "-o", gen + ".tmp",
          os.path.join(CSRC, "gen", "rfgen.cpp")], verbose)
    os.replace(gen + ".tmp", gen)
    cli_src = os.path.join(CSRC, "cli", "cuvhs_cli.cpp")
    if os.path.exists(cli_src):
        exe = os.path.join(LIB, "cuvhs")
        _run([HIPCC, "-O2", "-std=c++17", "-o", exe + ".tmp", cli_src, os.path.join(CSRC, "cli", "tbc_io.cpp"),
              "-L" + LIB, "-lcuvhs", "-Wl,-rpath,$ORIGIN"], verbose)
        os.replace(exe + ".tmp", exe)
    return so


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB)
