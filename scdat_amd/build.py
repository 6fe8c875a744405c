"""Build the gfx950 shared library libscdat_amd.so in-tree.

Each HIP source is compiled to an object with hipcc for gfx950, the host-only
mesh code with g++, then hipcc links the shared library.

Usage: python -m scdat_amd.build
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libscdat_amd.so")

HIP_SOURCES = ["context.hip", "push.hip", "field.hip", "sort.hip", "emit.hip", "adjoint.hip"]
CXX_SOURCES = ["mesh_host.cpp", "host_tools.cpp", "circuit.cpp", "sampler.cpp", "mesh_io.cpp"]

ARCH = "gfx950"
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    # packed FP32 instructions disabled for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose=False, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    cmds = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        cmds.append([hipcc()] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in CXX_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        cmds.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for c in cmds:
            if verbose:
                print(" ".join(c), flush=True)
        list(ex.map(_run, cmds))
    tmp = OUT + ".tmp"
    link = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(link), flush=True)
    _run(link)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    try:
        print(build(verbose="-v" in sys.argv))
    except RuntimeError as e:
        sys.stderr.write(str(e) + "\n")
        sys.exit(1)
