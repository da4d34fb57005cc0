"""Build the HIP transport library in-tree (gegs_amd/libgegs_hip.so) for gfx950,
with its host-only file formats (csrc/host_io.cpp, built by the host compiler),
and the C++ run driver gegs_amd/gegs_run linked against it."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgegs_hip.so")
RUNNER = os.path.join(HERE, "gegs_run")
SOURCES = ["gegs_capi.hip"]
HOST_SOURCES = ["host_io.cpp", "host_input.cpp", "host_ct.cpp"]
HEADERS = ["host_io.cpp", "host_input.cpp", "host_ct.cpp", "gegs_run.cpp", "transport.hpp", "gmath.hpp", "plain.hpp", "event.hpp", "mlc.hpp", "jaws.hpp", "blockwedge.hpp", "ctimport.hpp"]

HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",                              # no FMA contraction: arithmetic as written
    "-fhip-fp32-correctly-rounded-divide-sqrt",      # IEEE division and sqrt
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


# host-only C++: no device code, binary32 arithmetic as written
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    objs = []
    for s in HOST_SOURCES:
        o = os.path.join(CSRC, os.path.splitext(s)[0] + ".o")
        _run(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o], verbose)
        objs.append(o)
    tmp = LIB + ".tmp"
    # the host objects go to the linker as they are (hipcc would read them as HIP source)
    _run([hipcc()] + HIPCC_FLAGS + ["-o", tmp] + srcs + ["-Wl," + o for o in objs], verbose)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    tmp = RUNNER + ".tmp"
    _run(["g++"] + HOST_FLAGS + [os.path.join(CSRC, "gegs_run.cpp"), "-o", tmp, "-L" + HERE, "-lgegs_hip",
                                 "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib", "-L/opt/rocm/lib", "-lamdhip64"],
         verbose)
    os.replace(tmp, RUNNER)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
