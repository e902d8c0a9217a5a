"""Build the HIP library ``libmqamd.so`` in-tree for gfx950 (hipcc, no cmake).

``python -m mindquantum_amd._build`` or ``__graft_entry__.build()`` runs it; every call
recompiles.  Sources: ``csrc/kernels.hip`` (device code), ``csrc/capi.cpp`` and
``csrc/schedule.cpp`` (host code).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmqamd.so")
ARCH = os.environ.get("MQAMD_ARCH", "gfx950")

COMMON = ["-std=c++20", "-O3", "-fPIC", "-Wall", "-Wno-unused-function"]
# Packed FP32 instructions are disabled for all device code.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    jobs = [
        [hipcc, *COMMON, *DEVICE, "-c", os.path.join(CSRC, "kernels.hip"), "-o", os.path.join(BUILD, "kernels.o")],
        [hipcc, *COMMON, *DEVICE, "-c", os.path.join(CSRC, "capi.cpp"), "-o", os.path.join(BUILD, "capi.o")],
        [hipcc, *COMMON, *DEVICE, "-c", os.path.join(CSRC, "schedule.cpp"), "-o", os.path.join(BUILD, "schedule.o")],
    ]
    with ThreadPoolExecutor(max_workers=3) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([hipcc, *DEVICE, "-shared", "-fPIC", "-o", tmp,
          os.path.join(BUILD, "kernels.o"), os.path.join(BUILD, "capi.o"), os.path.join(BUILD, "schedule.o")])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
