"""Build the HIP library ``libmqamd.so`` in-tree for gfx950 (hipcc, no cmake).

``python -m mindquantum_amd._build`` or ``__graft_entry__.build()`` runs it; every call
recompiles.  Sources: ``csrc/kernels.hip``, ``csrc/aux_kernels.hip`` and ``csrc/dense_sweep.hip`` (device code), ``csrc/capi.cpp``,
``csrc/schedule.cpp`` and ``csrc/jit.cpp`` (host code).  The text of ``csrc/program.h`` and
``csrc/device_ops.h`` is embedded into the library (``_build/jit_headers.inc``) for the
per-circuit kernels that ``jit.cpp`` compiles with hipRTC at run time.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmqamd.so")
JITC = os.path.join(HERE, "mqamd_jitc")
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


JIT_HEADERS = ["program.h", "device_ops.h"]


def _write_jit_headers():
    """Emit the headers the generated kernels include, as C++ raw string literals."""
    names, bodies = [], []
    for h in JIT_HEADERS:
        with open(os.path.join(CSRC, h)) as f:
            text = f.read()
        assert ")MQAJIT\"" not in text
        names.append('"%s"' % h)
        bodies.append('R"MQAJIT(' + text + ')MQAJIT"')
    out = ("// generated by mindquantum_amd/_build.py from csrc/" + ", csrc/".join(JIT_HEADERS) + "\n"
           "static const char* const kJitHeaderNames[] = {" + ", ".join(names) + "};\n"
           "static const char* const kJitHeaders[] = {\n" + ",\n".join(bodies) + "};\n")
    path = os.path.join(BUILD, "jit_headers.inc")
    with open(path, "w") as f:
        f.write(out)


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    _write_jit_headers()
    hipcc = _hipcc()
    jobs = [
        [hipcc, *COMMON, *DEVICE, "-c", os.path.join(CSRC, "kernels.hip"), "-o", os.path.join(BUILD, "kernels.o")],
        [hipcc, *COMMON, *DEVICE, "-c", os.path.join(CSRC, "aux_kernels.hip"), "-o", os.path.join(BUILD, "aux_kernels.o")],
        [hipcc, *COMMON, *DEVICE, "-c", os.path.join(CSRC, "dense_sweep.hip"), "-o", os.path.join(BUILD, "dense_sweep.o")],
        [hipcc, *COMMON, *DEVICE, "-c", os.path.join(CSRC, "capi.cpp"), "-o", os.path.join(BUILD, "capi.o")],
        [hipcc, *COMMON, *DEVICE, "-c", os.path.join(CSRC, "schedule.cpp"), "-o", os.path.join(BUILD, "schedule.o")],
        [hipcc, *COMMON, *DEVICE, "-I", BUILD, "-c", os.path.join(CSRC, "jit.cpp"), "-o", os.path.join(BUILD, "jit.o")],
        [hipcc, *COMMON, "-c", os.path.join(CSRC, "jitc_main.cpp"), "-o", os.path.join(BUILD, "jitc_main.o")],
    ]
    with ThreadPoolExecutor(max_workers=7) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([hipcc, *DEVICE, "-shared", "-fPIC", "-o", tmp,
          os.path.join(BUILD, "kernels.o"), os.path.join(BUILD, "aux_kernels.o"), os.path.join(BUILD, "dense_sweep.o"),
          os.path.join(BUILD, "capi.o"), os.path.join(BUILD, "schedule.o"),
          os.path.join(BUILD, "jit.o"), "-lhiprtc"])
    os.replace(tmp, LIB)
    # compile helper of the code generator: hipRTC runs one compile at a time per process
    tmp = JITC + ".tmp"
    _run([hipcc, "-o", tmp, os.path.join(BUILD, "jitc_main.o"), os.path.join(BUILD, "jit.o"), "-lhiprtc"])
    os.replace(tmp, JITC)
    return LIB


SANITIZE_DIR = os.path.join(BUILD, "sanitize")
# Host-side AddressSanitizer + UndefinedBehaviorSanitizer; device code is built as usual.
SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer", "-g"]


def build_sanitized(verbose=False):
    """Build ``_build/sanitize/libmqamd.so`` (and its compile helper) with the host code - the C ABI,
    the tile-pass scheduler and the code generator - under ASan/UBSan.  It reuses the device objects
    of the regular build (run ``build()`` first).  Load it by setting ``MQAMD_LIBRARY`` to its path,
    with the ASan runtime preloaded, e.g.::

        LD_PRELOAD=$(hipcc -print-file-name=libclang_rt.asan-x86_64.so) ASAN_OPTIONS=detect_leaks=0 \\
        MQAMD_LIBRARY=mindquantum_amd/_build/sanitize/libmqamd.so python -m pytest tests -m "not gpu"
    """
    for obj in ("kernels.o", "aux_kernels.o", "dense_sweep.o"):
        if not os.path.exists(os.path.join(BUILD, obj)):
            raise RuntimeError("run build() first: the sanitized library reuses " + obj)
    os.makedirs(SANITIZE_DIR, exist_ok=True)
    _write_jit_headers()
    hipcc = _hipcc()
    flags = [f for f in COMMON if f != "-O3"] + ["-O1"] + SANITIZE
    out = lambda name: os.path.join(SANITIZE_DIR, name)  # noqa: E731
    jobs = [
        [hipcc, *flags, *DEVICE, "-c", os.path.join(CSRC, "capi.cpp"), "-o", out("capi.o")],
        [hipcc, *flags, *DEVICE, "-c", os.path.join(CSRC, "schedule.cpp"), "-o", out("schedule.o")],
        [hipcc, *flags, *DEVICE, "-I", BUILD, "-c", os.path.join(CSRC, "jit.cpp"), "-o", out("jit.o")],
        [hipcc, *flags, "-c", os.path.join(CSRC, "jitc_main.cpp"), "-o", out("jitc_main.o")],
    ]
    with ThreadPoolExecutor(max_workers=4) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    lib = out("libmqamd.so")
    _run([hipcc, *DEVICE, *SANITIZE, "-shared", "-fPIC", "-shared-libsan", "-o", lib,
          os.path.join(BUILD, "kernels.o"), os.path.join(BUILD, "aux_kernels.o"), os.path.join(BUILD, "dense_sweep.o"),
          out("capi.o"), out("schedule.o"),
          out("jit.o"), "-lhiprtc"])
    _run([hipcc, *SANITIZE, "-shared-libsan", "-o", out("mqamd_jitc"), out("jitc_main.o"), out("jit.o"), "-lhiprtc"])
    return lib


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        print(build_sanitized(verbose="-v" in sys.argv))
    else:
        print(build(verbose="-v" in sys.argv))
