"""Build libjurassic_gpu.so for gfx950 with hipcc (in-tree, no JIT cache).

Every HIP source is compiled with
  -Xclang -target-feature -Xclang -packed-fp32-ops   (no packed FP32 math)
  -fhip-fp32-correctly-rounded-divide-sqrt           (IEEE float division)
and the FP64 translation units with -ffp-contract=off so that each multiply
and add rounds separately, as in the reference's build.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libjurassic_gpu.so")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("JURASSIC_GPU_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-rdc",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-I" + CSRC, "-I" + os.path.join(os.path.dirname(HERE), "include")]

# (source, extra flags)
SOURCES = [
    ("jurassic_gpu.cpp", ["-x", "hip", "-ffp-contract=off"]),
    ("raytrace.hip", ["-ffp-contract=off"]),
    ("lookup_f64.hip", ["-ffp-contract=off"]),
    ("lookup_f32.hip", ["-ffp-contract=fast"]),
    ("radiance_f64.hip", ["-ffp-contract=off"]),
    ("radiance_f32.hip", ["-ffp-contract=fast"]),
    ("aux_kernels.hip", ["-ffp-contract=off"]),
    ("adjoint.hip", ["-ffp-contract=fast"]),
    ("lm_kernels.hip", ["-ffp-contract=off"]),
    ("retrieval.cpp", ["-x", "hip", "-ffp-contract=off"]),
    ("tblgen.hip", ["-ffp-contract=off"]),
]


def _compile(src, flags):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda a: _compile(*a), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
