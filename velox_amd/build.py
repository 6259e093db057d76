"""Builds libvxhip.so (the HIP device step) in-tree for gfx950.

Every HIP source is compiled with contraction off and correctly rounded FP32
division and square root, so device arithmetic is bitwise the reference's
strict CPU arithmetic, and with packed FP32 instructions disabled.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_build")
SOURCES = ["vx_step.hip", "vx_sort.hip"]
HOST_SOURCES = ["scene_build.cpp"]
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I" + os.path.join(ROOT, "include")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
         "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-I" + os.path.join(ROOT, "include")]


def build(jobs=None):
    """Compile every HIP translation unit (in parallel) and link libvxhip.so."""
    os.makedirs(OUT, exist_ok=True)
    objs = [os.path.join(OUT, s.replace(".hip", ".o")) for s in SOURCES]
    hobjs = [os.path.join(OUT, s.replace(".cpp", ".o")) for s in HOST_SOURCES]
    procs = [subprocess.Popen([HIPCC] + FLAGS + ["-c", os.path.join(HERE, "csrc", s), "-o", o])
             for s, o in zip(SOURCES, objs)]
    procs += [subprocess.Popen([CXX] + HOST_FLAGS + ["-c", os.path.join(HERE, "csrc", s), "-o", o])
              for s, o in zip(HOST_SOURCES, hobjs)]
    codes = [p.wait() for p in procs]
    if any(codes):
        raise RuntimeError(f"compile failed: {dict(zip(SOURCES + HOST_SOURCES, codes))}")
    objs += hobjs
    lib = os.path.join(OUT, "libvxhip.so")
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", lib + ".tmp"])
    os.replace(lib + ".tmp", lib)
    return lib


if __name__ == "__main__":
    print(build())
