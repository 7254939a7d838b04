"""Build libd2bi_gpu.so (gfx950) in-tree with hipcc.

    python -m simulated_tem_amd.build

The library holds the HIP kernels, the C-ABI and the host setup code.
Flags:
  --offload-arch=gfx950            MI355X only
  -ffp-contract=off                FP64 kernels keep the reference's unfused
                                   operation order; FAST kernels write fmaf
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                   no packed FP32 (v_pk_*_f32) instructions
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libd2bi_gpu.so")

SOURCES = ["d2bi_gpu.hip", "d2bi_host.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-result",
          "-I" + os.path.join(ROOT, "include")]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def build(verbose=True):
    os.makedirs(OUT_DIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
        if src.endswith(".hip"):
            cmd = [HIPCC, "-c", path, "-o", obj] + COMMON + DEVICE
        else:
            cmd = [HIPCC, "-x", "c++", "-c", path, "-o", obj] + COMMON + ["-D__HIP_PLATFORM_AMD__"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-o", tmp] + objs + ["--offload-arch=gfx950", "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
