"""Build libtwopcf_hip.so in-tree (gfx950).

    python -m two_pcf_amd.build

The host pieces that must round exactly as the reference's g++ build
(axis thresholds with glibc log1pf, cosmology, Galaxy::set) are compiled with
g++; the kernels and the C ABI with hipcc for gfx950.  Everything is built
with -ffp-contract=off: the exact pair pipeline must not fuse multiply-adds,
and the fast path asks for its fused multiply-adds explicitly (fmaf).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libtwopcf_hip.so")
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("TWOPCF_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    # packed FP32 instructions are kept out of the code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose_isa=False):
    os.makedirs(BUILD, exist_ok=True)
    host_o = os.path.join(BUILD, "host_math.o")
    run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
         "-c", os.path.join(CSRC, "host_math.cpp"), "-o", host_o])
    dev_o = os.path.join(BUILD, "twopcf_hip.o")
    cmd = [os.path.join(ROCM, "bin", "hipcc"), *HIP_FLAGS, "-I", os.path.join(ROOT, "include"),
           "-c", os.path.join(CSRC, "twopcf_hip.hip"), "-o", dev_o]
    if verbose_isa:
        cmd += ["-save-temps=obj"]
    run(cmd)
    tmp = LIB + ".tmp"
    run([os.path.join(ROCM, "bin", "hipcc"), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
         dev_o, host_o, "-lm"])
    os.replace(tmp, LIB)
    # the reference-interface host layer and the parameter-file driver
    host = os.path.join(CSRC, "host")
    # twopcf: the plain build; twopcf_invp: the inverse-p build
    for name, inv in (("twopcf", 0), ("twopcf_invp", 1)):
        drv = os.path.join(HERE, name)
        run(["g++", "-O2", "-std=c++17", "-pthread", "-ffp-contract=off", "-fno-fast-math", "-Wall",
             f"-DTWOPCF_INVERSE_P={inv}",
             os.path.join(host, "twopcf_ref_api.cpp"), os.path.join(host, "h5lib.cpp"),
             os.path.join(host, "main.cpp"),
             "-L", HERE, "-ltwopcf_hip", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN", "-ldl", "-o", drv + ".tmp"])
        os.replace(drv + ".tmp", drv)
    return LIB


if __name__ == "__main__":
    build(verbose_isa="--isa" in sys.argv)
