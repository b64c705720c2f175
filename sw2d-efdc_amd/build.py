"""Build libsw2d.so (HIP, gfx950) and the sw2d-efdc driver in place.

Run as a script or call build().  hipcc cross-compiles for gfx950 without a
GPU.  Flags:
  -ffp-contract=off  no fused multiply-add: FP64 results follow the written
                     evaluation order, as the CPU FP64 check does;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                     no packed FP32 instructions.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libsw2d.so")
DRIVER = os.path.join(HERE, "sw2d-efdc")

# (source, extra flags); the FP32 kernel may fuse multiply-adds within one
# source expression, nothing else may fuse any.  The fused kernels are
# scheduled for instruction-level parallelism (LLVM's AMDGPU scheduler
# strategies; the register counts, hence the waves per SIMD, stay as they
# are): measured per launch on the 256-scenario batch and the 2048^2 storm,
# "max-ilp" makes the FP32 form ~1% faster, "iterative-ilp" the FP64 and
# mixed forms 0.3-2%.  Scheduling does not change what is rounded where.
ILP_MAX = ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]
ILP_ITER = ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"]
HIP_SOURCES = [("sw2d_api.hip", []), ("sw2d_fused64.hip", ILP_ITER),
               ("sw2d_fused32.hip", ["-ffp-contract=on", *ILP_MAX]),
               ("sw2d_fusedmix.hip", ["-ffp-contract=on", *ILP_ITER]), ("sw2d_plain.hip", []),
               ("sw2d_batch.hip", []), ("sw2d_server.hip", []), ("sw2d_output.hip", []),
               ("sw2d_host.cpp", [])]
DRIVER_SOURCES = ["sw2d_driver.cpp", "sw2d_io.cpp", "sw2d_efdc_config.cpp", "sw2d_gefdc.cpp", "sw2d_lorp.cpp"]
ARCH = os.environ.get("SW2D_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False):
    """Compile every HIP/C++ source of the package (always recompiles)."""
    env = dict(os.environ)
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src, extra in HIP_SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), f"--offload-arch={ARCH}", *COMMON, *extra, "-c", "-o", obj,
               os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd))
        procs.append((subprocess.Popen(cmd, env=env), cmd))
        objs.append(obj)
    for p, cmd in procs:
        if p.wait() != 0:
            raise subprocess.CalledProcessError(p.returncode, cmd)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True, env=env)
    os.replace(tmp, LIB)
    dsrcs = [os.path.join(CSRC, s) for s in DRIVER_SOURCES]
    if all(os.path.exists(s) for s in dsrcs):
        tmp = DRIVER + ".tmp"
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-pthread", "-o", tmp, *dsrcs,
               f"-L{HERE}", "-lsw2d", f"-Wl,-rpath,$ORIGIN"]
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True, env=env)
        os.replace(tmp, DRIVER)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
