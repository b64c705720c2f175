"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host code, for
running the CPU tests under the sanitizers (no GPU involved).

Builds, with g++/gcc only:
  - _build_asan/sw2d-efdc: the driver (case-file parsing, the EFDC writers,
    the EFDC.INP card update) with the host-only C-ABI functions of
    csrc/sw2d_host.cpp (coupling records, catchment smoothing, evaporation
    series) compiled into it; the remaining entry points resolve to the
    regular libsw2d.so, so the GPU paths are not sanitized;
  - oracle/_asan/liboracle.so (make -C oracle asan).
Then, from the repository root:

  ASAN_OPTIONS=detect_leaks=0 LD_PRELOAD=$(gcc -print-file-name=libasan.so) \\
  SW2D_DRIVER=sw2d-efdc_amd/_build_asan/sw2d-efdc \\
  SW2D_ORACLE_LIB=oracle/_asan/liboracle.so python -m pytest tests -m "not gpu"

Run `python sw2d-efdc_amd/build.py` first (the driver links libsw2d.so).
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["sw2d_driver.cpp", "sw2d_io.cpp", "sw2d_efdc_config.cpp", "sw2d_gefdc.cpp", "sw2d_lorp.cpp",
           "sw2d_host.cpp"]
FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=undefined", "-std=c++17", "-Wall", "-pthread"]


def build():
    out = os.path.join(HERE, "_build_asan")
    os.makedirs(out, exist_ok=True)
    exe = os.path.join(out, "sw2d-efdc")
    cmd = ["g++", *FLAGS, "-o", exe, *[os.path.join(CSRC, s) for s in SOURCES],
           f"-L{HERE}", "-lsw2d", f"-Wl,-rpath,{HERE}"]
    print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    return exe


if __name__ == "__main__":
    build()
