"""Builds tfhe_amd/libtfhe_amd.so (HIP for gfx950 + host C++), in-tree.

    python -m tfhe_amd.build            # or tfhe_amd.build.build()
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "libtfhe_amd.so")
BUILD = os.path.join(PKG, "_build")

ARCH = os.environ.get("TFHE_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX_HOST", shutil.which("g++") or "g++")

# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    # keep per-thread arrays in registers (never in LDS, which the kernels budget explicitly)
    "-mllvm", "-disable-promote-alloca-to-lds",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose_resources=False):
    os.makedirs(BUILD, exist_ok=True)
    eng_o = os.path.join(BUILD, "engine.o")
    net_o = os.path.join(BUILD, "netlist.o")
    host_o = os.path.join(BUILD, "host.o")
    io_o = os.path.join(BUILD, "io.o")
    srv_o = os.path.join(BUILD, "serve.o")
    cli_o = os.path.join(BUILD, "client.o")
    extra = ["-Rpass-analysis=kernel-resource-usage"] if verbose_resources else []
    _run([HIPCC] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, "engine.hip"), "-o", eng_o])
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "netlist.hip"), "-o", net_o])
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "serve.hip"), "-o", srv_o])
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "client.hip"), "-o", cli_o])
    _run([CXX] + HOST_FLAGS + ["-c", os.path.join(CSRC, "host.cpp"), "-o", host_o])
    _run([CXX] + HOST_FLAGS + ["-c", os.path.join(CSRC, "io.cpp"), "-o", io_o])
    tmp = OUT + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, eng_o, net_o, srv_o, cli_o, host_o, io_o,
          "-Wl,-Bsymbolic", "-Wl,--no-undefined", "-lpthread"])
    os.replace(tmp, OUT)
    return OUT


SAN = os.path.join(PKG, "_build_san")
SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g", "-O1"]


def build_sanitized():
    """The same library with its host C++ (host.cpp, io.cpp) built under
    AddressSanitizer and UndefinedBehaviorSanitizer by g++, linked by g++ against
    the HIP objects of the normal build (GPU code is not instrumented), into
    tfhe_amd/_build_san/libtfhe_amd.so; and the CPU oracle likewise into
    tfhe_amd/_build_san/liboracle.so. Load them with TFHE_AMD_LIB /
    TFHE_AMD_ORACLE_LIB and gcc's libasan and libubsan preloaded."""
    os.makedirs(SAN, exist_ok=True)
    objs = [os.path.join(BUILD, f) for f in ("engine.o", "netlist.o", "serve.o", "client.o")]
    if not all(os.path.exists(o) for o in objs):
        build()
    host_o, io_o = os.path.join(SAN, "host.o"), os.path.join(SAN, "io.o")
    _run([CXX] + HOST_FLAGS + SAN_FLAGS + ["-c", os.path.join(CSRC, "host.cpp"), "-o", host_o])
    _run([CXX] + HOST_FLAGS + SAN_FLAGS + ["-c", os.path.join(CSRC, "io.cpp"), "-o", io_o])
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(HIPCC)))
    _run([CXX, "-shared", "-fPIC"] + SAN_FLAGS + ["-o", os.path.join(SAN, "libtfhe_amd.so")] + objs +
         [host_o, io_o, "-L" + os.path.join(rocm, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(rocm, "lib"),
          "-lpthread"])
    _run([CXX, "-std=c++17", "-fPIC", "-shared"] + SAN_FLAGS + ["-o", os.path.join(SAN, "liboracle.so"),
                                                                os.path.join(ROOT, "oracle", "tfhe_oracle.cpp"),
                                                                "-lpthread"])
    return SAN


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        build_sanitized()
    else:
        build(verbose_resources="-v" in sys.argv)
