"""Build recipe of the native library (hipcc for gfx950, no JIT, in-tree output).

The library is compiled with floating-point contraction off so that FP64/FP32
results keep the reference's rounding, and with packed FP32 instructions
disabled (-packed-fp32-ops) for gfx950.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libpalabos_amd.so")
SOURCES = [os.path.join(HERE, "csrc", "engine.hip")]
HEADERS = [os.path.join(HERE, "csrc", "d3q19_device.h"), os.path.join(ROOT, "include", "palabos_amd.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value",
]


def build_library(verbose=False):
    """Compiles palabos_amd/libpalabos_amd.so (always recompiles)."""
    cmd = [HIPCC] + FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", LIB + ".tmp"] + SOURCES
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print("built", LIB)
    return LIB


CPP_DIR = os.path.join(HERE, "cpp")
CPP_BIN = os.path.join(CPP_DIR, "bin")
CPP_EXAMPLES = ["cavity3d", "coprocessor"]


def build_cpp_examples(verbose=False):
    """Compiles the C++ programs written against the Palabos-compatible header
    (palabos_amd/cpp/palabos_amd.hh) with g++, linked to the native library."""
    os.makedirs(CPP_BIN, exist_ok=True)
    cxx = os.environ.get("CXX", "g++")
    for name in CPP_EXAMPLES:
        src = os.path.join(CPP_DIR, "examples", name + ".cpp")
        out = os.path.join(CPP_BIN, name)
        cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), "-I", CPP_DIR,
               src, "-L", HERE, "-lpalabos_amd", "-Wl,-rpath,$ORIGIN/../..", "-o", out]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"building {name} failed:\n" + res.stdout + res.stderr)
        if verbose:
            print("built", out)


if __name__ == "__main__":
    build_library(verbose=True)
    build_cpp_examples(verbose=True)
