"""Build recipe of the native library (hipcc for gfx950, no JIT, in-tree output).

The library is compiled with floating-point contraction off so that FP64/FP32
results keep the reference's rounding, and with packed FP32 instructions
disabled (-packed-fp32-ops) for gfx950.

The kernels are templates with many variants, so the library is built from
several translation units compiled in parallel: the step launches once per
storage type and dynamics (csrc/step.hip), the other typed operations once
per storage type (csrc/storage_ops.hip), the adjoint (csrc/adjoint.hip), the
allocation-free moments entry point (csrc/noalloc.hip) and the set-up and C ABI
(csrc/engine.hip). Objects go to palabos_amd/_build/.
"""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpalabos_amd.so")
OBJ_DIR = os.path.join(HERE, "_build")
HEADERS = [os.path.join(CSRC, "d3q19_device.h"), os.path.join(ROOT, "include", "palabos_amd.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value",
]

STORAGES = ["double", "float", "f16s", "bf16s"]


def units():
    """(object name, source, extra defines) of every translation unit."""
    out = [("engine", "engine.hip", []), ("adjoint", "adjoint.hip", []), ("noalloc", "noalloc.hip", []),
           ("bandwidth", "bandwidth.hip", [])]
    for s in STORAGES:
        out.append((f"ops_{s}", "storage_ops.hip", [f"-DPLBAMD_S={s}"]))
        for dyn in (0, 1):
            out.append((f"step_{s}_{dyn}", "step.hip", [f"-DPLBAMD_S={s}", f"-DPLBAMD_DYN={dyn}"]))
    return out


def _jobs():
    env = os.environ.get("MAX_JOBS")
    n = int(env) if env and env.isdigit() else (os.cpu_count() or 1)
    return max(1, min(16, n))


def build_library(verbose=False):
    """Compiles palabos_amd/libpalabos_amd.so (always recompiles)."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    include = ["-I", os.path.join(ROOT, "include")]

    def compile_one(unit):
        name, src, defs = unit
        obj = os.path.join(OBJ_DIR, name + ".o")
        cmd = [HIPCC] + FLAGS + include + defs + ["-c", os.path.join(CSRC, src), "-o", obj]
        res = subprocess.run(cmd, capture_output=True, text=True)
        return obj, res, cmd

    with ThreadPoolExecutor(_jobs()) as pool:
        results = list(pool.map(compile_one, units()))
    for obj, res, cmd in results:
        if res.returncode != 0:
            raise RuntimeError("hipcc failed: " + " ".join(cmd) + "\n" + res.stdout + res.stderr)
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp"] + [o for o, _, _ in results]
    res = subprocess.run(link, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n" + res.stdout + res.stderr)
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print("built", LIB)
    return LIB


CPP_DIR = os.path.join(HERE, "cpp")
CPP_BIN = os.path.join(CPP_DIR, "bin")
CPP_EXAMPLES = ["cavity3d", "coprocessor", "units", "cylinder3d"]


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


def resource_usage(storage, dyn, match=None):
    """The compiler's figures for the step kernels of one translation unit
    (gfx950 device pass with -Rpass-analysis=kernel-resource-usage): per
    kernel, VGPRs, SGPRs, spills, static LDS and the occupancy in waves per SIMD
    that the register count allows. Returns [(demangled name, {field: value})]."""
    import re
    include = ["-I", os.path.join(ROOT, "include")]
    cmd = ([HIPCC] + FLAGS + include + [f"-DPLBAMD_S={storage}", f"-DPLBAMD_DYN={dyn}", "--cuda-device-only",
            "-Rpass-analysis=kernel-resource-usage", "-c", os.path.join(CSRC, "step.hip"), "-o", os.devnull])
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed: " + " ".join(cmd) + "\n" + res.stderr)
    fields = {"VGPRs": "vgpr", "TotalSGPRs": "sgpr", "VGPRs Spill": "vgpr_spill",
              "LDS Size [bytes/block]": "static_lds", "Occupancy [waves/SIMD]": "waves_per_simd"}
    out = []
    for block in res.stderr.split("remark: Function Name: ")[1:]:
        name = block.split()[0]
        demangled = subprocess.run(["c++filt", name], capture_output=True, text=True).stdout.strip() or name
        if match and not re.search(match, demangled):
            continue
        vals = {}
        for key, short in fields.items():
            m = re.search(re.escape(key) + r": (\S+)", block)
            vals[short] = m.group(1) if m else "?"
        out.append((demangled, vals))
    return out


if __name__ == "__main__":
    import sys
    if len(sys.argv) >= 4 and sys.argv[1] == "--resource-usage":
        # python -m palabos_amd._build --resource-usage <storage> <dyn> [regex]
        for name, v in resource_usage(sys.argv[2], sys.argv[3], sys.argv[4] if len(sys.argv) > 4 else None):
            print(f"{name:110s} vgpr {v['vgpr']:>4s} sgpr {v['sgpr']:>4s} spill {v['vgpr_spill']:>3s} "
                  f"static_lds {v['static_lds']:>5s} waves/SIMD {v['waves_per_simd']}")
    else:
        build_library(verbose=True)
        build_cpp_examples(verbose=True)
