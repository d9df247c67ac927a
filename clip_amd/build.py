"""Builds libclip_amd.so (HIP kernels for gfx950 + host C++ behind the C-ABI)
in-tree, next to this file.

    python -m clip_amd.build            # or: python clip_amd/build.py
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libclip_amd.so")
DRIVER = os.path.join(HERE, "clip")

# (source, extra flags): the FP64 instantiation and the C-ABI unit (whose launchers
# work on FP64 arrays) are built without FP contraction
HIP_SOURCES = [("solver.hip", ["-ffp-contract=off"]), ("inst_f64.hip", ["-ffp-contract=off"]), ("inst_f32.hip", []),
               ("inst_mixed.hip", []), ("noalloc.hip", [])]
CPP_SOURCES = ["config.cpp"]
ARCH = os.environ.get("CLIP_AMD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-I", os.path.join(ROOT, "include"), "-I", CSRC]
# Packed FP32 instructions are disabled for all device code.
DEVICE_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-munsafe-fp-atomics"]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    objs, cmds = [], []
    for src, extra in HIP_SOURCES:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        cmds.append([hipcc(), "-c", *COMMON, *DEVICE_FLAGS, *extra, os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CPP_SOURCES:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        cmds.append(["g++", "-c", *COMMON, "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
                     "-I", "/opt/rocm/include", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # translation units compile concurrently (each is one process)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as pool:
        list(pool.map(run, cmds))
    tmp = LIB + ".tmp"
    run([hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, *objs, "-o", tmp])
    os.replace(tmp, LIB)
    # C++ driver with CLIP's main-loop shape (config.txt, -resume)
    drv_src = os.path.join(CSRC, "clip_main.cpp")
    if os.path.exists(drv_src):
        run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
             "-I", "/opt/rocm/include", drv_src, "-L", HERE, "-lclip_amd", "-L", "/opt/rocm/lib", "-lamdhip64",
             "-Wl,-rpath,$ORIGIN", "-o", DRIVER])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
