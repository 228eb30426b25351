"""Builds lc-backend-hip.so (gfx950) in-tree with hipcc.

Usage: python -m luisacompute_amd.build [--force]
Objects go to luisacompute_amd/_build/, the library to luisacompute_amd/lc-backend-hip.so.
"""
import argparse
import concurrent.futures
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT_DIR = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lc-backend-hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LC_HIP_ARCH", "gfx950")

# -ffp-contract=off: the watertight test must round exactly like the CPU oracle (no FMA
# contraction); box tests call __builtin_fmaf explicitly where fusion is wanted.
# -fgpu-flush-denormals-to-zero: f32 denormals flushed, like the reference CPU backend's
# FTZ/DAZ worker threads (src/rust/luisa_compute_backend_impl/src/cpu/mod.rs:503-516).
# -target-feature -packed-fp32-ops: no v_pk_*_f32 generation.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-fgpu-flush-denormals-to-zero",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]

SOURCES = ["bvh_build.hip", "trace.hip", "raysort.hip", "pathtracer.hip", "instances.hip", "query.hip", "serve.cpp", "runtime.cpp",
           "workload.cpp"]
HEADERS = ["lc_hip_bvh.h", "lc_hip_internal.h", "traverse.h", "ray_query.h"]


def _newer(src, dst):
    return not os.path.exists(dst) or os.path.getmtime(src) > os.path.getmtime(dst)


def _compile(src, force):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OUT_DIR, src + ".o")
    deps = [path] + [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(ROOT, "include", "lc_hip.h")]
    if not force and not any(_newer(d, obj) for d in deps):
        return obj
    cmd = [HIPCC] + COMMON + ["-c", path, "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = HIP_FLAGS
    subprocess.run(cmd, check=True)
    return obj


def build(force=True):
    os.makedirs(OUT_DIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), SOURCES))
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs + ["-lpthread"]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    a = ap.parse_args()
    print(build(force=a.force))
    sys.exit(0)
