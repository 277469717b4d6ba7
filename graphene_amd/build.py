"""Builds libgraphene_amd.so (gfx950) in-tree with hipcc.

Every translation unit is compiled with hipcc for gfx950 and linked into
graphene_amd/libgraphene_amd.so. Usage: python -m graphene_amd.build [-j N]
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgraphene_amd.so")
CLI = os.path.join(HERE, "benchmark")           # drop-in benchmark tool
CLI_SOURCE = "benchmark_main.cpp"

SOURCES = [
    "host_matrix.cpp",
    "config.cpp",
    "device_matrix.cpp",
    "solvers.cpp",
    "capi.cpp",
    "server.cpp",
    "batch.cpp",
    "ilu_pack.cpp",
    "gs_pack.cpp",
    "k_spmv.hip",
    "k_vec.hip",
    "k_precond.hip",
    "k_ilu_packed.hip",
    "k_ilu_level.hip",
    "k_ilu_rows.hip",
    "k_batch.hip",
    "k_noalloc.hip",
    "k_multi.hip",
    "k_ilu_block.hip",
    "k_gs.hip",
    "k_dword.hip",
]

COMMON = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # bit-reproducible sums: no FMA contraction
    "-ffp-contract=off",
    # no packed FP32 instructions (v_pk_*_f32): some of their forms have given
    # wrong results on gfx950 parts
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [hipcc()] + COMMON + ["-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                     if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{errs}")
    return obj, errs


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, len(SOURCES))
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, errs in ex.map(_compile, SOURCES):
            objs.append(obj)
            if verbose and errs.strip():
                print(errs, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    # the drop-in `benchmark` command-line tool, linked against the library
    obj, _ = _compile(CLI_SOURCE)
    tmp = CLI + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-o", tmp, obj, "-L", HERE, "-lgraphene_amd",
           "-Wl,-rpath,$ORIGIN"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed (benchmark)\n" + p.stdout + p.stderr)
    os.replace(tmp, CLI)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.v))
