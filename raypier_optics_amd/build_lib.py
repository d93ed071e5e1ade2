"""Compile the gfx950 HIP library (librpx.so) in-tree.

Usage: python -m raypier_optics_amd.build_lib   (or build() from Python)

Every HIP source is compiled with packed-FP32 instructions disabled.  The
float64 sources use -ffp-contract=off so that FP64 expressions round as
written (the FP64 tracer must reproduce the reference's topology bit for
bit).  The FP32-split source is not bound to the reference's rounding: it
allows fused multiply-adds and the faster float32 divide / square root.
"""
from concurrent.futures import ThreadPoolExecutor
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librpx.so")
SOURCES = ["rpx_trace.hip", "rpx_query.hip", "rpx_fields.hip", "rpx_split.hip", "rpx_serve.hip", "rpx_adjoint.hip",
           "rpx_select.hip", "rpx_lineage.hip",
           "rpx_sources.hip"]
# per-source flags that replace -ffp-contract=off
SOURCE_FLAGS = {
    "rpx_split.hip": ["-ffp-contract=fast", "-fno-hip-fp32-correctly-rounded-divide-sqrt"],
}
ARCH = os.environ.get("RPX_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: cannot build librpx.so")


def flags(src=None):
    return [
        "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    ] + SOURCE_FLAGS.get(src, ["-ffp-contract=off"]) + [
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(HERE, "..", "include"),
    ]


def build(verbose=False):
    cc = hipcc()
    build_dir = os.path.join(HERE, "_objs")
    os.makedirs(build_dir, exist_ok=True)
    objs = [os.path.join(build_dir, src.replace(".hip", ".o")) for src in SOURCES]

    def compile_one(k):
        cmd = [cc] + flags(SOURCES[k]) + ["-c", os.path.join(CSRC, SOURCES[k]), "-o", objs[k]]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    with ThreadPoolExecutor(max_workers=min(6, len(SOURCES))) as pool:
        list(pool.map(compile_one, range(len(SOURCES))))
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
