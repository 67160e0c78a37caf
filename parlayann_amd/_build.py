"""Builds parlayann_amd/libpann.so in-tree with hipcc for gfx950.

Every HIP source is compiled with packed-FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops); the host pass warns
that it does not know that feature, which is harmless.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpann.so")
# the beam-search kernels are spread over one translation unit per store (bs_*.hip)
SOURCES = ["beam_search.hip", "bs_f32.hip", "bs_f32_inorder.hip", "bs_u8.hip", "bs_i8.hip",
           "bs_bf16.hip", "bs_fp8.hip", "bs_fp8r.hip", "quantize.hip", "prune.hip", "hnsw.hip",
           "bruteforce.hip", "pann_capi.cpp"]
# the device's random row-gather ceiling, for bench.py's roofline (not part of libpann)
CALIB_SRC = "calib.hip"
CALIB_LIB = os.path.join(HERE, "libpann_calib.so")
ARCH = os.environ.get("PANN_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = SOURCES + [CALIB_SRC]
    with ThreadPoolExecutor(max_workers=min(len(srcs), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, srcs))
    for lib, o, libs in ((LIB, objs[:-1], []),
                         (CALIB_LIB, objs[-1:], [])):
        tmp = lib + ".tmp"
        cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *o, *libs]
        subprocess.run(cmd, check=True, capture_output=not verbose)
        os.replace(tmp, lib)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
