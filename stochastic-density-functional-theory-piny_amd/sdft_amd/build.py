"""Build libsdft_hip.so (gfx950) and libsdft_host.so in-tree with hipcc / gcc.

Every HIP source is compiled with packed-FP32 instruction selection disabled
(`-Xclang -target-feature -Xclang -packed-fp32-ops`).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(PKG_DIR, "csrc")
HOST = os.path.join(PKG_DIR, "host")
LIB_DIR = os.path.join(PKG_DIR, "lib")
OBJ_DIR = os.path.join(PKG_DIR, "_build")
REPO = os.path.dirname(PKG_DIR)
INCLUDE = os.path.join(REPO, "include")

HIP_SOURCES = ["sdft_kernels.hip", "sdft_gemm.hip", "sdft_cheby.hip", "sdft_api.hip",
               "sdft_batch.hip", "sdft_ks.hip", "sdft_rs.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}"]


def lib_path() -> str:
    return os.path.join(LIB_DIR, "libsdft_hip.so")


def host_lib_path() -> str:
    return os.path.join(LIB_DIR, "libsdft_host.so")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build(verbose: bool = False) -> str:
    """Compile all HIP sources (always; no staleness check) and link."""
    os.makedirs(LIB_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(HIP_SOURCES), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, HIP_SOURCES))
    out = lib_path()
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    host_src = [os.path.join(HOST, f) for f in sorted(os.listdir(HOST)) if f.endswith(".c")] \
        if os.path.isdir(HOST) else []
    if host_src:
        htmp = host_lib_path() + ".tmp"
        # exports limited to the sdft_* entry points (version script), the
        # library's own calls bound inside it (-Bsymbolic-functions)
        _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c99", "-Wall", f"-I{INCLUDE}", "-o", htmp]
             + host_src + [f"-L{LIB_DIR}", "-lsdft_hip", "-Wl,-rpath,$ORIGIN",
                           "-Wl,-Bsymbolic-functions",
                           "-Wl,--version-script=" + os.path.join(HOST, "libsdft_host.map")])
        os.replace(htmp, host_lib_path())
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
