"""Compile the HIP sources of eggpu_amd into ``eggpu_amd/libegx.so`` (gfx950)
and the host extension ``eggpu_amd/_egx_host*.so`` (gcc, CPython C API).

Run ``python -m eggpu_amd.build``.  Every call recompiles; objects go to
``eggpu_amd/_build/`` and the libraries are written in-tree, beside the
package that loads them.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD_DIR = os.path.join(HERE, "_build")
LIB_PATH = os.path.join(HERE, "libegx.so")
ARCH = os.environ.get("EGX_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    "egx_prim.hip",
    "egx_build.hip",
    "egx_gen.hip",
    "egx_pagerank.hip",
    "egx_spmm.hip",
    "egx_katz.hip",
    "egx_eigen.hip",
    "egx_api.hip",
    "egx_serve.hip",
    "egx_ingest.hip",
    "egx_noalloc.hip",
]

# -packed-fp32-ops: no packed FP32 instructions are generated.
COMMON_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unknown-warning-option",
]


def _hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; ROCm is required to build eggpu_amd")
    return exe


def _compile(src):
    obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), *COMMON_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def build_host(verbose=False):
    """The CPython extension that reads dict adjacencies into CSR arrays
    (csrc/egx_host.c), built with the host C compiler beside the package."""
    import sysconfig

    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        raise RuntimeError("a C compiler (gcc) is needed for eggpu_amd's host extension")
    out = os.path.join(HERE, "_egx_host" + sysconfig.get_config_var("EXT_SUFFIX"))
    tmp = out + ".tmp"
    cmd = [cc, "-O2", "-Wall", "-Wextra", "-shared", "-fPIC", "-pthread", "-I", sysconfig.get_paths()["include"],
           os.path.join(CSRC, "egx_host.c"), "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"host extension build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


def build(verbose=False):
    build_host(verbose)
    os.makedirs(BUILD_DIR, exist_ok=True)
    workers = min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB_PATH + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH}")
    return LIB_PATH


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
