"""Build the gfx950 knn_rms engine in-tree.

Produces
  mdsctk_amd/lib/libmdsctk_knn.so   C ABI (include/mdsctk_knn.h, include/mdsctk_io.h)
  mdsctk_amd/bin/knn_rms            drop-in for the reference knn_rms executable

with hipcc for --offload-arch=gfx950.  Packed FP32 instructions are disabled
for all device code (-target-feature -packed-fp32-ops).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
BIN_DIR = os.path.join(HERE, "bin")
OBJ_DIR = os.path.join(ROOT, "_build", "obj")
LIB_PATH = os.path.join(LIB_DIR, "libmdsctk_knn.so")
CLI_PATH = os.path.join(BIN_DIR, "knn_rms")
MAKE_SYSPARSE_PATH = os.path.join(BIN_DIR, "make_sysparse")
DECOMP_SPARSE_PATH = os.path.join(BIN_DIR, "decomp_sparse")
MAKE_GESPARSE_PATH = os.path.join(BIN_DIR, "make_gesparse")
KNN_DATA_PATH = os.path.join(BIN_DIR, "knn_data")
DECOMP_SPARSE_NYSTROM_PATH = os.path.join(BIN_DIR, "decomp_sparse_nystrom")

ARCH = os.environ.get("MDSCTK_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

DEVICE_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
                "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
          f"-I{os.path.join(ROOT, 'include')}"]

LIB_SOURCES = ["knn_kernels.hip", "xtc_kernels.hip", "api.cpp", "serve.cpp", "xtc.cpp",
               "xtc_device.cpp", "topology.cpp", "spectral.cpp"]
# executables: (name, main source)
TOOLS = [("knn_rms", "knn_rms_main.cpp"), ("make_sysparse", "make_sysparse_main.cpp"),
         ("decomp_sparse", "decomp_sparse_main.cpp"), ("make_gesparse", "make_gesparse_main.cpp"),
         ("decomp_sparse_nystrom", "decomp_sparse_nystrom_main.cpp"),
         ("knn_data", "knn_data_main.cpp")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that note is harmless.
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def _compile(src, verbose):
    os.makedirs(OBJ_DIR, exist_ok=True)
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + DEVICE_FLAGS + COMMON + lang + ["-c", src, "-o", obj]
    _run(cmd, verbose)
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every source and link the library and the CLI.  Returns the library path."""
    srcs = [os.path.join(CSRC, s) for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(len(srcs), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    _run([HIPCC] + DEVICE_FLAGS + ["-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB_PATH)
    os.makedirs(BIN_DIR, exist_ok=True)

    def link_tool(tool):
        name, main = tool
        exe = os.path.join(BIN_DIR, name)
        tmp = exe + ".tmp"
        _run([HIPCC] + COMMON + [os.path.join(CSRC, main), "-o", tmp, f"-L{LIB_DIR}",
                                  "-lmdsctk_knn", "-Wl,-rpath,$ORIGIN/../lib"], verbose)
        os.replace(tmp, exe)

    with cf.ThreadPoolExecutor(max_workers=len(TOOLS)) as ex:
        list(ex.map(link_tool, TOOLS))
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
