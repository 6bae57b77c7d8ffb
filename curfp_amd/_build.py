"""Build the gfx950 library and its Python extension in-tree.

Produces, next to this file:
  libcurfp_amd.so   — the C ABI of include/curfp.h (hipcc, --offload-arch=gfx950)
  _curfp_C*.so      — the pybind11 extension over that ABI (g++)

Run as ``python curfp_amd/_build.py`` (a plain script: it must not import the
package, whose import needs the built extension) or through
``__graft_entry__.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
BUILD = os.path.join(HERE, "_objs")

HIP_SOURCES = ["blas.hip", "gemm.hip", "gemm_bf16.hip", "gemm_bf16_dma.hip", "potrf.hip", "pack.hip", "refine.hip", "trsv_block.hip", "trsm_panel.hip", "trsm_panel64.hip", "lansf.hip", "pfcon.hip", "rfp_drivers.hip"]
LIB = os.path.join(HERE, "libcurfp_amd.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE,
]


# Sources whose MFMA kernels keep their accumulators in VGPRs.  By default
# the compiler places MFMA accumulators in AGPRs and, for accumulators that
# live across a loop with branches in it (the guarded slice loads), copies
# all of them AGPR -> VGPR -> AGPR on every iteration: the copies wait for the
# MFMA pipeline to drain and the doubled registers cut occupancy (the FP64
# 128 x 128 view GEMM: 336 registers per lane, 1 wave per SIMD; 218 and 2
# waves in VGPR form).  Kernels whose accumulators exceed the 256
# architectural VGPRs (the 256 x 256 BF16 tiles) keep the default.
VGPR_FORM_SOURCES = {"gemm.hip", "potrf.hip", "trsm_panel.hip", "trsm_panel64.hip", "trsv_block.hip"}


def ext_path() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "_curfp_C" + suffix)


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in res.stdout.splitlines() if "packed-fp32-ops" not in line
    )
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{out}")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    extra = ["-mllvm", "-amdgpu-mfma-vgpr-form"] if src in VGPR_FORM_SOURCES else []
    _run([os.path.join(ROCM, "bin", "hipcc"), *HIPCC_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build(verbose: bool = True) -> None:
    os.makedirs(BUILD, exist_ok=True)
    workers = min(len(HIP_SOURCES), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, HIP_SOURCES))
    tmp = LIB + ".tmp"
    _run([os.path.join(ROCM, "bin", "hipcc"), "--offload-arch=gfx950", "-shared", "-fPIC", *objs,
          "-L" + os.path.join(ROCM, "lib"), "-lhipblaslt", "-lrocblas", "-o", tmp])
    os.replace(tmp, LIB)
    if verbose:
        print(f"[curfp_amd] built {LIB}")

    import pybind11  # noqa: PLC0415

    py_inc = sysconfig.get_paths()["include"]
    ext = ext_path()
    tmp = ext + ".tmp"
    _run([
        "g++", "-O2", "-shared", "-fPIC", "-std=c++17",
        "-D__HIP_PLATFORM_AMD__",
        "-I" + pybind11.get_include(), "-I" + py_inc, "-I" + os.path.join(ROCM, "include"), "-I" + INCLUDE,
        os.path.join(CSRC, "pyext.cpp"),
        "-L" + HERE, "-lcurfp_amd", "-Wl,-rpath,$ORIGIN",
        "-o", tmp,
    ])
    os.replace(tmp, ext)
    if verbose:
        print(f"[curfp_amd] built {ext}")


if __name__ == "__main__":
    build()
    sys.exit(0)
