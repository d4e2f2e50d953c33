"""Builds the native parts in-tree:

* pqdts_amd/libpqdts_amd.so  -- HIP kernels + C ABI, hipcc for gfx950
* pqdts_amd/bin/pqdts_omp.x  -- the Fortran drop-in driver (amdflang), linked
                                 against libpqdts_amd.so through iso_c_binding
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FC = os.path.join(ROCM, "bin", "amdflang")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", "-shared",
    # no packed-FP32 instructions in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, **kw):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, **kw)
    lines = [ln for ln in r.stdout.splitlines() if "packed-fp32-ops" not in ln]
    if r.returncode != 0:
        sys.stderr.write("\n".join(lines) + "\n")
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    return "\n".join(lines)


def build_lib() -> str:
    src = os.path.join(HERE, "csrc", "pqdts_amd.hip")
    out = os.path.join(HERE, "libpqdts_amd.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, src])
    os.replace(tmp, out)
    return out


def build_driver() -> str | None:
    src = os.path.join(HERE, "driver", "pqdts_main.f90")
    if not os.path.exists(src) or not os.path.exists(FC):
        return None
    bindir = os.path.join(HERE, "bin")
    moddir = os.path.join(bindir, "mod")
    os.makedirs(moddir, exist_ok=True)
    out = os.path.join(bindir, "pqdts_omp.x")
    _run([FC, "-O2", "-module-dir", moddir, "-o", out, src,
          "-L", HERE, "-lpqdts_amd", "-Wl,-rpath,$ORIGIN/..", "-lstdc++"])
    shutil.rmtree(moddir, ignore_errors=True)
    return out


def build_all() -> None:
    build_lib()
    build_driver()


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(HERE, "libpqdts_amd.so"))
