"""Build the gfx950 HIP library in-tree: vapor_amd/_lib/libvapor_mac3d.so.

    python -m vapor_amd.build

Compiles every source under vapor_amd/csrc with hipcc for gfx950. Float
contraction is off (-ffp-contract=off) so each kernel rounds exactly as the
reference's CPU code does, and packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
CSRC = PACKAGE_DIR / "csrc"
LIB_DIR = PACKAGE_DIR / "_lib"
LIBRARY = LIB_DIR / "libvapor_mac3d.so"
SOURCES = ["vgx_kernels.hip", "vgx_api.hip", "vgx_batch.hip", "vgx_shadow.hip", "vgx_unet_kernels.hip", "vgx_unet.hip",
           "vgx2_kernels.hip", "vgx2_api.hip", "vsx_kernels.hip", "vsx_api.hip"]
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for candidate in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if candidate and Path(candidate).exists():
            return candidate
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, jobs: int = 6) -> Path:
    LIB_DIR.mkdir(exist_ok=True)
    commands = []
    objects = []
    for name in SOURCES:
        src = CSRC / name
        obj = LIB_DIR / (src.stem + ".o")
        commands.append([hipcc(), *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)])
        objects.append(str(obj))
    # The translation units are independent: compile them side by side.
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
        def run(cmd):
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)
        for future in [pool.submit(run, cmd) for cmd in commands]:
            future.result()
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objects, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIBRARY)
    for obj in objects:
        Path(obj).unlink()
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
