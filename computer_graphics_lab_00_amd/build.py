"""Build the gfx950 HIP library in-tree (``computer_graphics_lab_00_amd/libsph_hip.so``).

Every source is compiled by hipcc for gfx950 only. Floating-point flags matter for
parity with the reference's CPU path:
  -ffp-contract=off                 no a*b+c fusion (the reference's x86 build has none)
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE division and square root
  -fno-fast-math, denormals kept    (no -fgpu-flush-denormals-to-zero)
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
LIB = PKG / "libsph_hip.so"

SOURCES = [
    "sph_host.cpp",
    "sph_mc_table.cpp",
    "sph_grid.hip",
    "sph_forces.hip",
    "sph_engine.hip",
    "sph_api.hip",
    "sph_mlp.hip",
    "sph_mlp_fused.hip",
    "sph_surrogate_kernels.hip",
    "sph_surrogate.hip",
    "sph_surface.hip",
    "sph_output.hip",
    "sph_noalloc.hip",
    "sph_arith_check.hip",
]

COMMON_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Wall",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]

DEVICE_FLAGS = [
    "--offload-arch=gfx950",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libsph_hip.so")
    return exe


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build(verbose: bool = False) -> Path:
    """Compile every source for gfx950 and link the shared library. Always rebuilds."""
    obj_dir = PKG / "_build"
    obj_dir.mkdir(exist_ok=True)
    objects = []
    cc = hipcc()
    for src in SOURCES:
        path = CSRC / src
        obj = obj_dir / (path.stem + ".o")
        cmd = [cc, *COMMON_FLAGS]
        if path.suffix == ".hip":
            cmd += ["-x", "hip", *DEVICE_FLAGS]
        cmd += ["-c", str(path), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objects.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    link = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objects, "-lamdhip64"]
    if verbose:
        print(" ".join(link))
    _run(link)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
