"""Build the native library (libsing_hip.so) and the hip_cli drop-in.

Sources are compiled in-tree for gfx950 with hipcc (device code) and g++
(host-only code); outputs land next to this file so they travel with the
source tree:

    singularity_amd/libsing_hip.so   the C-ABI library (include/sing_abi.h)
    singularity_amd/hip_cli          flag-compatible replacement of cuda_cli
    singularity_amd/hip_diag         the CPU driver's kerr-geometry / photon-orbit / benchmark modes
    singularity_amd/hip_backend_check  the HipBackend : RenderBackend adapter (include/sing_hip_backend.hpp)
                                     built against the reference's render_backend.hpp, where that tree exists

Usage: python -m singularity_amd.build  (or import and call build()).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libsing_hip.so"
CLI = PKG / "hip_cli"
DIAG = PKG / "hip_diag"
ADAPTER = PKG / "hip_backend_check"  # include/sing_hip_backend.hpp driven through the reference's RenderBackend

ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    # packed FP32 instructions stay disabled on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]
CXX_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-Wall"]

HIP_SOURCES = ["trace.hip", "post.hip", "context.hip", "adjoint.hip", "batch.hip", "diag.hip", "noalloc.hip"]
CXX_SOURCES = ["host_scene.cpp"]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build singularity_amd")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        msg = "\n".join(
            line for line in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in line
        )
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{msg}")


def build(verbose: bool = False) -> Path:
    """Compile every source and link libsing_hip.so and hip_cli; returns the library path."""
    hipcc = _hipcc()
    OBJ.mkdir(exist_ok=True)
    jobs: list[list[str]] = []
    objs: list[Path] = []
    for src in HIP_SOURCES:
        path = CSRC / src
        if not path.exists():
            continue
        obj = OBJ / (src + ".o")
        objs.append(obj)
        jobs.append([hipcc, *HIP_FLAGS, "-c", str(path), "-o", str(obj)])
    for src in CXX_SOURCES:
        obj = OBJ / (src + ".o")
        objs.append(obj)
        jobs.append(["g++", *CXX_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(workers) as ex:
        for cmd in jobs:
            if verbose:
                print(" ".join(cmd), flush=True)
        list(ex.map(_run, jobs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    cli_src = CSRC / "hip_cli.cpp"
    if cli_src.exists():
        tmpc = CLI.with_suffix(".tmp")
        _run(
            [
                "g++", "-O2", "-std=c++20", "-Wall", "-o", str(tmpc), str(cli_src),
                f"-L{PKG}", "-lsing_hip", "-Wl,-rpath,$ORIGIN",
                "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib",
            ]
        )
        os.replace(tmpc, CLI)
    diag_src = CSRC / "hip_diag.cpp"
    if diag_src.exists():
        tmpd = DIAG.with_suffix(".tmp")
        _run(
            [
                "g++", "-O2", "-std=c++20", "-Wall", "-o", str(tmpd), str(diag_src),
                f"-L{PKG}", "-lsing_hip", "-Wl,-rpath,$ORIGIN",
                "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib",
            ]
        )
        os.replace(tmpd, DIAG)
    # the RenderBackend adapter (include/sing_hip_backend.hpp) compiled against the
    # reference's own interface header, where the reference tree is present
    ref_hdr = Path("/root/reference/render/include")
    if (ref_hdr / "render_backend.hpp").exists():
        tmpa = ADAPTER.with_suffix(".tmp")
        _run(
            [
                "g++", "-O2", "-std=c++20", "-Wall", "-Wextra", "-o", str(tmpa), str(ROOT / "tests" / "hip_backend_check.cpp"),
                f"-I{ref_hdr}", f"-I{ROOT / 'include'}",
                f"-L{PKG}", "-lsing_hip", "-Wl,-rpath,$ORIGIN",
                "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib",
            ]
        )
        os.replace(tmpa, ADAPTER)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(f"built {LIB}")
