"""Compile the native parts of epseon_backend_amd in-tree.

Artefacts (all git-ignored, all shipped with the source tree):

* ``epseon_backend_amd/lib/libepseon_hip.so`` -- gfx950 kernels + the C ABI of
  ``include/epseon_hip.h`` (hipcc);
* ``epseon_backend_amd/device/gpu/_libepseon_gpu<ext>`` -- the pybind11 module
  with the reference's Python surface, linked against libepseon_hip.so;
* ``epseon_backend_amd/device/cpu/_libepseon_cpu<ext>`` -- the CPU device
  module (``greet``), as in the reference.

Run ``python -m epseon_backend_amd._build`` or call :func:`build_all`.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
import sysconfig
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
LIB_DIR = PKG / "lib"
HIP_LIB = LIB_DIR / "libepseon_hip.so"
EXT_SUFFIX = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
GPU_MODULE = PKG / "device" / "gpu" / f"_libepseon_gpu{EXT_SUFFIX}"
CPU_MODULE = PKG / "device" / "cpu" / f"_libepseon_cpu{EXT_SUFFIX}"

OFFLOAD_ARCH = os.environ.get("EPSEON_OFFLOAD_ARCH", "gfx950")

# Device code: FP64 results must be bit-identical to the C oracle, so no
# contraction anywhere (every fused op is an explicit fma) and no packed FP32.
HIPCC_FLAGS = [
    "-O3",
    "-std=c++20",
    "-fPIC",
    f"--offload-arch={OFFLOAD_ARCH}",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    # Schedule for instruction-level parallelism: the heavy single-potential
    # scan runs one wavefront per SIMD, so FP64 latency must be hidden inside
    # the wavefront (dependent ops >= 4 instructions apart).
    "-mllvm",
    "-amdgpu-sched-strategy=max-ilp",
    "-Wall",
]


def _hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = Path(rocm) / "bin" / "hipcc"
    return str(cand) if cand.exists() else (shutil.which("hipcc") or "hipcc")


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, check=False)
    if proc.returncode != 0:
        msg = f"command failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}"
        raise RuntimeError(msg)


def _pybind_include() -> str:
    import pybind11

    return pybind11.get_include()


def build_hip_library() -> Path:
    """hipcc: kernels + C ABI -> lib/libepseon_hip.so."""
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    tmp = HIP_LIB.with_suffix(".so.tmp")
    srcs = [CSRC / "numerov_kernels.hip", CSRC / "eigenfunction_kernels.hip", CSRC / "epseon_capi.hip"]
    cmd = [
        _hipcc(),
        *HIPCC_FLAGS,
        "-shared",
        f"-I{REPO / 'include'}",
        f"-I{CSRC}",
        *map(str, srcs),
        "-o",
        str(tmp),
    ]
    _run(cmd)
    tmp.replace(HIP_LIB)
    return HIP_LIB


def build_gpu_module() -> Path:
    """g++: pybind11 module _libepseon_gpu (host only), linked to libepseon_hip.so."""
    GPU_MODULE.parent.mkdir(parents=True, exist_ok=True)
    tmp = GPU_MODULE.with_name(GPU_MODULE.name + ".tmp")
    py_inc = sysconfig.get_paths()["include"]
    cmd = [
        "g++",
        "-O2",
        "-std=c++20",
        "-fPIC",
        "-shared",
        "-fvisibility=hidden",
        "-ffp-contract=off",
        f"-I{py_inc}",
        f"-I{_pybind_include()}",
        f"-I{REPO / 'include'}",
        f"-I{CSRC}",
        str(CSRC / "gpu_module.cpp"),
        f"-L{LIB_DIR}",
        "-lepseon_hip",
        "-Wl,-rpath,$ORIGIN/../../lib",
        "-pthread",
        "-o",
        str(tmp),
    ]
    _run(cmd)
    tmp.replace(GPU_MODULE)
    return GPU_MODULE


def build_cpu_module() -> Path:
    """g++: CPython module _libepseon_cpu."""
    CPU_MODULE.parent.mkdir(parents=True, exist_ok=True)
    tmp = CPU_MODULE.with_name(CPU_MODULE.name + ".tmp")
    py_inc = sysconfig.get_paths()["include"]
    cmd = [
        "g++",
        "-O2",
        "-std=c++17",
        "-fPIC",
        "-shared",
        f"-I{py_inc}",
        str(CSRC / "cpu_module.cpp"),
        "-o",
        str(tmp),
    ]
    _run(cmd)
    tmp.replace(CPU_MODULE)
    return CPU_MODULE


def build_all(verbose: bool = True) -> list[Path]:
    """Build every native artefact (the HIP library first: the GPU module links it)."""
    out: list[Path] = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=2) as pool:
        hip = pool.submit(build_hip_library)
        cpu = pool.submit(build_cpu_module)
        out.append(hip.result())
        out.append(cpu.result())
    out.append(build_gpu_module())
    if verbose:
        for p in out:
            sys.stdout.write(f"built {p.relative_to(REPO)}\n")
    return out


if __name__ == "__main__":
    build_all()
