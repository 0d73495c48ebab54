"""Ahead-of-time build of libpbf_amd.so for gfx950 with hipcc.

Every call recompiles (the build check must exercise the compiler).  Objects
go to position-based-fluids_amd/_build/, the library to
position-based-fluids_amd/lib/libpbf_amd.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parent.parent          # position-based-fluids_amd/
REPO_ROOT = PKG_ROOT.parent
CSRC = PKG_ROOT / "csrc"
BUILD = PKG_ROOT / "_build"
LIB_DIR = PKG_ROOT / "lib"
LIB_PATH = LIB_DIR / "libpbf_amd.so"
HOST = PKG_ROOT / "host"
SIM_LIB_PATH = LIB_DIR / "libpbf_sim.so"
BIN_DIR = PKG_ROOT / "bin"
RUN_PATH = BIN_DIR / "pbf_run"
TORCH_EXT_PATH = PKG_ROOT / "pbf_amd" / "pbf_torch.so"

ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # IEEE binary32 as written: no implicit FMA contraction, only the fmaf
    # calls in the source (results reproducible and identical to the CPU
    # oracle's, which fuses at the same points)
    "-ffp-contract=off",
    # never emit packed-fp32 VALU ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
    f"-I{REPO_ROOT / 'include'}",
]

SOURCES = ["pbf_sort.hip", "pbf_forward.hip", "pbf_capi.hip", "pbf_adjoint.hip", "pbf_probe.hip", "pbf_noalloc.hip",
           "pbf_params.cpp"]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return exe


def _compile(src: Path, obj: Path) -> None:
    cmd = [hipcc(), *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [hipcc(), *HIP_FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    # surface warnings other than the known host-side packed-fp32 notice
    warn = [ln for ln in r.stderr.splitlines() if "warning" in ln and "packed-fp32" not in ln]
    if warn:
        print("\n".join(warn))


def build(jobs: int | None = None, verbose: bool = True) -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    objs = [BUILD / (Path(s).stem + ".o") for s in SOURCES]
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, CSRC / s, o) for s, o in zip(SOURCES, objs)]
        for f in futs:
            f.result()
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-soname,libpbf_amd.so", *map(str, objs),
           "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB_PATH)
    build_host()
    build_torch_ext()
    if verbose:
        print(f"built {LIB_PATH}, {SIM_LIB_PATH}, {RUN_PATH}, {TORCH_EXT_PATH}")
    return LIB_PATH


def build_host() -> None:
    """C++ Simulation drop-in (host only, g++) over the C ABI, and pbf_run."""
    BIN_DIR.mkdir(parents=True, exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    common = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", f"-I{REPO_ROOT / 'include'}"]
    tmp = SIM_LIB_PATH.with_suffix(".so.tmp")
    cmd = [cxx, *common, "-shared", str(HOST / "Simulation.cpp"), "-o", str(tmp), f"-L{LIB_DIR}", "-lpbf_amd",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, SIM_LIB_PATH)
    tmp = RUN_PATH.with_suffix(".tmp")
    cmd = [cxx, *common, str(HOST / "pbf_run.cpp"), "-o", str(tmp), f"-L{LIB_DIR}", "-lpbf_sim", "-lpbf_amd",
           "-Wl,-rpath,$ORIGIN/../lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"pbf_run build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, RUN_PATH)


def build_torch_ext() -> None:
    """The libtorch autograd Function (csrc/pbf_torch.cpp) as the Python
    module pbf_amd.pbf_torch, linked against libpbf_amd.so (host code only:
    the GPU work is in the library)."""
    import sysconfig

    import torch
    from torch.utils import cpp_extension as ce

    cxx = shutil.which("g++") or "g++"
    tlib = ce.library_paths()[0]
    inc = [f"-I{p}" for p in ce.include_paths()] + [f"-I{sysconfig.get_paths()['include']}",
                                                    f"-I{Path(hipcc()).resolve().parent.parent / 'include'}",
                                                    f"-I{REPO_ROOT / 'include'}"]
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    tmp = TORCH_EXT_PATH.with_suffix(".so.tmp")
    cmd = [cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-w", *inc, "-DTORCH_EXTENSION_NAME=pbf_torch",
           "-DTORCH_API_INCLUDE_EXTENSION_H", f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-D__HIP_PLATFORM_AMD__=1",
           "-DUSE_ROCM=1", str(CSRC / "pbf_torch.cpp"), "-o", str(tmp), f"-L{LIB_DIR}", "-lpbf_amd", f"-L{tlib}",
           "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_python", f"-Wl,-rpath,$ORIGIN/../lib",
           f"-Wl,-rpath,{tlib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"torch extension build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, TORCH_EXT_PATH)


if __name__ == "__main__":
    build()
