"""Build the native libraries in-tree (no cmake, plain compiler commands).

Outputs
  cuda_hlt_amd/lib/libvelo_hlt.so   HIP kernels for gfx950 + C-ABI launchers +
                                    the C++ host layer (gear, stream, visitors)
  cuda_hlt_amd/lib/libvelo_gen.so   synthetic event generator (C ABI)
  oracle/_build/libvelo_oracle.so   CPU oracle (test infrastructure)
  oracle/_ref/libref_sp.so          the reference's own SP-pattern builder,
                                    compiled from /root/reference when present
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
LIB = PKG / "lib"
ORACLE = ROOT / "oracle"
INCLUDE = ROOT / "include"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = os.environ.get("VELO_OFFLOAD_ARCH", "gfx950")

# Kernel flags: strict IEEE float (no FMA contraction, no denormal flush) so
# that float results equal the CPU oracle bit for bit; packed-fp32 disabled.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off", "-Wall",
              "-D__HIP_PLATFORM_AMD__"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _hipcc() -> str:
    p = ROCM / "bin" / "hipcc"
    return str(p) if p.exists() else "hipcc"


def kernel_sources() -> list[Path]:
    return sorted((CSRC / "kernels").glob("*.hip"))


def host_sources() -> list[Path]:
    return sorted((CSRC / "host").rglob("*.cpp"))


def build_velo_hlt(jobs: int = 8) -> Path:
    out_dir = PKG / "_build" / "obj"
    out_dir.mkdir(parents=True, exist_ok=True)
    LIB.mkdir(exist_ok=True)
    incs = [f"-I{INCLUDE}", f"-I{CSRC / 'common'}", f"-I{CSRC / 'host'}", f"-I{ROCM / 'include'}"]
    cmds, objs = [], []
    for src in kernel_sources():
        obj = out_dir / (src.stem + ".hip.o")
        cmds.append([_hipcc(), *HIP_FLAGS, *incs, "-c", str(src), "-o", str(obj)])
        objs.append(obj)
    for src in host_sources():
        obj = out_dir / (src.stem + ".cpp.o")
        cmds.append(["g++", *HOST_FLAGS, *incs, "-c", str(src), "-o", str(obj)])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    target = LIB / "libvelo_hlt.so"
    tmp = LIB / "libvelo_hlt.so.tmp"
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs), f"-L{ROCM / 'lib'}", "-lamdhip64", "-lz",
          f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, target)
    return target


def build_generator() -> Path:
    LIB.mkdir(exist_ok=True)
    target = LIB / "libvelo_gen.so"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", f"-I{INCLUDE}",
          str(CSRC / "gen" / "velo_gen.cpp"), "-o", str(target)])
    return target


def build_oracle() -> Path:
    d = ORACLE / "_build"
    d.mkdir(exist_ok=True)
    target = d / "libvelo_oracle.so"
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off",
          "-fno-fast-math", str(ORACLE / "velo_oracle.c"), "-o", str(target), "-lm"])
    return target


def build_reference_sp(reference: Path = Path("/root/reference")) -> Path | None:
    """Compile the reference's own cache_sp_patterns (x86/velo/clustering/src/
    ClusteringCommon.cpp) with a small C-ABI wrapper into oracle/_ref/."""
    src = reference / "x86" / "velo" / "clustering" / "src" / "ClusteringCommon.cpp"
    if not src.exists():
        return None
    d = ORACLE / "_ref"
    d.mkdir(exist_ok=True)
    target = d / "libref_sp.so"
    _run(["g++", "-O2", "-std=c++14", "-fPIC", "-shared", "-ffp-contract=off",
          str(src), str(ORACLE / "ref_sp_wrapper.cpp"), "-o", str(target)])
    return target


def build_all(verbose: bool = True) -> None:
    steps = [("generator", build_generator), ("oracle", build_oracle),
             ("reference SP tables", build_reference_sp), ("velo_hlt", build_velo_hlt)]
    for name, fn in steps:
        out = fn()
        if verbose:
            print(f"[build] {name}: {out}", flush=True)


if __name__ == "__main__":
    build_all()
    sys.exit(0)
