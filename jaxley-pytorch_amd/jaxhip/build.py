"""Build `libjxhip.so` (all HIP kernels of the package) for gfx950 with hipcc, and the
host library `libjxhost.so` (schedule construction, `csrc/host/*.cpp`) with g++.

Compiles every `csrc/*.hip` to an object in parallel and links one shared library into
`jaxhip/_lib/`. Called by `__graft_entry__.build()`; a user can also run
`python -m jaxhip.build` from `jaxley-pytorch_amd/`.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
ROOT = PKG_DIR.parent                      # jaxley-pytorch_amd/
CSRC = ROOT / "csrc"
LIB_DIR = PKG_DIR / "_lib"
LIB_PATH = LIB_DIR / "libjxhip.so"
HOST_LIB_PATH = LIB_DIR / "libjxhost.so"
INCLUDE = ROOT.parent / "include"
ARCH = os.environ.get("JAXHIP_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # Packed FP32 instructions stay off (they are generated from ordinary float code).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast",
    # Per-thread item arrays indexed by a loop counter stay in VGPRs (indexed moves)
    # instead of scratch memory.
    "-mllvm", "-amdgpu-promote-alloca-to-vector-max-regs=128",
    "-mllvm", "-amdgpu-promote-alloca-to-vector-vgpr-ratio=1",
    f"-I{INCLUDE}",
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the jaxhip kernels need ROCm's hipcc")


def build(verbose: bool = False, jobs: int = 8) -> Path:
    """Compile all sources and link `libjxhip.so`. Always recompiles."""
    hipcc = _hipcc()
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources in {CSRC}")
    obj_dir = ROOT / "_build"
    obj_dir.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)

    def compile_one(src: Path) -> Path:
        obj = obj_dir / (src.stem + ".o")
        cmd = [hipcc, *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
        return obj

    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(sources)))) as ex:
        objs = list(ex.map(compile_one, sources))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    build_host(verbose)
    return LIB_PATH


def build_host(verbose: bool = False) -> Path:
    """Compile the host library `libjxhost.so` (schedule construction; plain C++, no
    HIP, so that request-preparation processes can load it without the GPU runtime)."""
    sources = sorted((CSRC / "host").glob("*.cpp"))
    LIB_DIR.mkdir(exist_ok=True)
    tmp = HOST_LIB_PATH.with_suffix(".so.tmp")
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        raise RuntimeError("g++ not found; the jaxhip host library needs a C++ compiler")
    cmd = [cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{INCLUDE}", *map(str, sources),
           "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"host library build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, HOST_LIB_PATH)
    return HOST_LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
