"""Build the engine's shared library for gfx950 with hipcc, in-tree.

Output: winds-master_amd/lib/libwinds_bs.so (git-ignored, shipped with the
tree).  Every HIP translation unit is compiled with packed FP32 disabled and
FP contraction off: the FP64 kernels must round exactly as written, and the
FP32 kernels spell out every fused multiply-add they use.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parents[1]          # winds-master_amd/
REPO_ROOT = PKG_ROOT.parent
CSRC = PKG_ROOT / "csrc"
LIBDIR = PKG_ROOT / "lib"
OBJDIR = PKG_ROOT / "_build"
LIB = LIBDIR / "libwinds_bs.so"

SOURCES = [
    "bs_kernels.hip",
    "bs_api.cpp",
    "bs_batch.hip",
    "march.hip",
    "mlp.hip",
    "surrogate.hip",
    "ground.hip",
]

ARCH = "gfx950"


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def common_flags() -> list[str]:
    return [
        "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        f"-I{REPO_ROOT / 'include'}", f"-I{CSRC}",
        "-Wall", "-Wno-unused-function",
    ]


def _quiet(stderr: str) -> str:
    # The host pass does not know the device-only target feature; that
    # warning is expected and harmless.
    keep = [ln for ln in stderr.splitlines()
            if "packed-fp32-ops" not in ln and ln.strip()]
    return "\n".join(keep)


def build(verbose: bool = False) -> Path:
    OBJDIR.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = OBJDIR / (Path(src).stem + ".o")
        objs.append(obj)
        cmd = [hipcc(), *common_flags(), "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
    failed = []
    for src, p in procs:
        out, err = p.communicate()
        msg = _quiet(err)
        if p.returncode != 0:
            failed.append((src, msg))
        elif msg and verbose:
            print(msg, file=sys.stderr)
    if failed:
        raise RuntimeError("hipcc failed:\n" + "\n".join(f"[{s}]\n{m}" for s, m in failed))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *[str(o) for o in objs]]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + _quiet(r.stderr))
    os.replace(tmp, LIB)
    build_mex_mock(verbose)
    return LIB


MEX_DIR = PKG_ROOT / "mex"
MEX_MOCK_LIB = LIBDIR / "libwinds_mexmock.so"


def build_mex_mock(verbose: bool = False) -> Path:
    """The MEX gateway (mex/BiotSavartMex.cpp) linked against the in-process
    mock of MATLAB's MEX runtime (mex/mock), for tests: host code only, g++."""
    tmp = MEX_MOCK_LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
           f"-I{MEX_DIR / 'mock'}", f"-I{REPO_ROOT / 'include'}",
           str(MEX_DIR / "BiotSavartMex.cpp"), str(MEX_DIR / "mock" / "mock_mex.cpp"),
           "-o", str(tmp), f"-L{LIBDIR}", "-lwinds_bs", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("mock MEX build failed:\n" + r.stderr)
    os.replace(tmp, MEX_MOCK_LIB)
    return MEX_MOCK_LIB


if __name__ == "__main__":
    print(build(verbose=True))
