"""Build recipe for libplsparse_amd.so (gfx950) and the workload generator.

Every source under csrc/ is compiled by hipcc for gfx950 into _build/, then
linked in-tree into lib/libplsparse_amd.so.  The host-only workload generator
(workloads/gen.cpp, synthetic matrices of the benchmark configurations) is
built with g++ into lib/libplsparse_workloads.so.  Run `python build.py`.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "lib"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd[1:1] = ["-x", "hip"]
    _run(cmd)
    return obj


def build_library(jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    jobs = jobs or min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    out = LIB / "libplsparse_amd.so"
    tmp = LIB / "libplsparse_amd.so.tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_workloads() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libplsparse_workloads.so"
    tmp = LIB / "libplsparse_workloads.so.tmp"
    _run(["g++", "-O3", "-fPIC", "-shared", "-fopenmp", "-std=c++17",
          str(PKG / "workloads" / "gen.cpp"), str(PKG / "workloads" / "mtx.cpp"),
          "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_apps() -> Path:
    """The C driver apps/mtx_roofline.c, linked against both in-tree libraries
    (and the HIP runtime, for its device-resident vectors) with an $ORIGIN
    rpath so it runs from lib/."""
    LIB.mkdir(exist_ok=True)
    out = LIB / "mtx_roofline"
    tmp = LIB / "mtx_roofline.tmp"
    _run(["gcc", "-O2", "-std=c99", "-Wall", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT / 'include'}",
          f"-I{PKG / 'workloads'}", "-I/opt/rocm/include",
          str(PKG / "apps" / "mtx_roofline.c"), f"-L{LIB}", "-lplsparse_amd",
          "-lplsparse_workloads", "-L/opt/rocm/lib", "-lamdhip64", "-lm",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib", "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def main() -> None:
    print(build_library(), flush=True)
    print(build_workloads(), flush=True)
    print(build_apps(), flush=True)


if __name__ == "__main__":
    sys.exit(main())
