"""Build libqe_gpu.so (HIP, gfx950) and the C++ host operator library in-tree.

    python query-engine_amd/build.py            # everything
    python query-engine_amd/build.py --gpu-only

hipcc cross-compiles for gfx950 without a GPU. Objects go to
query-engine_amd/_build/, shared libraries to query-engine_amd/lib/.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
BUILD = PKG / "_build"
LIB = PKG / "lib"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed-FP32 instructions stay disabled on this fleet (hipcc's host pass warns
# that it does not know the feature; that is expected).
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value", f"-I{INCLUDE}", f"-I{CSRC}",
]
HIP_SOURCES = ["context.hip", "groupby.hip", "join.hip", "hashtable.hip", "sort.hip", "batched.hip",
               "asof.hip", "gen.hip", "filter.hip", "strings.hip"]
HOST_SOURCES = ["operators.cpp", "plan.cpp", "expr.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise SystemExit(f"build failed: {cmd[-1]}")


def _newer(src: list[Path], dst: Path) -> bool:
    if not dst.exists():
        return True
    t = dst.stat().st_mtime
    return any(s.stat().st_mtime > t for s in src)


def build_gpu(force: bool = True, jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    headers = list(CSRC.glob("*.hpp")) + [INCLUDE / "qe_gpu.h"]
    srcs = [CSRC / s for s in HIP_SOURCES if (CSRC / s).exists()]
    objs = []
    todo = []
    for s in srcs:
        o = BUILD / (s.stem + ".o")
        objs.append(o)
        if force or _newer([s] + headers, o):
            todo.append([HIPCC, *HIP_FLAGS, "-c", str(s), "-o", str(o)])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, todo))
    out = LIB / "libqe_gpu.so"
    tmp = LIB / "libqe_gpu.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, out)
    return out


def build_host(force: bool = True) -> Path | None:
    srcs = [HOST / s for s in HOST_SOURCES if (HOST / s).exists()]
    if not srcs:
        return None
    LIB.mkdir(exist_ok=True)
    out = LIB / "libqe_host.so"
    deps = srcs + list(HOST.glob("*.h")) + [INCLUDE / "qe_gpu.h"]
    if force or _newer(deps, out):
        tmp = LIB / "libqe_host.so.tmp"
        _run(["g++", "-O2", "-std=c++20", "-fPIC", "-shared", f"-I{INCLUDE}", f"-I{HOST}",
              "-o", str(tmp), *map(str, srcs), f"-L{LIB}", "-lqe_gpu",
              f"-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, out)
    tests = HOST / "tests"
    if tests.exists():
        exe = LIB / "qe_host_tests"
        tsrc = sorted(tests.glob("*.cpp"))
        if force or _newer(tsrc + [out], exe):
            _run(["g++", "-O2", "-std=c++20", f"-I{INCLUDE}", f"-I{HOST}", "-o", str(exe),
                  *map(str, tsrc), f"-L{LIB}", "-lqe_host", "-lqe_gpu",
                  f"-Wl,-rpath,$ORIGIN"])
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpu-only", action="store_true")
    ap.add_argument("--incremental", action="store_true")
    a = ap.parse_args()
    build_gpu(force=not a.incremental)
    if not a.gpu_only:
        build_host(force=not a.incremental)


if __name__ == "__main__":
    main()
