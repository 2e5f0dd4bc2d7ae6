"""Build recipe for the HIP library sphcode_amd/lib/libsphgpu.so (gfx950),
the C++ host layer sphcode_amd/lib/libsphhost.so (Module / Simulation /
Solver over the C ABI) and its driver sphcode_amd/bin/sphgpu_run.

Each source is compiled to an object in parallel and linked.  Run
``python -m sphcode_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
SRC = PKG / "csrc"
OBJ = ROOT / "_build" / "obj"
LIB = PKG / "lib" / "libsphgpu.so"
HOST_LIB = PKG / "lib" / "libsphhost.so"
HOST_BIN = PKG / "bin" / "sphgpu_run"
HOST_SRC = PKG / "host" / "src"
HOST_OBJ = ROOT / "_build" / "host"
CXX = os.environ.get("CXX", "g++")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SPHGPU_ARCH", "gfx950")

HOST_SOURCES = ["json.cpp", "parameters.cpp", "simulation.cpp", "modules.cpp", "solver.cpp", "snapshot.cpp", "output.cpp"]
HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
              "-I", str(ROOT / "include"), "-I", str(PKG / "host" / "include")]

SOURCES = ["api.hip", "tree.hip", "search.hip", "pre.hip", "force.hip", "gravity.hip", "step.hip", "adjoint.hip"]

FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled throughout
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FP32 division and square root (mixed-precision pair terms only) use the
    # hardware reciprocal / square root (about 1 ulp) instead of the
    # correctly rounded sequences; FP64 arithmetic is unaffected
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", str(ROOT / "include"),
]


def _compile(src: pathlib.Path) -> pathlib.Path:
    out = OBJ / (src.stem + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and "target-feature" not in l)
    if warn.strip():
        sys.stderr.write(warn + "\n")
    return out


def _run(cmd: list[str], what: str) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        sys.stderr.write(r.stderr)


def _compile_host(src: pathlib.Path) -> pathlib.Path:
    out = HOST_OBJ / (src.stem + ".o")
    _run([CXX, *HOST_FLAGS, "-c", str(src), "-o", str(out)], "compile")
    return out


def build_host(jobs: int | None = None) -> pathlib.Path:
    """Build libsphhost.so and sphgpu_run against an already built libsphgpu.so."""
    if not LIB.exists():
        raise RuntimeError(f"{LIB} is missing: build the HIP library first")
    HOST_OBJ.mkdir(parents=True, exist_ok=True)
    HOST_BIN.parent.mkdir(parents=True, exist_ok=True)
    srcs = [HOST_SRC / s for s in HOST_SOURCES] + [HOST_SRC / "main.cpp"]
    with cf.ThreadPoolExecutor(max_workers=jobs or min(len(srcs), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile_host, srcs))
    tmp = HOST_LIB.with_suffix(".so.tmp")
    _run([CXX, "-shared", "-o", str(tmp), *map(str, objs[:-1]), "-L", str(LIB.parent), "-lsphgpu",
          "-Wl,-rpath,$ORIGIN"], "link")
    os.replace(tmp, HOST_LIB)
    tmpb = HOST_BIN.with_suffix(".tmp")
    _run([CXX, "-o", str(tmpb), str(objs[-1]), "-L", str(LIB.parent), "-lsphhost", "-lsphgpu",
          "-Wl,-rpath,$ORIGIN/../lib"], "link")
    os.replace(tmpb, HOST_BIN)
    return HOST_LIB


def build(jobs: int | None = None) -> pathlib.Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    LIB.parent.mkdir(parents=True, exist_ok=True)
    srcs = [SRC / s for s in SOURCES]
    with cf.ThreadPoolExecutor(max_workers=jobs or min(len(srcs), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_host(jobs)
    return LIB


if __name__ == "__main__":
    print(build())
