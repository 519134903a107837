"""Build recipe for the gfx950 Box2D step libraries (in-tree, no JIT cache).

    python -m box2dpp_amd.build            # libb2g.so + libbox2d.so + CPU checkers

Outputs (git-ignored, shipped with the tree):
    box2dpp_amd/lib/libb2g.so    HIP kernels + C ABI (include/b2g.h), hipcc --offload-arch=gfx950
    box2dpp_amd/lib/libbox2d.so  host box2d:: API + scene runner, g++, links libb2g.so
    oracle/_build/liborc.so      CPU restatement (test infrastructure, via oracle/Makefile)
    oracle/_ref/libb2ref.so      reference build (only where the reference tree exists)

All code is compiled with FP contraction off; HIP code additionally with
packed FP32 disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "lib", "obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PYTORCH_ROCM_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{ROOT}/include", f"-I{PKG}/csrc", f"-I{PKG}",
]
HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall",
              f"-I{ROOT}/include", f"-I{PKG}"]

HIP_SOURCES = ["csrc/prims.hip", "csrc/b2g.hip", "csrc/mlp.hip"]
HOST_SOURCES = ["host/shapes.cpp", "host/world.cpp", "host/runner.cpp"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in res.stdout.splitlines() if "packed-fp32-ops" not in line)
    if res.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}")


def _compile_hip(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(PKG, src), "-o", obj])
    return obj


def _compile_host(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run(["g++", *HOST_FLAGS, "-c", os.path.join(PKG, src), "-o", obj])
    return obj


def build_libs() -> None:
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(PKG, s))]
    with ThreadPoolExecutor(max_workers=4) as ex:
        hip_objs = list(ex.map(_compile_hip, srcs))
        host_objs = list(ex.map(_compile_host, HOST_SOURCES))
    b2g = os.path.join(LIB, "libb2g.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *hip_objs, "-o", b2g + ".tmp",
          "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(b2g + ".tmp", b2g)
    host = os.path.join(LIB, "libbox2d.so")
    _run(["g++", "-shared", "-fPIC", *host_objs, "-o", host + ".tmp", f"-L{LIB}", "-lb2g",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(host + ".tmp", host)


def build_oracle() -> None:
    make = shutil.which("make") or "make"
    _run([make, "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"])
    ref_src = os.path.join(os.environ.get("REFERENCE_ROOT", "/root/reference"), "Regression", "box2d-ref")
    if os.path.isdir(ref_src):
        _run([make, "-s", "-j8", "-C", os.path.join(ROOT, "oracle"), "ref"])


def build_all() -> None:
    build_libs()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", os.listdir(LIB))
