"""Builds libmsq_hip.so (HIP kernels + C ABI) for gfx950 with hipcc.

Usage: python mixed-mlx_amd/build.py [--jobs N]
The library is written next to the Python binding, mixed-mlx_amd/msq/.
"""
from __future__ import annotations

import argparse
import concurrent.futures
import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "msq" / "libmsq_hip.so"
BUILD = HERE / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MSQ_OFFLOAD_ARCH", "gfx950")
# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 (they have returned wrong
# results on this hardware); f32 denormals kept (codecs are bit-exact).
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-flush-denormals-to-zero", "-Wall", "-Werror", "-Wno-unused-function",
]

SOURCES = ["kernels_codec.hip", "kernels_jacobi.hip", "kernels_tile_jacobi.hip", "kernels_gemm.hip",
           "kernels_misc.hip", "kernels_mlp.hip", "kernels_blocked.hip", "noalloc.hip",
           "capi.cpp"]


def compile_one(src: str) -> pathlib.Path:
    obj = BUILD / (src + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    if src.endswith(".cpp"):
        cmd.insert(1, "-x")
        cmd.insert(2, "hip")
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    warnings = [l for l in res.stderr.splitlines()
                if "warning" in l and "packed-fp32-ops" not in l]
    if warnings:
        sys.stderr.write("\n".join(warnings) + "\n")
    return obj


def build(jobs: int = 4) -> pathlib.Path:
    BUILD.mkdir(exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, OUT)
    build_cli()
    return OUT


CLI = HERE / "msq" / "msq_cli"
MIRROR = HERE / "msq" / "libmsq_mirror.so"
MIRROR_SOURCES = ["hip_runtime.cpp", "host_reference.cpp"]
CLI_SOURCES = ["msq_cli.cpp"]
HOST_FLAGS = ["-O2", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fPIC"]


def host_lapack() -> list[str]:
    """Link arguments of the host's binary64 LAPACK/BLAS: the OpenBLAS build
    that scipy ships (its symbols carry a scipy_ prefix)."""
    import glob
    import scipy
    libs = pathlib.Path(scipy.__file__).resolve().parent.parent / "scipy.libs"
    found = sorted(glob.glob(str(libs / "libscipy_openblas*.so")))
    if not found:
        raise RuntimeError(f"no OpenBLAS library under {libs}")
    lib = pathlib.Path(found[0])
    return ["-L", str(lib.parent), f"-l:{lib.name}", f"-Wl,-rpath,{lib.parent}"]


def _run(cmd: list[str], what: str) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"{what} build failed")


def build_cli(extra: list[str] | None = None) -> pathlib.Path:
    """The C++ mirror of the solver entry points (libmsq_mirror.so:
    hip_runtime.cpp + host_reference.cpp) and its command-line driver: host
    code over the C ABI, built with g++ (-Werror).  `extra` adds compiler
    flags (a sanitizer build passes its own)."""
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cxx = os.environ.get("CXX", "g++")
    flags = HOST_FLAGS + list(extra or [])
    _run([cxx, *flags, "-shared", *[str(CSRC / s) for s in MIRROR_SOURCES],
          "-o", str(MIRROR) + ".tmp", "-L", str(OUT.parent), "-lmsq_hip", "-Wl,-rpath,$ORIGIN",
          f"-Wl,-rpath-link,{rocm}/lib", *host_lapack()], "libmsq_mirror.so")
    os.replace(str(MIRROR) + ".tmp", MIRROR)
    _run([cxx, *flags, *[str(CSRC / s) for s in CLI_SOURCES], "-o", str(CLI) + ".tmp",
          "-L", str(OUT.parent), "-lmsq_mirror", "-lmsq_hip", "-Wl,-rpath,$ORIGIN",
          f"-Wl,-rpath-link,{rocm}/lib", *host_lapack()], "msq_cli")
    os.replace(str(CLI) + ".tmp", CLI)
    return CLI


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    print(build(ap.parse_args().jobs))
