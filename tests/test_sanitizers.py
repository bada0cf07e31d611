"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU
only): the library's C ABI layer (capi.cpp, host side instrumented; device
code as built) and the C++ mirror, driven by tests/cxx_sanitize_check.cpp,
and the command-line driver's argument checks.  Any sanitizer report fails
the test."""
import os
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
CSRC = ROOT / "mixed-mlx_amd" / "csrc"
BUILD = ROOT / "mixed-mlx_amd" / "_build"
CLANG = "/opt/rocm/llvm/bin/clang++"
HIPCC = "/opt/rocm/bin/hipcc"
SAN = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=all"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


@pytest.fixture(scope="module")
def sanitized(tmp_path_factory):
    out = tmp_path_factory.mktemp("san")
    objs = [BUILD / f"{s}.o" for s in ("kernels_codec.hip", "kernels_jacobi.hip",
                                       "kernels_gemm.hip", "kernels_misc.hip", "kernels_mlp.hip",
                                       "kernels_blocked.hip", "noalloc.hip")]
    if not all(o.exists() for o in objs):
        pytest.skip("build the library first (mixed-mlx_amd/build.py)")
    capi = out / "capi.o"
    # Host pass instrumented; each -fsanitize= applies to the host only.
    _run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
          "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=all",
          "-c", str(CSRC / "capi.cpp"), "-o", str(capi)])
    lib = out / "libmsq_hip.so"
    # Host link (the objects carry their device code): the sanitizer runtimes
    # come with the executable.
    _run([CLANG, "-shared", "-fPIC", "-o", str(lib), str(capi), *map(str, objs),
          "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    import sys
    sys.path.insert(0, str(ROOT / "mixed-mlx_amd"))
    import build as B
    exe = out / "check"
    _run([CLANG, "-O1", "-g", "-std=c++17", *SAN, "-I", str(CSRC),
          str(ROOT / "tests" / "cxx_sanitize_check.cpp"), str(CSRC / "hip_runtime.cpp"),
          str(CSRC / "host_reference.cpp"), "-o", str(exe), "-L", str(out), "-lmsq_hip",
          f"-Wl,-rpath,{out}", "-Wl,-rpath-link,/opt/rocm/lib", *B.host_lapack()])
    cli = out / "msq_cli"
    _run([CLANG, "-O1", "-g", "-std=c++17", *SAN, "-I", str(CSRC), str(CSRC / "msq_cli.cpp"),
          str(CSRC / "hip_runtime.cpp"), str(CSRC / "host_reference.cpp"), "-o", str(cli),
          "-L", str(out), "-lmsq_hip", f"-Wl,-rpath,{out}", "-Wl,-rpath-link,/opt/rocm/lib",
          *B.host_lapack()])
    return exe, cli


def test_host_code_is_clean_under_asan_and_ubsan(sanitized):
    exe, _ = sanitized
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=ENV, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


def test_cli_argument_checks_under_asan_and_ubsan(sanitized):
    _, cli = sanitized
    for args in (["--help"], ["--orders", "x"], ["--paper-test-set", "I"],
                 ["--rounding-seed", "5"], ["--native-stagnation-window", "0"],
                 ["--precision", "fp7"], ["64", "128"], ["--kernel-path"]):
        r = subprocess.run([str(cli), *args], capture_output=True, text=True, env=ENV, timeout=60)
        assert r.returncode in (0, 1), (args, r.returncode, r.stderr)
        assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, args
