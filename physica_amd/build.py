"""Build recipe for libphysica_amd.so (HIP, gfx950) and the oracle library.

Run ``python -m physica_amd.build`` (or ``__graft_entry__.build()``). The
shared library is written in-tree next to this file so it travels with the
repository snapshot.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
LIBRARY = PACKAGE / "libphysica_amd.so"

SOURCES = ["grid.hip", "masks.hip", "dfsph.hip", "dfsph_adjoint.hip", "sph.hip", "wcsph_adjoint.hip", "pbf.hip",
           "pbf_adjoint.hip", "capi.hip"]

# -ffp-contract=off keeps every multiply and add a separate IEEE operation
# (the kernels reproduce the reference's expressions operation by operation);
# FP32 division and sqrt are correctly rounded by HIP's default. Packed FP32
# instructions are disabled for the device pass.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return path


def build_library(verbose: bool = False) -> Path:
    """Compile every HIP source for gfx950 and link libphysica_amd.so."""
    objdir = ROOT / "_build" / "physica_amd"
    objdir.mkdir(parents=True, exist_ok=True)
    include = ["-I", str(ROOT / "include"), "-I", str(CSRC)]
    commands = []
    for name in SOURCES:
        src = CSRC / name
        obj = objdir / (src.stem + ".o")
        commands.append([hipcc(), *HIPCC_FLAGS, *include, "-c", str(src), "-o", str(obj)])
    objects = [cmd[-1] for cmd in commands]

    def compile_one(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    workers = max(1, min(len(commands), int(os.environ.get("PHX_BUILD_JOBS", "4"))))
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as pool:
        for future in [pool.submit(compile_one, cmd) for cmd in commands]:
            future.result()
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objects, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIBRARY)
    return LIBRARY


CPP_TESTS = ROOT / "tests" / "cpp"


def build_cpp_tests(verbose: bool = False) -> list:
    """Compile the C++ API programs under tests/cpp against the library."""
    outdir = CPP_TESTS / "_bin"
    outdir.mkdir(parents=True, exist_ok=True)
    built = []
    for src in sorted(CPP_TESTS.glob("*.cpp")):
        exe = outdir / src.stem
        cmd = [hipcc(), "-std=c++20", "-O2", "--offload-arch=gfx950", "-I", str(ROOT / "include"), str(src),
               "-L", str(PACKAGE), "-lphysica_amd", "-Wl,-rpath," + str(PACKAGE), "-Wl,-rpath,$ORIGIN/../../../physica_amd",
               "-o", str(exe)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        built.append(exe)
    return built


MATH_CHECK = ROOT / "tests" / "native" / "libphx_math_check.so"


def build_math_check(verbose: bool = False) -> Path:
    """Compile the rounding-helper checks (tests/native/math_check.hip)."""
    src = ROOT / "tests" / "native" / "math_check.hip"
    tmp = MATH_CHECK.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, "-shared", "-I", str(CSRC), str(src), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, MATH_CHECK)
    return MATH_CHECK


def build_oracle(verbose: bool = False) -> Path:
    """Compile the C oracle (test infrastructure) with its Makefile."""
    cmd = ["make", "-C", str(ROOT / "oracle"), "-j4"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, stdout=None if verbose else subprocess.DEVNULL)
    return ROOT / "oracle" / "_build" / "liboracle.so"


def main() -> int:
    build_library(verbose=True)
    build_cpp_tests(verbose=True)
    build_math_check(verbose=True)
    build_oracle(verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
