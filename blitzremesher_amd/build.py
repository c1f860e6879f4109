"""In-tree build of the gfx950 library (libblitz_hip.so) and the test oracle.

Every csrc/*.hip (kernels, host controller, the extern "C" layer) and
csrc/*.cpp (host-only code) is compiled to its own object under _build/ in
parallel, then hipcc links them into libblitz_hip.so. Kernels are launched
only from the translation unit that defines them, so no relocatable device
code is needed.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
LIBRARY = PACKAGE / "libblitz_hip.so"
OBJECTS = ROOT / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off keeps every FP64 multiply and add separately rounded, as
# in the reference CPU audit (built with -ffp-contract=off). Packed FP32 is
# disabled for all device code.
COMPILE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
HIP_FLAGS = [*COMPILE_FLAGS, "-shared"]


def sources() -> list[Path]:
    csrc = PACKAGE / "csrc"
    return sorted(csrc.glob("*.hip")) + sorted(csrc.glob("*.cpp"))


def _jobs() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def build_library(verbose: bool = False) -> Path:
    OBJECTS.mkdir(exist_ok=True)
    srcs = sources()
    objs = [OBJECTS / (s.name + ".o") for s in srcs]

    def compile_one(pair: tuple[Path, Path]) -> None:
        src, obj = pair
        cmd = [HIPCC, *COMPILE_FLAGS, "-I", str(ROOT / "include"), "-c", str(src), "-o", str(obj) + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(str(obj) + ".tmp", obj)

    with ThreadPoolExecutor(_jobs()) as pool:
        list(pool.map(compile_one, zip(srcs, objs)))
    link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(LIBRARY) + ".tmp", *map(str, objs)]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(str(LIBRARY) + ".tmp", LIBRARY)
    return LIBRARY


def build_oracle(verbose: bool = False) -> None:
    """Builds oracle/liboracle.so and, when the reference checkout exists,
    oracle/_ref/libblitz_ref.so (test infrastructure only)."""
    make = shutil.which("make")
    if make is None:
        raise RuntimeError("make is required to build the oracle")
    targets = ["liboracle", "ref"]
    subprocess.run([make, "-s", "-j8", "-f", str(ROOT / "oracle" / "Makefile"), *targets], check=True,
                   stdout=None if verbose else subprocess.DEVNULL)


def main() -> int:
    build_library(verbose=True)
    build_oracle(verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
