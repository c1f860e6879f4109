"""In-tree build of the gfx950 library (libblitz_hip.so) and the test oracle.

The library is one hipcc invocation over csrc/*.hip (kernels, host controller
and the extern "C" layer) plus csrc/*.cpp (host-only code).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
LIBRARY = PACKAGE / "libblitz_hip.so"
# Counter calibration probe (a plain copy kernel), loaded only by bench.py.
COPY_PROBE = PACKAGE / "calib" / "libblitz_copy_probe.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off keeps every FP64 multiply and add separately rounded, as
# in the reference CPU audit (built with -ffp-contract=off). Packed FP32 is
# disabled for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources() -> list[Path]:
    csrc = PACKAGE / "csrc"
    return sorted(csrc.glob("*.hip")) + sorted(csrc.glob("*.cpp"))


def build_library(verbose: bool = False) -> Path:
    cmd = [HIPCC, *HIP_FLAGS, "-I", str(ROOT / "include"), "-o", str(LIBRARY) + ".tmp",
           *[str(s) for s in sources()]]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(str(LIBRARY) + ".tmp", LIBRARY)
    probe = [HIPCC, *HIP_FLAGS, "-o", str(COPY_PROBE) + ".tmp", str(PACKAGE / "calib" / "copy_probe.hip")]
    if verbose:
        print(" ".join(probe), flush=True)
    subprocess.run(probe, check=True)
    os.replace(str(COPY_PROBE) + ".tmp", COPY_PROBE)
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
