"""MI355X (gfx950) HIP backend for BlitzRemesher's audit + scorer hot path.

The compute lives in libblitz_hip.so (hand-written HIP kernels behind the C ABI
of include/blitz_hip.h); this package is its Python view for tests and
benchmarks. Every entry point loads the library first and raises if it is not
built (there is no other implementation to fall back to); importing the
package alone does not, so that blitzremesher_amd.build can run before the
library exists.
"""
from . import _lib
from .audit import (PROFILE_ATTRIBUTES, PROFILE_COVERAGE, PROFILE_NORMALS, Audit, BlitzError, DeviceMesh,
                    EvalSettings, Measurement, Mesh, bounds, cameras)


__all__ = ["Audit", "DeviceMesh", "EvalSettings", "Measurement", "Mesh", "bounds", "cameras", "BlitzError",
           "PROFILE_COVERAGE", "PROFILE_NORMALS", "PROFILE_ATTRIBUTES"]
