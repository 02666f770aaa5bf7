"""MI355X-native bispectrum-multipole estimator (hot path of BARBIE/BEAM).

The estimator runs in libbarbie_amd.so (hand-written HIP kernels for gfx950
behind the C ABI of include/barbie.h); this package holds the Python side of
that boundary and the seeded input generator.
"""
from .native import (Geometry, Plan, Estimator, Server, BarbieError, normalise, lib, LIB_PATH, BIN_PATH,
                     NPOLES, PRECISIONS, RINGS_REFERENCE, RINGS_UNIFORM, RINGS_EXPLICIT,
                     power_multipoles, generate_field, estimate_mixed)
from .field import make_field, make_field_device, write_field, read_field, seed_for

__all__ = ["Geometry", "Plan", "Estimator", "Server", "BarbieError", "normalise", "lib", "LIB_PATH",
           "BIN_PATH", "NPOLES", "PRECISIONS", "RINGS_REFERENCE", "RINGS_UNIFORM",
           "RINGS_EXPLICIT", "make_field", "make_field_device", "write_field", "read_field", "seed_for",
           "power_multipoles", "generate_field", "estimate_mixed"]
