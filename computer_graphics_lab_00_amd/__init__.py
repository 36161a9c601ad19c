"""MI355X-native SPH fluid step (drop-in for the reference's src/physics fluid interface).

The compute path is ``libsph_hip.so`` (hand-written HIP kernels for gfx950 behind the C
ABI of ``include/sph_c_api.h``); this package is its Python host mirror. Importing the
package does not load the library; the first call does, and fails loudly if it is
missing.
"""
from .capi import SphCollider, SphError, SphForceField, SphParams, SphSettings  # noqa: F401
from .fluid import (FluidSolver, Workload, cell_count_for, derive_fluid_params, emit_lattice,  # noqa: F401
                    fluid_substeps, lattice_jitter, m1, r4m, s16k, spatial_hash, translation4)
