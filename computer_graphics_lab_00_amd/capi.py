"""ctypes mirror of ``include/sph_c_api.h`` and the loader of the in-tree HIP library.

The structures below are field-for-field copies of the C declarations; the C header
cites the reference type each one mirrors (src/physics/fluid_sph.h:28-67,
src/physics/force_field.h:20-28, src/scene/scene.h:186-203).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libsph_hip.so"

SPH_MAX_COLLIDERS = 8
SPH_MAX_FORCE_FIELDS = 8
SPH_CELL_CAPACITY = 32

SPH_OK = 0
STATUS_NAMES = {
    0: "SPH_OK",
    1: "SPH_ERR_INVALID_ARGUMENT",
    2: "SPH_ERR_NO_DEVICE",
    3: "SPH_ERR_OUT_OF_MEMORY",
    4: "SPH_ERR_DEVICE",
    5: "SPH_ERR_CAPACITY",
    6: "SPH_ERR_NOT_POWER_OF_TWO",
    7: "SPH_ERR_STATE",
    8: "SPH_ERR_DEGENERATE",
}
ERR_DEGENERATE = 8
MAX_BIN_OCCUPANCY = 16384

SHAPE_SPHERE, SHAPE_BOX, SHAPE_PLANE, SHAPE_CYLINDER, SHAPE_CAPSULE, SHAPE_MESH = range(6)
FIELD_WIND, FIELD_VORTEX, FIELD_POINT = range(3)
MODE_REFERENCE, MODE_UNCAPPED, MODE_DEDUPE, MODE_FAST_MATH = 0, 1, 2, 4
EOS_LINEAR, EOS_TAIT = 0, 1
CFG_KEEP_FORCES = 1
CFG_NO_GRAPH = 2
CFG_BUCKET_WALK = 4
CFG_TILED_DENSITY = 8
CHECK_COMPRESSION, CHECK_RAREFACTION = 1, 2
FEATURES_MOMENTS, FEATURES_KNN8 = 0, 1


class SphCollider(C.Structure):
    _fields_ = [
        ("shape", C.c_uint32),
        ("radius", C.c_float),
        ("halfExtents", C.c_float * 3),
        ("pad0", C.c_float),
        ("world", C.c_float * 16),
        ("inverseWorld", C.c_float * 16),
    ]


class SphForceField(C.Structure):
    _fields_ = [
        ("position", C.c_float * 3),
        ("axis", C.c_float * 3),
        ("strength", C.c_float),
        ("radius", C.c_float),
        ("falloff", C.c_float),
        ("type", C.c_uint32),
    ]


class SphParams(C.Structure):
    _fields_ = [
        ("emitterWorld", C.c_float * 16),
        ("emitterHalfExtents", C.c_float * 3),
        ("spacing", C.c_float),
        ("containerMin", C.c_float * 3),
        ("particleRadius", C.c_float),
        ("containerMax", C.c_float * 3),
        ("wallRestitution", C.c_float),
        ("gravity", C.c_float * 3),
        ("particleMass", C.c_float),
        ("smoothingRadius", C.c_float),
        ("restDensity", C.c_float),
        ("stiffness", C.c_float),
        ("viscosity", C.c_float),
        ("lattice", C.c_uint32 * 3),
        ("cellCount", C.c_uint32),
        ("colliderCount", C.c_uint32),
        ("colliders", SphCollider * SPH_MAX_COLLIDERS),
        ("fieldCount", C.c_uint32),
        ("fields", SphForceField * SPH_MAX_FORCE_FIELDS),
    ]


class SphSettings(C.Structure):
    """The simulation part of scene::Fluid, with the reference's defaults
    (src/scene/scene.h:186-203)."""

    _fields_ = [
        ("emitterHalfExtents", C.c_float * 3),
        ("particleCount", C.c_uint32),
        ("particleRadius", C.c_float),
        ("smoothingRadius", C.c_float),
        ("restDensity", C.c_float),
        ("stiffness", C.c_float),
        ("viscosity", C.c_float),
        ("containerMin", C.c_float * 3),
        ("containerMax", C.c_float * 3),
        ("gravity", C.c_float * 3),
    ]

    def __init__(self, **kw):
        super().__init__()
        self.emitterHalfExtents[:] = [0.5, 0.5, 0.5]
        self.particleCount = 8192
        self.particleRadius = 0.025
        self.smoothingRadius = 0.1
        self.restDensity = 1000.0
        self.stiffness = 50.0
        self.viscosity = 0.5
        self.containerMin[:] = [-1.0, 0.0, -1.0]
        self.containerMax[:] = [1.0, 2.0, 1.0]
        self.gravity[:] = [0.0, -9.81, 0.0]
        for k, v in kw.items():
            cur = getattr(self, k)
            if isinstance(cur, C.Array):
                cur[:] = list(v)
            else:
                setattr(self, k, v)


class SphSolverConfig(C.Structure):
    _fields_ = [
        ("maxParticles", C.c_uint32),
        ("maxCellCount", C.c_uint32),
        ("device", C.c_int32),
        ("flags", C.c_uint32),
    ]


class SphSceneCapacity(C.Structure):
    _fields_ = [("maxParticles", C.c_uint32), ("maxCellCount", C.c_uint32)]


class SphGpuInstance(C.Structure):
    _fields_ = [("model", C.c_float * 16), ("previousModel", C.c_float * 16), ("normalMatrix", C.c_float * 16),
                ("meshIndex", C.c_uint32), ("bucket", C.c_uint32), ("bucketBase", C.c_uint32),
                ("jointOffset", C.c_uint32), ("skinnedVertexOffset", C.c_uint32),
                ("previousSkinnedVertexOffset", C.c_uint32), ("skinnedMeshletOffset", C.c_uint32),
                ("visibilityBase", C.c_uint32)]


class SphTlasInstance(C.Structure):
    _fields_ = [("transform", (C.c_float * 4) * 3), ("customIndexMask", C.c_uint32), ("sbtOffsetFlags", C.c_uint32),
                ("accelerationStructureReference", C.c_uint64)]


class SphInstanceConfig(C.Structure):
    _fields_ = [("instanceBase", C.c_uint32), ("tlasBase", C.c_uint32), ("sphereMesh", C.c_uint32),
                ("fluidIndex", C.c_uint32), ("sphereBlas", C.c_uint64), ("surfaceBlas", C.c_uint64),
                ("resetHistory", C.c_uint32), ("pad", C.c_uint32)]


class SphSurfaceVertex(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("normal", C.c_uint32)]


class SphRolloutStats(C.Structure):
    _fields_ = [
        ("macroSteps", C.c_uint64),
        ("fallbacks", C.c_uint64),
        ("plainSubsteps", C.c_uint64),
        ("backedOff", C.c_uint64),
        ("maxDeviation", C.c_float),
        ("lastDeviation", C.c_float),
        ("maxRarefactionGrowth", C.c_float),
        ("pad0", C.c_uint32),
        ("rarefactionFallbacks", C.c_uint64),
    ]


class SphError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


_lib = None

_FP = C.POINTER(C.c_float)
_UP = C.POINTER(C.c_uint32)
_DP = C.POINTER(C.c_double)
_HP = C.POINTER(C.c_uint16)


def _declare(lib):
    P = C.c_void_p
    sigs = {
        "sph_last_error": (C.c_char_p, []),
        "sph_derive_params": (C.c_int, [C.POINTER(SphSettings), _FP, C.c_uint32, C.c_uint32,
                                        C.POINTER(SphCollider), C.c_uint32, C.POINTER(SphForceField), C.c_uint32,
                                        C.POINTER(SphParams)]),
        "sph_substeps": (C.c_uint32, [C.POINTER(SphParams), C.c_float]),
        "sph_cell_count_for": (C.c_uint32, [C.c_uint32]),
        "sph_hash_cell": (C.c_uint32, [C.c_int32, C.c_int32, C.c_int32, C.c_uint32]),
        "sph_spatial_cell": (None, [_FP, C.c_float, C.POINTER(C.c_int32)]),
        "sph_lattice_jitter": (None, [_FP, C.c_uint32, C.c_uint32, C.c_float]),
        "sph_create": (C.c_int, [C.POINTER(SphSolverConfig), C.POINTER(P)]),
        "sph_destroy": (C.c_int, [P]),
        "sph_emit": (C.c_int, [P, C.POINTER(SphParams), C.c_uint32]),
        "sph_step": (C.c_int, [P, C.POINTER(SphParams), C.c_float]),
        "sph_substep": (C.c_int, [P, C.POINTER(SphParams), C.c_float]),
        "sph_synchronize": (C.c_int, [P]),
        "sph_max_bin_occupancy": (C.c_int, [P, C.POINTER(C.c_uint32)]),
        "sph_particle_count": (C.c_uint32, [P]),
        "sph_read_particles": (C.c_int, [P, _FP, _FP]),
        "sph_read_previous": (C.c_int, [P, _FP]),
        "sph_keep_rendered": (C.c_int, [P]),
        "sph_load_state": (C.c_int, [P, _FP, _FP, C.c_uint32]),
        "sph_read_grid": (C.c_int, [P, _UP, _UP, _UP]),
        "sph_read_sorted": (C.c_int, [P, _UP, _UP, _UP]),
        "sph_read_forces": (C.c_int, [P, _FP]),
        "sph_set_mode": (C.c_int, [P, C.c_uint32]),
        "sph_set_eos": (C.c_int, [P, C.c_uint32]),
        "sph_conservation": (C.c_int, [P, C.POINTER(SphParams), _DP]),
        "sph_profile_enable": (C.c_int, [P, C.c_int]),
        "sph_profile_read": (C.c_int, [P, C.POINTER(C.c_char_p), _DP, _UP, C.c_uint32]),
        "sph_batch_create": (C.c_int, [C.POINTER(SphSceneCapacity), C.c_uint32, C.c_int32, C.c_uint32,
                                       C.POINTER(P)]),
        "sph_batch_destroy": (C.c_int, [P]),
        "sph_batch_scene_count": (C.c_uint32, [P]),
        "sph_batch_set_params": (C.c_int, [P, C.c_uint32, C.POINTER(SphParams)]),
        "sph_batch_set_layout": (C.c_int, [P, C.c_uint32, _UP, C.c_uint32]),
        "sph_batch_set_mode": (C.c_int, [P, C.c_uint32, C.c_uint32, C.c_uint32]),
        "sph_batch_set_schedule": (C.c_int, [P, _UP, C.c_uint32]),
        "sph_batch_emit": (C.c_int, [P, C.c_uint32, C.c_uint32]),
        "sph_batch_load_state": (C.c_int, [P, C.c_uint32, _FP, _FP, C.c_uint32]),
        "sph_batch_step": (C.c_int, [P, C.c_float]),
        "sph_batch_synchronize": (C.c_int, [P]),
        "sph_batch_max_bin_occupancy": (C.c_int, [P, C.POINTER(C.c_uint32)]),
        "sph_batch_particle_count": (C.c_uint32, [P, C.c_uint32]),
        "sph_batch_read_particles": (C.c_int, [P, C.c_uint32, _FP, _FP]),
        "sph_batch_profile_enable": (C.c_int, [P, C.c_int]),
        "sph_batch_profile_read": (C.c_int, [P, C.POINTER(C.c_char_p), _DP, _UP, C.c_uint32]),
        "sph_debug_step_constants": (C.c_int, [C.POINTER(SphParams), C.c_float, _FP]),
        "sph_debug_copy": (C.c_int, [C.c_uint64]),
        "sph_debug_gemm": (C.c_int, [_HP, _HP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _FP]),
        "sph_debug_forward": (C.c_int, [P, _HP, C.c_uint32, C.c_uint32, _HP, _FP]),
        "sph_mc_table": (C.c_int, [C.POINTER(C.c_int8)]),
        "sph_write_instances": (C.c_int, [P, C.POINTER(SphParams), C.POINTER(SphInstanceConfig), C.c_void_p,
                                          C.c_void_p]),
        "sph_surface_field": (C.c_int, [P, C.POINTER(SphParams), C.c_uint32, _FP]),
        "sph_march_field": (C.c_int, [_FP, C.c_uint32, _FP, _FP, C.c_float, C.c_void_p, C.c_uint32, _UP]),
        "sph_surface_mesh": (C.c_int, [P, C.POINTER(SphParams), C.c_uint32, C.c_float, C.c_void_p, C.c_uint32,
                                       _UP]),
        "sph_net_param_count": (C.c_uint32, []),
        "sph_net_create": (C.c_int, [C.c_int32, C.c_uint64, C.POINTER(P)]),
        "sph_net_destroy": (C.c_int, [P]),
        "sph_net_get_params": (C.c_int, [P, _FP]),
        "sph_net_set_params": (C.c_int, [P, _FP]),
        "sph_net_save": (C.c_int, [P, C.c_char_p]),
        "sph_net_load": (C.c_int, [P, C.c_char_p]),
        "sph_net_forward": (C.c_int, [P, _HP, C.c_uint32, _FP]),
        "sph_net_gradients": (C.c_int, [P, _HP, _FP, C.c_uint32, _FP, _FP]),
        "sph_net_train_batch": (C.c_int, [P, _HP, _FP, C.c_uint32, C.c_float, _FP]),
        "sph_dataset_create": (C.c_int, [C.c_int32, C.c_uint32, C.POINTER(P)]),
        "sph_dataset_destroy": (C.c_int, [P]),
        "sph_dataset_size": (C.c_uint32, [P]),
        "sph_dataset_record": (C.c_int, [P, P, C.POINTER(SphParams), C.c_float, C.c_uint32, C.c_uint64]),
        "sph_dataset_read": (C.c_int, [P, C.c_uint32, C.c_uint32, _HP, _FP, _FP]),
        "sph_net_train": (C.c_int, [P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64,
                                    _FP]),
        "sph_net_evaluate": (C.c_int, [P, P, C.c_uint32, C.c_uint32, _DP]),
        "sph_features": (C.c_int, [P, C.POINTER(SphParams), C.c_float, _HP]),
        "sph_rollout": (C.c_int, [P, P, C.POINTER(SphParams), C.c_float, C.c_uint32, C.c_float,
                                  C.POINTER(SphRolloutStats)]),
        "sph_net_inference_time": (C.c_int, [P, _DP, C.POINTER(C.c_uint64)]),
        "sph_net_set_smoothing": (C.c_int, [P, C.c_float]),
        "sph_net_set_backoff": (C.c_int, [P, C.c_uint32]),
        "sph_net_set_check": (C.c_int, [P, C.c_uint32]),
        "sph_net_set_features": (C.c_int, [P, C.c_uint32]),
        "sph_net_features": (C.c_uint32, [P]),
        "sph_dataset_set_features": (C.c_int, [P, C.c_uint32]),
        "sph_knn_features": (C.c_int, [P, C.POINTER(SphParams), C.c_float, _HP]),
        "sph_density_deviation": (C.c_int, [P, P, C.POINTER(SphParams), C.c_float, _DP]),
        "sph_check_exact_arithmetic": (C.c_int, [C.c_int32, C.c_uint64, C.POINTER(C.c_uint64)]),
        "sph_short_division": (C.c_int, [P, _UP]),
        "sph_substep_device": (C.c_int, [C.POINTER(SphParams), C.c_float, C.c_uint32, C.c_uint32, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def load(path: str | os.PathLike | None = None):
    """Load libsph_hip.so (built in-tree by ``build.py``). Raises if it is missing:
    there is no fallback path."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(f"{p} is missing: build it with `python -m computer_graphics_lab_00_amd.build`")
    lib = _declare(C.CDLL(str(p)))
    if path is None:
        _lib = lib
    return lib


def check(status: int) -> None:
    if status != SPH_OK:
        msg = load().sph_last_error().decode(errors="replace")
        raise SphError(status, msg)
