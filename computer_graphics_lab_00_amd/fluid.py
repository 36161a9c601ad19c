"""Host-side mirror of the reference's fluid interface over the C ABI.

``FluidSolver`` has the members of ``physics::FluidSolver`` (src/physics/fluid_sph.h:101-127):
``emit``, ``step``, ``particle_count``, ``particles``, ``previous_particles``,
``keep_rendered``; ``derive_fluid_params`` / ``fluid_substeps`` / ``cell_count_for``
mirror ``physics::deriveFluidParams``, ``physics::fluidSubsteps`` and
``gfx::cellCountFor``. Arrays are numpy float32 (N, 4) in particle-index order,
the layout of ``particles()`` (xyz position, w density).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import capi
from .capi import SphCollider, SphForceField, SphParams, SphSettings, SphSolverConfig, check

_FP = C.POINTER(C.c_float)
_UP = C.POINTER(C.c_uint32)


def _fptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_FP)


def _uptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_UP)


def identity4() -> list[float]:
    return [1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0]


def translation4(x: float, y: float, z: float) -> list[float]:
    m = identity4()
    m[12], m[13], m[14] = x, y, z
    return m


def derive_fluid_params(settings: SphSettings, emitter_world=None, particle_count: int | None = None,
                        cell_count: int = 1024, colliders=(), fields=()) -> SphParams:
    """physics::deriveFluidParams (src/physics/fluid_sph.cpp:65-100)."""
    lib = capi.load()
    world = (C.c_float * 16)(*(emitter_world if emitter_world is not None else identity4()))
    n = settings.particleCount if particle_count is None else particle_count
    col = (SphCollider * max(1, len(colliders)))(*colliders)
    fld = (SphForceField * max(1, len(fields)))(*fields)
    out = SphParams()
    check(lib.sph_derive_params(C.byref(settings), world, n, cell_count, col, len(colliders), fld, len(fields),
                                C.byref(out)))
    return out


def fluid_substeps(params: SphParams, dt: float) -> int:
    """physics::fluidSubsteps (src/physics/fluid_sph.cpp:128-136)."""
    return int(capi.load().sph_substeps(C.byref(params), dt))


def cell_count_for(particles: int) -> int:
    """gfx::cellCountFor (src/gfx/fluid.cpp:50-52)."""
    return int(capi.load().sph_cell_count_for(particles))


def spatial_hash(cx: int, cy: int, cz: int, cell_count: int) -> int:
    return int(capi.load().sph_hash_cell(cx, cy, cz, cell_count))


def lattice_jitter(pos4: np.ndarray, seed: int, amplitude: float) -> None:
    """In-place seeded jitter of pos4[:, :3] (LCG of tests/marching_cubes_test.cpp:265-269)."""
    assert pos4.dtype == np.float32 and pos4.flags.c_contiguous and pos4.shape[1] == 4
    capi.load().sph_lattice_jitter(_fptr(pos4), pos4.shape[0], seed & 0xFFFFFFFF, amplitude)


class FluidSolver:
    """physics::FluidSolver on one MI355X (src/physics/fluid_sph.h:101-127)."""

    def __init__(self, max_particles: int, max_cell_count: int = 1 << 16, device: int = 0,
                 keep_forces: bool = False, flags: int = 0):
        self._lib = capi.load()
        cfg = SphSolverConfig(max_particles, max_cell_count, device,
                              flags | (capi.CFG_KEEP_FORCES if keep_forces else 0))
        h = C.c_void_p()
        check(self._lib.sph_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.max_particles = max_particles
        self.max_cell_count = max_cell_count

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- the reference's members ---
    def emit(self, params: SphParams, count: int) -> None:
        check(self._lib.sph_emit(self._h, C.byref(params), count))

    def step(self, params: SphParams, dt: float) -> None:
        check(self._lib.sph_step(self._h, C.byref(params), dt))

    def particle_count(self) -> int:
        return int(self._lib.sph_particle_count(self._h))

    def particles(self) -> np.ndarray:
        out = np.empty((self.particle_count(), 4), np.float32)
        check(self._lib.sph_read_particles(self._h, _fptr(out), None))
        return out

    def previous_particles(self) -> np.ndarray:
        out = np.empty((self.particle_count(), 4), np.float32)
        check(self._lib.sph_read_previous(self._h, _fptr(out)))
        return out

    def keep_rendered(self) -> None:
        check(self._lib.sph_keep_rendered(self._h))

    # --- extensions ---
    def substep(self, params: SphParams, dt: float) -> None:
        check(self._lib.sph_substep(self._h, C.byref(params), dt))

    def synchronize(self) -> None:
        check(self._lib.sph_synchronize(self._h))

    def max_bin_occupancy(self) -> int:
        """Particles in the fullest spatial bin of the last grid build (synchronous); steps
        are refused with SPH_ERR_DEGENERATE once it reaches capi.MAX_BIN_OCCUPANCY."""
        out = C.c_uint32()
        check(self._lib.sph_max_bin_occupancy(self._h, C.byref(out)))
        return int(out.value)

    def short_division(self) -> bool:
        """Whether the last substep's force pass used its short division and square root
        (the same bits as the IEEE sequences, inside their exact operand ranges)."""
        out = C.c_uint32()
        check(self._lib.sph_short_division(self._h, C.byref(out)))
        return bool(out.value)

    def velocities(self) -> np.ndarray:
        out = np.empty((self.particle_count(), 4), np.float32)
        check(self._lib.sph_read_particles(self._h, None, _fptr(out)))
        return out

    def state(self) -> tuple[np.ndarray, np.ndarray]:
        n = self.particle_count()
        pos = np.empty((n, 4), np.float32)
        vel = np.empty((n, 4), np.float32)
        check(self._lib.sph_read_particles(self._h, _fptr(pos), _fptr(vel)))
        return pos, vel

    def load_state(self, pos4: np.ndarray, vel4: np.ndarray | None = None) -> None:
        pos4 = np.ascontiguousarray(pos4, np.float32)
        vel4 = None if vel4 is None else np.ascontiguousarray(vel4, np.float32)
        check(self._lib.sph_load_state(self._h, _fptr(pos4), _fptr(vel4), pos4.shape[0]))

    def grid(self, cell_count: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(cellOf, cellCounts, cellParticles[C, 32]) of the last substep's FluidGrid."""
        n = self.particle_count()
        cell_of = np.empty(n, np.uint32)
        counts = np.empty(cell_count, np.uint32)
        slots = np.empty((cell_count, capi.SPH_CELL_CAPACITY), np.uint32)
        check(self._lib.sph_read_grid(self._h, _uptr(cell_of), _uptr(counts), _uptr(slots)))
        return cell_of, counts, slots

    def sorted_layout(self, cell_count: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(keys, ids) per device slot and bucketStarts[cellCount] (0xFFFFFFFF if empty)."""
        n = self.particle_count()
        keys = np.empty(n, np.uint32)
        ids = np.empty(n, np.uint32)
        starts = np.empty(cell_count, np.uint32)
        check(self._lib.sph_read_sorted(self._h, _uptr(keys), _uptr(ids), _uptr(starts)))
        return keys, ids, starts

    def forces(self) -> np.ndarray:
        out = np.empty((self.particle_count(), 4), np.float32)
        check(self._lib.sph_read_forces(self._h, _fptr(out)))
        return out

    def set_mode(self, flags: int) -> None:
        check(self._lib.sph_set_mode(self._h, flags))

    def set_eos(self, eos: int) -> None:
        check(self._lib.sph_set_eos(self._h, eos))

    def conservation(self, params: SphParams) -> tuple[float, float, float]:
        out = (C.c_double * 3)()
        check(self._lib.sph_conservation(self._h, C.byref(params), out))
        return out[0], out[1], out[2]

    def profile_enable(self, on: bool = True) -> None:
        check(self._lib.sph_profile_enable(self._h, 1 if on else 0))

    def profile_read(self) -> dict[str, tuple[float, int]]:
        cap = 16
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        launches = (C.c_uint32 * cap)()
        w = self._lib.sph_profile_read(self._h, names, ms, launches, cap)
        if w < 0:
            check(w)
        return {names[i].decode(): (ms[i], launches[i]) for i in range(w)}


@dataclass
class Workload:
    """A dam-break scene: the lattice in the low corner of an axis-aligned box."""

    name: str
    lattice: tuple[int, int, int]
    spacing: float
    box_min: tuple[float, float, float]
    box_max: tuple[float, float, float]
    smoothing: float
    stiffness: float = 2000.0
    viscosity: float = 0.5
    cell_count: int = 1024
    jitter_seed: int = 12345
    jitter: float | None = None   # amplitude in metres; default 0.05 * spacing
    corner: tuple[float, float, float] | None = None   # lowest lattice point; default box_min + r

    @property
    def count(self) -> int:
        return self.lattice[0] * self.lattice[1] * self.lattice[2]

    def params(self) -> SphParams:
        r = self.spacing * 0.5
        nx, ny, nz = self.lattice
        # Half-extents half a spacing past the lattice so floor(2 half / spacing) == n exactly.
        settings = SphSettings(
            emitterHalfExtents=[(nx + 0.5) * self.spacing * 0.5, (ny + 0.5) * self.spacing * 0.5,
                                (nz + 0.5) * self.spacing * 0.5],
            particleCount=self.count, particleRadius=r, smoothingRadius=self.smoothing, restDensity=1000.0,
            stiffness=self.stiffness, viscosity=self.viscosity, containerMin=self.box_min,
            containerMax=self.box_max, gravity=[0.0, -9.81, 0.0])
        low = self.corner if self.corner is not None else [self.box_min[a] + r for a in range(3)]
        center = [low[a] + (self.lattice[a] - 1) * self.spacing * 0.5 for a in range(3)]
        return derive_fluid_params(settings, translation4(*center), self.count, self.cell_count)

    def initial_state(self, params: SphParams | None = None) -> tuple[np.ndarray, np.ndarray]:
        """Lattice positions (FluidSolver::emit's rule, computed on the host in float32)
        plus the seeded jitter; zero velocities."""
        p = params or self.params()
        pos = emit_lattice(p, self.count)
        amp = self.jitter if self.jitter is not None else 0.05 * self.spacing
        if amp > 0:
            lattice_jitter(pos, self.jitter_seed, amp)
        return pos, np.zeros_like(pos)


def emit_lattice(params: SphParams, count: int) -> np.ndarray:
    """FluidSolver::emit (src/physics/fluid_sph.cpp:138-154) for a translation-only or
    general emitter matrix, in float32 with glm's mat4*vec4 order."""
    f = np.float32
    cx, cy, cz = (max(1, int(v)) for v in params.lattice)
    sp = f(params.spacing)
    ext = [f(f(float(c - 1)) * sp) * f(0.5) for c in (cx, cy, cz)]
    i = np.arange(count, dtype=np.int64)
    idx = [i % cx, (i // cx) % cy, i // (cx * cy)]
    local = [(idx[a].astype(np.float32) * sp - ext[a]).astype(np.float32) for a in range(3)]
    m = np.array(params.emitterWorld, dtype=np.float32)
    out = np.zeros((count, 4), np.float32)
    for r in range(3):
        a0 = (m[0 + r] * local[0] + m[4 + r] * local[1]).astype(np.float32)
        a1 = (m[8 + r] * local[2] + m[12 + r] * f(1.0)).astype(np.float32)
        out[:, r] = a0 + a1
    out[:, 3] = f(params.restDensity)
    return out


# The workloads of the benchmark configurations.
def s16k(cell_count: int = 32768) -> Workload:
    """16,384 particles, 32x32x16 at 0.02 m, h = 0.04 m, in the corner of a 1 m box."""
    return Workload("dam_break_16k", (32, 32, 16), 0.02, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.04,
                    cell_count=cell_count)


def m1(cell_count: int = 1 << 21) -> Workload:
    """1,048,576 particles, 128x128x64 at 0.0078 m, h = 0.0156 m, in a 2 x 1 x 1 m box."""
    return Workload("dam_break_1m", (128, 128, 64), 0.0078, (0.0, 0.0, 0.0), (2.0, 1.0, 1.0), 0.0156,
                    cell_count=cell_count)


def r4m(cell_count: int = 1 << 23) -> Workload:
    """4,194,304 particles, 256x128x128 at 0.0049 m, h = 0.0098 m, in a 2.5 x 0.64 x 1.25 m
    box: the 1M dam break's proportions (half the length, the full height, half the
    depth) at a finer resolution."""
    return Workload("dam_break_4m", (256, 128, 128), 0.0049, (0.0, 0.0, 0.0), (2.5, 0.64, 1.25), 0.0098,
                    cell_count=cell_count)
