"""Batches of independent SPH scenes on one GPU, the workload generator of the batched
benchmark, and the layout autotuner.

``SceneBatch`` wraps the batch C ABI: every kernel of a substep is one launch over all
scenes and a whole step is replayed from a hipGraph. Each scene's results equal those of
the same scene run alone (``FluidSolver``) bit for bit.

The autotuner searches, per scene, over the knobs that change speed but not the physics
of a given configuration: the hash-table size (cellCount, 1x / 2x / 4x the reference
rule bit_ceil(2N)), the spatial-bin tile shape, and neighbour list vs bucket walk. Each
candidate is timed on that scene alone; the choice is cached in JSON keyed by the
scene's signature, so a second run reproduces it without timing.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import sys
import time
from dataclasses import asdict, dataclass, field
from pathlib import Path

import numpy as np

from . import capi
from .capi import SphParams, SphSceneCapacity, SphSettings, check
from .fluid import derive_fluid_params, emit_lattice, lattice_jitter, translation4

_FP = C.POINTER(C.c_float)
_UP = C.POINTER(C.c_uint32)


def _fptr(a):
    return None if a is None else a.ctypes.data_as(_FP)


def bit_ceil(v: int) -> int:
    return 1 << max(0, int(v - 1).bit_length())


@dataclass
class SceneSpec:
    """One scene of a batch: a lattice block in a cubic box."""

    index: int
    count: int
    box: float
    kind: str                      # "dam" (block in the low corner) or "drop" (block raised in the middle)
    lattice: tuple[int, int, int]
    spacing: float
    seed: int
    cell_count: int = 0            # 0: bit_ceil(2 count)
    tile: tuple[int, int, int] = (2, 2, 2)
    neighbour_list: bool = True
    stiffness: float = 2000.0
    viscosity: float = 0.5

    def signature(self) -> str:
        return f"{self.kind}:{self.count}:{self.box:.6f}:{self.spacing:.6e}"

    def cells(self) -> int:
        return self.cell_count or bit_ceil(2 * self.count)

    def params(self) -> SphParams:
        r = self.spacing * 0.5
        nx, ny, nz = self.lattice
        s = SphSettings(emitterHalfExtents=[(nx + 0.5) * self.spacing * 0.5, (ny + 0.5) * self.spacing * 0.5,
                                            (nz + 0.5) * self.spacing * 0.5],
                        particleCount=self.count, particleRadius=r, smoothingRadius=2.0 * self.spacing,
                        restDensity=1000.0, stiffness=self.stiffness, viscosity=self.viscosity,
                        containerMin=[0.0, 0.0, 0.0], containerMax=[self.box] * 3, gravity=[0.0, -9.81, 0.0])
        half = [(n - 1) * self.spacing * 0.5 for n in self.lattice]
        if self.kind == "dam":
            center = [r + half[0], r + half[1], r + half[2]]
        else:
            center = [0.5 * self.box, 0.6 * self.box, 0.5 * self.box]
        return derive_fluid_params(s, translation4(*center), self.count, self.cells())

    def initial_state(self, params: SphParams | None = None):
        p = params or self.params()
        pos = emit_lattice(p, self.count)
        lattice_jitter(pos, self.seed, 0.05 * self.spacing)
        return pos, np.zeros_like(pos)


def batch64(seed: int = 2024, scenes: int = 64, n_min: int = 2000, n_max: int = 200000) -> list[SceneSpec]:
    """The batched workload: particle counts log-uniform in [2k, 200k], box edges uniform
    in [0.5, 2] m, h = 2 x the lattice spacing, alternating dam-break and droplet scenes."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(scenes):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        box = float(rng.uniform(0.5, 2.0))
        side = int(math.ceil(n ** (1.0 / 3.0)))
        nz = int(math.ceil(n / (side * side)))
        spacing = 0.45 * box / side
        out.append(SceneSpec(index=i, count=n, box=box, kind="dam" if i % 2 == 0 else "drop",
                             lattice=(side, side, nz), spacing=spacing, seed=seed * 1000 + i))
    return out


class SceneBatch:
    """Independent scenes advanced together (sph_batch_*)."""

    def __init__(self, specs: list[SceneSpec], device: int = 0, flags: int = 0, schedule: list[int] | None = None):
        self._lib = capi.load()
        self.specs = specs
        caps = (SphSceneCapacity * len(specs))(*[SphSceneCapacity(s.count, s.cells()) for s in specs])
        h = C.c_void_p()
        check(self._lib.sph_batch_create(caps, len(specs), device, flags, C.byref(h)))
        self._h = h
        self.params = []
        for i, s in enumerate(specs):
            p = s.params()
            self.params.append(p)
            check(self._lib.sph_batch_set_params(self._h, i, C.byref(p)))
            tile = (C.c_uint32 * 3)(*s.tile)
            check(self._lib.sph_batch_set_layout(self._h, i, tile, 1 if s.neighbour_list else 0))
        if schedule:
            order = (C.c_uint32 * len(schedule))(*schedule)
            check(self._lib.sph_batch_set_schedule(self._h, order, len(schedule)))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sph_batch_destroy(self._h)
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

    def load_initial(self):
        for i, s in enumerate(self.specs):
            pos, vel = s.initial_state(self.params[i])
            check(self._lib.sph_batch_load_state(self._h, i, _fptr(pos), _fptr(vel), s.count))

    def load_state(self, scene: int, pos4: np.ndarray, vel4: np.ndarray | None = None):
        pos4 = np.ascontiguousarray(pos4, np.float32)
        vel4 = None if vel4 is None else np.ascontiguousarray(vel4, np.float32)
        check(self._lib.sph_batch_load_state(self._h, scene, _fptr(pos4), _fptr(vel4), pos4.shape[0]))

    def set_mode(self, scene: int, mode: int, eos: int = 0):
        """The scene's neighbour-search / arithmetic mode (capi.MODE_*) and equation of state."""
        check(self._lib.sph_batch_set_mode(self._h, scene, mode, eos))

    def step(self, dt: float):
        check(self._lib.sph_batch_step(self._h, dt))

    def synchronize(self):
        check(self._lib.sph_batch_synchronize(self._h))

    def max_bin_occupancy(self) -> int:
        """Particles in the fullest spatial bin of the last grid build, over all scenes."""
        out = C.c_uint32()
        check(self._lib.sph_batch_max_bin_occupancy(self._h, C.byref(out)))
        return int(out.value)

    def state(self, scene: int):
        n = self._lib.sph_batch_particle_count(self._h, scene)
        pos = np.empty((n, 4), np.float32)
        vel = np.empty((n, 4), np.float32)
        check(self._lib.sph_batch_read_particles(self._h, scene, _fptr(pos), _fptr(vel)))
        return pos, vel

    def profile_enable(self, on: bool = True):
        check(self._lib.sph_batch_profile_enable(self._h, 1 if on else 0))

    def profile_read(self):
        cap = 16
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        launches = (C.c_uint32 * cap)()
        w = self._lib.sph_batch_profile_read(self._h, names, ms, launches, cap)
        if w < 0:
            check(w)
        return {names[i].decode(): (ms[i], launches[i]) for i in range(w)}

    @property
    def particles(self) -> int:
        return sum(s.count for s in self.specs)


# ---------------------------------------------------------------------------------------
# autotuner
# ---------------------------------------------------------------------------------------

TILES = [(2, 2, 2), (3, 3, 0), (4, 2, 0)]
CELL_FACTORS = [1, 2, 4]


@dataclass
class TuneResult:
    cell_count: int
    tile: tuple[int, int, int]
    neighbour_list: bool
    ms_per_step: float
    candidates: list = field(default_factory=list)


def _time_scene(spec: SceneSpec, dt: float, steps: int, warmup: int) -> float:
    with SceneBatch([spec]) as b:
        b.load_initial()
        for _ in range(warmup):
            b.step(dt)
        b.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.step(dt)
        b.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps


SCHEDULES = ("scene order", "largest first", "smallest first")


def schedule_order(specs: list[SceneSpec], name: str) -> list[int] | None:
    """The block schedule `name` as a scene order (None: scene order)."""
    if name == "largest first":
        return sorted(range(len(specs)), key=lambda i: (-specs[i].count, i))
    if name == "smallest first":
        return sorted(range(len(specs)), key=lambda i: (specs[i].count, i))
    return None


def _time_batch(specs: list[SceneSpec], schedule: list[int] | None, dt: float, steps: int, warmup: int) -> float:
    with SceneBatch(specs, schedule=schedule) as b:
        b.load_initial()
        for _ in range(warmup):
            b.step(dt)
        b.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.step(dt)
        b.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps


def autotune_schedule(specs: list[SceneSpec], cache_path: str | Path, dt: float = 1.0 / 600.0, steps: int = 10,
                      warmup: int = 3, retune: bool = False) -> list[int] | None:
    """Choose the batch's block schedule (the order of the scenes' workgroups in every
    launch) by timing the whole batch with each of SCHEDULES; cached by the batch's
    signature. Returns the scene order (None: scene order)."""
    cache_path = Path(cache_path)
    cache = json.loads(cache_path.read_text()) if cache_path.exists() else {}
    key = "batch:" + "|".join(f"{s.signature()}:{s.cell_count}:{s.tile}:{s.neighbour_list}" for s in specs)
    if key not in cache or retune:
        timed = {name: _time_batch(specs, schedule_order(specs, name), dt, steps, warmup) for name in SCHEDULES}
        best = min(timed, key=timed.get)
        cache[key] = {"schedule": best, "ms_per_step": timed}
        print(f"[autotune] block schedule: {best} ({timed})", file=sys.stderr, flush=True)
        cache_path.parent.mkdir(parents=True, exist_ok=True)
        cache_path.write_text(json.dumps(cache, indent=1, sort_keys=True) + "\n")
    return schedule_order(specs, cache[key]["schedule"])


def autotune(specs: list[SceneSpec], cache_path: str | Path, dt: float = 1.0 / 600.0, steps: int = 6,
             warmup: int = 2, retune: bool = False) -> list[SceneSpec]:
    """Choose cellCount, tile and list/bucket-walk per scene (each timed alone; the batch's
    block schedule is chosen by autotune_schedule). Cached by scene signature in
    `cache_path`; with a cache hit nothing is timed and the cached choice is returned."""
    cache_path = Path(cache_path)
    cache = json.loads(cache_path.read_text()) if cache_path.exists() else {}
    out = []
    changed = False
    for s in specs:
        key = s.signature()
        if key in cache and not retune:
            c = cache[key]
            out.append(SceneSpec(**{**asdict(s), "cell_count": c["cell_count"], "tile": tuple(c["tile"]),
                                    "neighbour_list": c["neighbour_list"]}))
            continue
        best = None
        tried = []
        base = bit_ceil(2 * s.count)
        for f in CELL_FACTORS:
            for tile in TILES:
                for lst in (True, False):
                    cand = SceneSpec(**{**asdict(s), "cell_count": base * f, "tile": tile, "neighbour_list": lst})
                    ms = _time_scene(cand, dt, steps, warmup)
                    tried.append({"cell_count": base * f, "tile": list(tile), "neighbour_list": lst, "ms": ms})
                    if best is None or ms < best[0]:
                        best = (ms, cand)
        cache[key] = {"cell_count": best[1].cell_count, "tile": list(best[1].tile),
                      "neighbour_list": best[1].neighbour_list, "ms_per_step": best[0], "candidates": tried}
        print(f"[autotune] scene {s.index} ({s.count} particles): cells {best[1].cell_count} tile {best[1].tile} "
              f"list {best[1].neighbour_list} {best[0]:.3f} ms/step", file=sys.stderr, flush=True)
        out.append(best[1])
        changed = True
    if changed:
        cache_path.parent.mkdir(parents=True, exist_ok=True)
        cache_path.write_text(json.dumps(cache, indent=1, sort_keys=True) + "\n")
    return out
