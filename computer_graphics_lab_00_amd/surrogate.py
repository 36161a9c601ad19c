"""The learned surrogate of the SPH step, host side.

No counterpart in the reference: this sits on top of ``physics::FluidSolver::step``
(src/physics/fluid_sph.cpp:151-303) and keeps that path as its baseline and as the fallback it checks against.

* ``SurrogateNet``: the MLP 64 -> 128 -> 128 -> 128 -> 3 (ReLU) that maps a particle's
  neighbourhood features to its SPH acceleration; trained and evaluated on the GPU by
  the library's BF16 MFMA kernels (FP32 accumulate, FP32 master weights, Adam).
* ``TrainingSet``: device-resident (features, target) rows recorded from plain-SPH
  trajectories; the target is the mean SPH acceleration over a 4-substep macro-step.
* ``training_scene`` / ``generate``: the seeded scenes the trajectories come from —
  random lattices with velocities drawn from N(0, 0.5 m/s).
* ``rollout``: ``FluidSolver.step`` with every 4 substeps replaced by one network
  evaluation, checked by the next density pass and redone by the plain path when the
  mean density excess is above the threshold (1 % by default).
"""
from __future__ import annotations

import ctypes as C
import math
import sys
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import capi
from .capi import SphParams, SphRolloutStats, check
from .fluid import FluidSolver, Workload, cell_count_for, fluid_substeps

FEATURES = 64
OUTPUTS = 3
# MFMA work of one forward pass per particle row, the padded output layer included:
# 2 (64 x 128 + 128 x 128 + 128 x 128 + 128 x 32). The rollout's kernel builds its rows
# from the neighbourhood sums and skips the two k-blocks of layer 0 that only see the
# zero inputs 20..63: 2 (32 x 128 + 128 x 128 + 128 x 128 + 128 x 32) per row.
MLP_FLOP_PER_ROW = 2 * (64 * 128 + 128 * 128 + 128 * 128 + 128 * 32)
ROLLOUT_MLP_FLOP_PER_ROW = 2 * (32 * 128 + 128 * 128 + 128 * 128 + 128 * 32)
# The useful part: 20 live inputs and 3 live outputs of the moment network
# (2 (20 x 128 + 128 x 128 + 128 x 128 + 128 x 3) per row); the rest is padding to MFMA tiles.
USEFUL_MLP_FLOP_PER_ROW = 2 * (20 * 128 + 128 * 128 + 128 * 128 + 128 * 3)
# MFMA work of one training step per batch row: the forward pass, the four weight gradients
# (the same products, reduced over rows) and the input gradients of layers 3, 2 and 1
# (layer 0's input gradient is not needed): 2 MLP_FLOP_PER_ROW + 2 (32 x 128 + 2 x 128 x 128).
TRAIN_FLOP_PER_ROW = 2 * MLP_FLOP_PER_ROW + 2 * (32 * 128 + 128 * 128 + 128 * 128)
MACRO = 4
DEFAULT_THRESHOLD = 0.01
# Rollout policy of the measured configuration, chosen on unseen dam columns of the training
# family (100 steps each), not on the benchmark scene.
DEFAULT_SMOOTHING = 1.0   # XSPH factor
# Dam columns of the training family that training never uses (indices from 5000 on, one in
# three is a column): the held-out scenes the policy was chosen on.
VALIDATION_SCENES = (5000, 5003, 5006, 5009)
DEFAULT_BACKOFF = 8       # plain macro-steps after repeated failed checks, at most
_FP = C.POINTER(C.c_float)
_HP = C.POINTER(C.c_uint16)
_DP = C.POINTER(C.c_double)

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def _hp(a):
    return None if a is None else a.ctypes.data_as(_HP)


def _fp(a):
    return None if a is None else a.ctypes.data_as(_FP)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """The first `count` outputs of splitmix64 (Steele, Lea & Flood 2014) from `seed`."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + _GOLDEN * np.arange(1, count + 1, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def gaussian(seed: int, count: int, sigma: float = 1.0) -> np.ndarray:
    """`count` N(0, sigma^2) samples: Box-Muller (1958) on pairs of splitmix64 uniforms."""
    pairs = (count + 1) // 2
    z = splitmix64(seed, 2 * pairs)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log1p(-u1))   # 1 - u1 is in (0, 1]
    g = np.empty(2 * pairs)
    g[0::2] = r * np.cos(2.0 * math.pi * u2)
    g[1::2] = r * np.sin(2.0 * math.pi * u2)
    return (g[:count] * sigma).astype(np.float32)


class SurrogateNet:
    """The network on one GPU; parameters W[out][in] then b[out] per layer."""

    def __init__(self, seed: int = 1, device: int = 0):
        self._lib = capi.load()
        h = C.c_void_p()
        check(self._lib.sph_net_create(device, seed, C.byref(h)))
        self._h = h
        self.param_count = int(self._lib.sph_net_param_count())
        self.max_backoff = 0

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sph_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        out = np.empty(self.param_count, np.float32)
        check(self._lib.sph_net_get_params(self._h, _fp(out)))
        return out

    @params.setter
    def params(self, values) -> None:
        v = np.ascontiguousarray(values, dtype=np.float32)
        if v.shape != (self.param_count,):
            raise ValueError(f"expected {self.param_count} parameters")
        check(self._lib.sph_net_set_params(self._h, _fp(v)))

    def save(self, path) -> None:
        check(self._lib.sph_net_save(self._h, str(path).encode()))

    def load(self, path) -> None:
        check(self._lib.sph_net_load(self._h, str(path).encode()))

    def forward(self, x_bits: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x_bits, dtype=np.uint16)
        y = np.empty((x.shape[0], OUTPUTS), np.float32)
        check(self._lib.sph_net_forward(self._h, _hp(x), x.shape[0], _fp(y)))
        return y

    def gradients(self, x_bits: np.ndarray, targets: np.ndarray) -> tuple[np.ndarray, float]:
        x = np.ascontiguousarray(x_bits, dtype=np.uint16)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        g = np.empty(self.param_count, np.float32)
        loss = C.c_float()
        check(self._lib.sph_net_gradients(self._h, _hp(x), _fp(t), x.shape[0], _fp(g), C.byref(loss)))
        return g, float(loss.value)

    def train_batch(self, x_bits: np.ndarray, targets: np.ndarray, lr: float = 1e-3) -> float:
        x = np.ascontiguousarray(x_bits, dtype=np.uint16)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        loss = C.c_float()
        check(self._lib.sph_net_train_batch(self._h, _hp(x), _fp(t), x.shape[0], lr, C.byref(loss)))
        return float(loss.value)

    def train(self, data: "TrainingSet", begin: int, end: int, batch: int, steps: int, lr: float = 1e-3,
              seed: int = 7, losses: bool = False) -> np.ndarray | None:
        out = np.empty(steps, np.float32) if losses else None
        check(self._lib.sph_net_train(self._h, data._h, begin, end, batch, steps, lr, seed, _fp(out)))
        return out

    def evaluate(self, data: "TrainingSet", begin: int, end: int) -> dict[str, float]:
        out = (C.c_double * 3)()
        check(self._lib.sph_net_evaluate(self._h, data._h, begin, end, out))
        return {"network": out[0], "gravity_only": out[1], "hold_first_substep": out[2]}

    def set_smoothing(self, smoothing: float) -> None:
        """XSPH velocity smoothing at the start of each rollout macro-step (0: off)."""
        check(self._lib.sph_net_set_smoothing(self._h, smoothing))

    @property
    def feature_kind(self) -> int:
        """capi.FEATURES_MOMENTS or capi.FEATURES_KNN8 (kept by save / load)."""
        return int(self._lib.sph_net_features(self._h))

    @feature_kind.setter
    def feature_kind(self, kind: int) -> None:
        check(self._lib.sph_net_set_features(self._h, kind))

    def set_check(self, sides: int) -> None:
        """Sides of the rollout's density check: capi.CHECK_COMPRESSION, CHECK_RAREFACTION
        or both (the default)."""
        check(self._lib.sph_net_set_check(self._h, sides))

    def set_backoff(self, max_backoff: int) -> None:
        """After a failed check, run 1, 2, 4, ... (at most max_backoff) macro-steps plain."""
        check(self._lib.sph_net_set_backoff(self._h, max_backoff))
        self.max_backoff = max_backoff

    def inference_time(self) -> tuple[float, int]:
        ms = C.c_double()
        rows = C.c_uint64()
        check(self._lib.sph_net_inference_time(self._h, C.byref(ms), C.byref(rows)))
        return ms.value, rows.value


class TrainingSet:
    """(feature row, target) pairs resident on the GPU."""

    def __init__(self, capacity: int, device: int = 0, feature_kind: int = capi.FEATURES_MOMENTS):
        self._lib = capi.load()
        h = C.c_void_p()
        check(self._lib.sph_dataset_create(device, capacity, C.byref(h)))
        self._h = h
        self.capacity = capacity
        check(self._lib.sph_dataset_set_features(self._h, feature_kind))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sph_dataset_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self) -> int:
        return int(self._lib.sph_dataset_size(self._h))

    def record(self, solver: FluidSolver, params: SphParams, dt_sub: float, samples: int, seed: int) -> None:
        """One macro-step of the plain path (advances the solver by 4 substeps)."""
        check(self._lib.sph_dataset_record(self._h, solver._h, C.byref(params), dt_sub, samples, seed))

    def read(self, begin: int = 0, count: int | None = None):
        count = len(self) - begin if count is None else count
        x = np.empty((count, FEATURES), np.uint16)
        t = np.empty((count, OUTPUTS), np.float32)
        t0 = np.empty((count, OUTPUTS), np.float32)
        check(self._lib.sph_dataset_read(self._h, begin, count, _hp(x), _fp(t), _fp(t0)))
        return x, t, t0


def features(solver: FluidSolver, params: SphParams, dt_sub: float, kind: int = capi.FEATURES_MOMENTS) -> np.ndarray:
    """Feature rows (BF16 bits) of every particle of the current state, particle order."""
    n = solver.particle_count()
    x = np.empty((n, FEATURES), np.uint16)
    fn = capi.load().sph_knn_features if kind == capi.FEATURES_KNN8 else capi.load().sph_features
    check(fn(solver._h, C.byref(params), dt_sub, _hp(x)))
    return x


def density_deviation(solver: FluidSolver, net: SurrogateNet, params: SphParams, dt_sub: float) -> tuple[float, float]:
    """(mean max(rho / rho0 - 1, 0), mean |rho / rho0 - 1|) of the current state."""
    out = (C.c_double * 2)()
    check(capi.load().sph_density_deviation(solver._h, net._h, C.byref(params), dt_sub, out))
    return out[0], out[1]


def rollout(solver: FluidSolver, net: SurrogateNet, params: SphParams, dt: float, steps: int = 1,
            threshold: float = DEFAULT_THRESHOLD, stats: SphRolloutStats | None = None) -> SphRolloutStats:
    """`steps` calls of FluidSolver.step(dt) with the surrogate; returns the counters
    (accumulated into `stats` when given)."""
    st = stats if stats is not None else SphRolloutStats()
    check(capi.load().sph_rollout(solver._h, net._h, C.byref(params), dt, steps, threshold, C.byref(st)))
    return st


# ---- training scenes ----

LATTICES = ((40, 40, 40), (64, 32, 32), (32, 64, 32), (32, 32, 64))
RADII = (0.0098, 0.0156)   # kernel radii of the 4M and 1M dam breaks; spacing = h / 2
DAM_COLUMNS = ((16, 128, 32), (32, 128, 16), (24, 96, 28))


@dataclass
class TrainingScene:
    workload: Workload
    velocity_seed: int
    warmup_steps: int
    at_rest: bool = False

    def initial_state(self):
        p = self.workload.params()
        pos, _ = self.workload.initial_state(p)
        vel = np.zeros_like(pos)
        if not self.at_rest:
            vel[:, :3] = gaussian(self.velocity_seed, pos.shape[0] * 3, 0.5).reshape(-1, 3)
        return p, pos, vel


def training_scene(index: int, seed: int = 2024, dam_every: int = 3, dam_warmup: int = 48) -> TrainingScene:
    """Scene `index` of the seeded training family, at one of the two resolutions. One scene
    in `dam_every` is a dam break at rest — one of three 64k columns (16 x 128 x 32,
    32 x 128 x 16, 24 x 96 x 28) in the corner of a box 4-8x its length and 1-2x its depth —
    so the collapsing, hydrostatically loaded columns the benchmarks run are in the data;
    the others are one of four 64k lattices placed at a random spot of a box 1.2-2x its
    size (resting on the floor half of the time) with velocities from N(0, 0.5 m/s).
    Recorded after 0-47 plain steps."""
    u = (splitmix64(seed * 1000003 + index, 8) >> np.uint64(11)).astype(np.float64) / 9007199254740992.0
    h = RADII[(index // len(LATTICES)) % len(RADII)]
    spacing = h * 0.5
    r = spacing * 0.5
    dam = index % dam_every == dam_every - 1
    if dam:
        lattice = DAM_COLUMNS[(index // dam_every) % len(DAM_COLUMNS)]
        extent = [n * spacing for n in lattice]
        box = [extent[0] * (4.0 + 4.0 * u[0]), extent[1] * 1.05, extent[2] * (1.0 + 1.0 * u[2])]
        corner = [r, r, r]
    else:
        lattice = LATTICES[index % len(LATTICES)]
        extent = [n * spacing for n in lattice]
        box = [extent[a] * (1.2 + 0.8 * u[a]) for a in range(3)]
        corner = []
        for a in range(3):
            slack = box[a] - extent[a] - 2 * r
            lo = r + slack * u[3 + a]
            if a == 1 and u[6] < 0.5:
                lo = r
            corner.append(lo)
    count = lattice[0] * lattice[1] * lattice[2]
    wl = Workload(f"train_{index}", lattice, spacing, (0.0, 0.0, 0.0), tuple(box), h,
                  cell_count=max(1024, 1 << int(math.ceil(math.log2(2 * count)))),
                  jitter_seed=(seed + index) & 0xFFFFFFFF, corner=tuple(corner))
    warm = dam_warmup if dam else 48
    return TrainingScene(wl, velocity_seed=seed * 7919 + index, warmup_steps=int(u[7] * warm), at_rest=dam)


# Position noise (fraction of the lattice spacing) and velocity noise (m/s) added before
# the records of a scene, cycled record by record: the network also sees states a little
# off the plain trajectory, with the plain path's response to them as the target, so a
# rollout that drifts is pulled back instead of running away (noise injection,
# Sanchez-Gonzalez et al. 2020).
NOISE = ((0.0, 0.0), (0.005, 0.02), (0.01, 0.05), (0.02, 0.1))


def perturb(solver: FluidSolver, spacing: float, pos_noise: float, vel_noise: float, seed: int) -> None:
    if pos_noise <= 0.0 and vel_noise <= 0.0:
        return
    pos, vel = solver.state()
    n = pos.shape[0]
    pos[:, :3] += gaussian(seed, 3 * n, pos_noise * spacing).reshape(n, 3)
    vel[:, :3] += gaussian(seed ^ 0x5DEECE66D, 3 * n, vel_noise).reshape(n, 3)
    solver.load_state(pos, vel)


def generate(data: TrainingSet, scenes: int, records: int = 12, samples: int = 4096, dt: float = 1.0 / 600.0,
             gap_steps: int = 2, seed: int = 2024, first: int = 0, noise=NOISE, net: "SurrogateNet | None" = None,
             dam_every: int = 3, dam_warmup: int = 48, log=None, solver_flags: int = 0) -> list[tuple[int, int]]:
    """Records `records` macro-steps of each of scenes first..first+scenes-1 into `data`;
    returns each scene's row range. Between records the scene runs `gap_steps` steps so
    the rows of one scene are not all alike; before record r the state is perturbed by
    noise[r % len(noise)]. With `net`, the gap steps are rollout steps of that network
    with the check off, so the rows come from the states the surrogate itself reaches
    (data aggregation, Ross et al. 2011) with the plain path's response as target.
    `solver_flags` are SPH_CFG_* flags of the recording solver (e.g. CFG_NO_GRAPH)."""
    ranges = []
    solver = None
    cap = 0
    for i in range(first, first + scenes):
        sc = training_scene(i, seed, dam_every, dam_warmup)
        p, pos, vel = sc.initial_state()
        n = pos.shape[0]
        if solver is None or n > cap or sc.workload.cell_count > solver.max_cell_count:
            if solver is not None:
                solver.close()
            cap = max(n, 65536)
            solver = FluidSolver(cap, max(sc.workload.cell_count, 1 << 18), keep_forces=True, flags=solver_flags)
        solver.load_state(pos, vel)
        sub = fluid_substeps(p, dt)
        dt_sub = min(dt, 1.0 / 30.0) / sub
        for _ in range(sc.warmup_steps):
            solver.step(p, dt)
        begin = len(data)
        for r in range(records):
            if noise:
                pn, vn = noise[r % len(noise)]
                perturb(solver, sc.workload.spacing, pn, vn, seed=(seed << 24) + i * 1000 + r)
            data.record(solver, p, dt_sub, min(samples, n), seed=(seed << 20) + i * 1000 + r)
            if net is not None:
                rollout(solver, net, p, dt, gap_steps, float("inf"))
            else:
                for _ in range(gap_steps):
                    solver.step(p, dt)
        ranges.append((begin, len(data)))
        if log is not None and (i - first) % 20 == 19:
            print(f"[surrogate] recorded {i - first + 1}/{scenes} scenes, {len(data)} rows", file=log, flush=True)
    if solver is not None:
        solver.close()
    return ranges


def default_weights() -> Path:
    return Path(__file__).resolve().parent / "tuned" / "surrogate_mlp.bin"


def knn_weights() -> Path:
    """The network trained on the 8-nearest-neighbour rows."""
    return Path(__file__).resolve().parent / "tuned" / "surrogate_knn8.bin"


def record_workload(data: TrainingSet, workload: Workload, records: int, samples: int, warmup_steps: int = 0,
                    gap_steps: int = 2, dt: float = 1.0 / 600.0, seed: int = 99,
                    solver_flags: int = 0) -> tuple[int, int]:
    """Rows recorded from any workload's plain trajectory (e.g. the dam breaks the
    rollout is measured on), for evaluating the network on that distribution."""
    p = workload.params()
    pos, vel = workload.initial_state(p)
    n = pos.shape[0]
    solver = FluidSolver(n, workload.cell_count, keep_forces=True, flags=solver_flags)
    solver.load_state(pos, vel)
    sub = fluid_substeps(p, dt)
    dt_sub = float(np.float32(np.float32(min(dt, 1.0 / 30.0)) / np.float32(sub)))
    for _ in range(warmup_steps):
        solver.step(p, dt)
    begin = len(data)
    for r in range(records):
        data.record(solver, p, dt_sub, min(samples, n), seed=seed * 1000 + r)
        for _ in range(gap_steps):
            solver.step(p, dt)
    solver.close()
    return begin, len(data)


def train_default(net: SurrogateNet, scenes: int = 200, held_out: int = 20, records: int = 12,
                  samples: int = 4096, steps: int = 6000, batch: int = 8192, lr: float = 1e-3,
                  eval_workloads=(), dagger_rounds: int = 0, dam_every: int = 3, dam_warmup: int = 48,
                  log=sys.stderr, solver_flags: int = 0) -> dict:
    """The training recipe: `scenes` trajectories, the last `held_out` kept for evaluation;
    Adam at lr for `steps` batches, then lr / 10 for a quarter as many. Each of
    `dagger_rounds` rounds then adds the same number of scenes recorded along the
    network's own rollouts (seeds not used before) and trains again on all rows. Each
    workload in `eval_workloads` contributes 8 records of 16384 rows from its own plain
    trajectory, used for evaluation only. The rows are of the network's feature kind."""
    per_round = scenes * records * samples
    data = TrainingSet(per_round * (1 + dagger_rounds) + len(eval_workloads) * 8 * 16384,
                       feature_kind=net.feature_kind)
    ranges = generate(data, scenes, records, samples, dam_every=dam_every, dam_warmup=dam_warmup, log=log,
                      solver_flags=solver_flags)
    split = ranges[scenes - held_out][0] if held_out else len(data)
    end = len(data)
    evals = [(wl.name, record_workload(data, wl, 8, 16384, warmup_steps=4, gap_steps=4, solver_flags=solver_flags))
             for wl in eval_workloads]
    t0 = time.perf_counter()
    losses = net.train(data, 0, split, batch, steps, lr, losses=True)
    tail = net.train(data, 0, split, batch, max(1, steps // 4), lr * 0.1, seed=11, losses=True)
    seconds = time.perf_counter() - t0   # every training step ends in a synchronisation
    report = {"rows": len(data), "train_rows": split, "loss_first": float(losses[0]),
              "loss_last": float(tail[-1])}
    # Whole training steps (gather, products, reductions, Adam and the per-step loss
    # readback) against the MFMA work of the padded batch.
    train_steps = steps + max(1, steps // 4)
    rows_padded = (batch + 31) // 32 * 32
    report["train_steps"] = train_steps
    report["train_ms_per_step"] = seconds / train_steps * 1e3
    report["train_tflops"] = TRAIN_FLOP_PER_ROW * rows_padded * train_steps / seconds / 1e12
    train_ranges = [(0, split)]
    for rnd in range(dagger_rounds):
        b0 = len(data)
        generate(data, scenes, records, samples, first=scenes * (rnd + 1), net=net, dam_every=dam_every,
                 dam_warmup=dam_warmup, log=log, solver_flags=solver_flags)
        train_ranges.append((b0, len(data)))
        # rows of all rounds: train on a window ending at the newest rows and covering them all
        for k, (r0, r1) in enumerate(train_ranges):
            net.train(data, r0, r1, batch, steps // (2 * len(train_ranges)), lr * 0.3, seed=100 * rnd + k)
        tail = net.train(data, b0, len(data), batch, max(1, steps // 8), lr * 0.1, seed=1000 + rnd, losses=True)
        report[f"dagger_{rnd}_loss_last"] = float(tail[-1])
        report[f"dagger_{rnd}_rows"] = len(data) - b0
    if held_out:
        report["held_out"] = net.evaluate(data, split, end)
    report["train"] = net.evaluate(data, 0, min(split, 1 << 20))
    for name, (b0, b1) in evals:
        report["eval_" + name] = net.evaluate(data, b0, b1)
    data.close()
    return report


# ---- evaluation: plain path against the surrogate on one scene ----

def densities(probe: FluidSolver, params: SphParams, pos4: np.ndarray, vel4: np.ndarray, dt_sub: float) -> np.ndarray:
    """Per-particle densities of a state as the plain density pass computes them (the density
    a substep from that state starts with), in particle order."""
    probe.load_state(pos4, vel4)
    probe.substep(params, dt_sub)
    return probe.particles()[:, 3].copy()


def step_error(solver: FluidSolver, net: "SurrogateNet", probe: FluidSolver, params: SphParams, dt: float,
               dt_sub: float, threshold: float, stats: SphRolloutStats) -> dict:
    """One rollout step of `solver` together with the plain step from the identical starting
    state: per-particle |rho_surrogate - rho_plain| / rho0 of the two resulting states."""
    pos0, vel0 = solver.state()
    rollout(solver, net, params, dt, 1, threshold, stats)
    ps, vs = solver.state()
    probe.load_state(pos0, vel0)
    probe.step(params, dt)
    pp, vp = probe.state()
    rs = densities(probe, params, ps, vs, dt_sub)
    rp = densities(probe, params, pp, vp, dt_sub)
    e = np.abs(rs.astype(np.float64) - rp.astype(np.float64)) / float(params.restDensity)
    return {"mean": float(e.mean()), "p99": float(np.percentile(e, 99)), "max": float(e.max())}


def _energy(pos, vel, params: SphParams) -> float:
    """Sum of 1/2 m |v|^2 - m g.x in FP64 (the conservation sum of the plain path)."""
    m = float(params.particleMass)
    g = np.array(params.gravity, np.float64)
    v = vel[:, :3].astype(np.float64)
    x = pos[:, :3].astype(np.float64)
    return float(0.5 * m * np.sum(v * v) - m * np.sum(x @ g))


def compare(workload: Workload, net: SurrogateNet, steps: int, dt: float = 1.0 / 600.0,
            threshold: float = DEFAULT_THRESHOLD, stats_every: int = 1, unchecked: bool = True, log=None,
            progress_every: int = 20, error_every: int = 5) -> dict:
    """Runs `steps` FluidSolver.step(dt) calls of `workload` with the plain path and with
    the surrogate from the same initial state and returns timing and density/energy
    statistics. The timed runs carry no statistics work; the statistics come from
    separate, untimed runs."""
    p = workload.params()
    pos0, vel0 = workload.initial_state(p)
    n = pos0.shape[0]
    sub = fluid_substeps(p, dt)
    dt_sub = float(np.float32(np.float32(min(dt, 1.0 / 30.0)) / np.float32(sub)))
    solver = FluidSolver(n, workload.cell_count)
    result = {"particles": n, "steps": steps, "substeps_per_step": sub, "threshold": threshold}

    def say(msg):
        if log is not None:
            print(f"[surrogate] {msg}", file=log, flush=True)

    probe = None
    step_errors = []
    occupancy = [0]   # fullest spatial bin seen at the untimed runs' progress points
    done = [0]        # steps the current run has completed

    def run(use_net: bool, timed: bool, label: str, errors: bool = False):
        nonlocal threshold, probe
        solver.load_state(pos0, vel0)
        net.set_backoff(net.max_backoff)   # every run starts with no back-off pending
        st = SphRolloutStats()
        series = []
        solver.synchronize()
        t0 = time.perf_counter()
        done[0] = 0
        for i in range(steps):
            done[0] = i
            if use_net and errors and (i + 1) % error_every == 0:
                # this step and the plain step from the same state, compared particle by particle
                if probe is None:
                    probe = FluidSolver(n, workload.cell_count)
                step_errors.append({"step": i + 1, **step_error(solver, net, probe, p, dt, dt_sub, threshold, st)})
            elif use_net:
                rollout(solver, net, p, dt, 1, threshold, st)
            else:
                solver.step(p, dt)
            if not timed and (i + 1) % stats_every == 0:
                series.append(density_deviation(solver, net, p, dt_sub))
            # progress of the untimed runs (the host waits for the device here, so a slow run
            # shows its pace instead of going quiet); timed runs report once at their end
            if not timed and (i + 1) % progress_every == 0:
                occupancy[0] = max(occupancy[0], solver.max_bin_occupancy())
                say(f"{label}: step {i + 1}/{steps}, {time.perf_counter() - t0:.1f} s")
        solver.synchronize()
        return time.perf_counter() - t0, st, series

    run(False, False, "warm-up (plain)")   # plans, graphs; untimed, so it reports progress
    say("warm-up done")
    t_plain, _, _ = run(False, True, "plain")
    say(f"plain: {1e3 * t_plain / steps:.3f} ms/step")
    net.inference_time()
    t_sur, st, _ = run(True, True, "surrogate")
    say(f"surrogate: {1e3 * t_sur / steps:.3f} ms/step")
    mlp_ms, mlp_rows = net.inference_time()
    if unchecked:
        # the same rollout with the check never failing: the cost of the surrogate alone
        saved = threshold
        threshold = float("inf")
        try:
            t_free, _, _ = run(True, True, "surrogate, check off")
            say(f"surrogate, check off: {1e3 * t_free / steps:.3f} ms/step")
            result["unchecked_ms_per_step"] = 1e3 * t_free / steps
            result["unchecked_speedup"] = t_plain / t_free
            _, _, free_series = run(True, False, "surrogate statistics, check off")
            result["unchecked_density_excess"] = [float(v[0]) for v in free_series]
        except capi.SphError as e:
            if e.status != capi.ERR_DEGENERATE:
                raise
            # unchecked, the network's trajectory collapsed: the solver refused the state
            result["unchecked_degenerate_at_step"] = done[0] + 1
            result["unchecked_degenerate"] = str(e)
            say(f"surrogate, check off: refused at step {done[0] + 1}: {e}")
        finally:
            threshold = saved
            net.inference_time()
    result["max_bin_occupancy"] = occupancy[0]
    result["plain_ms_per_step"] = 1e3 * t_plain / steps
    result["surrogate_ms_per_step"] = 1e3 * t_sur / steps
    result["plain_particle_steps_per_s"] = n * steps / t_plain
    result["surrogate_particle_steps_per_s"] = n * steps / t_sur
    result["speedup"] = t_plain / t_sur
    result["macro_steps"] = int(st.macroSteps)
    result["fallbacks"] = int(st.fallbacks)
    result["fallback_rate"] = st.fallbacks / max(1, st.macroSteps)
    result["backed_off"] = int(st.backedOff)
    result["network_macro_steps"] = int(st.macroSteps - st.fallbacks)
    result["network_share"] = (st.macroSteps - st.fallbacks) / max(1, st.macroSteps + st.backedOff)
    result["mlp_ms"] = mlp_ms
    result["mlp_rows"] = mlp_rows
    if mlp_ms > 0:
        flop = ROLLOUT_MLP_FLOP_PER_ROW * mlp_rows
        result["mlp_tflops"] = flop / (mlp_ms * 1e-3) / 1e12
        # one network pass per macro-step; the rollout times one pass in eight
        result["mlp_ms_per_macro_step"] = mlp_ms * n / mlp_rows
    if log is not None:
        print(f"[surrogate] timed: plain {result['plain_ms_per_step']:.3f} ms/step, surrogate "
              f"{result['surrogate_ms_per_step']:.3f} ms/step, fallbacks {st.fallbacks}/{st.macroSteps}",
              file=log, flush=True)
    _, _, plain_series = run(False, False, "plain statistics")
    e_plain = _energy(*solver.state(), p)
    _, st2, sur_series = run(True, False, "surrogate statistics", errors=True)
    e_sur = _energy(*solver.state(), p)
    e0 = _energy(pos0, vel0, p)
    ps = np.array(plain_series)
    ss = np.array(sur_series)
    result["plain_density_excess_mean"] = float(ps[:, 0].mean())
    result["plain_density_abs_mean"] = float(ps[:, 1].mean())
    result["surrogate_density_excess_mean"] = float(ss[:, 0].mean())
    result["surrogate_density_abs_mean"] = float(ss[:, 1].mean())
    result["density_error_vs_plain"] = float(np.mean(np.abs(ss[:, 1] - ps[:, 1])))
    result["energy_initial"] = e0
    result["energy_drift_plain"] = (e_plain - e0) / abs(e0)
    result["energy_drift_surrogate"] = (e_sur - e0) / abs(e0)
    result["series_plain"] = ps.tolist()
    result["series_surrogate"] = ss.tolist()
    # mean deficit mean(max(1 - rho / rho0, 0)) = mean |dev| - mean excess
    result["plain_density_deficit_mean"] = float((ps[:, 1] - ps[:, 0]).mean())
    result["surrogate_density_deficit_mean"] = float((ss[:, 1] - ss[:, 0]).mean())
    result["fallbacks_rarefaction"] = int(st.rarefactionFallbacks)
    result["max_rarefaction_growth"] = float(st.maxRarefactionGrowth)
    if step_errors:
        result["step_error_vs_plain"] = {
            "sampled_steps": len(step_errors), "every": error_every,
            "mean": float(np.mean([e["mean"] for e in step_errors])),
            "p99_mean": float(np.mean([e["p99"] for e in step_errors])),
            "max_of_means": float(max(e["mean"] for e in step_errors)),
            "max": float(max(e["max"] for e in step_errors)),
            "series": step_errors,
        }
    if probe is not None:
        probe.close()
    solver.close()
    return result


def time_rollout(workload: Workload, net: SurrogateNet, steps: int, plain_steps: int = 5, warm_steps: int = 5,
                 dt: float = 1.0 / 600.0, threshold: float = DEFAULT_THRESHOLD) -> dict:
    """Wall time of `steps` surrogate steps of `workload` taken after `plain_steps` plain steps
    and `warm_steps` surrogate steps from its initial state (a window of the rollout further
    into the flow, for traces of the steady macro-step), with the rollout's statistics."""
    p = workload.params()
    pos0, vel0 = workload.initial_state(p)
    solver = FluidSolver(pos0.shape[0], workload.cell_count)
    try:
        solver.load_state(pos0, vel0)
        for _ in range(plain_steps):
            solver.step(p, dt)
        net.set_backoff(net.max_backoff)
        if warm_steps:
            rollout(solver, net, p, dt, warm_steps, threshold, SphRolloutStats())
        solver.synchronize()
        st = SphRolloutStats()
        t0 = time.perf_counter()
        rollout(solver, net, p, dt, steps, threshold, st)
        solver.synchronize()
        t = time.perf_counter() - t0
    finally:
        solver.close()
    return {"particles": int(pos0.shape[0]), "first_step": plain_steps + warm_steps + 1, "steps": steps,
            "ms_per_step": 1e3 * t / steps, "macro_steps": int(st.macroSteps), "fallbacks": int(st.fallbacks),
            "backed_off": int(st.backedOff)}


def main(argv=None) -> int:
    import argparse
    import json
    ap = argparse.ArgumentParser(prog="python -m computer_graphics_lab_00_amd.surrogate")
    sp = ap.add_subparsers(dest="cmd", required=True)
    t = sp.add_parser("train", help="record trajectories and train the network on the GPU")
    t.add_argument("--scenes", type=int, default=240)
    t.add_argument("--held-out", type=int, default=20)
    t.add_argument("--records", type=int, default=12)
    t.add_argument("--samples", type=int, default=4096)
    t.add_argument("--steps", type=int, default=24000)
    t.add_argument("--batch", type=int, default=16384)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--out", default=str(default_weights()))
    t.add_argument("--dagger", type=int, default=0, help="rounds of training on the network's own rollouts")
    t.add_argument("--dam-every", type=int, default=3, help="one training scene in this many is a dam-break column")
    t.add_argument("--dam-warmup", type=int, default=48, help="dam columns are recorded after 0 .. this many - 1 steps")
    t.add_argument("--eval", nargs="*", default=["m1"], choices=("m1", "r4m", "s16k"),
                   help="dam breaks whose own trajectories the network is also evaluated on")
    t.add_argument("--features", choices=("moments", "knn8"), default="moments",
                   help="network input: neighbourhood moments or the 8 nearest neighbours")
    t.add_argument("--eager", action="store_true",
                   help="record with kernels launched one by one instead of replayed step graphs")
    e = sp.add_parser("compare", help="plain path against the surrogate on a dam break")
    e.add_argument("--weights", default=str(default_weights()))
    e.add_argument("--workload", choices=("m1", "r4m", "s16k"), default="m1")
    e.add_argument("--steps", type=int, default=20)
    e.add_argument("--threshold", type=float, default=DEFAULT_THRESHOLD)
    e.add_argument("--json", default=None)
    e.add_argument("--smoothing", type=float, default=DEFAULT_SMOOTHING)
    e.add_argument("--backoff", type=int, default=DEFAULT_BACKOFF)
    e.add_argument("--no-unchecked", action="store_true", help="skip the runs with the check switched off")
    e.add_argument("--stats-every", type=int, default=1, help="density statistics every this many steps")
    e.add_argument("--error-every", type=int, default=5,
                   help="per-particle error against plain from the same state every this many steps")
    e.add_argument("--progress-every", type=int, default=20)
    r = sp.add_parser("time", help="wall time of a window of the surrogate rollout further into a dam break")
    r.add_argument("--weights", default=str(default_weights()))
    r.add_argument("--workload", choices=("m1", "r4m", "s16k"), default="m1")
    r.add_argument("--steps", type=int, default=10)
    r.add_argument("--plain-steps", type=int, default=5, help="plain steps before the window")
    r.add_argument("--warm-steps", type=int, default=5, help="surrogate steps before the window")
    r.add_argument("--smoothing", type=float, default=DEFAULT_SMOOTHING)
    r.add_argument("--backoff", type=int, default=DEFAULT_BACKOFF)
    v = sp.add_parser("validate", help="checked rollouts of held-out dam columns of the training family")
    v.add_argument("--weights", default=str(default_weights()))
    v.add_argument("--scenes", type=int, nargs="*", default=list(VALIDATION_SCENES))
    v.add_argument("--steps", type=int, default=100)
    v.add_argument("--smoothing", type=float, default=DEFAULT_SMOOTHING)
    v.add_argument("--backoff", type=int, default=DEFAULT_BACKOFF)
    a = ap.parse_args(argv)
    if a.cmd == "validate":
        net = SurrogateNet()
        net.load(a.weights)
        net.set_smoothing(a.smoothing)
        net.set_backoff(a.backoff)
        speedups = []
        for index in a.scenes:
            sc = training_scene(index)
            r = compare(sc.workload, net, a.steps, unchecked=False)
            speedups.append(r["speedup"])
            print(json.dumps({"scene": index, "dam_column": sc.at_rest, "lattice": list(sc.workload.lattice),
                              "h": sc.workload.smoothing, "speedup": r["speedup"], "fallbacks": r["fallbacks"],
                              "macro_steps": r["macro_steps"], "density_error_vs_plain": r["density_error_vs_plain"]}),
                  flush=True)
        print(json.dumps({"smoothing": a.smoothing, "backoff": a.backoff, "steps": a.steps,
                          "mean_speedup": sum(speedups) / max(1, len(speedups))}), flush=True)
        return 0
    if a.cmd == "train":
        net = SurrogateNet(seed=a.seed)
        net.feature_kind = capi.FEATURES_KNN8 if a.features == "knn8" else capi.FEATURES_MOMENTS
        t0 = time.perf_counter()
        from . import fluid
        wls = [{"m1": fluid.m1, "r4m": fluid.r4m, "s16k": fluid.s16k}[w]() for w in a.eval]
        rep = train_default(net, a.scenes, a.held_out, a.records, a.samples, a.steps, a.batch, a.lr, wls,
                            a.dagger, a.dam_every, a.dam_warmup,
                            solver_flags=capi.CFG_NO_GRAPH if a.eager else 0)
        rep["seconds"] = time.perf_counter() - t0
        Path(a.out).parent.mkdir(parents=True, exist_ok=True)
        net.save(a.out)
        print(json.dumps(rep), flush=True)
        return 0
    from . import fluid
    net = SurrogateNet()
    net.load(a.weights)
    net.set_smoothing(a.smoothing)
    net.set_backoff(a.backoff)
    wl = {"m1": fluid.m1, "r4m": fluid.r4m, "s16k": fluid.s16k}[a.workload]()
    if a.cmd == "time":
        print(json.dumps(time_rollout(wl, net, a.steps, a.plain_steps, a.warm_steps)), flush=True)
        return 0
    res = compare(wl, net, a.steps, threshold=a.threshold, stats_every=a.stats_every, unchecked=not a.no_unchecked,
                  log=sys.stderr, progress_every=a.progress_every, error_every=a.error_every)
    if a.json:
        Path(a.json).write_text(json.dumps(res))
    res.pop("series_plain")
    res.pop("series_surrogate")
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
