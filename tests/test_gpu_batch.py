"""GPU: batched scenes equal the same scenes run alone, bit for bit; the hipGraph replay
equals eager launches; the autotuner's choice comes back from its cache."""
import sys

import numpy as np
import pytest

from computer_graphics_lab_00_amd import batch, capi, fluid

pytestmark = pytest.mark.gpu
DT = 1.0 / 600.0


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def small_specs():
    specs = batch.batch64(11, scenes=6, n_min=2000, n_max=20000)
    # different stiffness -> different substep counts inside one batch
    specs[1].stiffness = 200.0
    specs[2].stiffness = 20000.0
    specs[3].tile = (3, 3, 0)
    specs[4].neighbour_list = False
    specs[5].cell_count = 4 * batch.bit_ceil(2 * specs[5].count)
    return specs


def run_alone(spec, steps):
    p = spec.params()
    pos, vel = spec.initial_state(p)
    flags = 0 if spec.neighbour_list else capi.CFG_BUCKET_WALK
    with fluid.FluidSolver(spec.count, max(p.cellCount, 1024), flags=flags) as s:
        s.load_state(pos, vel)
        for _ in range(steps):
            s.step(p, DT)
        return s.state()


@pytest.mark.parametrize("flags", [0, capi.CFG_TILED_DENSITY])
@pytest.mark.parametrize("tile", [(2, 2, 2), (4, 2, 0), (6, 0, 0), (0, 0, 6), (1, 2, 3)])
def test_tile_shapes_match_the_oracle(tile, flags, oracle_lib):
    """The spatial-bin tile shape only moves buckets in memory; the tiled density pass
    keeps each half-tile's neighbourhood table by that shape (a shape whose neighbourhood
    is too large makes the launch use the plain walk). One step(1/600) of a scene with
    each shape equals the oracle bit for bit, with either density pass."""
    spec = batch.batch64(5, scenes=1, n_min=12000, n_max=12001)[0]
    spec.tile = tile
    say(f"case: tile {tile}, {spec.count} particles, flags {flags}")
    p = spec.params()
    pos, vel = spec.initial_state(p)
    with batch.SceneBatch([spec], flags=flags) as b:
        b.load_state(0, pos, vel)
        b.step(DT)
        gp, gv = b.state(0)
    ora = oracle_lib.OracleSolver(32)
    ora.load_state(pos, vel)
    ora.step(p, DT, 8)
    op, ov = ora.state()
    assert np.array_equal(bits(gp), bits(op)) and np.array_equal(bits(gv), bits(ov))


@pytest.mark.parametrize("flags", [0, capi.CFG_TILED_DENSITY])
def test_batch_equals_scenes_run_alone(flags):
    specs = small_specs()
    subs = [fluid.fluid_substeps(s.params(), DT) for s in specs]
    say(f"case: batch of {len(specs)} scenes, substeps {subs}")
    assert len(set(subs)) > 1, "the batch should mix substep counts"
    with batch.SceneBatch(specs, flags=flags) as b:
        b.load_initial()
        for _ in range(4):
            b.step(DT)
        got = [b.state(i) for i in range(len(specs))]
    for i, s in enumerate(specs):
        alone = run_alone(s, 4)
        assert np.array_equal(bits(got[i][0]), bits(alone[0])), f"scene {i} positions"
        assert np.array_equal(bits(got[i][1]), bits(alone[1])), f"scene {i} velocities"


def test_graph_replays_equal_eager_launches():
    """The step's hipGraph replayed several times, with reads and synchronisations in
    between, against the same batch launched eagerly."""
    specs = small_specs()[:4]
    say("case: graph vs eager")
    with batch.SceneBatch(specs) as g, batch.SceneBatch(specs, flags=capi.CFG_NO_GRAPH) as e:
        g.load_initial()
        e.load_initial()
        for k in range(5):
            g.step(DT)
            e.step(DT)
            for i in range(len(specs)):
                a, b = g.state(i), e.state(i)
                assert np.array_equal(bits(a[0]), bits(b[0])), f"replay {k} scene {i}"
                assert np.array_equal(bits(a[1]), bits(b[1])), f"replay {k} scene {i}"


def test_solver_graph_equals_eager():
    say("case: solver graph vs eager")
    w = fluid.s16k()
    p = w.params()
    pos, vel = w.initial_state(p)
    outs = []
    for flags in (0, capi.CFG_NO_GRAPH):
        with fluid.FluidSolver(w.count, 1 << 16, flags=flags) as s:
            s.load_state(pos, vel)
            for _ in range(7):          # 7 steps x 5 substeps: odd and even id parities
                s.step(p, DT)
                s.particles()
            outs.append(s.state())
    assert np.array_equal(bits(outs[0][0]), bits(outs[1][0]))
    assert np.array_equal(bits(outs[0][1]), bits(outs[1][1]))


def test_autotune_then_cache(tmp_path):
    specs = batch.batch64(13, scenes=2, n_min=2000, n_max=4000)
    say("case: autotune 2 scenes")
    cache = tmp_path / "tune.json"
    first = batch.autotune(specs, cache, steps=2, warmup=1)
    assert cache.exists()
    again = batch.autotune(specs, cache)
    assert [(a.cell_count, a.tile, a.neighbour_list) for a in first] == \
           [(b.cell_count, b.tile, b.neighbour_list) for b in again]
    # the tuned scenes still run and stay finite
    with batch.SceneBatch(again) as b:
        b.load_initial()
        b.step(DT)
        assert np.isfinite(b.state(0)[0]).all()


def test_block_schedule_does_not_change_results():
    """The block schedule only reorders workgroups: largest-first equals scene order bit
    for bit."""
    specs = small_specs()
    outs = []
    for name in ("scene order", "largest first"):
        with batch.SceneBatch(specs, schedule=batch.schedule_order(specs, name)) as b:
            b.load_initial()
            for _ in range(3):
                b.step(DT)
            outs.append([b.state(i) for i in range(len(specs))])
    for i in range(len(specs)):
        assert np.array_equal(bits(outs[0][i][0]), bits(outs[1][i][0]))
        assert np.array_equal(bits(outs[0][i][1]), bits(outs[1][i][1]))


def test_batch64_at_config_size(oracle_lib):
    """The batched workload at its configured size: 64 scenes of 2k-200k particles (2.29M),
    one step(1/600) of every scene in one graph replay. The smallest, the median and the
    largest scene equal the same scene stepped alone and the oracle's step, bit for bit."""
    specs = batch.batch64(2024)
    total = sum(s.count for s in specs)
    say(f"case: batch64, {len(specs)} scenes, {total} particles")
    assert len(specs) == 64 and total > 2_000_000
    order = sorted(range(len(specs)), key=lambda i: specs[i].count)
    picks = [order[0], order[len(order) // 2], order[-1]]
    with batch.SceneBatch(specs) as b:
        b.load_initial()
        b.step(DT)
        got = {i: b.state(i) for i in picks}
    for i in picks:
        s = specs[i]
        alone = run_alone(s, 1)
        p = s.params()
        pos, vel = s.initial_state(p)
        ora = oracle_lib.OracleSolver(32)
        ora.load_state(pos, vel)
        ora.step(p, DT, 16)
        op, ov = ora.state()
        say(f"  scene {i}: {s.count} particles, {fluid.fluid_substeps(p, DT)} substeps")
        assert np.array_equal(bits(got[i][0]), bits(alone[0])) and np.array_equal(bits(got[i][1]), bits(alone[1]))
        assert np.array_equal(bits(got[i][0]), bits(op)) and np.array_equal(bits(got[i][1]), bits(ov))


def test_batch_bin_occupancy_covers_every_scene():
    """A scene hashed into one bucket (cellCount 1) puts all its particles into one spatial
    bin: the batch's fullest bin is exactly that scene's particle count, beside a scene
    hashed normally; a state loaded afterwards starts the figure over."""
    specs = batch.batch64(13, scenes=2, n_min=2000, n_max=3000)
    specs[1].cell_count = 1
    say(f"case: batch occupancy, {specs[0].count} + {specs[1].count} particles (the second in one bucket)")
    with batch.SceneBatch(specs) as b:
        b.load_initial()
        assert b.max_bin_occupancy() == 0
        b.step(DT)
        assert b.max_bin_occupancy() == specs[1].count
        p = specs[0].params()
        pos, vel = specs[0].initial_state(p)
        b.load_state(0, pos, vel)
        assert b.max_bin_occupancy() == 0


def test_scenes_of_different_modes_in_one_batch_equal_each_alone():
    """A launch takes its kernel instances from the modes of all its scenes together (the
    de-duplicating and fast-math instances carry a per-scene branch), so a reference-mode
    scene can run inside another mode's instance. Each scene of a batch mixing the reference
    mode, uncapped buckets, de-duplicated buckets with the Tait EOS and fast math equals the
    same scene stepped alone in its own mode, bit for bit, over three steps."""
    specs = batch.batch64(17, scenes=4, n_min=3000, n_max=9000)
    modes = [(capi.MODE_REFERENCE, capi.EOS_LINEAR), (capi.MODE_UNCAPPED, capi.EOS_LINEAR),
             (capi.MODE_DEDUPE, capi.EOS_TAIT), (capi.MODE_FAST_MATH, capi.EOS_LINEAR)]
    say(f"case: mixed modes, {[s.count for s in specs]} particles")
    with batch.SceneBatch(specs) as b:
        for i, (mode, eos) in enumerate(modes):
            b.set_mode(i, mode, eos)
        b.load_initial()
        for _ in range(3):
            b.step(DT)
        for i, (spec, (mode, eos)) in enumerate(zip(specs, modes)):
            p = spec.params()
            pos, vel = spec.initial_state(p)
            with fluid.FluidSolver(spec.count, max(p.cellCount, 1024)) as s:
                s.set_mode(mode)
                s.set_eos(eos)
                s.load_state(pos, vel)
                for _ in range(3):
                    s.step(p, DT)
                ap, av = s.state()
            bp, bv = b.state(i)
            assert np.array_equal(bits(bp), bits(ap)), f"scene {i} (mode {mode}) positions differ"
            assert np.array_equal(bits(bv), bits(av)), f"scene {i} (mode {mode}) velocities differ"
