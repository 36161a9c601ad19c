"""The committed fixtures still come out of the oracle bit for bit (CPU)."""
import numpy as np
import pytest

from tests.golden import make_golden as G


@pytest.mark.parametrize("name", G.CASES)
def test_oracle_reproduces_fixture(name, oracle_lib):
    ref = np.load(G.HERE / f"sph_{name}.npz")
    p, pos, vel, dt = G.case_inputs(name)
    assert np.array_equal(pos, ref["pos0"]) and np.float32(dt) == ref["dt"]
    s = oracle_lib.OracleSolver(32)
    s.load_state(ref["pos0"], ref["vel0"])
    s.substep(p, float(ref["dt"]))
    cell_of, counts, slots = s.grid(p.cellCount)
    pos1, vel1 = s.state()
    assert np.array_equal(cell_of, ref["cell_of"])
    assert np.array_equal(counts, ref["counts"])
    assert np.array_equal(slots, ref["slots"])
    assert np.array_equal(pos1.view(np.uint32), ref["pos1"].view(np.uint32))
    assert np.array_equal(vel1.view(np.uint32), ref["vel1"].view(np.uint32))
    assert np.array_equal(s.forces().view(np.uint32), ref["force"].view(np.uint32))


def test_fixture_exercises_the_quirks():
    one = np.load(G.HERE / "sph_n40_c1.npz")
    assert one["counts"][0] == 40 and list(one["slots"][0]) == list(range(32))
    big = np.load(G.HERE / "sph_n512_c1024.npz")
    assert big["counts"].sum() == 512 and big["counts"].max() > 1
