"""The rocprofv3 trace summariser on a small hand-made database (CPU only): per-kernel
statistics keyed by name and template arguments, and busy time against span."""
import sqlite3

import pytest

from computer_graphics_lab_00_amd import rocpd_summary as R


@pytest.fixture
def trace_db(tmp_path):
    path = tmp_path / "t.db"
    c = sqlite3.connect(path)
    c.execute("create table kernels (name text, start integer, end integer)")
    rows = [
        # (name, start ns, end ns): two density launches of one instance, one of another,
        # a force launch, and a runtime copy kernel
        ("void sphdev::(anonymous namespace)::k_density<false, false, false>(sphdev::Launch, int)", 0, 1000),
        ("sphdev::(anonymous namespace)::k_force<false>(sphdev::Launch)", 1100, 1600),
        ("void sphdev::(anonymous namespace)::k_density<false, false, false>(sphdev::Launch, int)", 1700, 3700),
        ("void sphdev::(anonymous namespace)::k_density<true, false, false>(sphdev::Launch, int)", 3700, 4000),
        ("__amd_rocclr_copyBuffer", 5000, 5100),
    ]
    c.executemany("insert into kernels values (?, ?, ?)", rows)
    c.commit()
    c.close()
    return str(path)


def test_short_names_keep_template_arguments():
    assert R._short("void a::k_density<false, true, false>(int)") == "k_density<false,true,false>"
    assert R._short("sphdev::k_keys(sphdev::Launch)") == "k_keys"
    assert R._short("_ZN6sphdev12_GLOBAL__N_19k_densityILb1ELb0ELb0EEEvNS_6LaunchE") == "k_density<1,0,0>"
    assert R._short("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"


def test_kernel_stats(trace_db):
    stats = {k: (n, avg, mn, mx, tot) for k, n, avg, mn, mx, tot in R.kernel_stats(trace_db)}
    assert stats["k_density<false,false,false>"] == (2, 1.5, 1.0, 2.0, 0.003)
    assert stats["k_density<true,false,false>"] == (1, 0.3, 0.3, 0.3, 0.0003)
    assert stats["k_force<false>"][0] == 1
    # largest total first
    assert R.kernel_stats(trace_db)[0][0] == "k_density<false,false,false>"


def test_busy_vs_span_counts_only_the_library_kernels(trace_db):
    g = R.busy_vs_span(trace_db, 3)   # the last three k_ dispatches: force, density, density<true>
    assert g["dispatches"] == 3
    assert g["busy_ns"] == 500 + 2000 + 300
    assert g["span_ns"] == 4000 - 1100
    assert g["max_gap_ns"] == 100
    assert g["idle_fraction"] == pytest.approx(100 / 2900)


def test_kernel_stats_over_a_launch_range(trace_db):
    stats = {k: (n, avg) for k, n, avg, mn, mx, tot in R.kernel_stats(trace_db, 1, 5)}
    # the second density launch only; kernels with a single launch drop out
    assert stats == {"k_density<false,false,false>": (1, 2.0)}


def test_kernel_stats_over_the_last_launches(trace_db):
    stats = {k: (n, avg) for k, n, avg, mn, mx, tot in R.kernel_stats(trace_db, tail=1)}
    # the last launch of each kernel: the second density launch, the others' only one
    assert stats["k_density<false,false,false>"] == (1, 2.0)
    assert stats["k_force<false>"] == (1, 0.5)
    assert stats["k_density<true,false,false>"] == (1, 0.3)
    assert R.kernel_stats(trace_db, tail=0) == []
