"""The counter summarizer behind the committed traffic and instruction profiles, on synthetic
rocprofv3 counter CSVs: calibration scaling, per-kernel averages, kernel classes."""
import csv

import pytest

from computer_graphics_lab_00_amd import counter_summary as CS

FIELDS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size",
          "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
          "Accum_VGPR_Count", "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]

DENSITY = "void sphdev::(anonymous namespace)::k_density<false, false, false>(sphdev::Launch, HIP_vector_type<float, 4u>*)"
FORCE = "void sphdev::(anonymous namespace)::k_force<false>(sphdev::Launch, HIP_vector_type<float, 4u> const*)"
COPY = "sphdev::(anonymous namespace)::k_calibration_copy(HIP_vector_type<float, 4u> const*, unsigned long)"


def _csv(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for i, (name, counter, value) in enumerate(rows):
            r = {k: 0 for k in FIELDS}
            r.update({"Dispatch_Id": i // 8 if counter.startswith("SQ") or counter.startswith("GRBM") else i,
                      "Kernel_Name": name, "Counter_Name": counter, "Counter_Value": value})
            w.writerow(r)


def test_traffic_scales_by_the_calibration_copy(tmp_path):
    f, w = tmp_path / "f.csv", tmp_path / "w.csv"
    # the copy reads 256 MiB but FETCH_SIZE shows half of it; WRITE_SIZE shows all
    _csv(f, [(COPY, "FETCH_SIZE", 128 * 1024), (DENSITY, "FETCH_SIZE", 1000.0), (DENSITY, "FETCH_SIZE", 3000.0),
             (FORCE, "FETCH_SIZE", 500.0)])
    _csv(w, [(COPY, "WRITE_SIZE", 256 * 1024), (DENSITY, "WRITE_SIZE", 100.0), (DENSITY, "WRITE_SIZE", 300.0),
             (FORCE, "WRITE_SIZE", 50.0)])
    t = CS.traffic(str(f), str(w))
    assert t["fetch_scale"] == pytest.approx(2.0) and t["write_scale"] == pytest.approx(1.0)
    d = t["kernels"]["density"]
    assert d["fetch_bytes"] == 2000 * 1024 * 2 and d["write_bytes"] == 200 * 1024
    assert d["bytes_per_launch"] == d["fetch_bytes"] + d["write_bytes"] and d["launches"] == 2
    assert t["kernels"]["force_integrate"]["fetch_bytes"] == 500 * 1024 * 2
    assert "k_calibration_copy" not in t["kernels"]


def test_traffic_needs_the_calibration(tmp_path):
    f, w = tmp_path / "f.csv", tmp_path / "w.csv"
    _csv(f, [(DENSITY, "FETCH_SIZE", 1.0)])
    _csv(w, [(DENSITY, "WRITE_SIZE", 1.0)])
    with pytest.raises(ValueError):
        CS.traffic(str(f), str(w))


def test_sq_per_wave_figures_and_valu_busy(tmp_path):
    s = tmp_path / "sq.csv"
    rows = []
    for name, waves, valu in ((DENSITY, 16384, 16384 * 6000), (FORCE, 16384, 16384 * 2000)):
        rows += [(name, "SQ_WAVES", waves), (name, "SQ_INSTS_VALU", valu), (name, "SQ_INSTS_SALU", waves * 100),
                 (name, "SQ_INSTS_VMEM_RD", waves * 300), (name, "GRBM_GUI_ACTIVE", 8 * 500000),
                 (name, "SQ_WAVE_CYCLES", 1), (name, "SQ_BUSY_CYCLES", 1), (name, "GRBM_COUNT", 1)]
    _csv(s, rows)
    q = CS.sq(str(s))["kernels"]
    assert q["density"]["valu_instructions_per_wave"] == 6000.0
    assert q["density"]["vmem_reads_per_wave"] == 300.0 and q["density"]["gpu_cycles_per_xcd"] == 500000
    assert q["density"]["valu_busy"] == pytest.approx(16384 * 6000 * 4 / (256 * 4 * 500000), abs=1e-3)
    assert q["force_integrate"]["salu_instructions_per_wave"] == 100.0


def test_kernel_classes():
    assert CS.kernel_class(DENSITY) == "density"
    assert CS.kernel_class("void sphdev::(anonymous namespace)::k_density<false, false, true>(x)") == "density_fast_math"
    assert CS.kernel_class("void sphdev::(anonymous namespace)::k_force<true>(x)") == "force_integrate_fast_math"
    assert CS.kernel_class("sphdev::(anonymous namespace)::k_iota(unsigned int*, unsigned int)") is None
