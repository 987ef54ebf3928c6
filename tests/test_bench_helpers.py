"""CPU checks of bench.py's measurement helpers: the per-kernel byte model,
the profile-summary parsers and the one-GPU contract."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_stage_bytes_add_up_to_the_substep_minimum():
    # 204 B per particle + 84 B per cell per substep on either path: clearGrid
    # is charged as the p2g_1 flush, copyPosition as g2p's PosVel write
    def part(n, g, kernels, binned):
        return sum(bench.stage_bytes(k, n, g, binned_path=binned) for k in kernels)
    ref = ("clear_grid", "p2g_1", "p2g_2", "update_grid", "g2p", "copy_position")
    prod = ("bin_gather", "bin_scan", "bin_scatter", "p2g_1", "reduce_1", "p2g_2", "reduce_2",
            "g2p")
    for kernels, binned in ((ref, False), (prod, True)):
        assert (part(1, 0, kernels, binned), part(0, 1, kernels, binned)) == (204, 84)
    # G2P: x in; x, v, C and PosVel out; the 27-node gather of 12-B velocities
    assert bench.stage_bytes("g2p", 1, 0) == 12 + 60 + 24
    assert bench.stage_bytes("g2p", 0, 1) == 12


def test_every_production_slot_has_a_kernel_and_the_binning_is_overhead():
    for slot in ("bin_gather", "bin_scan", "bin_scatter", "p2g_1", "reduce_1", "p2g_2",
                 "reduce_2", "g2p"):
        assert slot in bench.SLOT_KERNEL
        assert bench.stage_bytes(slot, 300_000, 512_000) + \
            bench.overhead_bytes(slot, 300_000, 512_000) > 0
    for slot in ("bin_gather", "bin_scan", "bin_scatter"):
        assert bench.stage_bytes(slot, 300_000, 512_000) == 0


def test_roofline_entry_is_bytes_over_launch_time():
    e = bench.roofline_entry("g2p", 300_000, 512_000, {"g2p": 20.0}, True)
    want = (96 * 300_000 + 12 * 512_000) / 20e-6 / 1e9
    assert e["achieved"] == pytest.approx(want)
    assert e["frac"] == pytest.approx(want / 8000.0)
    assert e["overhead_bytes_per_launch"] == 24 * 300_000


def test_committed_profile_gives_the_dominant_kernels_traffic():
    t = bench.measured_traffic("g2p")
    assert t is not None and t > bench.stage_bytes("g2p", 300_000, 512_000)


def _write_counters(path, rows):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "run_counter_collection.csv"), "w") as f:
        f.write("Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value\n")
        for i, row in enumerate(rows):
            k, c, v = row[:3]
            d = row[3] if len(row) > 3 else i
            f.write(f'{d},"{k}",{c},{v}\n')


def _write_trace(path, rows):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "run_kernel_trace.csv"), "w") as f:
        f.write("Dispatch_Id,Kernel_Name,Start_Timestamp,End_Timestamp\n")
        for d, k, ns in rows:
            f.write(f'{d},"{k}",{1000 * d},{1000 * d + ns}\n')


G2P = "void mpm::(anonymous namespace)::k_g2p<2>(mpm::BatchArgs)"
COPY = "__amd_rocclr_copyBuffer"


def test_summary_selects_dispatches_and_scales_counters(tmp_path):
    # four g2p dispatches of 10, 20, 30, 40 us; dispatches 1..2 selected
    _write_trace(str(tmp_path / "trace"), [(1, G2P, 10000), (2, G2P, 20000), (3, G2P, 30000),
                                           (4, G2P, 40000), (5, COPY, 5000)])
    kib = bench.CALIBRATION_BYTES // 1024
    _write_counters(str(tmp_path / "fetch"), [
        (COPY, "FETCH_SIZE", kib // 2, 5),
        (G2P, "FETCH_SIZE", 100, 1), (G2P, "FETCH_SIZE", 200, 2), (G2P, "FETCH_SIZE", 400, 3),
        (G2P, "FETCH_SIZE", 800, 4)])
    _write_counters(str(tmp_path / "write"), [
        (COPY, "WRITE_SIZE", kib, 5),
        (G2P, "WRITE_SIZE", 300, 1), (G2P, "WRITE_SIZE", 300, 2), (G2P, "WRITE_SIZE", 300, 3),
        (G2P, "WRITE_SIZE", 300, 4)])
    _write_counters(str(tmp_path / "req"), [
        (G2P, "TCC_EA0_RDREQ_32B_sum", 10, 2), (G2P, "TCC_EA0_RDREQ_64B_sum", 20, 2),
        (G2P, "TCC_EA0_RDREQ_128B_sum", 30, 2), (G2P, "TCC_EA0_RDREQ_sum", 60, 2),
        (G2P, "TCC_EA0_RDREQ_DRAM_sum", 15, 2), (G2P, "TCC_EA0_WRREQ_sum", 50, 2),
        (G2P, "TCC_EA0_WRREQ_64B_sum", 40, 2), (G2P, "TCC_EA0_WRREQ_DRAM_sum", 25, 2)])
    dirs = [str(tmp_path / d) for d in ("fetch", "write", "req")]
    out = bench.summarize_profiles(str(tmp_path / "trace"), dirs, 1, 3)
    row = [l for l in out.splitlines() if l.startswith("k_g2p")][0].split()
    assert row[1] == "2" and float(row[2]) == pytest.approx(25.0)   # dispatches 2, 3
    # FETCH reads half the true bytes on the copy, WRITE all of them
    want = ((200 + 400) / 2 / 0.5 + 300) * 1024
    line = [l for l in out.splitlines() if l.startswith("TRAFFIC")][0]
    assert float(line.split()[-1]) == pytest.approx(want, rel=1e-3)
    # the request-size counters exist only for dispatch 2: all dispatches used
    assert float(row[5]) == pytest.approx(32 * 10 + 64 * 20 + 128 * 30)
    assert float(row[6]) == pytest.approx(32 * 10 + 64 * 40)
    assert float(row[7]) == pytest.approx(0.25) and float(row[8]) == pytest.approx(0.5)


def test_counter_table_averages_per_kernel(tmp_path):
    _write_counters(str(tmp_path / "c"), [("k_a", "SQ_WAVES", 4), ("k_a", "SQ_WAVES", 8),
                                          ("k_b", "SQ_WAVES", 1)])
    rows = bench.counter_table([str(tmp_path / "c")]).splitlines()
    assert rows[0].split() == ["kernel", "SQ_WAVES"]
    assert float(rows[1].split()[1]) == 6.0 and rows[1].split()[0] == "k_a"


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in r.stderr


def test_host_transfers_report_rates_and_the_readback_inclusive_value():
    import time
    import numpy as np

    class FakeSim:  # each transfer takes ~2 ms of host time
        calls = []

        def synchronize(self):
            pass

        def read_posvel(self):
            self.calls.append("read")
            time.sleep(0.002)

        def write_particles(self, rec):
            self.calls.append("write")
            time.sleep(0.002)

    rec = np.zeros((1000, 20), dtype=np.float32)
    h = bench.host_transfers(FakeSim(), rec, frame_ms=8.0, reps=3)
    assert FakeSim.calls == ["read"] * 3 + ["write"] * 3
    assert 2.0 <= h["read_posvel_ms"] < 8.0 and 2.0 <= h["write_particles_ms"] < 8.0
    assert h["read_posvel_GBs"] == pytest.approx(32 * 1000 / (h["read_posvel_ms"] * 1e-3) / 1e9)
    assert h["write_particles_GBs"] == pytest.approx(80 * 1000 / (h["write_particles_ms"] * 1e-3) / 1e9)
    # two substeps per frame, the readback added to every frame
    assert h["value_with_posvel_readback"] == pytest.approx(
        1000 * 2 / ((8.0 + h["read_posvel_ms"]) * 1e-3))
