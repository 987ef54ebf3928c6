"""Benchmark of the MLS-MPM substep on one MI355X.

One "step" is one reference frame: execute() = 2 substeps
(mls-mpm/mls-mpm.ts:245-268).  The workload is the 300,000-particle dam break
on an 80^3 grid: the timed region replays its frames from the initial state,
resizing the box to 64 cells in x at frame min(500, frames / 2).  Inputs are
resident in HBM before timing starts.

Prints ONE JSON line with the throughput (particle-substeps/s), the roofline
of the slowest kernel and of G2P (algorithmic bytes / HIP-event kernel time,
averaged over an eager replay of the timed frames), the 2M-particle scene,
the batched and adjoint comparisons, and the CPU baseline (the oracle's
OpenMP restatement on the host cores, and on one core).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "webgpu-ocean_amd"))
sys.path.insert(0, ROOT)

METRIC = "particle-substeps/s @300k MLS-MPM, % HBM peak; batched & adjoint vs plain"
HBM_PEAK_GBS = 8000.0
SUBSTEPS_PER_FRAME = 2
RESIZE_FRAME = 500


# Algorithmic HBM bytes per launch, per kernel: the minimum traffic of the
# four barrier-separated phases of the reference's substep, 204 B per particle
# + 84 B per cell in total (particle fields: x 12 B, v 12 B, C 36 B; cells
# 16 B = vx, vy, vz, mass):
#   p2g_1     reads x, v, C                          60 / particle
#   p2g_1 flush writes the cells (clear fused)       16 / cell
#   p2g_2     reads x, C                             48 / particle
#             reads masses, RMW of vx, vy, vz         4 + 24 / cell
#   updateGrid reads 16, writes 12                   28 / cell
#   g2p       reads x, writes x, v, C, PosVel(24)    96 / particle
#             gathers updated velocities             12 / cell
# Each production kernel is charged the phases it performs; the binning and
# the intermediates (permutation, ids, tile partials) carry no algorithmic
# bytes: they are listed apart as overhead_bytes.
def stage_bytes(name: str, n: int, g: int, binned_path: bool = True) -> int:
    """Algorithmic HBM bytes of one launch of kernel slot `name`."""
    reference_shaped = {
        "clear_grid": 16 * g,
        "p2g_1": 60 * n,
        "p2g_2": 48 * n + 28 * g,
        "update_grid": 28 * g,
        "g2p": (12 + 60) * n + 12 * g,
        "copy_position": 24 * n,
    }
    binned = {
        "bin_gather": 0,
        "bin_scan": 0,
        "bin_scatter": 0,
        "p2g_1": 60 * n,
        "reduce_1": 16 * g,                 # the p2g_1 flush
        "p2g_2": 48 * n + 4 * g,
        "reduce_2": 24 * g + 28 * g,        # p2g_2's momentum RMW + updateGrid
        "g2p": 96 * n + 12 * g,
    }
    return (binned if binned_path else reference_shaped).get(name, 0)


def overhead_bytes(name: str, n: int, g: int) -> int:
    """Bytes of the production path's intermediates per launch (not part of
    the algorithmic bytes): per particle the binning key/rank/source and
    permutation, the level order, the caller ids; per 4^3 tile its count,
    offset and chunk range.  Tile partial windows are not included (their
    size depends on the particle distribution)."""
    tiles = g // 64
    table = {
        "bin_gather": 8 * tiles,
        "bin_scan": 16 * tiles,
        "bin_scatter": 16 * n,
        "p2g_1": 8 * n,        # permutation in, level order out
        "p2g_2": 4 * n,        # level order in
        "g2p": 24 * n,         # permutation, id in; id, key, rank, source out
    }
    return table.get(name, 0)


G2P_BYTES_DOC = "G2P: 96 B/particle + 12 B/cell (x in; x, v, C, PosVel out; 27-node gather)"


CALIBRATION_BYTES = 64 << 20


def calibration_copy(nbytes: int = CALIBRATION_BYTES) -> None:
    """One device-to-device copy of a known size through the HIP runtime, so a
    profiler's FETCH_SIZE / WRITE_SIZE for it calibrate the counters (the copy
    reads and writes exactly nbytes).  A 512 MiB fill between writing the
    source and copying it evicts the source from the L2 caches."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b, ev = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    evict = 512 << 20
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or \
            hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) or \
            hip.hipMalloc(ctypes.byref(ev), ctypes.c_size_t(evict)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipMemset(ev, 2, ctypes.c_size_t(evict))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    hip.hipFree(ev)


def _short_kernel(name: str) -> str:
    if "rocprim" in name:  # library kernels: named by the algorithm they serve
        for part in ("onesweep_config", "radix_sort_block_sort", "merge_sort", "scan_config",
                     "init_lookback", "histogram", "transform"):
            if part in name:
                return "rocprim " + part.replace("_config", "")
        return "rocprim"
    n = name.replace("(anonymous namespace)::", "").replace("mpm::", "")
    return n.split("(")[0].replace("void ", "").strip()


def _csv_rows(d: str, suffix: str):
    import csv
    import glob

    files = sorted(f for f in glob.glob(os.path.join(d, "**", "*"), recursive=True)
                   if f.endswith(suffix))
    if not files:
        raise SystemExit(f"no *{suffix} under {d}")
    rows = []
    for f in files:
        rows.extend(csv.DictReader(open(f)))
    return rows


def _select(per_kernel: dict, lo: int, hi: int) -> dict:
    """Dispatches lo..hi-1 (in dispatch order) of every kernel that has them,
    else all of its dispatches."""
    out = {}
    for k, items in per_kernel.items():
        items = sorted(items)
        out[k] = items[lo:hi] if len(items) >= hi > lo else items
    return out


def trace_durations(trace_dir: str, lo: int = 0, hi: int = 0) -> dict:
    """{kernel: (average us, dispatches)} of a rocprofv3 --kernel-trace run
    (csv), over each kernel's dispatches lo..hi-1 (hi = 0: all)."""
    per = {}
    for r in _csv_rows(trace_dir, "kernel_trace.csv"):
        k = _short_kernel(r["Kernel_Name"])
        us = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        per.setdefault(k, []).append((int(r["Dispatch_Id"]), us))
    sel = _select(per, lo, hi) if hi else {k: sorted(v) for k, v in per.items()}
    return {k: (sum(u for _, u in v) / len(v), len(v)) for k, v in sel.items() if v}


def counter_values(dirs, lo: int = 0, hi: int = 0, reduce=None) -> dict:
    """{kernel: {counter: average per dispatch}} of rocprofv3 --pmc runs (csv),
    over each kernel's dispatches lo..hi-1 (hi = 0: all); `reduce` replaces
    the average (e.g. max)."""
    out = {}
    for d in dirs:
        per = {}
        for r in _csv_rows(d, "counter_collection.csv"):
            k = _short_kernel(r["Kernel_Name"])
            per.setdefault((k, r["Counter_Name"]), []).append(
                (int(r["Dispatch_Id"]), float(r["Counter_Value"])))
        for (k, c), v in per.items():
            v = sorted(v)
            if hi and len(v) >= hi > lo:
                v = v[lo:hi]
            xs = [x for _, x in v]
            out.setdefault(k, {})[c] = reduce(xs) if reduce else sum(xs) / len(xs)
    return out


def summarize_profiles(trace_dir: str, counter_dirs=(), lo: int = 0, hi: int = 0) -> str:
    """Per-kernel table of one profiled command: average duration from the
    kernel trace and bytes per dispatch from counter runs of the same command,
    both over each kernel's dispatches lo..hi-1.

    - FETCH_SIZE / WRITE_SIZE (KiB), each divided by what the same counter
      reads for the calibration copy of CALIBRATION_BYTES (bench.py
      --calibrate): the `traffic` column, which bench.py reports;
    - request sizes: reads 32 n32 + 64 n64 + 128 n128 (TCC_EA0_RDREQ_32B/64B/
      128B), writes 32 (n - n64) + 64 n64 (TCC_EA0_WRREQ / _64B);
    - the share of read and write requests that went to DRAM rather than the
      MALL (TCC_EA0_RDREQ_DRAM / TCC_EA0_RDREQ, TCC_EA0_WRREQ_DRAM / TCC_EA0_WRREQ)."""
    dur = trace_durations(trace_dir, lo, hi)
    cnt = counter_values(counter_dirs, lo, hi) if counter_dirs else {}
    # the calibration copy is the largest copy dispatch of the run
    copies = [v for k, v in (counter_values(counter_dirs, reduce=max) if counter_dirs
                             else {}).items() if "copyBuffer" in k]

    def calib(name):
        vals = [c[name] for c in copies if name in c]
        return max(vals) * 1024.0 / CALIBRATION_BYTES if vals else None

    fcal, wcal = calib("FETCH_SIZE"), calib("WRITE_SIZE")
    out = []
    if fcal and wcal:
        out.append(f"calibration: the {CALIBRATION_BYTES >> 20} MiB copy reads FETCH_SIZE x{fcal:.3f} "
                   f"and WRITE_SIZE x{wcal:.3f} of its true size")
    out.append(f"{'kernel':24s} {'disp':>5s} {'avg_us':>8s} {'traffic_B':>11s} {'GB/s':>6s} "
               f"{'rd_req_B':>10s} {'wr_req_B':>10s} {'rd_DRAM':>7s} {'wr_DRAM':>7s}")
    for k, (us, n) in sorted(dur.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        if k.startswith("__amd"):
            continue
        c = cnt.get(k, {})
        traffic = rate = rd = wr = rdd = wrd = ""
        if "FETCH_SIZE" in c and "WRITE_SIZE" in c and fcal and wcal:
            t = (c["FETCH_SIZE"] / fcal + c["WRITE_SIZE"] / wcal) * 1024
            traffic, rate = f"{t:.4g}", f"{t / us / 1e3:.0f}"
        if all(x in c for x in ("TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum",
                                "TCC_EA0_RDREQ_128B_sum")):
            rd = f"{32 * c['TCC_EA0_RDREQ_32B_sum'] + 64 * c['TCC_EA0_RDREQ_64B_sum'] + 128 * c['TCC_EA0_RDREQ_128B_sum']:.4g}"
        if "TCC_EA0_WRREQ_sum" in c and "TCC_EA0_WRREQ_64B_sum" in c:
            n64 = c["TCC_EA0_WRREQ_64B_sum"]
            wr = f"{32 * (c['TCC_EA0_WRREQ_sum'] - n64) + 64 * n64:.4g}"
        if c.get("TCC_EA0_RDREQ_sum") and "TCC_EA0_RDREQ_DRAM_sum" in c:
            rdd = f"{c['TCC_EA0_RDREQ_DRAM_sum'] / c['TCC_EA0_RDREQ_sum']:.2f}"
        if c.get("TCC_EA0_WRREQ_sum") and "TCC_EA0_WRREQ_DRAM_sum" in c:
            wrd = f"{c['TCC_EA0_WRREQ_DRAM_sum'] / c['TCC_EA0_WRREQ_sum']:.2f}"
        out.append(f"{k[:24]:24s} {n:5d} {us:8.2f} {traffic:>11s} {rate:>6s} {rd:>10s} {wr:>10s} "
                   f"{rdd:>7s} {wrd:>7s}")
        if traffic:
            out.append(f"TRAFFIC {k} {traffic}")
    return "\n".join(out)


def counter_table(dirs, lo: int = 0, hi: int = 0) -> str:
    """Per-kernel averages (per dispatch) of every counter of one or more
    rocprofv3 --pmc runs (csv output), one column per counter."""
    vals = counter_values(dirs, lo, hi)
    names = sorted({c for per in vals.values() for c in per})
    out = [f"{'kernel':24s}" + "".join(f"{c[:18]:>19s}" for c in names)]
    for k, per in vals.items():
        if k.startswith("__amd") or "copyBuffer" in k or "fill" in k.lower():
            continue
        out.append(f"{k[:24]:24s}" + "".join(
            f"{per.get(c, float('nan')):19.4g}" for c in names))
    return "\n".join(out)


DEFAULT_TUNING = {"path": 0, "particles_per_thread": 2, "grid_blocks": 0, "xcd_remap": 0,
                  "tile_edge": 4}

TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "substep_kernels_dambreak_300k.txt")


# kernel behind each production time_stages slot (default tuning: 2 particles per thread;
# G2P in the gather counting mode, which the 300k / 80^3 scene uses)
SLOT_KERNEL = {"bin_gather": "k_bin_gather", "bin_scan": "k_scan_keys",
               "bin_scatter": "k_bin_scatter", "p2g_1": "k_p2g1<2, 2>",
               "reduce_1": "k_reduce<1, 2>", "p2g_2": "k_p2g2<2, 2>", "reduce_2": "k_reduce<2, 2>",
               "g2p": "k_g2p<2, 2>"}


def measured_traffic(slot: str):
    """HBM bytes per launch of the kernel behind `slot`, from the committed
    profile summary of the default 300k dam-break run, or None."""
    kernel = SLOT_KERNEL.get(slot)
    if kernel is None:
        return None
    try:
        for line in open(TRAFFIC_FILE):
            t = line.split()
            if len(t) >= 3 and t[0] == "TRAFFIC" and " ".join(t[1:-1]) == kernel:
                return float(t[-1])
    except OSError:
        return None
    return None


def host_transfers(sim, rec, frame_ms: float, reps: int = 5) -> dict:
    """Host <-> device transfers at the boundary, outside the metric: the
    PosVel readback (N x 32 B, what a host-side consumer of every frame would
    add) and the particle upload (N x 80 B records, as at construction and
    reset).  Medians of `reps`; the upload replaces the simulator's state."""
    def med(fn):
        ts = []
        for _ in range(reps):
            sim.synchronize()
            t0 = time.perf_counter()
            fn()
            sim.synchronize()
            ts.append(time.perf_counter() - t0)
        return 1e3 * sorted(ts)[len(ts) // 2]

    n = rec.shape[0]
    read_ms = med(sim.read_posvel)
    write_ms = med(lambda: sim.write_particles(rec))
    return {
        "read_posvel_ms": read_ms,
        "read_posvel_GBs": 32 * n / (read_ms * 1e-3) / 1e9,
        "write_particles_ms": write_ms,
        "write_particles_GBs": 80 * n / (write_ms * 1e-3) / 1e9,
        # particle-substeps/s if every frame's PosVel also went to the host
        "value_with_posvel_readback": n * SUBSTEPS_PER_FRAME / ((frame_ms + read_ms) * 1e-3),
    }


def cpu_threads() -> tuple[int, int]:
    """(threads used, host CPUs this process may run on).  All CPUs of the
    affinity mask, capped by OMP_NUM_THREADS where the host sets it (a
    shared GPU host grants each GPU job a CPU share and says so there)."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    used = min(avail, int(cap)) if cap.isdigit() and int(cap) > 0 else avail
    return max(1, used), avail


def cpu_baseline(rec, box, seconds: float = 10.0, reps: int = 5) -> dict:
    """Oracle (CPU restatement of the reference WGSL) timed on the host cores:
    median of `reps` runs of k substeps each, on every allowed core and on
    one core, about `seconds` of work per leg."""
    import statistics

    import oracle

    n = rec.shape[0]
    threads, avail = cpu_threads()

    def leg(th):
        sc = oracle.Scene(rec, box)
        t0 = time.perf_counter()
        sc.substep(threads=th)
        one = time.perf_counter() - t0
        k = max(1, min(200, int(seconds / reps / max(one, 1e-4))))
        rates = []
        for _ in range(reps):
            t0 = time.perf_counter()
            for _ in range(k):
                sc.substep(threads=th)
            rates.append(n * k / (time.perf_counter() - t0))
        return statistics.median(rates), k

    value, k = leg(threads)
    value1, k1 = leg(1)
    return {
        "value": value,
        "unit": "particle-substeps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"median of {reps} runs of {k} substeps ({threads} threads) and of {k1} "
                  f"substeps (1 thread) of the same {n}-particle 80^3 dam break from its "
                  "initial state (OpenMP over particles, atomic i32 P2G)",
        "value_1_thread": value1,
        "host_cpus_visible": avail,
    }


def bench_batched(substeps: int = 500) -> dict:
    """64 uneven scenes in one launch per stage (autotuned) vs the same scenes
    advanced one after another by single simulators; results are identical
    per scene (tests/test_batch_gpu.py)."""
    from mlsmpm import MLSMPMBatch, MLSMPMSimulator, Params, scenes

    sc = scenes.config3()
    params = []
    for _, _, phys in sc:
        p = Params(3.0, 4.0, 0.1, 0.2, 1e7)
        p.stiffness, p.dynamic_viscosity = phys["stiffness"], phys["dynamic_viscosity"]
        params.append(p)
    total = sum(r.shape[0] for r, _, _ in sc)
    with MLSMPMBatch([(r, b, p) for (r, b, _), p in zip(sc, params)]) as bt:
        tuning = bt.autotune(12)  # 36 candidates x 12 substeps: stable against noise
        bt.step(2)
        bt.synchronize()
        t0 = time.perf_counter()
        bt.step(substeps)
        bt.synchronize()
        batched_s = time.perf_counter() - t0
    # sequential plain launches: each scene advanced alone, one after another
    sims = [MLSMPMSimulator(params=p, particles=r, init_box_size=b)
            for (r, b, _), p in zip(sc, params)]
    for sim in sims:
        sim.execute(2)
        sim.synchronize()
    t0 = time.perf_counter()
    for sim in sims:
        sim.execute(substeps)
        sim.synchronize()
    seq_s = time.perf_counter() - t0
    for sim in sims:
        sim.close()
    return {
        "workload": "64 uneven scenes (5k-200k particles, edges 32-96), %d substeps" % substeps,
        "particles": total,
        "value": total * substeps / batched_s,
        "sequential_value": total * substeps / seq_s,
        "speedup": seq_s / batched_s,
        "unit": "particle-substeps/s",
        "tuning": tuning,
    }


def bench_adjoint(substeps: int = 64, every: int = 8) -> dict:
    """100k-particle 64^3 differentiable rollout: backward (forward with
    checkpoints + recomputation + adjoint) vs the plain forward rollout."""
    from mlsmpm import MLSMPMSimulator, scenes

    rec, box, target = scenes.config4()
    with MLSMPMSimulator(particles=rec, init_box_size=box) as sim:
        sim.execute(4)
        sim.reset(rec.shape[0], box, particles=rec)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.execute(substeps)
        sim.synchronize()
        fwd = time.perf_counter() - t0
        sim.reset(rec.shape[0], box, particles=rec)
        sim.backward(4, target, checkpoint_every=2)   # warm the adjoint kernels
        sim.reset(rec.shape[0], box, particles=rec)
        sim.synchronize()
        t0 = time.perf_counter()
        g = sim.backward(substeps, target, checkpoint_every=every)
        bwd = time.perf_counter() - t0
    return {
        "workload": "100k particles, 64^3, %d substeps, checkpoint every %d" % (substeps, every),
        "forward_ms": 1e3 * fwd,
        "backward_ms": 1e3 * bwd,
        "cost_ratio": bwd / fwd,
        "loss": g["loss"],
    }


def timed_rollout(sim, frames: int, resize_frame: int, resize_box) -> float:
    """Wall seconds of `frames` frames (hipGraph-replayed substeps), the box
    changed at `resize_frame`; inputs are already on the device."""
    sim.synchronize()
    t0 = time.perf_counter()
    for f in range(frames):
        if f == resize_frame:
            sim.change_box_size(resize_box)
        sim.execute()
    sim.synchronize()
    return time.perf_counter() - t0


def kernel_times(sim, substeps: int, resize_frame: int, resize_box):
    """Per-kernel HIP-event times of `substeps` eager substeps (frames of 2,
    the box changed at `resize_frame`): ({slot: us per launch},
    {slot: us per substep})."""
    stage = {}
    for f in range((substeps + SUBSTEPS_PER_FRAME - 1) // SUBSTEPS_PER_FRAME):
        if f == resize_frame:
            sim.change_box_size(resize_box)
        k = min(SUBSTEPS_PER_FRAME, substeps - f * SUBSTEPS_PER_FRAME)
        for name, (ms, c) in sim.time_stages(k).items():
            acc = stage.setdefault(name, [0.0, 0])
            acc[0] += ms
            acc[1] += c
    per_launch = {k: 1e3 * ms / max(c, 1) for k, (ms, c) in stage.items()}
    per_substep = {k: 1e3 * ms / substeps for k, (ms, c) in stage.items()}
    return per_launch, per_substep


def roofline_entry(name, n, g, per_launch_us, binned, traffic=lambda k: None) -> dict:
    """Algorithmic bytes of one launch of `name` over its average launch time."""
    b = stage_bytes(name, n, g, binned)
    us = per_launch_us[name]
    achieved = b / (us * 1e-6) / 1e9 if us > 0 else 0.0
    return {
        "kernel": name,
        "achieved": achieved,
        "frac": achieved / HBM_PEAK_GBS,
        "avg_launch_us": us,
        "algorithmic_bytes_per_launch": b,
        "overhead_bytes_per_launch": overhead_bytes(name, n, g) if binned else 0,
        "traffic": traffic(name),
    }


def bench_config2(frames: int = 300) -> dict:
    """2,000,000 particles on 160^3 (two colliding blocks): a working set
    (about 230 MB of particles and cells, ~500 MB with the buffers of the
    binned path) beyond the 256 MB MALL, so the roofline is read where the
    traffic goes to HBM.  Graph-replayed frames, then the same frames eagerly
    with per-kernel events."""
    from mlsmpm import MLSMPMSimulator, scenes

    rec, box = scenes.config2()
    n = rec.shape[0]
    with MLSMPMSimulator(particles=rec, init_box_size=box) as sim:
        g = sim.grid_count
        sim.execute(4)
        sim.reset(n, box, particles=rec)
        elapsed = timed_rollout(sim, frames, frames, box)
        sim.reset(n, box, particles=rec)
        per_launch, per_substep = kernel_times(sim, 2 * frames, frames, box)
    dom = max(per_substep, key=per_substep.get)
    return {
        "workload": f"2M particles, 160^3 grid, two colliding blocks, {frames} frames",
        "value": n * 2 * frames / elapsed,
        "unit": "particle-substeps/s",
        "wall_us_per_substep": 1e6 * elapsed / (2 * frames),
        "roofline": roofline_entry(dom, n, g, per_launch, True),
        "g2p": roofline_entry("g2p", n, g, per_launch, True),
        "stages_us_per_launch": per_launch,
    }


# Algorithmic bytes per launch of the SPH kernels (sph.hip): what each kernel
# must move for its own inputs and outputs, per particle (n) and per cell of
# the 62^3 grid (g); neighbour reads are reuse of the sorted copies (cache),
# not counted.
SPH_BYTES = {
    "count": (28, 0),          # pos 16 in; cell 4 + rank 4 out (+ the count atomics, 4)
    "scan_reduce": (0, 4),     # counts in
    "scan": (0, 12),           # counts in, prefix out, counts cleared
    "scatter": (16, 0),        # cell, rank, cell end in; slot out
    "sort_gather": (84, 0),    # slot, cell, 2 prefix entries, pos, vel in; order, 2 copies out
    "density": (36, 0),        # order, own position in; densities + sorted copy out
    "force": (124, 0),         # order, position, velocity, densities in; force, state, posvel out
}


def sph_roofline(stage: str, n: int, g: int, us: float) -> dict:
    per_p, per_c = SPH_BYTES[stage]
    b = per_p * n + per_c * g
    ach = b / (us * 1e-6) / 1e9 if us > 0 else 0.0
    return {"kernel": stage, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None,
            "bytes_per_launch": b, "us_per_launch": us}


def sph_cpu_baseline(rec, half, seconds: float = 6.0, reps: int = 3) -> dict:
    """The SPH oracle (CPU restatement of sph/*.wgsl, OpenMP over particles)
    on the host cores: median of `reps` runs, all allowed threads and one."""
    import statistics

    from oracle import sph as osph

    n = rec.shape[0]
    threads, _ = cpu_threads()

    def leg(th):
        sc = osph.Scene(rec, half)
        t0 = time.perf_counter()
        sc.substep(threads=th)
        one = time.perf_counter() - t0
        k = max(1, min(100, int(seconds / reps / max(one, 1e-4))))
        rates = []
        for _ in range(reps):
            t0 = time.perf_counter()
            for _ in range(k):
                sc.substep(threads=th)
            rates.append(n * k / (time.perf_counter() - t0))
        return statistics.median(rates), k

    value, k = leg(threads)
    value1, k1 = leg(1)
    return {"value": value, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"median of {reps} runs of {k} substeps ({threads} threads) and of {k1} "
                      f"substeps (1 thread) of the same {n}-particle dam break from its "
                      "initial state",
            "value_1_thread": value1}


def bench_sph(frames: int = 500, cpu: bool = True) -> dict:
    """SPH path (sph/sph.ts): the reference's largest setting (40,000
    particles, half box 1.4 x 2 x 1.4, main.ts:131-132) and its particle cap
    (200,000, common.ts:11) in the same box, `frames` frames x 2 substeps from
    the initial state with the z wall moved in to 0.7 of its extent at the
    middle frame (the box slider, main.ts:253-260)."""
    from sph import SPHSimulator, scenes

    out = {}
    half = scenes.INIT_HALF_BOXES[3]
    narrow = [half[0], half[1], half[2] * 0.7]
    for label, count in (("40k", 40000), ("200k", 200000)):
        rec = scenes.dam_break(3, seed=0, num_particles=count)
        with SPHSimulator(particles=rec, init_half_box_size=half) as sim:
            g = sim.grid_count
            sim.execute(10)
            sim.reset(count, half, particles=rec)
            sim.synchronize()
            t0 = time.perf_counter()
            for f in range(frames):
                if f == frames // 2:
                    sim.change_box_size(narrow)
                sim.execute()
            sim.synchronize()
            elapsed = time.perf_counter() - t0
            sim.reset(count, half, particles=rec)
            stages_ms = sim.time_stages(2 * frames)
        per_launch = {k: 1e3 * v / (2 * frames) for k, v in stages_ms.items()}
        dom = max(per_launch, key=per_launch.get)
        out[label] = {
            "workload": f"SPH dam break, {count} particles, 62^3 neighbour grid, {frames} "
                        f"frames x 2 substeps, z wall at 0.7 from frame {frames // 2}",
            "value": count * 2 * frames / elapsed,
            "unit": "particle-substeps/s",
            "wall_us_per_substep": 1e6 * elapsed / (2 * frames),
            "stages_us_per_launch": per_launch,
            "sum_kernel_us": sum(per_launch.values()),
            "roofline": sph_roofline(dom, count, g, per_launch[dom]),
        }
        if cpu and label == "40k":
            out[label]["cpu_baseline"] = sph_cpu_baseline(rec, half)
    return out


def bench_render(frames: int = 200, cpu: bool = True) -> dict:
    """Screen-space renderer (render/fluidRender.ts): the reference's startup
    scene (MLS-MPM dam break of 70,000 particles in a 40 x 30 x 60 box, camera
    at distance 70 looking at (20, 7.5, 30), fov 45 deg, main.ts:128-196) and
    the 300k dam break of the headline workload (camera at distance 140), each
    after 100 simulated frames, rendered `frames` times from the simulator's
    device PosVel buffer (fluid and sphere modes)."""
    import math

    import numpy as np

    import render
    from mlsmpm import MLSMPMSimulator, scenes

    fov = 45 * math.pi / 180
    cube = render.synthetic_cubemap(64)
    out = {}
    cases = (("ref70k", 70000, [40.0, 30.0, 60.0], 70.0), ("dam300k", 300000, None, 140.0))
    for label, count, box, dist in cases:
        if box is None:
            rec, box = scenes.config1(n=count)
        else:
            rec = scenes.init_dambreak(box, count)
        with MLSMPMSimulator(particles=rec, init_box_size=box) as sim:
            for _ in range(100):
                sim.execute()
            sim.synchronize()
            ptr = sim.posvel_device_ptr()
            pv = sim.read_posvel() if (cpu and label == "ref70k") else None
            for w, h in ((1280, 720), (1920, 1080)):
                cam = render.Camera()
                cam.reset(w, h, dist, [box[0] / 2, box[1] / 4, box[2] / 2], fov)
                u = cam.uniforms(w, h, 1.2)
                key = f"{label} {w}x{h}"
                with render.FluidRenderer(w, h, 0.6, fov, cube) as r:
                    entry = {}
                    for mode, sphere in (("fluid", False), ("spheres", True)):
                        r.time_frames(ptr, count, u, 10, sphere)
                        ms = r.time_frames(ptr, count, u, frames, sphere) / frames
                        entry[mode] = {"ms_per_frame": ms, "frames_per_s": 1e3 / ms}
                    r.execute(ptr, count, u)
                    d = r.read_image("depth")
                    entry["covered_pixel_fraction"] = float(np.count_nonzero(d) / d.size)
                    out[key] = entry
                if pv is not None and (w, h) == (1280, 720):
                    from oracle import render as orender
                    c = orender.consts(w, h, 0.6, fov)
                    threads, _ = cpu_threads()
                    reps = []
                    for _ in range(3):
                        t0 = time.perf_counter()
                        orender.render_frame(pv, count, u, c, cube, False, threads)
                        reps.append(time.perf_counter() - t0)
                    sec = sorted(reps)[1]
                    out[key]["cpu_baseline"] = {
                        "value": 1.0 / sec, "unit": "frames/s", "cores": threads, "kind": "port",
                        "sample": f"median of 3 oracle frames of the same scene at {w}x{h} "
                                  f"(fluid mode; the impostor splat serial, the filters and "
                                  f"shading on {threads} threads)"}
    out["workload"] = ("frames of FluidRenderer.execute from the MLS-MPM PosVel buffer: tiled "
                       "impostor raster (depth test + r16float thickness blend in draw order), "
                       "4 x bilateral, 1 x gaussian, fluid shading; synthetic cube map")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed frames (2 substeps each)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--particles", type=int, default=300_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched and adjoint comparisons")
    ap.add_argument("--extras-only", action="store_true",
                    help="run only the batched and adjoint comparisons (one JSON line; "
                         "for profiling them)")
    ap.add_argument("--sph-only", action="store_true",
                    help="run only the SPH section (--steps frames; one JSON line)")
    ap.add_argument("--render-only", action="store_true",
                    help="run only the renderer section (--steps frames; one JSON line)")
    ap.add_argument("--stage-substeps", type=int, default=0,
                    help="substeps of the per-kernel timing replay (default: all of the timed "
                         "frames)")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a 64 MiB device copy (counter calibration under a profiler)")
    ap.add_argument("--summarize", metavar="TRACE_DIR",
                    help="print the per-kernel table of a rocprofv3 kernel-trace run (with "
                         "--pmc-dirs: bytes from counter runs of the same command) and exit")
    ap.add_argument("--pmc-dirs", nargs="*", default=[], metavar="DIR",
                    help="rocprofv3 --pmc output directories for --summarize / --counters")
    ap.add_argument("--counters", action="store_true",
                    help="print per-kernel averages of every counter in --pmc-dirs and exit")
    ap.add_argument("--dispatches", type=str, default="",
                    help="LO:HI -- summarize each kernel's dispatches LO..HI-1 only")
    ap.add_argument("--tuning", type=str, default="",
                    help='JSON launch configuration, e.g. {"particles_per_thread": 1}')
    args = ap.parse_args()
    lo, hi = (int(x) for x in args.dispatches.split(":")) if args.dispatches else (0, 0)
    if args.summarize:
        print(summarize_profiles(args.summarize, args.pmc_dirs, lo, hi))
        return 0
    if args.counters:
        print(counter_table(args.pmc_dirs, lo, hi))
        return 0

    if args.gpus != 1:
        print(f"bench.py: the library runs on one GPU; --gpus {args.gpus} is not supported "
              "(a scene would shard by x-slabs of the grid, exchanging the boundary tiles' "
              "partial windows and halo velocities with the neighbouring GPU)",
              file=sys.stderr)
        return 2

    from mlsmpm import MLSMPMSimulator, scenes

    if args.calibrate:
        calibration_copy()

    if args.sph_only:
        print(json.dumps({"sph": bench_sph(frames=args.steps, cpu=not args.no_cpu_baseline)}),
              flush=True)
        return 0
    if args.render_only:
        print(json.dumps({"render": bench_render(frames=args.steps,
                                                 cpu=not args.no_cpu_baseline)}), flush=True)
        return 0
    if args.extras_only:
        print(json.dumps({"batched": bench_batched(), "adjoint": bench_adjoint()}), flush=True)
        return 0

    rec, box = scenes.config1(n=args.particles)
    n = rec.shape[0]
    sim = MLSMPMSimulator(particles=rec, init_box_size=box)
    g = sim.grid_count
    if args.tuning:
        sim.set_tuning(**json.loads(args.tuning))
    # the box change lands inside the timed frames whatever their number
    resize_frame = min(RESIZE_FRAME, args.steps // 2)
    resize_box = [64.0, box[1], box[2]]

    # warmup on the same scene, then restart from the initial state
    for _ in range(args.warmup):
        sim.execute()
    sim.synchronize()
    sim.reset(n, box, particles=rec)
    sim.synchronize()

    elapsed = timed_rollout(sim, args.steps, resize_frame, resize_box)
    substeps = args.steps * SUBSTEPS_PER_FRAME
    value = n * substeps / elapsed

    # per-kernel HIP-event timing of the same path on the same stream: the
    # timed frames replayed from the initial state, box change included, with
    # events bound to every kernel
    sim.reset(n, box, particles=rec)
    replay = min(args.stage_substeps or substeps, substeps)
    per_launch_us, per_substep_us = kernel_times(sim, replay, resize_frame, resize_box)
    tuning = sim.tuning
    binned = tuning["path"] == 0
    host = host_transfers(sim, rec, 1e3 * elapsed / args.steps)
    sim.close()

    # the dominant kernel is the slowest one per substep; G2P, the kernel the
    # HBM target is set for, is reported beside it
    dom = max(per_substep_us, key=per_substep_us.get)
    traffic = (lambda k: measured_traffic(k)) if tuning == DEFAULT_TUNING else (lambda k: None)
    roof = roofline_entry(dom, n, g, per_launch_us, binned, traffic)
    roof.update({"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s"})
    roof["g2p"] = roofline_entry("g2p", n, g, per_launch_us, binned, traffic)
    roof["bytes_basis"] = ("algorithmic bytes of the reference's phases (204 B/particle + "
                           "84 B/cell per substep) charged per kernel; " + G2P_BYTES_DOC)
    substep_us = sum(per_substep_us.values())
    total_bytes = 204 * n + 84 * g

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (i32 fixed-point grid)",
        "data": "synthetic",
        "config": {
            "workload": f"MLS-MPM dam break, 300k particles, 80^3 grid, 2 substeps/frame, "
                        f"{args.steps} frames from the initial state, box x resized to 64 at "
                        f"frame {resize_frame}",
            "particles": n,
            "grid": [int(b) for b in box],
            "grid_cells": g,
            "substeps_per_step": SUBSTEPS_PER_FRAME,
            "timed_frames": args.steps,
            "resize_frame": resize_frame,
        },
        "roofline": roof,
        "stages_us_per_launch": per_launch_us,
        "stages_us_per_substep": per_substep_us,
        "kernels_per_substep": len([k for k, v in per_substep_us.items() if v > 0]),
        "stages_roofline": {k: roofline_entry(k, n, g, per_launch_us, binned, traffic)
                            for k in per_launch_us},
        "tuning": tuning,
        "substep_roofline": {
            "bytes": total_bytes,
            "sum_kernel_us": substep_us,
            "achieved_GBs": total_bytes / (substep_us * 1e-6) / 1e9,
            "frac": total_bytes / (substep_us * 1e-6) / 1e9 / HBM_PEAK_GBS,
            "wall_us_per_substep": 1e6 * elapsed / substeps,
            "wall_frac": total_bytes / (elapsed / substeps) / 1e9 / HBM_PEAK_GBS,
        },
        "host_transfers": host,
    }
    if not args.no_extras:
        out["config2"] = bench_config2()
        out["batched"] = bench_batched()
        out["adjoint"] = bench_adjoint()
        out["sph"] = bench_sph(cpu=not args.no_cpu_baseline)
        out["render"] = bench_render(cpu=not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(rec, box)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
