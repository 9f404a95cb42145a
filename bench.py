#!/usr/bin/env python3
"""Substep throughput of the MI355X PB-MPM solver on the 262,144-particle slab.

Workload: a 1.0 x 0.2 x 1.0 m Neo-Hookean slab of 128 x 16 x 128 = 262,144
particles on a 128 x 64 x 128-cell grid (129 x 65 x 129 nodes, dx = 1/128),
PB-MPM with 20 solver iterations per substep, dt = 2e-3, a kinematic sphere
indenter (r = 0.1 m) moving down at 0.05 m/s with friction 0.3, its target
updated every 10 substeps.  Inputs are synthetic (SplitMix64-seeded) and
already resident in HBM when timing starts.  A step is one substep.

Prints one JSON line: particle-substeps/s, the roofline of the dominant work
(a PB iteration after the first: its per-particle kernel, G2P and the material
constraint, and its per-pblock P2G scatter, 19 of the 20 iterations) measured
with HIP events on the solver stream over the timed region (the events ride
on the dispatches of every 10th substep), and a CPU baseline timed on this
host.
"""
import argparse
import copy
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cressim-mpm_amd", "python"))

import numpy as np  # noqa: E402

import crmpm_scenes as S  # noqa: E402

METRIC = "particle-substeps/s at 262k PB-MPM; batched & fwd+adj vs plain; HBM GB/s"
HBM_PEAK_GBS = 8000.0
# Minimal FP32 state traffic per particle for one G2P + one P2G of PB-MPM:
# P2G reads x, v, F, C, (alpha, r) = 12+12+36+36+8 = 104 B; G2P reads x and
# writes v, C = 12+12+36 = 60 B.  Grid: P2G writes and G2P reads 16 B per
# touched node.  (The iteration's two kernels also pass v and C between them
# through memory; that traffic is the implementation's, not the algorithm's.)
BYTES_PER_PARTICLE_ITERATION = 164
BYTES_PER_TOUCHED_NODE_ITERATION = 32
# The same split by kernel: the per-particle kernel reads x, F, (alpha, r)
# and the grid, and writes v, C; the scatter reads x, m, v, C and writes the
# momentum of the touched nodes.
BYTES_DENSE = (12 + 36 + 8 + 12 + 36, 16)
BYTES_SCATTER = (12 + 4 + 12 + 36, 12)
PROFILES = os.path.join(ROOT, "profiles")
# what the library's autotuner picked: launch sizes, the P2G sub-tile split
# and the G2P (gather) split
TUNED_KEYS = ("tile_waves", "update_waves", "sub_size", "sub_max", "gather_sub_size", "gather_sub_max", "scatter_waves")
PROFILE_EVERY = 10
COPY_CAL_BYTES = 256 << 20
LAUNCH = None  # --launch: a fixed launch configuration in place of the autotuner's


def copy_calibration():
    """Three device-to-device copies of 256 MiB: the known-size kernel that
    scales rocprofv3's FETCH_SIZE / WRITE_SIZE readings to bytes."""
    import torch

    x = torch.ones(COPY_CAL_BYTES // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copy calibration: 3 copies of {COPY_CAL_BYTES} bytes", flush=True)


def _short_name(name):
    import re

    # kernel name with its template arguments, e.g. k_pb_tile<1, false>
    m = re.search(r"(k_\w+(?:<[^<>()]*>)?|__amd_rocclr_\w+)", name)
    return m.group(1) if m else name.split("(")[0][-60:]


def _read_csv(directory, suffix):
    import csv
    import glob

    files = sorted(glob.glob(os.path.join(directory, "**", "*" + suffix), recursive=True))
    if not files:
        raise SystemExit(f"no *{suffix} under {directory}")
    with open(files[0]) as f:
        return list(csv.DictReader(f))


def _counter_per_launch(directory, counter=None):
    # kernel -> mean counter value per dispatch (FETCH_SIZE / WRITE_SIZE: KiB)
    acc = {}
    for r in _read_csv(directory, "counter_collection.csv"):
        if counter is not None and r["Counter_Name"] != counter:
            continue
        k = _short_name(r["Kernel_Name"])
        acc.setdefault(k, []).append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}, acc


# VALU issue model: a wave64 VALU instruction occupies its SIMD for 4 cycles;
# MI355X has 256 CUs x 4 SIMDs at up to 2.4 GHz.
SIMDS = 1024
CLOCK_HZ = 2.4e9


def valu_issue_us(instructions):
    return instructions * 4.0 / SIMDS / CLOCK_HZ * 1e6


# What --summarize covers per profiled workload: the bench invocation that
# rocprofv3 wraps (with the launch configuration the timed run's autotuner
# picked, so no trial dispatch enters the statistics), the summary's title and
# file name, and whether the counter passes belong to it or the kernel trace alone.
PB_LAUNCH = "0,4096,1048576,1,1048576,1"
MLS_CUTTING_LAUNCH = "16384,4096,1048576,1,1048576,0"
MLS_CUBE_LAUNCH = "4096,0,128,8,1048576,0"
PB_ADJOINT_LAUNCH = "4096,2048,1048576,1,1048576,1,4"
ADJOINT_LAUNCH = "0,4096,1048576,1,1048576,0,4"
PROFILE_TARGETS = {
    "pb_slab": dict(run=f"python3 bench.py --workload pb_slab --cpu-substeps 0 --launch {PB_LAUNCH}", counters=True,
                    title="Kernel trace and counters: PB-MPM slab, 262,144 particles, 20 iterations, 510 substeps",
                    md="pb_slab262k_kernel_trace.md"),
    "mls_cutting": dict(run=f"python3 bench.py --workload mls_cutting --cpu-substeps 0 --launch {MLS_CUTTING_LAUNCH}",
                        counters=True,
                        title="Kernel trace and counters: MLS-MPM cutting block, 2,097,152 particles, 510 substeps",
                        md="mls_cutting2m_kernel_trace.md"),
    "mls_cube": dict(run=f"python3 bench.py --workload mls_cube --cpu-substeps 0 --launch {MLS_CUBE_LAUNCH}",
                     counters=True,
                     title="Kernel trace and counters: MLS-MPM elastic cube, 32,768 particles, 205 substeps",
                     md="mls_cube32k_kernel_trace.md"),
    "pb_adjoint": dict(run=f"python3 bench.py --workload pb_adjoint --launch {PB_ADJOINT_LAUNCH}", counters=False,
                       title="Kernel trace: differentiable PB block, 65,536 particles, 20 iterations, 20 substeps, "
                             "forward + adjoint",
                       md="pb_adjoint65k_kernel_trace.md"),
    "batched": dict(run="python3 bench.py --workload batched", counters=False,
                    title="Kernel trace: 64 batched scenes (MLS 100 substeps, PB 20 substeps) and the same scenes "
                          "as single solvers",
                    md="batched64_kernel_trace.md"),
    "meshsdf": dict(run="python3 bench.py --workload meshsdf", counters=True,
                    title="Kernel trace and counters: mesh -> SDF table, 20,480-triangle icosphere on 128^3 nodes",
                    md="meshsdf_icosphere128_kernel_trace.md"),
    "engine": dict(run="python3 bench.py --workload engine", counters=False,
                   title="Kernel trace: engine advanceAll, 64 MLS scenes in frames of 10 substeps with host readback, "
                         "then the same batch stepped alone",
                   md="engine64_kernel_trace.md"),
    "adjoint": dict(run=f"python3 bench.py --workload adjoint --cpu-substeps 0 --launch {ADJOINT_LAUNCH}",
                    counters=False,
                    title="Kernel trace: differentiable MLS block, 65,536 particles, 100 substeps, forward + adjoint",
                    md="adjoint65k_kernel_trace.md"),
}

# Counter passes: (directory, counters, calibrated by the copy kernel).  Each
# pass fits gfx950's per-block limits (at most 4 TCC counters: FETCH_SIZE
# takes 3, WRITE_SIZE 2; at most 8 SQ counters).
COUNTER_PASSES = (
    ("fetch", "FETCH_SIZE", True),
    ("write", "WRITE_SIZE", True),
    ("sq", "SQ_INSTS_VALU SQ_INSTS_LDS SQ_INSTS_SALU SQ_WAVES SQ_BUSY_CYCLES SQ_WAVE_CYCLES SQ_ACTIVE_INST_VALU "
           "SQ_INSTS_VMEM_RD", False),
    ("tcc", "TCC_HIT_sum TCC_MISS_sum TCC_ATOMIC_sum TCC_EA0_ATOMIC_sum", False),
    ("dram", "TCC_EA0_RDREQ_DRAM_32B_sum TCC_EA0_WRREQ_WRITE_DRAM_32B_sum", True),
    ("stall", "SQ_WAVE_CYCLES SQ_WAIT_ANY SQ_WAIT_INST_ANY SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_LDS "
              "SQ_LDS_BANK_CONFLICT SQ_INSTS_VMEM_WR SQ_INSTS_LDS", False),
)


def counters_file(target):
    return PROFILE_TARGETS[target]["md"].replace("_kernel_trace.md", "_counters.json")


def profiled_row(target, kernel):
    """The committed counter figures of one kernel of a profiled workload
    (written by --summarize), or None."""
    path = os.path.join(PROFILES, counters_file(target))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(kernel)


def roofline_counters(row, avg_s):
    """Counter figures of a profiled kernel for a roofline line: HBM-side
    traffic per launch, the L2 hit rate, DRAM-path bytes, atomics, and the
    VALU issue time over the measured duration."""
    if row is None:
        return {"traffic": None}
    vi = row.get("valu_instructions")
    rd, wr = row.get("dram_path_read_bytes"), row.get("dram_path_write_bytes")
    return {"traffic": row.get("bytes_per_launch"), "l2_hit_rate": row.get("l2_hit_rate"),
            "dram_path_bytes": None if rd is None or wr is None else rd + wr,
            "atomics_per_launch": row.get("atomics_l2"),
            "valu_issue_frac": None if vi is None else valu_issue_us(vi) / (avg_s * 1e6),
            "profile_avg_launch_us": row.get("avg_us")}


def iteration_counters(rows, avg_s):
    """roofline_counters over several kernels run back to back (a PB
    iteration): traffic, DRAM-path bytes, atomics and VALU instructions
    summed, the L2 hit rate from the summed hits and misses."""
    if any(r is None for r in rows):
        return {"traffic": None}
    tot = lambda key: None if any(r.get(key) is None for r in rows) else sum(r[key] for r in rows)  # noqa: E731
    h, m = tot("tcc_hit"), tot("tcc_miss")
    rd, wr = tot("dram_path_read_bytes"), tot("dram_path_write_bytes")
    vi = tot("valu_instructions")
    return {"traffic": tot("bytes_per_launch"),
            "l2_hit_rate": None if h is None or m is None or h + m == 0 else h / (h + m),
            "dram_path_bytes": None if rd is None or wr is None else rd + wr,
            "atomics_per_launch": tot("atomics_l2"),
            "valu_issue_frac": None if vi is None else valu_issue_us(vi) / (avg_s * 1e6),
            "profile_avg_launch_us": tot("avg_us")}


def profile_commands(target):
    t = PROFILE_TARGETS[target]
    run = t["run"]
    cmds = [f"rocprofv3 --kernel-trace --stats --output-format csv -d DIR/trace -o run -- {run}"]
    if t["counters"]:
        for d, counters, cal in COUNTER_PASSES:
            cmds.append(f"rocprofv3 --pmc {counters} --output-format csv -d DIR/{d} -o run -- {run}")
            if cal:
                cmds.append(f"rocprofv3 --pmc {counters} --output-format csv -d DIR/{d}_copy -o cp -- "
                            "python3 bench.py --copy-calibration")
    cmds.append(f"python3 bench.py --summarize DIR --profile-target {target}")
    return cmds


def _copy_reading(directory, counter):
    _, per = _counter_per_launch(directory, counter)
    return max(per["__amd_rocclr_copyBuffer"])


def summarize(out_dir, title, commands, counters=True):
    """Turn the rocprofv3 outputs of one profiling session into the committed
    summary: kernel-trace statistics and, with counters, per dispatch of each
    kernel: L2-to-memory bytes (FETCH_SIZE + WRITE_SIZE), bytes on the L2's
    DRAM path (requests the L2 sends towards memory: served by the MALL or by
    HBM), each scaled by what the same counter reads for a device copy of known
    size; the L2 hit rate; atomic requests at the L2 and beyond it; the VALU
    instruction count."""
    stats = _read_csv(os.path.join(out_dir, "trace"), "kernel_stats.csv")
    c = {}
    fac = {}
    if counters:
        kib = COPY_CAL_BYTES / 1024.0
        c["fetch"] = _counter_per_launch(os.path.join(out_dir, "fetch"))[0]
        c["write"] = _counter_per_launch(os.path.join(out_dir, "write"))[0]
        fac["fetch"] = kib / _copy_reading(os.path.join(out_dir, "fetch_copy"), "FETCH_SIZE")
        fac["write"] = kib / _copy_reading(os.path.join(out_dir, "write_copy"), "WRITE_SIZE")
        c["valu"] = _counter_per_launch(os.path.join(out_dir, "sq"), "SQ_INSTS_VALU")[0]
        for name in ("TCC_HIT_sum", "TCC_MISS_sum", "TCC_ATOMIC_sum", "TCC_EA0_ATOMIC_sum"):
            c[name] = _counter_per_launch(os.path.join(out_dir, "tcc"), name)[0]
        for name in ("TCC_EA0_RDREQ_DRAM_32B_sum", "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum"):
            c[name] = _counter_per_launch(os.path.join(out_dir, "dram"), name)[0]
            # 32-byte requests -> bytes, scaled so the copy reads (writes) its 256 MiB
            fac[name] = COPY_CAL_BYTES / (32.0 * _copy_reading(os.path.join(out_dir, "dram_copy"), name))
        for name in ("SQ_WAVE_CYCLES", "SQ_WAIT_ANY", "SQ_WAIT_INST_ANY", "SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_LDS",
                     "SQ_LDS_BANK_CONFLICT"):
            c["stall_" + name] = _counter_per_launch(os.path.join(out_dir, "stall"), name)[0]
    lines = [f"# {title}", "", "Commands (repository root, rocprofv3 on one MI355X):", "", "```"]
    lines += commands + ["```", ""]
    if counters:
        lines += ["| kernel | calls | avg µs | total % | FETCH_SIZE KiB | WRITE_SIZE KiB | L2↔memory MB (calibrated) "
                  "| DRAM-path read MB | DRAM-path write MB | L2 hit % | atomics at L2 | atomics past L2 "
                  "| SQ_INSTS_VALU | VALU issue µs | waiting % | issue-blocked % | VALU active % | LDS active % "
                  "| LDS bank conflict % |",
                  "|---|---|---|---|---|---|---|---|---|---|---|---|---|---|---|---|---|---|---|"]
    else:
        lines += ["| kernel | calls | avg µs | total % |", "|---|---|---|---|"]
    merged = {}
    for r in stats:
        k = _short_name(r["Name"])
        n_, t_ = merged.get(k, (0, 0.0))
        merged[k] = (n_ + int(r["Calls"]), t_ + float(r["TotalDurationNs"]))
    total = sum(t for _, t in merged.values())

    def per_kernel(k):
        if not counters:
            return {}
        g = lambda key: c[key].get(k)  # noqa: E731
        out = {"fetch_kib_raw": g("fetch"), "write_kib_raw": g("write"), "valu_instructions": g("valu")}
        if out["fetch_kib_raw"] is not None and out["write_kib_raw"] is not None:
            out["bytes_per_launch"] = 1024.0 * (out["fetch_kib_raw"] * fac["fetch"] + out["write_kib_raw"] * fac["write"])
        h, m = g("TCC_HIT_sum"), g("TCC_MISS_sum")
        out["tcc_hit"], out["tcc_miss"] = h, m
        out["l2_hit_rate"] = None if h is None or m is None or h + m == 0 else h / (h + m)
        out["atomics_l2"], out["atomics_past_l2"] = g("TCC_ATOMIC_sum"), g("TCC_EA0_ATOMIC_sum")
        rd, wr = g("TCC_EA0_RDREQ_DRAM_32B_sum"), g("TCC_EA0_WRREQ_WRITE_DRAM_32B_sum")
        out["dram_path_read_bytes"] = None if rd is None else 32.0 * rd * fac["TCC_EA0_RDREQ_DRAM_32B_sum"]
        out["dram_path_write_bytes"] = None if wr is None else 32.0 * wr * fac["TCC_EA0_WRREQ_WRITE_DRAM_32B_sum"]
        wc = g("stall_SQ_WAVE_CYCLES")
        for key, name in (("wait_frac", "SQ_WAIT_ANY"), ("issue_wait_frac", "SQ_WAIT_INST_ANY"),
                          ("valu_active_frac", "SQ_ACTIVE_INST_VALU"), ("lds_active_frac", "SQ_ACTIVE_INST_LDS"),
                          ("lds_bank_conflict_frac", "SQ_LDS_BANK_CONFLICT")):
            v = g("stall_" + name)
            out[key] = None if v is None or not wc else v / wc
        return out

    f = lambda v, fmt: "" if v is None else format(v, fmt)  # noqa: E731
    for k, (n_, t_) in sorted(merged.items(), key=lambda kv: -kv[1][1]):
        row = f"| `{k}` | {n_} | {t_ / n_ / 1e3:.1f} | {100 * t_ / total:.2f} |"
        if counters:
            q = per_kernel(k)
            mb = lambda key: None if q.get(key) is None else q[key] / 1e6  # noqa: E731
            row += (f" {f(q['fetch_kib_raw'], ',.0f')} | {f(q['write_kib_raw'], ',.0f')} | {f(mb('bytes_per_launch'), '.2f')}"
                    f" | {f(mb('dram_path_read_bytes'), '.2f')} | {f(mb('dram_path_write_bytes'), '.2f')}"
                    f" | {f(None if q['l2_hit_rate'] is None else 100 * q['l2_hit_rate'], '.1f')}"
                    f" | {f(q['atomics_l2'], ',.0f')} | {f(q['atomics_past_l2'], ',.0f')}"
                    f" | {f(q['valu_instructions'], ',.0f')}"
                    f" | {f(None if q['valu_instructions'] is None else valu_issue_us(q['valu_instructions']), '.1f')}"
                    + "".join(f" | {f(None if q[key] is None else 100 * q[key], '.1f')}"
                              for key in ("wait_frac", "issue_wait_frac", "valu_active_frac", "lds_active_frac",
                                          "lds_bank_conflict_frac")) + " |")
        lines.append(row)
    lines += ["", "Counter columns are per dispatch, averaged over the kernel's dispatches. Kernel names carry their "
              "template arguments (the tile kernels' last one selects the persistent launch form)."]
    if not counters:
        return "\n".join(lines) + "\n", None
    lines += ["", f"Calibration: a device-to-device copy of {COPY_CAL_BYTES >> 20} MiB reads FETCH_SIZE "
              f"{COPY_CAL_BYTES / 1024 / fac['fetch']:,.0f} KiB and WRITE_SIZE {COPY_CAL_BYTES / 1024 / fac['write']:,.0f} KiB "
              f"per dispatch, so L2↔memory bytes = FETCH_SIZE x {fac['fetch']:.3f} KiB + WRITE_SIZE x {fac['write']:.3f} KiB; "
              f"its TCC_EA0_RDREQ_DRAM_32B_sum and TCC_EA0_WRREQ_WRITE_DRAM_32B_sum read "
              f"{fac['TCC_EA0_RDREQ_DRAM_32B_sum']:.3f} and {fac['TCC_EA0_WRREQ_WRITE_DRAM_32B_sum']:.3f} of its bytes "
              "/ 32 inverted, the scale applied to the DRAM-path columns.",
              "", "DRAM-path bytes are the L2's requests towards memory (TCC_EA0_*_DRAM): the 256 MB MALL (Infinity "
              "Cache) sits behind the L2 and serves those it holds, so they bound HBM traffic from above; no counter "
              "here separates MALL hits from HBM.",
              "", "L2 hit % = TCC_HIT / (TCC_HIT + TCC_MISS). Atomics at L2 = TCC_ATOMIC_sum (atomic requests of all "
              "types reaching the L2); past L2 = TCC_EA0_ATOMIC_sum.",
              "", "VALU issue µs = SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x 2.4 GHz): the time the VALU pipes need "
              "for the kernel's vector instructions; a kernel whose duration is close to it is bound by VALU issue, not HBM.",
              "", "Wave-state columns are fractions of the kernel's wavefront cycles (SQ_WAVE_CYCLES): waiting = "
              "SQ_WAIT_ANY (at an s_waitcnt: memory or LDS results outstanding), issue-blocked = SQ_WAIT_INST_ANY "
              "(an instruction ready but not issued: its unit busy with other wavefronts or a dependency), VALU and "
              "LDS active = SQ_ACTIVE_INST_VALU and SQ_ACTIVE_INST_LDS, LDS bank conflict = SQ_LDS_BANK_CONFLICT "
              "cycles. Summed over the wavefronts resident on a SIMD, the VALU-active fraction approaches 1 when "
              "the SIMD's VALU is saturated."]
    result = {}
    for k in merged:
        q = per_kernel(k)
        n_, t_ = merged[k]
        q["calls"] = n_
        q["avg_us"] = t_ / n_ / 1e3
        result[k] = q
    result["_calibration"] = fac
    return "\n".join(lines) + "\n", result


_CPU_INFO = None


def host_cpu():
    """The host CPU's model name and the cores this process may use."""
    global _CPU_INFO
    if _CPU_INFO is None:
        model = platform.processor() or platform.machine()
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
        usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        _CPU_INFO = {"cpu_model": model, "nproc": usable}
    return _CPU_INFO


def cpu_baseline(scene, substeps, label=None):
    """The reference's own CPU solver (compiled from its sources, oracle/_ref)
    or, where that build is absent, the oracle's restatement of it, one
    thread as the reference runs, on `substeps` substeps of the scene (its
    first kinematic frame applied)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cpu_checkers as CC

    which = "ref" if CC.ref_available() else "oracle"
    sc = copy.deepcopy(scene)
    o = CC.CpuSolver(sc, which)
    o.compute_initial_data()
    if sc["frames"]:
        S.apply_frame(sc, sc["frames"][0])
        o.upload_shapes()
    t0 = time.perf_counter()
    o.step(substeps)
    dt = time.perf_counter() - t0
    o.close()
    n = len(scene["xm"])
    info = host_cpu()
    return {
        "value": n * substeps / dt,
        "unit": "particle-substeps/s",
        "cores": 1,
        "kind": "reference" if which == "ref" else "port",
        "sample": f"{substeps} substep(s) of the full {n}-particle scene{label or ''}, {dt:.1f} s on 1 core",
        **info,
    }


def batched(kind_name, steps, warmup, tune=True, graph=True, n_streams=4):
    """64 scenes of uneven size (config3) advanced as one batch, against the
    same scenes as 64 single-scene solvers (same kernels, each autotuned and
    graph-replayed like the batch):
      * sum of plain: each solver's substeps timed alone, summed;
      * advanceAll analogue: every solver's substeps enqueued solver by
        solver, as SimulationFactoryImpl::_advanceAllTaskFn steps them
        (src/engine/simulation_factory_impl.cpp:1106-1150), on one stream
        (the reference's default, numCudaStreams = 1) and on `n_streams`
        streams taken by load (getCudaStreamByLoad, :1093-1104), timed from
        the first enqueue to the last stream's end.
    Returns the batch throughput and the speedups."""
    import torch

    import crmpm_hip as H

    kind = {"mls": S.MLS, "pb": S.PB}[kind_name]
    scenes = S.config3(kind=kind, steps=steps + warmup)
    n_total = sum(len(sc["xm"]) for sc in scenes)
    b = H.HipBatch(copy.deepcopy(scenes))
    b.compute_initial_data()
    b.step(warmup)
    tuned = None
    if tune:
        tuned = b.autotune(steps_per_trial=2) + b.gather_split() + (b.scatter_waves(),)
    b.set_graph_mode(graph)
    b.step(1)  # capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    finite = all(np.isfinite(b.state(i)["xm"]).all() for i in (0, len(scenes) - 1))
    b.close()
    streams = [torch.cuda.Stream() for _ in range(n_streams)]
    solvers = []
    for sc in scenes:
        g = H.HipSolver(copy.deepcopy(sc), stream=streams[0])
        g.compute_initial_data()
        g.step(warmup)
        if tune:
            g.autotune(steps_per_trial=2)
        g.set_graph_mode(graph)
        g.step(1)  # capture
        solvers.append(g)
    torch.cuda.synchronize()
    t_plain = 0.0
    for g in solvers:
        t0 = time.perf_counter()
        g.step(steps)
        torch.cuda.synchronize()
        t_plain += time.perf_counter() - t0

    def advance_all():
        t0 = time.perf_counter()
        for g in solvers:
            g.step(steps)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    t_one = advance_all()
    # n streams by load: each solver takes the least-loaded stream when it is created
    load = [0] * n_streams
    for g in solvers:
        k = min(range(n_streams), key=lambda i: load[i])
        load[k] += 1
        g.set_stream(streams[k])
    torch.cuda.synchronize()
    t_multi = advance_all()
    for g in solvers:
        g.close()
    return {"workload": f"{len(scenes)} {kind_name.upper()} scenes, log-uniform 2k-200k particles, grids 32^3-128^3",
            "particles": n_total, "substeps": steps, "iterations": scenes[0]["iterations"] if kind == S.PB else None,
            "value": n_total * steps / t_batch, "unit": "particle-substeps/s",
            "batch_ms_per_substep": t_batch / steps * 1e3,
            "sum_of_plain_ms_per_substep": t_plain / steps * 1e3,
            "speedup_vs_sum_of_plain": t_plain / t_batch,
            "advance_all_1_stream_ms_per_substep": t_one / steps * 1e3,
            "speedup_vs_advance_all_1_stream": t_one / t_batch,
            f"advance_all_{n_streams}_streams_ms_per_substep": t_multi / steps * 1e3,
            f"speedup_vs_advance_all_{n_streams}_streams": t_multi / t_batch,
            "plain_solvers": "autotuned, hipGraph replay" if tune and graph else
                             ("autotuned" if tune else ("hipGraph replay" if graph else "eager")),
            "autotuned": None if tuned is None else dict(zip(TUNED_KEYS, tuned)),
            "hip_graph": graph, "finite": bool(finite)}


def engine(frames=10, substeps=10, warmup=1):
    """The engine runtime (advanceAll): the 64 MLS scenes of the batched
    workload held by one engine, advanced by frames of 10 substeps, each frame
    ending with positions, velocities and coupling forces back in pinned host
    memory; against the same batch stepped without the engine's per-frame
    uploads and readback."""
    import torch

    import crmpm_hip as H

    scenes = S.config3(kind=S.MLS, steps=1)
    n_total = sum(len(sc["xm"]) for sc in scenes)
    dt = scenes[0]["dt"]
    frame = (substeps - 0.5) * dt  # FP32 accumulation reaches it after `substeps` steps
    eng, _, _ = H.engine_from_scenes(copy.deepcopy(scenes))
    eng.advance_all(frame * warmup)
    eng.fetch_results_all()
    eng.autotune(steps_per_trial=2)
    eng.set_graph_mode(True)
    eng.advance_all(frame)
    eng.fetch_results_all()
    t0 = time.perf_counter()
    for _ in range(frames):
        eng.advance_all(frame)
        eng.fetch_results_all()
    t_eng = time.perf_counter() - t0
    k = eng.substeps(0)
    finite = bool(np.isfinite(eng.host(0)["xm"]).all() and np.isfinite(eng.host(len(scenes) - 1)["xm"]).all())
    eng.close()
    b = H.HipBatch(copy.deepcopy(scenes))
    b.compute_initial_data()
    b.step(substeps * warmup)
    b.autotune(steps_per_trial=2)
    b.set_graph_mode(True)
    b.step(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.step(frames * k)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    b.close()
    # the reference's CPU MLS solver (or the oracle) on the largest scene, a few substeps
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cpu_checkers as CC

    which = "ref" if CC.ref_available() else "oracle"
    big = max(scenes, key=lambda sc: len(sc["xm"]))
    o = CC.CpuSolver(copy.deepcopy(big), which)
    o.compute_initial_data()
    cpu_steps = 30
    t0 = time.perf_counter()
    o.step(cpu_steps)
    dt_cpu = time.perf_counter() - t0
    o.close()
    cpu = {"value": len(big["xm"]) * cpu_steps / dt_cpu, "unit": "particle-substeps/s", "cores": 1,
           "kind": "reference" if which == "ref" else "port",
           "sample": f"{cpu_steps} substeps of the largest scene ({len(big['xm'])} particles) in {dt_cpu:.1f} s on 1 core"}
    return {"workload": f"{len(scenes)} MLS scenes ({n_total} particles) in one engine, frames of {k} substeps with "
                        "host readback", "particles": n_total, "frames": frames, "substeps_per_frame": k,
            "cpu_baseline": cpu,
            "value": n_total * k * frames / t_eng, "unit": "particle-substeps/s",
            "ms_per_frame": t_eng / frames * 1e3, "batch_only_ms_per_frame": t_batch / frames * 1e3,
            "engine_overhead_frac": t_eng / t_batch - 1.0,
            # each frame moves 32 B per particle (position|mass and velocity) to the
            # host; the rate is over the engine's time beyond the batch's own
            "readback_GBps": 32.0 * n_total * frames / max(t_eng - t_batch, 1e-9) / 1e9, "finite": finite}


def adjoint(steps=100, every=10, cpu=True):
    """Differentiable MLS block (config4: 65,536 particles, 96^3 grid, GRF log E):
    time of a forward rollout with checkpoints, of forward + adjoint, and of the
    plain solver's forward, for the same substeps."""
    import torch

    import crmpm_hip as H

    sc, log_e = S.config4(steps=steps)
    hidden, log_e_true = S.config4(steps=steps, field_seed=S.HIDDEN_FIELD_SEED)
    sim = H.DifferentiableMls(sc, steps, checkpoint_every=every, tune=LAUNCH or True)
    dev = sim.solver.device
    nu = 0.3
    v0 = torch.as_tensor(sc["v"][:, :3], dtype=torch.float64, device=dev)

    def lame_t(le):
        E = torch.exp(torch.as_tensor(le, dtype=torch.float64, device=dev))
        return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))

    lt, mt = lame_t(log_e_true)
    # final positions under the hidden field plus 1 mm measurement noise
    noise = torch.as_tensor(S.gaussian_noise(len(log_e), S.TARGET_NOISE_SEED, S.TARGET_NOISE), device=dev)
    target = sim.forward(lt, mt, v0).double() + noise
    lam, mu = lame_t(log_e)
    n = sim.n

    def fwd():
        return sim.forward(lam, mu, v0)

    def fwd_adj():
        x = sim.forward(lam, mu, v0)
        gx = 2.0 * (x.double() - target) / n
        return sim.backward(gx)

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        best = 1e30
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    sim_tuned = sim.tuned
    t_f = timed(fwd)
    t_fa = timed(fwd_adj)
    x = fwd().double()
    loss = float(((x - target) ** 2).sum(1).mean())
    dlam, dmu, dv0 = fwd_adj()
    g = (dlam * lam + dmu * mu)
    finite = bool(torch.isfinite(g).all() and torch.isfinite(dv0).all())
    sim.close()
    plain = H.HipSolver(copy.deepcopy(sc))
    plain.compute_initial_data()
    if LAUNCH:
        plain.set_launch(LAUNCH)
    else:
        plain.autotune(steps_per_trial=2)
    plain.set_graph_mode(True)
    plain.step(5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain.step(steps)
    torch.cuda.synchronize()
    t_p = time.perf_counter() - t0
    plain.close()
    return {"workload": "MLS block 64x32x32 = 65,536 particles, 96^3 grid, GRF log E, loss MSE to a hidden-field rollout",
            "particles": n, "substeps": steps, "checkpoint_every": every,
            "plain_forward_ms_per_substep": t_p / steps * 1e3,
            "checkpointed_forward_ms_per_substep": t_f / steps * 1e3,
            "forward_adjoint_ms_per_substep": t_fa / steps * 1e3,
            "forward_adjoint_vs_plain_forward": t_fa / t_p, "plain_forward": "autotuned, hipGraph replay",
            "loss": loss, "autotuned": dict(zip(TUNED_KEYS, sim_tuned)),
            "grad_logE_abs_max": float(g.abs().max()), "grad_v0_abs_max": float(dv0.abs().max()),
            "cpu_baseline": cpu_baseline(sc, 20, label=" (forward only: the reference has no adjoint)") if cpu else None,
            "cpu_forward_adjoint": adjoint_cpu_rate(sc, 10) if cpu else None,
            "finite": finite}


def pb_block_scene(steps=20):
    """Differentiable PB block: the config-4 block's 65,536 particles, 96^3
    grid and initial velocities as a Neo-Hookean PB-MPM body, 20 iterations,
    dt = 2e-3, alpha = 0.5 / dt, ratio 0.9 (the slab's constraint
    parameters)."""
    sc, _ = S.config4(steps=steps)
    sc["kind"] = S.PB
    sc["iterations"] = 20
    sc["dt"] = 2e-3
    sc["params"][:, 0] = 0.5 / sc["dt"]
    sc["params"][:, 1] = 0.9
    sc["type"][:] = S.NEOHOOKEAN
    return sc


def pb_adjoint(steps=20, every=10):
    """Differentiable PB block (pb_block_scene): forward + adjoint (the
    constraint's adjoint through the SVD) against the plain solver's
    forward, for the same substeps; the loss is the MSE to the final
    positions plus 1 mm of Gaussian noise (the constraint's parameters
    barely move the positions over 20 substeps, so a rollout with other
    parameters would give a loss near zero)."""
    import torch

    import crmpm_hip as H

    sc = pb_block_scene(steps)
    sim = H.DifferentiablePb(sc, steps, checkpoint_every=every, tune=LAUNCH or True)
    dev = sim.solver.device
    alpha = torch.as_tensor(sc["params"][:, 0], dtype=torch.float64, device=dev)
    ratio = torch.as_tensor(sc["params"][:, 1], dtype=torch.float64, device=dev)
    v0 = torch.as_tensor(sc["v"][:, :3], dtype=torch.float64, device=dev)
    n = sim.n
    noise = torch.as_tensor(S.gaussian_noise(n, S.TARGET_NOISE_SEED, S.TARGET_NOISE), device=dev)
    target = sim.forward(alpha, ratio, v0).double() + noise

    def fwd():
        return sim.forward(alpha, ratio, v0)

    def fwd_adj():
        x = sim.forward(alpha, ratio, v0)
        return sim.backward(2.0 * (x.double() - target) / n)

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        best = 1e30
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    tuned = sim.tuned
    t_f = timed(fwd)
    t_fa = timed(fwd_adj)
    loss = float(((fwd().double() - target) ** 2).sum(1).mean())
    da, dr, dv0 = fwd_adj()
    finite = bool(torch.isfinite(da).all() and torch.isfinite(dr).all() and torch.isfinite(dv0).all())
    sim.close()
    plain = H.HipSolver(copy.deepcopy(sc))
    plain.compute_initial_data()
    if LAUNCH:
        plain.set_launch(LAUNCH)
    else:
        plain.autotune(steps_per_trial=2)
    plain.set_graph_mode(True)
    plain.step(2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain.step(steps)
    torch.cuda.synchronize()
    t_p = time.perf_counter() - t0
    plain.close()
    return {"workload": "PB block 64x32x32 = 65,536 particles, 96^3 grid, Neo-Hookean, 20 iterations, dt 2e-3, "
                        "loss MSE to the final positions plus 1 mm noise",
            "particles": n, "substeps": steps, "iterations": 20, "checkpoint_every": every,
            "plain_forward_ms_per_substep": t_p / steps * 1e3,
            "checkpointed_forward_ms_per_substep": t_f / steps * 1e3,
            "forward_adjoint_ms_per_substep": t_fa / steps * 1e3,
            "forward_adjoint_vs_plain_forward": t_fa / t_p, "plain_forward": "autotuned, hipGraph replay",
            "loss": loss, "autotuned": dict(zip(TUNED_KEYS, tuned)),
            "grad_alpha_abs_max": float(da.abs().max()), "grad_ratio_abs_max": float(dr.abs().max()),
            "grad_v0_abs_max": float(dv0.abs().max()), "finite": finite}


def adjoint_cpu_rate(sc, substeps):
    """The FP64 autodiff restatement of the MLS substep (oracle/mls_autodiff.py,
    torch on the host, the checker of the GPU adjoint) timed on `substeps`
    substeps of forward + reverse of the block: the nearest CPU counterpart
    of the differentiable form, which the reference lacks."""
    import torch

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cpu_checkers as CC
    import mls_autodiff as AD

    n = len(sc["xm"])
    o = CC.CpuSolver(copy.deepcopy(sc), "oracle", fp64=True)
    o.compute_initial_data()
    V0 = o.state()["V0"]
    o.close()
    t = lambda a: torch.as_tensor(np.asarray(a, np.float64))  # noqa: E731
    lam = t(sc["params"][:, 0]).requires_grad_(True)
    mu = t(sc["params"][:, 1]).requires_grad_(True)
    v0 = t(sc["v"][:, :3]).requires_grad_(True)
    threads = torch.get_num_threads()
    t0 = time.perf_counter()
    c = AD.consts_from_scene(sc, V0, lam=lam, mu=mu)
    x, _, _, _ = AD.rollout(t(sc["xm"][:, :3]), v0, torch.eye(3, dtype=torch.float64).repeat(n, 1, 1),
                            torch.zeros(n, 3, 3, dtype=torch.float64), c, substeps)
    (x ** 2).sum().backward()
    dt = time.perf_counter() - t0
    return {"value": n * substeps / dt, "unit": "particle-substeps/s (forward + adjoint)", "cores": threads,
            "kind": "port", "sample": f"{substeps} substeps of the {n}-particle block, forward + reverse, FP64 torch "
                                      f"autograd on {threads} threads, {dt:.1f} s", **host_cpu()}


def mls_run(scene, steps, warmup, tune=True):
    """Run an MLS scene (applying its kinematic frames) and time `steps`
    substeps; every 10th substep records HIP events on all its launches."""
    import torch

    import crmpm_hip as H

    solver = H.HipSolver(copy.deepcopy(scene))
    solver.compute_initial_data()
    frames = solver.scene["frames"]
    fsub = frames[0]["substeps"] if frames else 0

    def run(first, count, sample=False):
        for s in range(first, first + count):
            if frames and (s % fsub == 0 or s == first):
                S.apply_frame(solver.scene, frames[s // fsub])
                solver._upload_shapes()
            if sample:
                solver.profile_mask(H.PROF_CATEGORIES if (s - first) % PROFILE_EVERY == 0 else [])
            solver.step(1)

    run(0, warmup)
    if LAUNCH:
        tuned = solver.set_launch(LAUNCH)
    else:
        tuned = solver.autotune(steps_per_trial=2) + solver.gather_split() + (solver.scatter_waves(),) if tune else None
    # substeps replay a captured hipGraph (the sampled ones, which record events, run eagerly)
    solver.set_graph_mode(True)
    torch.cuda.synchronize()
    grid = solver.grid()
    touched = int(np.count_nonzero(grid[:, 3] > 0))
    solver.profile_begin(max_launches=(steps // PROFILE_EVERY + 1) * 24 + 64)
    solver.profile_mask([])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(warmup, steps, sample=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = solver.profile_end()
    st = solver.state()
    solver.close()
    return elapsed, prof, touched, tuned, st


def mls_kernel_roofline(n, touched, prof, slicer, target=None, tuned=None):
    """Achieved HBM rate of the MLS P2G and G2P kernels on their algorithmic
    bytes: P2G reads x, v, F, C, V0, (lambda, mu) = 108 B per particle and
    writes 16 B per touched node; G2P reads x, F (and x again per slicer test,
    12 B) and writes x, v, F, C = 96 B per particle and reads 16 B per node."""
    out = {}
    pers = "true" if tuned and tuned[0] > 0 else "false"
    # the kernels' names as the profiler lists them; gather split 0 is the
    # per-particle G2P
    rows = {"mls_p2g": f"k_mls_p2g<{pers}, 4>",
            "mls_g2p": "k_g2p_dense<0>" if tuned and tuned[5] == 0 else f"k_mls_g2p<{pers}, 4>"}
    for name, per_p in (("mls_p2g", 108), ("mls_g2p", 48 + (12 if slicer else 0) + 96)):
        ms, k = prof[name]
        if not k:
            continue
        avg = ms / k / 1e3
        achieved = (per_p * n + 16 * touched) / avg / 1e9
        row = rows[name]
        out[name] = {"kernel": row, "avg_launch_us": avg * 1e6, "launches": k, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "algorithmic_bytes_per_launch": per_p * n + 16 * touched,
                     **(roofline_counters(profiled_row(target, row), avg) if target else {})}
    return out


def mls_cutting(steps=500, warmup=10, cpu=True):
    """MLS at scale with cutting (config5: 2,097,152 particles, 257 x 129 x 257 nodes)."""
    sc = S.config5(steps=steps + warmup)
    n = len(sc["xm"])
    elapsed, prof, touched, tuned, st = mls_run(sc, steps, warmup)
    return {"workload": "MLS block 256x32x256 = 2,097,152 particles, 257x129x257 nodes, capsule tool + growing quad "
                        "slicer cutting at 0.02 m/s", "particles": n, "substeps": steps,
            "value": n * steps / elapsed, "unit": "particle-substeps/s", "ms_per_substep": elapsed / steps * 1e3,
            "touched_nodes": touched, "roofline": mls_kernel_roofline(n, touched, prof, slicer=True, target="mls_cutting",
                                            tuned=tuned),
            "autotuned": dict(zip(TUNED_KEYS, tuned)),
            "cpu_baseline": cpu_baseline(sc, 2) if cpu else None,
            "finite": bool(np.isfinite(st["xm"]).all())}


def lattice(reps=5):
    """Particle-object lattice on the device (ParticleObjectImpl::initialize):
    a sphere of radius 0.3 m at 2 mm spacing (304^3 lattice points), against
    the oracle's restatement of the engine's host loop over the same lattice."""
    import torch

    import crmpm_hip as H

    obj = S.lattice_object(S.SPHERE, (0.3, 0, 0, 0), 0.002, 8e-6, position=(0.5, 0.5, 0.5))
    n, pts = H.particle_lattice(obj)  # warm-up, sizes the output
    torch.cuda.synchronize()
    best = 1e30
    for _ in range(reps):
        t0 = time.perf_counter()
        m, _ = H.particle_lattice(obj)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    dims = [int((np.float32(0.6) + np.float32(0.008)) / np.float32(0.002))] * 3
    points = dims[0] * dims[1] * dims[2]
    out = {"workload": "sphere r = 0.3 m at 2 mm spacing: lattice points -> particles (float4 position|mass)",
           "lattice_points": points, "particles": n, "seconds": best, "points_per_s": points / best,
           "particles_per_s": n / best,
           # two SDF passes over the lattice, one 16-byte write per particle
           "hbm_write_GBps": 16.0 * n / best / 1e9}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cpu_checkers as CC

    t0 = time.perf_counter()
    k, _ = CC.particle_lattice(obj, cap=0)  # one pass: count only
    dt = time.perf_counter() - t0
    out["cpu_baseline"] = {"value": points / dt, "unit": "lattice points/s", "cores": 1, "kind": "port",
                           "sample": f"the full {points}-point lattice ({k} particles, count pass only) "
                                     f"in {dt:.2f} s on 1 core"}
    out["finite"] = bool(torch.isfinite(pts).all())
    return out


def mesh_sdf(reps=3):
    """Triangle mesh -> SDF table on the device (SdfBuilder's computeSdf +
    computeSdfGradient): a subdivided icosahedron of radius 0.1 m, 20,480
    triangles, on a 128^3 table, against the reference's host builder (or the
    oracle) on the same mesh at a coarser cell, per table node."""
    import torch

    import crmpm_hip as H

    pts, tris = S.icosphere(0.1, 5)
    cell, fatten = 0.0018, 0.0147

    def timed(rule):
        dims, _, out = H.mesh_sdf(pts, tris, cell, fatten, rule=rule)  # warm-up, loads the code objects
        torch.cuda.synchronize()
        best = 1e30
        for _ in range(reps):
            t0 = time.perf_counter()
            H.mesh_sdf(pts, tris, cell, fatten, rule=rule)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return dims, out, best

    dims, out, best = timed("reference")
    _, _, best_exact = timed("exact")
    nodes = int(np.prod(dims))
    pairs = nodes * len(tris)
    res = {"workload": "icosphere r = 0.1 m, 20,480 triangles -> SDF table of 128^3 nodes (float4 gradient|distance)",
           "rule": "reference (signs by the tree's winding approximation)",
           "dims": [int(d) for d in dims], "triangles": len(tris), "seconds": best, "nodes_per_s": nodes / best,
           "node_triangle_pairs_per_s": pairs / best, "seconds_exact_rule": best_exact,
           "finite": bool(torch.isfinite(out).all())}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cpu_checkers as CC

    which = "reference" if CC.ref_available() else "oracle"
    ccell = cell * 4
    t0 = time.perf_counter()
    cd, _, _ = CC.mesh_sdf(pts, tris, ccell, fatten, which=which)
    dt = time.perf_counter() - t0
    cn = int(np.prod(cd))
    res["cpu_baseline"] = {"value": cn / dt, "unit": "table nodes/s", "cores": 1,
                           "kind": "reference" if which == "reference" else "port",
                           "sample": f"the same mesh on a {tuple(int(d) for d in cd)} table (cell {ccell:.4f} m) in "
                                     f"{dt:.2f} s on 1 core"}
    res["speedup_vs_cpu"] = res["nodes_per_s"] / res["cpu_baseline"]["value"]
    return res


def mls_cube(steps=200, warmup=5, cpu=True):
    """The reference's CPU-runnable case (config1: 32^3 = 32,768 particles, 64^3 cells, 200 substeps)."""
    sc = S.config1(steps=steps + warmup)
    n = len(sc["xm"])
    elapsed, prof, touched, tuned, st = mls_run(sc, steps, warmup)
    return {"workload": "MLS elastic cube 32^3 = 32,768 particles on 65^3 nodes, ground plane", "particles": n,
            "substeps": steps, "value": n * steps / elapsed, "unit": "particle-substeps/s",
            "ms_per_substep": elapsed / steps * 1e3, "roofline": mls_kernel_roofline(n, touched, prof, slicer=False, target="mls_cube",
                                            tuned=tuned),
            "autotuned": dict(zip(TUNED_KEYS, tuned)),
            "cpu_baseline": cpu_baseline(sc, 50) if cpu else None,
            "finite": bool(np.isfinite(st["xm"]).all())}


def nodes_in_indenter(solver, grid):
    """Grid nodes (after a substep's update) holding mass whose position lies
    inside the scene's sphere indenter (shape 0)."""
    sc = solver.scene
    nx, ny, nz = solver.dims
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pos = np.stack([i.ravel(), j.ravel(), k.ravel()], 1) * sc["dx"] + np.asarray(sc["lo"], np.float64)
    sh = sc["shape_list"][0]
    r = sc["geoms"][sh["geom"]][1][0]
    inside = np.linalg.norm(pos - np.asarray(sh["pos"][:3]), axis=1) < r
    return int((inside & (grid[:, 3] > 0)).sum())


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-substeps", type=int, default=4,
                    help="substeps in the CPU baseline sample, about 3 s each on one core (0 = skip)")
    ap.add_argument("--workload", choices=["all", "pb_slab", "batched", "adjoint", "pb_adjoint", "mls", "mls_cutting", "mls_cube",
                                           "lattice", "meshsdf", "engine"], default="all",
                    help="all: the PB slab headline line plus the batched, adjoint and MLS workloads")
    ap.add_argument("--launch", metavar="T,U,S,M,GS,GM[,W]",
                    help="use this launch configuration (tile waves, update waves, sub-tile size, sub-tiles, gather "
                         "sub-tile size, gather sub-tiles, scatter wavefronts per tile (default 1)) instead of the "
                         "autotuner's (PB slab and MLS workloads): "
                         "profiled runs pass the one the timed run picked, so no autotuner trial enters their stats")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 256 MiB copy kernel that calibrates rocprofv3's byte counters")
    ap.add_argument("--summarize", metavar="DIR",
                    help="write the profile summary from rocprofv3 outputs under DIR/{trace,fetch,write,fetch_copy,write_copy,sq}")
    ap.add_argument("--profile-target", choices=sorted(PROFILE_TARGETS), default="pb_slab",
                    help="which profiled workload --summarize describes")
    ap.add_argument("--summary-out", metavar="DIR", default=os.path.join(ROOT, "profiles"),
                    help="where --summarize writes its .md and .json (default: profiles/)")
    args = ap.parse_args()
    global LAUNCH
    if args.launch:
        LAUNCH = tuple(int(x) for x in args.launch.split(","))
        assert len(LAUNCH) in (6, 7), "--launch takes 6 or 7 integers"
    if args.summarize:
        tgt = PROFILE_TARGETS[args.profile_target]
        md, traffic = summarize(args.summarize, tgt["title"], profile_commands(args.profile_target), tgt["counters"])
        os.makedirs(args.summary_out, exist_ok=True)
        with open(os.path.join(args.summary_out, tgt["md"]), "w") as f:
            f.write(md)
        if traffic is not None:
            # per-kernel counter figures, which bench.py's roofline lines quote
            with open(os.path.join(args.summary_out, counters_file(args.profile_target)), "w") as f:
                json.dump(traffic, f, indent=1, sort_keys=True)
        print(md)
        return
    if args.copy_calibration:
        copy_calibration()
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} (WORLD_SIZE {world}) is not supported: "
              "this solver has no multi-GPU path.", file=sys.stderr)
        sys.exit(2)

    import torch

    import crmpm_hip as H

    cpu = args.cpu_substeps > 0
    if args.workload in ("mls", "mls_cutting", "mls_cube"):
        out = {}
        if args.workload != "mls_cube":
            out["mls_cutting"] = mls_cutting(cpu=cpu)
        if args.workload != "mls_cutting":
            out["mls_cube"] = mls_cube(cpu=cpu)
        print(json.dumps(out), flush=True)
        return
    if args.workload == "lattice":
        print(json.dumps({"lattice": lattice()}), flush=True)
        return
    if args.workload == "engine":
        print(json.dumps({"engine": engine()}), flush=True)
        return
    if args.workload == "meshsdf":
        print(json.dumps({"mesh_sdf": mesh_sdf()}), flush=True)
        return
    if args.workload == "adjoint":
        print(json.dumps({"adjoint": adjoint(cpu=cpu)}), flush=True)
        return
    if args.workload == "pb_adjoint":
        print(json.dumps({"pb_adjoint": pb_adjoint()}), flush=True)
        return
    if args.workload == "batched":
        print(json.dumps({"batched_mls": batched("mls", 100, 5), "batched_pb": batched("pb", 20, 2)}), flush=True)
        return
    K, W = args.steps, args.warmup
    scene = S.config2(steps=K + W)
    n = len(scene["xm"])
    solver = H.HipSolver(copy.deepcopy(scene))
    solver.compute_initial_data()
    frames = solver.scene["frames"]
    fsub = frames[0]["substeps"]

    def run(first, count, sample=False):
        # substeps [first, first + count) of the scene, kinematic target per frame
        for s in range(first, first + count):
            if s % fsub == 0 or s == first:
                S.apply_frame(solver.scene, frames[s // fsub])
                solver._upload_shapes()
            if sample:
                solver.profile_mask(H.PROF_CATEGORIES if (s - first) % PROFILE_EVERY == 0 else [])
            solver.step(1)

    run(0, W)
    # the library's autotuner picks the tile split and launch sizes on this state (which it leaves unchanged)
    tuned = (solver.set_launch(LAUNCH) if LAUNCH else
             solver.autotune(steps_per_trial=2) + solver.gather_split() + (solver.scatter_waves(),))
    torch.cuda.synchronize()
    # touched nodes of one substep (grid nodes holding mass), for the grid term of the roofline
    grid = solver.grid()
    touched = int(np.count_nonzero(grid[:, 3] > 0))
    contact_start = nodes_in_indenter(solver, grid)

    # Every PROFILE_EVERY-th substep records HIP events around each of its
    # launches (a recorded launch costs a few microseconds of dispatch, so
    # recording every launch would slow the run it measures; recording whole
    # substeps keeps each kernel's start event behind its predecessor's end).
    solver.profile_begin(max_launches=(K // PROFILE_EVERY + 1) * (2 * scene["iterations"] + 16) + 64)
    solver.profile_mask([])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(W, K, sample=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = solver.profile_end()
    elapsed = t1 - t0
    value = n * K / elapsed

    # one iteration after the first = the per-particle kernel + the per-pblock scatter
    dense_ms, dense_n = prof["pb_fused"]
    scat_ms, scat_n = prof["pb_scatter"]
    avg_dense_s = dense_ms / max(dense_n, 1) / 1e3
    avg_scat_s = scat_ms / max(scat_n, 1) / 1e3
    avg_iter_s = avg_dense_s + avg_scat_s
    bytes_iter = BYTES_PER_PARTICLE_ITERATION * n + BYTES_PER_TOUCHED_NODE_ITERATION * touched
    achieved = bytes_iter / avg_iter_s / 1e9
    dense_row = "k_pb_dense<1>"
    # the names as the profiler lists them (the scatter runs one-wavefront workgroups)
    scat_row = f"k_pb_scatter<false, {'true' if tuned[0] > 0 else 'false'}, 1>"
    rows = (profiled_row("pb_slab", dense_row), profiled_row("pb_slab", scat_row))
    per_kernel = {}
    for name, row, avg_s, (bp, bn), launches in ((dense_row, rows[0], avg_dense_s, BYTES_DENSE, dense_n),
                                                 (scat_row, rows[1], avg_scat_s, BYTES_SCATTER, scat_n)):
        alg = bp * n + bn * touched
        per_kernel[name] = {"avg_launch_us": avg_s * 1e6, "launches": launches, "algorithmic_bytes_per_launch": alg,
                            "achieved": alg / avg_s / 1e9, "frac": alg / avg_s / 1e9 / HBM_PEAK_GBS,
                            **roofline_counters(row, avg_s)}
    counters = iteration_counters(rows, avg_iter_s)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "pb_mpm_neohookean_slab_sphere_indenter", "particles": n,
                   "grid_nodes": list(solver.dims), "dx": scene["dx"], "iterations": scene["iterations"],
                   "dt": scene["dt"], "frame_substeps": fsub, "parallelism": "single_gpu",
                   "autotuned": dict(zip(TUNED_KEYS, tuned))},
        "roofline": {"bound": "hbm", "kernel": f"{dense_row} + {scat_row} (one PB iteration: G2P and the "
                                                 "constraint per particle, P2G per pblock)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "avg_launch_us": avg_iter_s * 1e6, "launches": min(dense_n, scat_n),
                     "algorithmic_bytes_per_launch": bytes_iter, "touched_nodes": touched,
                     # HBM-side traffic, L2 hit rate, DRAM-path bytes, atomics and VALU issue of the
                     # same kernels from the committed rocprofv3 counter passes of this workload
                     # (profiles/pb_slab262k_counters.json), summed over the iteration's two kernels
                     **counters, "kernels": per_kernel},
    }
    st = solver.state()
    out["state_check"] = {"nonfinite_particles": int((~np.isfinite(st["xm"])).any(axis=1).sum()),
                          "max_speed": float(np.abs(st["v"]).max()),
                          "y_range": [float(st["xm"][:, 1].min()), float(st["xm"][:, 1].max())],
                          # grid nodes holding mass inside the sphere indenter at the first and the
                          # last substep of the timed window: the contact the window exercises
                          "indenter_contact_nodes": [contact_start, nodes_in_indenter(solver, solver.grid())],
                          "indenter_bottom_y": float(solver.scene["shape_list"][0]["pos"][1]
                                                     - solver.scene["geoms"][0][1][0])}
    solver.close()
    if args.cpu_substeps > 0:
        out["cpu_baseline"] = cpu_baseline(scene, args.cpu_substeps, label=", 20 iterations")
    else:
        out["cpu_baseline"] = None
    if args.workload == "all":
        # the batched form beside the plain solver (64 scenes, MLS 100 substeps, PB 20 substeps)
        for key, kind, steps, warm in (("batched_mls", "mls", 100, 5), ("batched_pb", "pb", 20, 2)):
            try:
                out[key] = batched(kind, steps, warm)
            except Exception as e:  # reported, not fatal: the headline line still prints
                out[key] = {"error": repr(e)}
        for key, fn in (("adjoint", adjoint), ("pb_adjoint", pb_adjoint), ("mls_cutting", mls_cutting), ("mls_cube", mls_cube),
                        ("lattice", lattice), ("mesh_sdf", mesh_sdf), ("engine", engine)):
            try:
                out[key] = fn()
            except Exception as e:
                out[key] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
