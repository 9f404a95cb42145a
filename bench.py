#!/usr/bin/env python3
"""Substep throughput of the MI355X PB-MPM solver on the 262,144-particle slab.

Workload: a 1.0 x 0.2 x 1.0 m Neo-Hookean slab of 128 x 16 x 128 = 262,144
particles on a 128 x 64 x 128-cell grid (129 x 65 x 129 nodes, dx = 1/128),
PB-MPM with 20 solver iterations per substep, dt = 2e-3, a kinematic sphere
indenter (r = 0.1 m) moving down at 0.05 m/s with friction 0.3, its target
updated every 10 substeps.  Inputs are synthetic (SplitMix64-seeded) and
already resident in HBM when timing starts.  A step is one substep.

Prints one JSON line: particle-substeps/s, the roofline of the dominant kernel
(the fused G2P+P2G iteration kernel) measured with HIP events on the solver
stream over the timed region (the events ride on the dispatches of every 10th
substep), and a CPU baseline timed on this host.
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
# touched node.
BYTES_PER_PARTICLE_FUSED = 164
BYTES_PER_TOUCHED_NODE_FUSED = 32
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pb_fused_slab262k_traffic.json")
# the fused iteration kernel (template Mode 1; either launch form)
FUSED_KERNEL = "k_pb_tile<1,"
# what the library's autotuner picked: launch sizes, the P2G sub-tile split
# and the G2P (gather) split
TUNED_KEYS = ("tile_waves", "update_waves", "sub_size", "sub_max", "gather_sub_size", "gather_sub_max")
PROFILE_EVERY = 10
COPY_CAL_BYTES = 256 << 20


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
# rocprofv3 wraps, the summary's title and file name, and whether the HBM
# counter passes (FETCH_SIZE, WRITE_SIZE with their copy calibration, SQ
# instruction counts) belong to it or the kernel trace alone.
PROFILE_TARGETS = {
    "pb_slab": dict(run="python3 bench.py --workload pb_slab --cpu-substeps 0", counters=True,
                    title="Kernel trace and HBM traffic: PB-MPM slab, 262,144 particles, 20 iterations, 500 substeps",
                    md="pb_slab262k_kernel_trace.md"),
    "mls": dict(run="python3 bench.py --workload mls", counters=True,
                title="Kernel trace and HBM traffic: MLS-MPM cutting block, 2,097,152 particles, 500 substeps; "
                      "then the 32,768-particle cube, 200 substeps",
                md="mls_cutting2m_kernel_trace.md"),
    "batched": dict(run="python3 bench.py --workload batched", counters=False,
                    title="Kernel trace: 64 batched scenes (MLS 100 substeps, PB 20 substeps) and the same scenes "
                          "one at a time",
                    md="batched64_kernel_trace.md"),
    "meshsdf": dict(run="python3 bench.py --workload meshsdf", counters=True,
                    title="Kernel trace and counters: mesh -> SDF table, 20,480-triangle icosphere on 128^3 nodes",
                    md="meshsdf_icosphere128_kernel_trace.md"),
    "engine": dict(run="python3 bench.py --workload engine", counters=False,
                   title="Kernel trace: engine advanceAll, 64 MLS scenes in frames of 10 substeps with host readback, "
                         "then the same batch stepped alone",
                   md="engine64_kernel_trace.md"),
    "adjoint": dict(run="python3 bench.py --workload adjoint", counters=False,
                    title="Kernel trace: differentiable MLS block, 65,536 particles, 100 substeps, forward + adjoint",
                    md="adjoint65k_kernel_trace.md"),
}


def profile_commands(target):
    t = PROFILE_TARGETS[target]
    run = t["run"]
    cmds = [f"rocprofv3 --kernel-trace --stats --output-format csv -d DIR/trace -o run -- {run}"]
    if t["counters"]:
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            d = c.split("_")[0].lower()
            cmds.append(f"rocprofv3 --pmc {c} --output-format csv -d DIR/{d} -o run -- {run}")
            cmds.append(f"rocprofv3 --pmc {c} --output-format csv -d DIR/{d}_copy -o cp -- "
                        "python3 bench.py --copy-calibration")
        cmds.append("rocprofv3 --pmc SQ_INSTS_VALU SQ_INSTS_LDS SQ_INSTS_SALU SQ_WAVES SQ_BUSY_CYCLES SQ_WAVE_CYCLES "
                    f"SQ_ACTIVE_INST_VALU SQ_INSTS_VMEM_RD --output-format csv -d DIR/sq -o run -- {run}")
    cmds.append(f"python3 bench.py --summarize DIR --profile-target {target}")
    return cmds


def summarize(out_dir, title, commands, counters=True):
    """Turn the rocprofv3 outputs of one profiling session into the committed
    summary: kernel-trace statistics and, with counters, FETCH_SIZE /
    WRITE_SIZE per dispatch of each kernel, the copy-kernel calibration that
    scales them to bytes, and the VALU instruction count."""
    stats = _read_csv(os.path.join(out_dir, "trace"), "kernel_stats.csv")
    fetch = write = valu = {}
    f_fac = w_fac = None
    if counters:
        fetch, _ = _counter_per_launch(os.path.join(out_dir, "fetch"))
        write, _ = _counter_per_launch(os.path.join(out_dir, "write"))
        _, fetch_cp = _counter_per_launch(os.path.join(out_dir, "fetch_copy"))
        _, write_cp = _counter_per_launch(os.path.join(out_dir, "write_copy"))
        valu = _counter_per_launch(os.path.join(out_dir, "sq"), "SQ_INSTS_VALU")[0]
        kib = COPY_CAL_BYTES / 1024.0
        f_fac = kib / max(fetch_cp["__amd_rocclr_copyBuffer"])
        w_fac = kib / max(write_cp["__amd_rocclr_copyBuffer"])
    lines = [f"# {title}", "", "Commands (repository root, rocprofv3 on one MI355X):", "", "```"]
    lines += commands + ["```", ""]
    if counters:
        lines += ["| kernel | calls | avg µs | total % | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch "
                  "| HBM MB/dispatch (calibrated) | SQ_INSTS_VALU/dispatch | VALU issue µs |",
                  "|---|---|---|---|---|---|---|---|---|"]
    else:
        lines += ["| kernel | calls | avg µs | total % |", "|---|---|---|---|"]
    merged = {}
    for r in stats:
        k = _short_name(r["Name"])
        c, t = merged.get(k, (0, 0.0))
        merged[k] = (c + int(r["Calls"]), t + float(r["TotalDurationNs"]))
    total = sum(t for _, t in merged.values())
    for k, (c, t) in sorted(merged.items(), key=lambda kv: -kv[1][1]):
        fr, wr = fetch.get(k), write.get(k)
        mb = "" if fr is None or wr is None else f"{(fr * f_fac + wr * w_fac) * 1024 / 1e6:.2f}"
        vi = valu.get(k)
        row = f"| `{k}` | {c} | {t / c / 1e3:.1f} | {100 * t / total:.2f} |"
        if counters:
            row += (f" {'' if fr is None else f'{fr:,.0f}'} | {'' if wr is None else f'{wr:,.0f}'} | {mb} | "
                    f"{'' if vi is None else f'{vi:,.0f}'} | {'' if vi is None else f'{valu_issue_us(vi):.1f}'} |")
        lines.append(row)
    lines += ["", "Kernel names carry their template arguments: the tile kernels' last one selects the persistent "
              "launch form, which appears in the autotuner's trial substeps."]
    if not counters:
        return "\n".join(lines) + "\n", None
    lines += ["", f"Calibration: a device-to-device copy of {COPY_CAL_BYTES >> 20} MiB reads FETCH_SIZE "
              f"{max(fetch_cp['__amd_rocclr_copyBuffer']):,.0f} KiB and WRITE_SIZE {max(write_cp['__amd_rocclr_copyBuffer']):,.0f} KiB "
              f"per dispatch, so bytes = FETCH_SIZE x {f_fac:.3f} KiB + WRITE_SIZE x {w_fac:.3f} KiB.",
              "", "VALU issue µs = SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x 2.4 GHz): the time the VALU pipes need "
              "for the kernel's vector instructions; a kernel whose duration is close to it is bound by VALU issue, not HBM."]
    fused = max((k for k in merged if k.startswith(FUSED_KERNEL)), key=lambda k: merged[k][0], default=None)
    result = {"kernel": fused, "fetch_kib_raw": fetch.get(fused), "write_kib_raw": write.get(fused),
              "fetch_factor": f_fac, "write_factor": w_fac, "valu_instructions": valu.get(fused)}
    if result["fetch_kib_raw"] is not None and result["write_kib_raw"] is not None:
        result["bytes_per_launch"] = 1024.0 * (result["fetch_kib_raw"] * f_fac + result["write_kib_raw"] * w_fac)
    return "\n".join(lines) + "\n", result


def cpu_baseline(scene, substeps):
    """Reference CPU solver (compiled from its sources) or the oracle, one thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cpu_checkers as CC

    which = "ref" if CC.ref_available() else "oracle"
    sc = copy.deepcopy(scene)
    o = CC.CpuSolver(sc, which)
    o.compute_initial_data()
    S.apply_frame(sc, sc["frames"][0])
    o.upload_shapes()
    t0 = time.perf_counter()
    o.step(substeps)
    dt = time.perf_counter() - t0
    o.close()
    n = len(scene["xm"])
    return {
        "value": n * substeps / dt,
        "unit": "particle-substeps/s",
        "cores": 1,
        "kind": "reference" if which == "ref" else "port",
        "sample": f"{substeps} substep(s) of the full 262,144-particle scene, 20 iterations, "
                  f"{dt:.1f} s on 1 core of {platform.processor() or platform.machine()}",
    }


def batched(kind_name, steps, warmup, tune=True, graph=True):
    """64 scenes of uneven size (config3) advanced as one batch, against the sum
    of the same scenes run one at a time through the single-scene solver
    (same kernels).  Returns the batch throughput and the speedup."""
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
        tuned = b.autotune(steps_per_trial=2) + b.gather_split()
    b.set_graph_mode(graph)
    b.step(1)  # capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    finite = all(np.isfinite(b.state(i)["xm"]).all() for i in (0, len(scenes) - 1))
    b.close()
    t_plain = 0.0
    for sc in scenes:
        g = H.HipSolver(copy.deepcopy(sc))
        g.compute_initial_data()
        g.step(warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.step(steps)
        torch.cuda.synchronize()
        t_plain += time.perf_counter() - t0
        g.close()
    return {"workload": f"{len(scenes)} {kind_name.upper()} scenes, log-uniform 2k-200k particles, grids 32^3-128^3",
            "particles": n_total, "substeps": steps, "iterations": scenes[0]["iterations"] if kind == S.PB else None,
            "value": n_total * steps / t_batch, "unit": "particle-substeps/s",
            "batch_ms_per_substep": t_batch / steps * 1e3, "sum_of_plain_ms_per_substep": t_plain / steps * 1e3,
            "speedup_vs_sum_of_plain": t_plain / t_batch,
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


def adjoint(steps=100, every=10):
    """Differentiable MLS block (config4: 65,536 particles, 96^3 grid, GRF log E):
    time of a forward rollout with checkpoints, of forward + adjoint, and of the
    plain solver's forward, for the same substeps."""
    import torch

    import crmpm_hip as H

    sc, log_e = S.config4(steps=steps)
    hidden, log_e_true = S.config4(steps=steps, field_seed=20261203 + 44)
    sim = H.DifferentiableMls(sc, steps, checkpoint_every=every, tune=True)
    dev = sim.solver.device
    nu = 0.3
    v0 = torch.as_tensor(sc["v"][:, :3], dtype=torch.float64, device=dev)

    def lame_t(le):
        E = torch.exp(torch.as_tensor(le, dtype=torch.float64, device=dev))
        return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))

    lt, mt = lame_t(log_e_true)
    target = sim.forward(lt, mt, v0).double()
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
    plain.autotune(steps_per_trial=2)
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
            "forward_adjoint_vs_plain_forward": t_fa / t_p, "loss": loss,
            "autotuned": dict(zip(TUNED_KEYS, sim_tuned)),
            "grad_logE_abs_max": float(g.abs().max()), "finite": finite}


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
    tuned = solver.autotune(steps_per_trial=2) + solver.gather_split() if tune else None
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


def mls_kernel_roofline(n, touched, prof, slicer):
    """Achieved HBM rate of the MLS P2G and G2P kernels on their algorithmic
    bytes: P2G reads x, v, F, C, V0, (lambda, mu) = 108 B per particle and
    writes 16 B per touched node; G2P reads x, F (and x again per slicer test,
    12 B) and writes x, v, F, C = 96 B per particle and reads 16 B per node."""
    out = {}
    for name, per_p in (("mls_p2g", 108), ("mls_g2p", 48 + (12 if slicer else 0) + 96)):
        ms, k = prof[name]
        if not k:
            continue
        avg = ms / k / 1e3
        achieved = (per_p * n + 16 * touched) / avg / 1e9
        out[name] = {"avg_launch_us": avg * 1e6, "launches": k, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "algorithmic_bytes_per_launch": per_p * n + 16 * touched}
    return out


def mls_cutting(steps=500, warmup=10):
    """MLS at scale with cutting (config5: 2,097,152 particles, 257 x 129 x 257 nodes)."""
    sc = S.config5(steps=steps + warmup)
    n = len(sc["xm"])
    elapsed, prof, touched, tuned, st = mls_run(sc, steps, warmup)
    return {"workload": "MLS block 256x32x256 = 2,097,152 particles, 257x129x257 nodes, capsule tool + growing quad "
                        "slicer cutting at 0.02 m/s", "particles": n, "substeps": steps,
            "value": n * steps / elapsed, "unit": "particle-substeps/s", "ms_per_substep": elapsed / steps * 1e3,
            "touched_nodes": touched, "roofline": mls_kernel_roofline(n, touched, prof, slicer=True),
            "autotuned": dict(zip(TUNED_KEYS, tuned)),
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
    dims, _, out = H.mesh_sdf(pts, tris, cell, fatten)  # warm-up, loads the code objects
    torch.cuda.synchronize()
    best = 1e30
    for _ in range(reps):
        t0 = time.perf_counter()
        H.mesh_sdf(pts, tris, cell, fatten)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    nodes = int(np.prod(dims))
    pairs = nodes * len(tris)
    res = {"workload": "icosphere r = 0.1 m, 20,480 triangles -> SDF table of 128^3 nodes (float4 gradient|distance)",
           "dims": [int(d) for d in dims], "triangles": len(tris), "seconds": best, "nodes_per_s": nodes / best,
           "node_triangle_pairs_per_s": pairs / best, "finite": bool(torch.isfinite(out).all())}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cpu_checkers as CC

    which = "reference" if CC.ref_available() else "oracle"
    ccell = cell * (4 if which == "reference" else 16)
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


def mls_cube(steps=200, warmup=5):
    """The reference's CPU-runnable case (config1: 32^3 = 32,768 particles, 64^3 cells, 200 substeps)."""
    sc = S.config1(steps=steps + warmup)
    n = len(sc["xm"])
    elapsed, prof, touched, tuned, st = mls_run(sc, steps, warmup)
    return {"workload": "MLS elastic cube 32^3 = 32,768 particles on 65^3 nodes, ground plane", "particles": n,
            "substeps": steps, "value": n * steps / elapsed, "unit": "particle-substeps/s",
            "ms_per_substep": elapsed / steps * 1e3, "roofline": mls_kernel_roofline(n, touched, prof, slicer=False),
            "finite": bool(np.isfinite(st["xm"]).all())}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-substeps", type=int, default=4,
                    help="substeps in the CPU baseline sample, about 3 s each on one core (0 = skip)")
    ap.add_argument("--workload", choices=["all", "pb_slab", "batched", "adjoint", "mls", "lattice", "meshsdf", "engine"], default="all",
                    help="all: the PB slab headline line plus the batched, adjoint and MLS workloads")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 256 MiB copy kernel that calibrates rocprofv3's byte counters")
    ap.add_argument("--summarize", metavar="DIR",
                    help="write the profile summary from rocprofv3 outputs under DIR/{trace,fetch,write,fetch_copy,write_copy,sq}")
    ap.add_argument("--profile-target", choices=sorted(PROFILE_TARGETS), default="pb_slab",
                    help="which profiled workload --summarize describes")
    ap.add_argument("--summary-out", metavar="DIR", default=os.path.join(ROOT, "profiles"),
                    help="where --summarize writes its .md and .json (default: profiles/)")
    args = ap.parse_args()
    if args.summarize:
        tgt = PROFILE_TARGETS[args.profile_target]
        md, traffic = summarize(args.summarize, tgt["title"], profile_commands(args.profile_target), tgt["counters"])
        os.makedirs(args.summary_out, exist_ok=True)
        with open(os.path.join(args.summary_out, tgt["md"]), "w") as f:
            f.write(md)
        if args.profile_target == "pb_slab":
            with open(os.path.join(args.summary_out, os.path.basename(TRAFFIC_FILE)), "w") as f:
                json.dump(traffic, f, indent=1)
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

    if args.workload == "mls":
        print(json.dumps({"mls_cutting": mls_cutting(), "mls_cube": mls_cube()}), flush=True)
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
        print(json.dumps({"adjoint": adjoint()}), flush=True)
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
    tuned = solver.autotune(steps_per_trial=2) + solver.gather_split()
    torch.cuda.synchronize()
    # touched nodes of one substep (grid nodes holding mass), for the grid term of the roofline
    grid = solver.grid()
    touched = int(np.count_nonzero(grid[:, 3] > 0))

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

    fused_ms, fused_n = prof["pb_fused"]
    avg_fused_s = fused_ms / max(fused_n, 1) / 1e3
    bytes_fused = BYTES_PER_PARTICLE_FUSED * n + BYTES_PER_TOUCHED_NODE_FUSED * touched
    achieved = bytes_fused / avg_fused_s / 1e9
    traffic = None
    valu_insts = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            prof_summary = json.load(f)
        traffic = prof_summary.get("bytes_per_launch")
        valu_insts = prof_summary.get("valu_instructions")
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
        "roofline": {"bound": "hbm", "kernel": "k_pb_tile<1, ...> (fused G2P+P2G)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "avg_launch_us": avg_fused_s * 1e6, "launches": fused_n,
                     "algorithmic_bytes_per_launch": bytes_fused, "touched_nodes": touched,
                     # the fused kernel's VALU issue time (rocprofv3 SQ_INSTS_VALU of this workload) over its duration
                     "valu_issue_frac": None if valu_insts is None else valu_issue_us(valu_insts) / (avg_fused_s * 1e6)},
    }
    st = solver.state()
    out["state_check"] = {"nonfinite_particles": int((~np.isfinite(st["xm"])).any(axis=1).sum()),
                          "max_speed": float(np.abs(st["v"]).max()),
                          "y_range": [float(st["xm"][:, 1].min()), float(st["xm"][:, 1].max())]}
    solver.close()
    if args.cpu_substeps > 0:
        out["cpu_baseline"] = cpu_baseline(scene, args.cpu_substeps)
    else:
        out["cpu_baseline"] = None
    if args.workload == "all":
        # the batched form beside the plain solver (64 scenes, MLS 100 substeps, PB 20 substeps)
        for key, kind, steps, warm in (("batched_mls", "mls", 100, 5), ("batched_pb", "pb", 20, 2)):
            try:
                out[key] = batched(kind, steps, warm)
            except Exception as e:  # reported, not fatal: the headline line still prints
                out[key] = {"error": repr(e)}
        for key, fn in (("adjoint", adjoint), ("mls_cutting", mls_cutting), ("mls_cube", mls_cube),
                        ("lattice", lattice), ("mesh_sdf", mesh_sdf), ("engine", engine)):
            try:
                out[key] = fn()
            except Exception as e:
                out[key] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
