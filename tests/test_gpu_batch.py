"""Batched scenes (crHipBatch*): one launch per phase over all scenes.

A scene in a batch runs the same per-tile arithmetic in the same order as
when it runs alone, so its trajectory must match the single-scene solver's;
the tolerance (relative 1e-6 after 10 substeps) only absorbs the float
atomics of computeInitialData's mass scatter, whose order is not fixed.
Graph replay and autotuned launch sizes must not change results at all.
The batched workload's generator at test size (8 scenes, particle counts
scaled down 50x, the same 32^3-128^3 grids, box and capsule obstacles in
contact) runs autotuned and graph-replayed and is checked scene by scene
against the oracle: relative error <= 1e-4 on x, v and F.
"""
import copy

import numpy as np
import pytest

import crmpm_scenes as S

pytestmark = pytest.mark.gpu

KINDS = {"mls": S.MLS, "pb": S.PB, "std": S.STD}
FIELDS = ("xm", "v", "F", "C", "V0")


def rel(a, b):
    return float(np.abs(a - b).max()) / max(float(np.abs(b).max()), 1e-30)


def mixed_scenes(kind):
    """Five scenes of different sizes, grids and shapes."""
    specs = [(6, 1 / 32, "plane"), (10, 1 / 24, "plane+sphere+box"), (4, 1 / 40, "plane+capsule"),
             (8, 1 / 32, "plane+slicer"), (12, 1 / 28, "plane")]
    out = []
    for i, (n_side, dx, shapes) in enumerate(specs):
        sc = S.small_scene(kind, 100 + i, n_side=n_side, dx=dx, shapes=shapes)
        out.append(sc)
    return out


def single_runs(scenes, steps):
    import crmpm_hip as H

    res = []
    for sc in scenes:
        g = H.HipSolver(copy.deepcopy(sc))
        g.compute_initial_data()
        g.step(steps)
        res.append(g.state())
        g.close()
    return res


@pytest.mark.parametrize("kind", list(KINDS))
def test_batch_matches_single_scenes(kind):
    import crmpm_hip as H

    scenes = mixed_scenes(KINDS[kind])
    print(f"batch {kind}: {len(scenes)} scenes, {sum(len(s['xm']) for s in scenes)} particles", flush=True)
    b = H.HipBatch(copy.deepcopy(scenes))
    b.compute_initial_data()
    b.step(10)
    alone = single_runs(scenes, 10)
    for i in range(len(scenes)):
        st = b.state(i)
        errs = {k: rel(st[k], alone[i][k]) for k in FIELDS}
        print(f"  scene {i}: " + " ".join(f"{k}={v:.1e}" for k, v in errs.items()), flush=True)
        for k in FIELDS:
            assert errs[k] <= 1e-6, (i, k, errs[k])
    b.close()


@pytest.mark.parametrize("kind", ["mls", "pb"])
def test_graph_replay_matches_eager(kind):
    """Each replay of the captured substep equals the eager substep, with eager
    work and synchronizations between replays."""
    import torch

    import crmpm_hip as H

    scenes = mixed_scenes(KINDS[kind])
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        eager = H.HipBatch(copy.deepcopy(scenes), stream=stream)
        graph = H.HipBatch(copy.deepcopy(scenes), stream=stream)
        eager.compute_initial_data()
        graph.compute_initial_data()
        graph.set_graph_mode(True)
        scratch = torch.zeros(1 << 16, device="cuda")
        for r in range(4):
            print(f"graph {kind} replay {r}", flush=True)
            eager.step(1)
            graph.step(1)
            scratch.add_(1.0)  # unrelated eager work between replays
            torch.cuda.synchronize()
            for i in range(len(scenes)):
                a, b = eager.state(i), graph.state(i)
                for k in FIELDS:
                    assert rel(b[k], a[k]) <= 1e-6, (r, i, k)
        eager.close()
        graph.close()


def test_autotune_keeps_state_and_results():
    import crmpm_hip as H

    scenes = mixed_scenes(S.PB)
    b = H.HipBatch(copy.deepcopy(scenes))
    b.compute_initial_data()
    b.step(2)
    before = [b.state(i) for i in range(len(scenes))]
    t, u, ss, sm = b.autotune(steps_per_trial=2)
    print(f"autotune picked tile waves {t}, update waves {u}, sub-tiles {sm} x {ss}", flush=True)
    after = [b.state(i) for i in range(len(scenes))]
    for i in range(len(scenes)):
        for k in FIELDS:
            assert np.array_equal(before[i][k], after[i][k]), (i, k)
    # every launch size gives the same arithmetic (the tile split and the
    # wavefronts per scatter tile fix the summation order): persistent
    # launches against one wavefront per sub-tile
    ref = H.HipBatch(copy.deepcopy(scenes))
    ref.compute_initial_data()
    ref.step(2)
    ref.set_tile_split(ss, sm)
    # wavefronts per scatter tile fix a summation order as the split does
    ref.set_scatter_waves(b.scatter_waves())
    b.set_launch_sizes(2048, 2048)
    b.step(3)
    ref.step(3)
    for i in range(len(scenes)):
        x, y = b.state(i), ref.state(i)
        for k in FIELDS:
            assert rel(x[k], y[k]) <= 1e-6, (i, k)
    b.close()
    ref.close()


@pytest.mark.parametrize("kind", ["mls", "std"])
def test_gather_split_is_bitwise_neutral(kind):
    """The G2P kernels' split only regroups particles into wavefronts: each
    particle's gather is the same arithmetic, so results are bitwise equal."""
    import crmpm_hip as H

    scenes = mixed_scenes(KINDS[kind])
    runs = []
    for split in ((1 << 20, 1), (64, 8), (128, 2)):
        b = H.HipBatch(copy.deepcopy(scenes))
        b.set_gather_split(*split)
        assert b.gather_split() == split
        b.compute_initial_data()
        b.step(6)
        runs.append([b.state(i) for i in range(len(scenes))])
        b.close()
    for r in runs[1:]:
        for i in range(len(scenes)):
            for k in FIELDS:
                assert np.array_equal(r[i][k], runs[0][i][k]), (i, k)


@pytest.mark.parametrize("kind", ["mls", "std"])
@pytest.mark.parametrize("shapes", ["plane+sphere+box", "plane+slicer", "plane+capsule+meshslicer"])
def test_per_particle_gather_is_bitwise_neutral(kind, shapes):
    """G2P per particle in particle index order (gather split 0, one scene)
    reads the same nodes in the same order as the tile kernels and takes the
    same particles (masked ones left alone): results bitwise equal, also
    replayed as a graph.  A batch of several scenes refuses that split."""
    import crmpm_hip as H

    sc = S.small_scene(KINDS[kind], 7, n_side=10, dx=1 / 32, shapes=shapes)
    sc["mask"][::7] = 0
    print(f"per-particle gather {kind} {shapes}: {len(sc['xm'])} particles", flush=True)
    runs = []
    for cfg, graph in (((0, 0, 1 << 20, 1, 1 << 20, 1), False), ((0, 0, 1 << 20, 1, 1 << 20, 0), False),
                       ((2048, 0, 1 << 20, 1, 1 << 20, 0), True)):
        g = H.HipSolver(copy.deepcopy(sc))
        g.set_launch(cfg)
        assert g.gather_split() == cfg[4:]
        g.set_graph_mode(graph)
        g.compute_initial_data()
        g.step(12)
        runs.append(g.state())
        g.close()
    moved = np.abs(runs[0]["xm"] - sc["xm"]).max()
    assert moved > 0.0
    for r in runs[1:]:
        for k in FIELDS:
            assert np.array_equal(r[k], runs[0][k]), k
    b = H.HipBatch(mixed_scenes(KINDS[kind])[:2])
    with pytest.raises(RuntimeError):
        b.set_gather_split(1 << 20, 0)
    b.close()


def test_single_solver_graph_with_kinematic_frames():
    """The single-scene solver replaying its substep as a hipGraph, with the
    indenter's target uploaded between frames and every third substep run
    eagerly with profiling events, matches eager stepping bitwise after every
    frame."""
    import crmpm_hip as H

    sc = S.config2(nx=24, ny=6, nz=24, steps=40)
    runs = {}
    for graph in (False, True):
        g = H.HipSolver(copy.deepcopy(sc))
        g.compute_initial_data()
        if graph:
            g.set_graph_mode(True)
            g.profile_begin(max_launches=4096)
        frames = g.scene["frames"]
        states = []
        for s in range(40):
            if s % frames[0]["substeps"] == 0:
                S.apply_frame(g.scene, frames[s // frames[0]["substeps"]])
                g._upload_shapes()
            if graph:
                g.profile_mask(H.PROF_CATEGORIES if s % 3 == 0 else [])
            g.step(1)
            if s % 10 == 9:
                states.append(g.state())
        if graph:
            g.profile_end()
        runs[graph] = states
        g.close()
    for a, b in zip(runs[False], runs[True]):
        for k in FIELDS:
            assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("kind,steps", [("mls", 100), ("pb", 20)])
def test_batched_workload_matches_oracle(kind, steps):
    """The 64-scene generator (crmpm_scenes.config3) at test size: its first 8
    scenes with 40-4,000 particles (log-uniform), grids of 32^3-128^3 cells,
    each body starting half a cell above its box or capsule obstacle so that
    contact begins within the run; autotuned, graph-replayed batch against
    the oracle run scene by scene (the reference's advanceAll steps each
    scene's solver in turn, src/engine/simulation_factory_impl.cpp:1106-1150).
    PB bodies come to rest on their obstacle within 20 substeps (max |v| from
    0.43 to 0.004 m/s), where a one-ulp change of the oracle's input positions
    alone moves the velocities by 2.7e-3 of their max: velocity errors are
    measured against the larger of the field's max and the bodies' initial
    speed, 0.5 m/s."""
    import crmpm_hip as H
    import cpu_checkers as CC
    from parity_bounds import TOL_100

    scenes = S.config3(n_scenes=8, kind=KINDS[kind], steps=steps, n_min=40, n_max=4000, gap=0.5)
    print(f"batched workload {kind}: {len(scenes)} scenes, {sum(len(s['xm']) for s in scenes)} particles, "
          f"{steps} substeps", flush=True)
    b = H.HipBatch(copy.deepcopy(scenes))
    b.compute_initial_data()
    b.step(1)
    b.autotune(steps_per_trial=2)
    b.set_graph_mode(True)
    b.step(steps - 1)
    for i, sc in enumerate(scenes):
        o = CC.CpuSolver(copy.deepcopy(sc), "oracle")
        o.compute_initial_data()
        o.step(steps)
        os_ = o.state()
        o.close()
        d = CC.CpuSolver(copy.deepcopy(sc), "oracle", fp64=True)
        d.compute_initial_data()
        d.step(steps)
        od = d.state()
        d.close()
        st = b.state(i)
        errs = {k: rel(st[k], os_[k]) for k in ("xm", "v", "F")}
        errs["v"] = float(np.abs(st["v"] - os_["v"]).max()) / max(float(np.abs(os_["v"]).max()), 0.5)
        print(f"  scene {i} ({len(sc['xm'])} particles, {round(1 / sc['dx'])}^3 cells, "
              f"{'box' if sc['geoms'][0][0] == S.BOX else 'capsule'}): gpu-vs-oracle "
              + " ".join(f"{k}={v:.2e}" for k, v in errs.items())
              + " | fp64 replay: gpu " + " ".join(f"{k}={rel(st[k], od[k]):.2e}" for k in ("xm", "v", "F"))
              + " cpu " + " ".join(f"{k}={rel(os_[k], od[k]):.2e}" for k in ("xm", "v", "F")), flush=True)
        for k in ("xm", "v", "F"):
            assert errs[k] <= TOL_100, (i, k, errs[k])
    b.close()
