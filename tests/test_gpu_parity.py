"""GPU parity: the HIP solver (through the C-ABI) against the CPU oracle.

Tolerances (FP32, fixed, from tests/parity_bounds.py):
  * binning keys and sort permutation: bit-exact against the oracle's base cells
    and a stable sort;
  * positions, velocities, F: relative error (max abs difference over the
    field's max magnitude) <= 5e-5 after 10 substeps, <= 1e-4 after 100 or
    more;
  * the HIP path must be bitwise reproducible run to run.
The error of both FP32 paths against the oracle's FP64 replay is printed for
the record; it does not widen the bounds.
"""
import copy

import numpy as np
import pytest

import cpu_checkers as CC
import crmpm_scenes as S
from parity_bounds import TOL_10, TOL_100, rel

pytestmark = pytest.mark.gpu


def hip_run(sc, steps):
    import crmpm_hip as H

    g = H.HipSolver(copy.deepcopy(sc))
    g.compute_initial_data()
    g.run(steps)
    st = g.state()
    g.close()
    return st


def cpu_run(sc, steps, fp64=False):
    o = CC.CpuSolver(copy.deepcopy(sc), "oracle", fp64=fp64)
    o.compute_initial_data()
    o.run(steps)
    st = o.state()
    o.close()
    return st


def report(label, gs, os_, od=None):
    errs = {k: rel(gs[k], os_[k]) for k in ("xm", "v", "F", "C", "V0")}
    line = f"{label}: gpu-vs-oracle " + " ".join(f"{k}={v:.2e}" for k, v in errs.items())
    if od is not None:
        line += " | fp64 replay: gpu " + " ".join(f"{k}={rel(gs[k], od[k]):.2e}" for k in ("xm", "v", "F"))
        line += " cpu " + " ".join(f"{k}={rel(os_[k], od[k]):.2e}" for k in ("xm", "v", "F"))
    print(line, flush=True)
    return errs


KINDS = {"mls": S.MLS, "pb": S.PB, "std": S.STD}


@pytest.mark.parametrize("kind", list(KINDS))
def test_binning_bit_exact(kind):
    import crmpm_hip as H

    sc = S.small_scene(KINDS[kind], 3, n_side=12, shapes="plane", jitter=0.45)
    # a few particles outside the grid and a masked one exercise the clamping and the sentinel
    sc["xm"][0, 0] = -0.01
    sc["xm"][1, 1] = 1.2
    sc["mask"][5] = 0
    print(f"binning {kind} n={len(sc['xm'])}", flush=True)
    g = H.HipSolver(copy.deepcopy(sc))
    keys, order = g.binning()
    o = CC.CpuSolver(copy.deepcopy(sc), "oracle")
    base = o.base_cells().astype(np.int64)
    dims = np.array(g.dims)
    nb = (dims + 3) // 4
    b = np.clip(base, 0, dims - 1) >> 2
    ref_keys = b[:, 0] + nb[0] * (b[:, 1] + nb[1] * b[:, 2])
    if KINDS[kind] != S.STD:
        ref_keys[sc["mask"] == 0] = np.prod(nb)
    ref_order = np.argsort(ref_keys, kind="stable")
    assert np.array_equal(order.astype(np.int64), ref_order)
    assert np.array_equal(keys.astype(np.int64), ref_keys[ref_order])
    g.close()


@pytest.mark.parametrize("kind", list(KINDS))
def test_initial_volume(kind):
    sc = S.small_scene(KINDS[kind], 11, n_side=10, shapes="plane")
    print(f"initial volume {kind}", flush=True)
    gs = hip_run(sc, 0)
    os_ = cpu_run(sc, 0)
    assert rel(gs["V0"], os_["V0"]) < 1e-6
    assert np.array_equal(gs["F"], os_["F"])


SHAPES = ["plane", "plane+sphere+box", "plane+capsule", "plane+slicer", "plane+mesh", "plane+meshslicer"]


@pytest.mark.parametrize("shapes", SHAPES)
@pytest.mark.parametrize("kind", list(KINDS))
def test_parity_10_substeps(kind, shapes):
    sc = S.small_scene(KINDS[kind], 7 + KINDS[kind], shapes=shapes)
    gs = hip_run(sc, 10)
    os_ = cpu_run(sc, 10)
    od = cpu_run(sc, 10, fp64=True)
    errs = report(f"{kind} {shapes} 10 steps", gs, os_, od)
    for k in ("xm", "v", "F"):
        assert errs[k] <= TOL_10, (k, errs[k])


def test_parity_mls_cube_200_substeps():
    """The elastic cube at full size: 32^3 = 32,768 particles on 65^3 nodes
    falling onto the ground plane, 200 substeps."""
    sc = S.config1()
    print("mls cube 32^3, 200 substeps", flush=True)
    gs = hip_run(sc, 200)
    os_ = cpu_run(sc, 200)
    od = cpu_run(sc, 200, fp64=True)
    errs = report("mls cube 32^3 200 steps", gs, os_, od)
    for k in ("xm", "v", "F"):
        assert errs[k] <= TOL_100, (k, errs[k])


def test_parity_pb_slab_100_substeps():
    """PB-MPM slab with the sphere indenter pressing into it from the first
    substep (tests/test_oracle.py checks the contact), 32x8x32 particles,
    20 iterations, 100 substeps."""
    sc = S.config2(nx=32, ny=8, nz=32, steps=100, iterations=20)
    print("pb slab 32x8x32, 100 substeps", flush=True)
    gs = hip_run(sc, 100)
    os_ = cpu_run(sc, 100)
    od = cpu_run(sc, 100, fp64=True)
    errs = report("pb slab 32x8x32 100 steps", gs, os_, od)
    for k in ("xm", "v", "F"):
        assert errs[k] <= TOL_100, (k, errs[k])


def test_parity_mls_cutting_30_substeps():
    """The cutting scene at reduced size: a 24 x 8 x 24 block, the capsule tool
    and the growing quad slicer sweeping at 1 m/s (so the cut crosses many
    particles within 30 substeps), on the full 257 x 129 x 257-node grid."""
    sc = S.config5(nx=24, ny=8, nz=24, steps=30, speed=1.0)
    print("mls cutting 24x8x24, 30 substeps", flush=True)
    gs = hip_run(sc, 30)
    os_ = cpu_run(sc, 30)
    od = cpu_run(sc, 30, fp64=True)
    errs = report("mls cutting 24x8x24 30 steps", gs, os_, od)
    for k in ("xm", "v", "F"):
        assert errs[k] <= TOL_10, (k, errs[k])


@pytest.mark.parametrize("kind", list(KINDS))
def test_bitwise_reproducible(kind):
    sc = S.small_scene(KINDS[kind], 5, shapes="plane+sphere+box")
    a = hip_run(sc, 5)
    b = hip_run(sc, 5)
    for k in ("xm", "v", "F", "C"):
        assert np.array_equal(a[k], b[k]), k


def test_grid_after_one_step():
    import crmpm_hip as H

    sc = S.small_scene(S.MLS, 21, shapes="plane+sphere")
    g = H.HipSolver(copy.deepcopy(sc))
    g.compute_initial_data()
    g.step(1)
    gg = g.grid()
    o = CC.CpuSolver(copy.deepcopy(sc), "oracle")
    o.compute_initial_data()
    o.step(1)
    og = o.grid()
    assert rel(gg[:, 3], og[:, 3]) < 1e-6
    assert rel(gg[:, :3], og[:, :3]) < 1e-5
    g.close()


def test_masked_and_partial_active():
    import crmpm_hip as H

    sc = S.small_scene(S.MLS, 31, shapes="plane")
    sc["mask"][::7] = 0
    n = len(sc["xm"])
    g = H.HipSolver(copy.deepcopy(sc))
    g.compute_initial_data()
    H.check(g.lib.crHipSolverSetNumActiveParticles(g.h, n - 40), "num active")
    g.step(5)
    gs = g.state()
    o = CC.CpuSolver(copy.deepcopy(sc), "oracle")
    o.compute_initial_data()
    o.fn["set_num_active"](o.h, n - 40)
    o.step(5)
    os_ = o.state()
    assert rel(gs["xm"], os_["xm"]) < TOL_10
    # inactive particles are untouched
    inactive = np.zeros(n, bool)
    inactive[::7] = True
    inactive[n - 40:] = True
    assert np.array_equal(gs["xm"][inactive], sc["xm"][inactive].astype(np.float64))
    g.close()


def test_cpp_solver_interface():
    """The C++ MpmSolverHip interface drives a scene like the reference engine does."""
    import os
    import subprocess

    exe = os.path.join(CC.ROOT, "tests", "cpp", "solver_api_test")
    assert os.path.exists(exe), "build() compiles tests/cpp/solver_api_test"
    print("cpp solver_api_test", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0


@pytest.mark.parametrize("kind", list(KINDS))
def test_reset_subset_mid_run(kind):
    """ParticleObject::reset semantics (src/engine/particle_object_impl.cpp:131-157):
    the engine restores x of the object's particles and recomputes their
    initial data (F = I, V0 from the mass already on the grid, Q10); v and C
    keep their values (Q19).  GPU and oracle run 12 substeps, reset every
    third particle, run 8 more."""
    import crmpm_hip as H

    sc = S.small_scene(KINDS[kind], 40 + KINDS[kind], shapes="plane")
    idx = np.arange(0, len(sc["xm"]), 3, dtype=np.uint32)
    x0 = np.asarray(sc["xm"], np.float32)
    g = H.HipSolver(copy.deepcopy(sc))
    g.compute_initial_data()
    g.step(12)
    o = CC.CpuSolver(copy.deepcopy(sc), "oracle")
    o.compute_initial_data()
    o.step(12)
    import torch

    g.xm[torch.as_tensor(idx.astype(np.int64), device=g.xm.device)] = torch.as_tensor(x0[idx], device=g.xm.device)
    g.compute_initial_data(idx)
    xo = o.state()["xm"].astype(np.float32)
    xo[idx] = x0[idx]
    o.fn["set_particles"](o.h, len(xo), CC.P(np.ascontiguousarray(xo), CC._f), None, None, None, None)
    o.compute_initial_data(idx)
    g.step(8)
    o.step(8)
    gs, os_ = g.state(), o.state()
    g.close()
    o.close()
    errs = report(f"{kind} reset 12+8 steps", gs, os_)
    # the reset particles restart from F = I
    assert np.abs(gs["F"][idx] - os_["F"][idx]).max() < 1e-3
    for k in ("xm", "v", "F", "V0"):
        assert errs[k] <= TOL_100, (k, errs[k])


def _coupled_scene(kind, contact_model):
    sc = S.small_scene(KINDS[kind], 77 + KINDS[kind], shapes="plane+capsule")
    sc["contact_model"] = contact_model
    sc["shape_list"][1]["comInvMass"] = (0.0, 0.01, 0.0, 2.0)
    sc["shape_list"][1]["inertia0"] = (50.0, 40.0, 50.0, 0.0)
    sc["shape_list"][1]["inertia1"] = (1.0, 0.5, 0.2, 0.0)
    return sc


@pytest.mark.parametrize("contact_model", [0, 1])
@pytest.mark.parametrize("kind", list(KINDS))
def test_dynamic_shape_coupling_force(kind, contact_model):
    """Force and torque a dynamic shape receives from the grid
    (addMomentumToRigidBody, src/solver/mpm_rigid_contact.h:219-348,378-389),
    accumulated over 10 substeps, against the oracle (which sums in FP64)."""
    import crmpm_hip as H

    sc = _coupled_scene(kind, contact_model)
    g = H.HipSolver(copy.deepcopy(sc))
    g.compute_initial_data()
    g.step(10)
    gf, gt = g.shape_force()
    gs = g.state()
    g.close()
    o = CC.CpuSolver(copy.deepcopy(sc), "oracle")
    o.compute_initial_data()
    o.step(10)
    of, ot = o.shape_force()
    os_ = o.state()
    o.close()
    assert np.abs(of[1, :3]).max() > 0, "the capsule receives no force"
    ef, et = rel(gf[1, :3], of[1, :3]), rel(gt[1, :3], ot[1, :3])
    print(f"{kind} contact model {contact_model}: force rel {ef:.2e} torque rel {et:.2e}", flush=True)
    assert ef <= 1e-4 and et <= 1e-4, (ef, et)
    assert rel(gs["v"], os_["v"]) <= TOL_100


@pytest.mark.parametrize("kind", ["mls", "pb"])
def test_empty_and_fully_masked(kind):
    """No active particle (numActive 0) and every particle masked: a step
    changes no particle, as in the reference, whose loops skip them."""
    import crmpm_hip as H

    sc = S.small_scene(KINDS[kind], 55, shapes="plane+sphere")
    n = len(sc["xm"])
    g = H.HipSolver(copy.deepcopy(sc))
    g.compute_initial_data()
    before = g.state()
    H.check(g.lib.crHipSolverSetNumActiveParticles(g.h, 0), "num active")
    g.step(3)
    mid = g.state()
    for k in ("xm", "v", "F", "C"):
        assert np.array_equal(mid[k], before[k]), k
    H.check(g.lib.crHipSolverSetNumActiveParticles(g.h, n), "num active")
    import torch

    g.mask[:] = 0
    torch.cuda.synchronize()  # the solver runs on its own stream
    g.step(3)
    after = g.state()
    for k in ("xm", "v", "F", "C"):
        assert np.array_equal(after[k], before[k]), k
    g.close()


def test_batch_with_an_empty_scene():
    """A scene with no active particle inside a batch leaves the other scenes'
    trajectories as they are when run alone."""
    import crmpm_hip as H

    scenes = [S.small_scene(S.MLS, 60 + i, n_side=6 + 2 * i, shapes="plane") for i in range(3)]
    b = H.HipBatch(copy.deepcopy(scenes))
    b.compute_initial_data()
    H.check(b.lib.crHipBatchSetSceneNumActiveParticles(b.h, 1, 0), "num active")
    b.step(5)
    for i in (0, 2):
        g = H.HipSolver(copy.deepcopy(scenes[i]))
        g.compute_initial_data()
        g.step(5)
        gs, bs = g.state(), b.state(i)
        g.close()
        for k in ("xm", "v", "F", "C"):
            assert rel(bs[k], gs[k]) <= 1e-6, (i, k)
    x1 = b.state(1)["xm"]
    assert np.array_equal(x1, np.asarray(scenes[1]["xm"], np.float64))
    b.close()


@pytest.mark.parametrize("kind", list(KINDS))
def test_parity_dense_cells(kind):
    """Cells of about 216 particles (a 12^3 lattice at a sixth of the cell
    spacing over 2^3 cells): the per-tile particle order takes ranks beyond its
    first 32 levels and beyond its 128-level table, where particles of one cell
    are ordered by the per-cell fallback.  A particle dropped or placed twice
    there shows in the mass and momentum the grid receives."""
    dx = 1.0 / 32
    sc = S.small_scene(KINDS[kind], 90 + KINDS[kind], n_side=12, shapes="plane", dx=dx)
    spacing = dx / 6
    rng = S.SplitMix64(91 + KINDS[kind])
    origin = (16 * dx, 8 * dx, 16 * dx)
    pts = S.lattice_block(rng, (12, 12, 12), origin, (spacing,) * 3, 0.25 * spacing)
    sc["xm"][:, :3] = pts
    sc["xm"][:, 3] = 1000.0 * spacing ** 3
    cells = np.floor(pts / dx).astype(np.int64)
    _, per_cell = np.unique(cells[:, 0] + 64 * (cells[:, 1] + 64 * cells[:, 2]), return_counts=True)
    assert per_cell.max() > 128, per_cell.max()
    gs = hip_run(sc, 10)
    os_ = cpu_run(sc, 10)
    od = cpu_run(sc, 10, fp64=True)
    errs = report(f"{kind} dense cells (max {per_cell.max()} per cell) 10 steps", gs, os_, od)
    for k in ("xm", "v", "F"):
        assert errs[k] <= TOL_10, (k, errs[k])
