"""The CPU oracle, pinned three ways (CPU only):

1. bit-exact against the reference's own CPU solvers compiled from its sources
   (oracle/_ref, present where the reference tree is), on seeded scenes with
   plane/sphere/box/capsule/slicer shapes, for MLS, PB and standard MPM;
2. bit-exact against the committed golden fixtures written by that build
   (tests/golden/make_golden.py), which travel where the reference does not;
3. against closed forms and invariants that do not depend on the reference:
   SVD reconstruction and orthogonality, the lattice initial volume, mass
   conservation of the scatter, free fall and uniform translation.
A planted one-ulp change in the oracle's input is caught by (1)/(2).
"""
import copy
import glob
import math
import os
import re

import numpy as np
import pytest

import cpu_checkers as CC
import crmpm_scenes as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("xm", "v", "F", "C", "V0")
needs_ref = pytest.mark.skipif(not CC.ref_available(), reason="reference CPU build (oracle/_ref) not present")


def run(sc, which, steps, fp64=False):
    s = CC.CpuSolver(copy.deepcopy(sc), which, fp64=fp64)
    s.compute_initial_data()
    v0 = s.state()["V0"]
    s.run(steps)
    st = s.state()
    st["V0_init"] = v0
    st["force"], st["torque"] = s.shape_force()
    s.close()
    return st


SHAPES = ["plane", "plane+sphere+box", "plane+capsule", "plane+slicer", "plane+mesh", "plane+meshslicer"]
KINDS = {"mls": S.MLS, "pb": S.PB, "std": S.STD}


@needs_ref
@pytest.mark.parametrize("shapes", SHAPES)
@pytest.mark.parametrize("kind", list(KINDS))
def test_oracle_equals_reference_build(kind, shapes):
    sc = S.small_scene(KINDS[kind], 100 + KINDS[kind], shapes=shapes, n_side=6)
    a = run(sc, "oracle", 8)
    b = run(sc, "ref", 8)
    for k in FIELDS + ("force", "torque"):
        assert np.array_equal(a[k], b[k]), k


@needs_ref
def test_oracle_equals_reference_effective_mass():
    sc = S.small_scene(S.MLS, 77, shapes="plane+capsule")
    sc["contact_model"] = 1
    sc["shape_list"][1]["comInvMass"] = (0.0, 0.01, 0.0, 2.0)
    sc["shape_list"][1]["inertia0"] = (50.0, 40.0, 50.0, 0.0)
    sc["shape_list"][1]["inertia1"] = (1.0, 0.5, 0.2, 0.0)
    a = run(sc, "oracle", 10)
    b = run(sc, "ref", 10)
    for k in FIELDS + ("force", "torque"):
        assert np.array_equal(a[k], b[k]), k
    assert np.abs(a["force"]).max() > 0


@needs_ref
def test_svd_stress_constraint_equal_reference():
    rng = np.random.default_rng(5)
    A = rng.standard_normal((500, 3, 3)).astype(np.float32)
    for which_a, which_b in (("oracle", "ref"),):
        ua, va, sa = CC.svd3(which_a, A)
        ub, vb, sb = CC.svd3(which_b, A)
        assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(sa, sb)
    n = 300
    F = (np.eye(3) + 0.3 * rng.standard_normal((n, 3, 3))).astype(np.float32)
    C = (3.0 * rng.standard_normal((n, 3, 3))).astype(np.float32)
    params = np.zeros((n, 4), np.float32)
    params[:, 0] = rng.uniform(100, 5000, n)
    params[:, 1] = rng.uniform(100, 5000, n)
    types = rng.integers(0, 3, n).astype(np.int32)
    cm = lambda M: np.ascontiguousarray(M.transpose(0, 2, 1).reshape(n, 9))
    for name in ("stress",):
        oa, ob = np.zeros((n, 9)), np.zeros((n, 9))
        CC.load("oracle")[name](n, CC.P(params, CC._f), CC.P(types, CC._i), CC.P(cm(F), CC._f), CC.P(oa, CC._d))
        CC.load("ref")[name](n, CC.P(params, CC._f), CC.P(types, CC._i), CC.P(cm(F), CC._f), CC.P(ob, CC._d))
        assert np.array_equal(oa, ob)
    params[:, 0] = 250.0
    params[:, 1] = rng.uniform(0, 1, n)
    oa, ob = np.zeros((n, 9)), np.zeros((n, 9))
    args = lambda out: (n, 2e-3, CC.P(params, CC._f), CC.P(types, CC._i), CC.P(cm(F), CC._f), CC.P(cm(C), CC._f),
                        CC.P(out, CC._d))
    CC.load("oracle")["pb_constraint"](*args(oa))
    CC.load("ref")["pb_constraint"](*args(ob))
    assert np.array_equal(oa, ob)


def _cases():
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


MG = _cases()
GOLDEN_CASES = [name for name, _, _ in MG.cases()]


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_matches_golden(name):
    sc, steps = None, None
    for n, s, k in MG.cases():
        if n == name:
            sc, steps = s, k
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    assert str(g["digest"]) == MG.input_digest(sc), "scene generator changed"
    st = run(sc, "oracle", steps)
    for k in FIELDS + ("V0_init", "force", "torque"):
        assert np.array_equal(st[k].astype(np.float32), g[k], equal_nan=True), k


def test_svd_matches_golden():
    g = np.load(os.path.join(GOLDEN, "svd3.npz"))
    U, Vt, Sg = CC.svd3("oracle", g["A"])
    assert np.array_equal(U.astype(np.float32), g["U"])
    assert np.array_equal(Vt.astype(np.float32), g["Vt"])
    assert np.array_equal(Sg.astype(np.float32), g["S"])


def test_planted_ulp_is_caught():
    """A one-ulp change of a single input position changes the FP32 trajectory."""
    sc = S.small_scene(S.MLS, 7, shapes="plane")
    g = np.load(os.path.join(GOLDEN, "mls_plane.npz"))
    sc["xm"][17, 0] = np.nextafter(sc["xm"][17, 0], np.float32(2))
    st = run(sc, "oracle", 10)
    assert not np.array_equal(st["xm"].astype(np.float32), g["xm"])


# ---------------- closed forms and invariants ----------------

def test_svd_reconstructs_transpose():
    """svd3 factors A^T: u diag(s) vt = A^T, u and vt orthogonal, |s| = numpy's singular
    values.  The algorithm runs a fixed 4 Jacobi sweeps, so even in FP64 it
    reconstructs to about 1e-4; a transposed or sign-flipped factor fails by O(1)."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((200, 3, 3))
    U, Vt, Sg = CC.svd3("oracle", A.astype(np.float32), fp64=True)
    R = U @ (Sg[:, :, None] * Vt)
    assert np.abs(R - A.astype(np.float32).transpose(0, 2, 1)).max() < 5e-4
    assert np.abs(R - A.astype(np.float32)).max() > 0.1  # it is A^T, not A
    I = np.eye(3)
    assert np.abs(U @ U.transpose(0, 2, 1) - I).max() < 1e-6
    assert np.abs(Vt @ Vt.transpose(0, 2, 1) - I).max() < 1e-6
    ref = np.linalg.svd(A.astype(np.float32).astype(np.float64), compute_uv=False)
    assert np.abs(np.sort(np.abs(Sg), 1)[:, ::-1] - ref).max() < 5e-4


def test_initial_volume_closed_form():
    """On a uniform lattice of spacing dx/2 the interpolated density is exact:
    interior particles get V0 = spacing^3."""
    dx = 1.0 / 32
    h = dx / 2
    sc = S.new_scene(S.MLS, dx, (0, 0, 0), (1, 1, 1), 20 ** 3, 1e-4)
    rng = S.SplitMix64(1)
    sc["xm"][:, :3] = S.lattice_block(rng, (20, 20, 20), (0.25, 0.25, 0.25), (h, h, h), 0.0)
    sc["xm"][:, 3] = 1000.0 * h ** 3
    o = CC.CpuSolver(sc, "oracle", fp64=True)
    o.compute_initial_data()
    st = o.state()
    x = st["xm"][:, :3]
    interior = np.all((x > 0.25 + 3 * dx) & (x < 0.25 + 20 * h - 3 * dx), axis=1)
    assert interior.sum() > 100
    assert np.abs(st["V0"][interior] / h ** 3 - 1).max() < 1e-12
    # mass conservation of the initial scatter
    assert abs(o.grid()[:, 3].sum() - sc["xm"][:, 3].astype(np.float64).sum()) < 1e-12


def test_free_fall_and_translation():
    """Away from walls and shapes, one MLS substep gives v = v0 + g dt for the
    interior of a uniformly translating undeformed block, F stays I and C 0.
    Near the block surface nodes below the 1e-4 kg cutoff hand their raw
    momentum back as velocity (src/solver/mls_mpm_algorithm.h:264-266), so
    the surface particles slow down instead."""
    sc = S.small_scene(S.MLS, 9, n_side=16, shapes="", jitter=0.0, v0=(0.3, -0.2, 0.1))
    sc["v"][:, :3] = (0.3, -0.2, 0.1)
    o = CC.CpuSolver(sc, "oracle", fp64=True)
    o.compute_initial_data()
    x0 = sc["xm"][:, :3].astype(np.float64)
    o.step(1)
    st = o.state()
    lo, hi = x0.min(0), x0.max(0)
    dx = sc["dx"]
    interior = np.all((x0 > lo + 2 * dx) & (x0 < hi - 2 * dx), axis=1)
    assert interior.sum() >= 27
    f32 = lambda x: np.float64(np.float32(x))  # the replay takes FP32 inputs
    expect = np.array([f32(0.3), f32(-0.2) + f32(-9.81) * f32(1e-4), f32(0.1)])
    assert np.abs(st["v"][interior] - expect).max() < 1e-12
    assert np.abs(st["F"][interior] - np.eye(3).reshape(1, 9)).max() < 1e-12
    assert np.abs(st["C"][interior]).max() < 1e-10
    surface = ~interior
    assert np.abs(st["v"][surface] - expect).max() > 1e-4


def test_p2g_momentum_conservation():
    """Total grid mass and momentum after P2G equal the particles' (no shapes, interior)."""
    sc = S.small_scene(S.PB, 13, shapes="")
    sc["gravity"] = (0.0, 0.0, 0.0)
    sc["iterations"] = 1
    o = CC.CpuSolver(sc, "oracle", fp64=True)
    o.compute_initial_data()
    # C = 0 and no constraint effect at F = I: grid momentum = sum m v
    o.step(1)
    g = o.grid()
    m = sc["xm"][:, 3].astype(np.float64)
    assert abs(g[:, 3].sum() - m.sum()) < 1e-12 * max(1.0, m.sum())


def test_splitmix64_known_answer():
    """SplitMix64 stream for seed 0 (Steele et al. 2014 reference values)."""
    r = S.SplitMix64(0)
    u = r.next_u64(3)
    assert [int(x) for x in u] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_grid_dims_match_reference_expression():
    assert S.grid_dims(S.config2(nx=4, ny=4, nz=4, steps=1)) == (129, 65, 129)
    assert S.grid_dims(S.config1(n_side=4, steps=1)) == (65, 65, 65)


# ---------------- the C-ABI library (no compute, no GPU) ----------------

def test_capi_library_exports_header():
    import ctypes

    import crmpm_hip as H

    hdr = open(os.path.join(CC.ROOT, "include", "crmpm_hip.h")).read()
    names = set(re.findall(r"\bint\s+(crHip\w+)\s*\(", hdr))
    assert names, "no declarations parsed"
    lib = ctypes.CDLL(H.LIB_PATH)
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert names == set(H.EXPORTS), names ^ set(H.EXPORTS)


def test_capi_argument_errors_without_gpu():
    """Calls that fail validation return an error code before touching the device."""
    import ctypes

    import crmpm_hip as H

    lib = H.load_library()
    h = ctypes.c_void_p()
    assert lib.crHipSolverCreate(7, 10, 0.1, H._f3(0, 0, 0), H._f3(1, 1, 1), ctypes.byref(h)) == 1
    assert lib.crHipSolverCreate(1, 10, -0.1, H._f3(0, 0, 0), H._f3(1, 1, 1), ctypes.byref(h)) == 1
    assert lib.crHipSolverCreate(1, 10, 0.1, H._f3(0, 0, 0), H._f3(1, 1, 1), ctypes.byref(h)) == 0
    dims = (ctypes.c_int * 3)()
    assert lib.crHipSolverGetGridDims(h, dims) == 0 and tuple(dims) == (11, 11, 11)
    assert lib.crHipSolverStep(h, None) == 2  # not initialized
    assert lib.crHipSolverSetNumActiveParticles(h, 11) == 1
    assert lib.crHipSolverSetShapeContactModel(h, 3) == 1
    assert lib.crHipSolverDestroy(h) == 0
    # the engine runtime refuses bad arguments before allocating anything
    e = ctypes.c_void_p()
    assert lib.crHipEngineCreate(0, 1, 1, 1, 1, ctypes.byref(e)) == 1
    assert lib.crHipEngineCreate(10, 1, 1, 1, 1, None) == 1
    assert lib.crHipEngineDestroy(None) == 1
    assert lib.crHipEngineAdvanceAll(None, 0.01) == 1
    assert lib.crHipEngineFetchResultsAll(None) == 1


def test_config3_generator():
    """Batched workload generator: deterministic, 64 scenes, counts in range, bodies inside the box."""
    a = S.config3()
    b = S.config3()
    assert len(a) == 64
    for x, y in zip(a, b):
        assert np.array_equal(x["xm"], y["xm"])
        n = len(x["xm"])
        assert 1000 <= n <= 200000
        assert x["xm"][:, :3].min() > 3 * x["dx"] and x["xm"][:, :3].max() < 1 - 3 * x["dx"]
        G = round(1 / x["dx"])
        assert G % 4 == 0 and 32 <= G <= 128


@pytest.mark.parametrize("kind", list(KINDS))
def test_planted_friction_error_exceeds_parity_bound(kind):
    """The fixed GPU parity bounds catch a contact bug: the capsule scene of
    the 10-substep parity test with the capsule's friction 10% too high (a
    mistake planted in the oracle's input) moves the velocities far beyond
    TOL_10."""
    from parity_bounds import TOL_10, rel

    sc = S.small_scene(KINDS[kind], 7 + KINDS[kind], shapes="plane+capsule")
    a = run(sc, "oracle", 10)
    p0 = list(sc["shape_list"][1]["params0"])
    p0[2] *= 1.1
    sc["shape_list"][1]["params0"] = tuple(p0)
    b = run(sc, "oracle", 10)
    assert rel(b["v"], a["v"]) > 20 * TOL_10, rel(b["v"], a["v"])


def _nodes_in_sphere(o, scene, shape=0):
    """Grid nodes holding mass whose position lies inside the scene's sphere."""
    dims, dx = o.dims, scene["dx"]
    k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    pos = np.stack([i.ravel(), j.ravel(), k.ravel()], 1) * dx + np.asarray(scene["lo"], np.float64)
    c = np.asarray(o.scene["shape_list"][shape]["pos"][:3])
    r = o.scene["geoms"][o.scene["shape_list"][shape]["geom"]][1][0]
    return int(((np.linalg.norm(pos - c, axis=1) < r) & (o.grid()[:, 3] > 0)).sum())


def test_config2_indenter_contacts_the_slab():
    """The PB slab's sphere indenter starts embedded and presses on the slab
    through the first 25 substeps (the bench's timed window and the parity
    test's start): grid nodes holding mass lie inside the sphere."""
    sc = S.config2(nx=32, ny=8, nz=32, steps=25, iterations=20)
    o = CC.CpuSolver(copy.deepcopy(sc), "oracle")
    o.compute_initial_data()
    counts = []
    for fr in sc["frames"]:
        S.apply_frame(o.scene, fr)
        o.upload_shapes()
        o.step(fr["substeps"])
        counts.append(_nodes_in_sphere(o, sc))
    o.close()
    print("nodes inside the indenter after substeps 10, 20, 25:", counts)
    assert min(counts) > 0, counts


def test_config5_generator():
    """Cutting scene: 2,097,152 particles on 257 x 129 x 257 nodes, the block
    inset from the lower faces, the slicer's half extent growing by the swept
    length each frame."""
    sc = S.config5()
    assert len(sc["xm"]) == 256 * 32 * 256
    assert S.grid_dims(sc) == (257, 129, 257)
    assert sc["xm"][:, 1].min() > 2 * sc["dx"]
    halves = [f["geoms"][0][1][0] for f in sc["frames"]]
    assert np.all(np.diff(halves) > 0)
    assert abs(halves[-1] - 0.02 * 500 * 1e-4 / 2) < 1e-9


# ---------------------------------------------------------------------------
# Particle lattice of a particle object (ParticleObjectImpl::initialize,
# src/engine/particle_object_impl.cpp:36-88), through the oracle.
# ---------------------------------------------------------------------------
def _axis_points(lo_b, hi_b, spacing, inv):
    """Lattice coordinates of one axis in FP32, as the engine forms them:
    bounds fattened by 2 spacings, (hi - lo) / spacing / invScale points,
    point i at i * spacing * invScale + lo."""
    f = np.float32
    s, inv = f(spacing), f(inv)
    lo, hi = f(lo_b) - f(2) * s, f(hi_b) + f(2) * s
    n = int(((hi - lo) / s) / inv)
    i = np.arange(n, dtype=np.float32)
    return (i * s) * inv + lo


def test_lattice_box_is_the_product_of_axis_lattices():
    h = (0.1, 0.05, 0.08)
    o = S.lattice_object(S.BOX, h + (0.0,), 0.01, 2e-3, position=(0.3, 0.2, 0.1))
    n, pts = CC.particle_lattice(o)
    axes = [_axis_points(-h[k], h[k], 0.01, 1.0) for k in range(3)]
    keep = [a[np.abs(a) <= np.float32(h[k])] for k, a in enumerate(axes)]
    # loop order: x outer, z inner
    X, Y, Z = np.meshgrid(*keep, indexing="ij")
    exp = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    assert n == len(exp) == len(keep[0]) * len(keep[1]) * len(keep[2])
    np.testing.assert_array_equal(pts[:, :3], (exp + np.asarray([0.3, 0.2, 0.1], np.float32)).astype(np.float32))
    assert (pts[:, 3] == np.float32(2e-3)).all()


def test_lattice_sphere_count_order_and_pose():
    r, s = 0.1, 0.0125
    inv = (0.5, 1.0, 2.0)
    a = math.radians(30.0)
    ax = np.array([1.0, 2.0, 2.0]) / 3.0
    quat = tuple(ax * math.sin(a / 2)) + (math.cos(a / 2),)
    pos = (0.5, 0.4, 0.3)
    o = S.lattice_object(S.SPHERE, (r, 0, 0, 0), s, 1e-3, position=pos, rotation=quat, inv_scale=inv)
    n, pts = CC.particle_lattice(o)
    # inside test in the geometry frame: |q| - r <= 0 in FP32
    axes = [_axis_points(-r, r, s, inv[k]) for k in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    q = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    nrm = np.sqrt(q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1] + q[:, 2] * q[:, 2])
    q = q[(nrm - np.float32(r)) <= 0]
    assert n == len(q)
    # world pose: rotate(q / invScale) + position, against a float64 rotation matrix
    x, y, z, w = quat
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    world = (q.astype(np.float64) / np.asarray(inv)) @ R.T + np.asarray(pos)
    np.testing.assert_allclose(pts[:, :3], world, rtol=0, atol=2e-6)


def test_lattice_mesh_table_matches_sphere_within_a_cell():
    # the sampled-sphere SDF table gives the sphere's particles up to the
    # trilinear interpolation error near the surface
    s = 0.01
    o = S.lattice_object(S.TRIMESH, (0, 0, 0, 0), s, 1e-3, sdf=S.sphere_sdf_grid(0.05, 0.01, 17))
    n, pts = CC.particle_lattice(o)
    rad = np.linalg.norm(pts[:, :3].astype(np.float64), axis=1)
    assert n > 0 and rad.max() <= 0.05 + 0.5 * s
    exact = S.lattice_object(S.SPHERE, (0.05, 0, 0, 0), s, 1e-3, position=(0, 0, 0))
    exact["bounds_min"], exact["bounds_max"] = o["bounds_min"], o["bounds_max"]
    m, _ = CC.particle_lattice(exact)
    assert abs(n - m) <= 0.1 * m
