"""Mesh -> SDF tables (SdfBuilder, src/geometry_utils/sdf_builder.cpp), CPU only:
the triangle-mesh SDF and the mesh-slicer table.

The oracle's restatement of the reference's builder (oracle/mesh_sdf_ref.hpp:
its AABB tree, the beta = 2 winding approximation for the sign, the
warm-started closest-point walks, the slicer's on-surface flag) is pinned
1. bit for bit against the reference's own host builder, compiled from its
   sources (oracle/_ref) and committed as fixtures (tests/golden/sdf_*.npz,
   slicer_*.npz), on the reference's test meshes (assets/sphere.obj at its
   test's cell size, the open assets/hemisphere.obj, the winding-number test
   cube, assets/plane.obj and assets/curved_plane.obj as slicers) and
   generated ones (subdivided icosahedron, torus, degenerate triangles, a
   sphere away from the origin whose table still spans the origin);
2. against closed forms: the signed distance of a box for the cube mesh and a
   brute-force sampled distance to a single triangle.
The oracle's exact rule (exact winding number; a slicer row dropped only by
its final closest point) is checked against the same fixtures up to the
documented differences.  A mesh with its winding flipped (a planted mistake in
the oracle's input) flips every sign and is caught by (1).  The library's
host-side entry points (table size, argument errors) are checked here too,
without a GPU.
"""
import ctypes
import os

import numpy as np
import pytest

import cpu_checkers as CC

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = ["sdf_sphere_obj", "sdf_sphere_obj_fine", "sdf_hemisphere_obj", "sdf_cube", "sdf_icosphere", "sdf_torus",
            "sdf_curved_plane_obj", "sdf_degenerate", "sdf_icosphere_offorigin"]
# The exact rule signs with the exact winding number where the reference
# thresholds its far-field approximation (beta = 2): on an open mesh they
# may disagree where the winding number is near 0.5.  Measured: 3 of 6,877
# nodes past the hemisphere's rim; 58 of 4,761 around the curved sheet,
# whose winding number crosses 0.5 away from the surface on its concave side.
OPEN_SIGN_MISMATCHES = {"sdf_hemisphere_obj": 3, "sdf_curved_plane_obj": 58}


def bit_equal(a, b):
    """Same float32 bits, NaN rows included."""
    return a.shape == b.shape and np.array_equal(np.asarray(a, np.float32).view(np.uint32),
                                                 np.asarray(b, np.float32).view(np.uint32))


def fixture(name):
    return dict(np.load(os.path.join(GOLDEN, f"{name}.npz")))


def node_points(dims, lower, cell):
    z, y, x = np.meshgrid(*(np.arange(d) for d in dims[::-1]), indexing="ij")
    ix = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float32)
    return lower.astype(np.float32) + ix * np.float32(cell)


def sign_stable(dist, dims):
    """Nodes whose own sign and their six neighbours' signs are given by
    `dist` alone: where two tables agree on these, their gradients agree."""
    s = np.sign(dist).reshape(dims[2], dims[1], dims[0])
    ok = np.ones_like(s, bool)
    for ax in range(3):
        for sh in (1, -1):
            ok &= np.roll(s, sh, axis=ax) == s
    return ok.ravel()


def compare_tables(name, g, dims, lower, grid, rel=1e-6):
    """Table against the fixture: same dims and lower bound, |distance| equal
    up to the rounding of ties between triangles, same sign (all nodes for a
    closed mesh), same gradient where the signs around a node agree."""
    assert np.array_equal(dims, g["dims"]) and np.array_equal(lower, g["lower"])
    ref = g["grid"]
    scale = max(1.0, float(np.abs(ref[:, 3]).max()))
    dd = np.abs(np.abs(grid[:, 3]) - np.abs(ref[:, 3]))
    assert dd.max() <= rel * scale, (name, dd.max())
    mism = np.sign(grid[:, 3]) != np.sign(ref[:, 3])
    assert mism.sum() <= OPEN_SIGN_MISMATCHES.get(name, 0), (name, mism.sum())
    stable = sign_stable(ref[:, 3], dims) & sign_stable(grid[:, 3], dims)
    stable &= ~mism
    gd = np.abs(grid[stable, :3] - ref[stable, :3]).max()
    assert gd <= 2e-4, (name, gd)
    return dd


needs_ref = pytest.mark.skipif(not CC.ref_available(), reason="reference CPU build (oracle/_ref) not present")


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_matches_reference_fixture(name):
    g = fixture(name)
    dims, lower, grid = CC.mesh_sdf(g["points"], g["tris"], float(g["cell"]), float(g["fatten"]))
    assert np.array_equal(dims, g["dims"]) and np.array_equal(lower, g["lower"])
    assert bit_equal(grid, g["grid"]), name


@pytest.mark.parametrize("name", FIXTURES)
def test_exact_rule_oracle_against_reference_fixture(name):
    g = fixture(name)
    dims, lower, grid = CC.mesh_sdf(g["points"], g["tris"], float(g["cell"]), float(g["fatten"]), rule="exact")
    dd = compare_tables(name, g, dims, lower, grid)
    # the same arithmetic: most distances are bit-identical
    assert np.mean(dd == 0) > 0.95, np.mean(dd == 0)


@needs_ref
def test_reference_build_reproduces_fixture():
    g = fixture("sdf_sphere_obj")
    dims, lower, grid = CC.mesh_sdf(g["points"], g["tris"], float(g["cell"]), float(g["fatten"]), which="reference")
    assert np.array_equal(grid, g["grid"]) and np.array_equal(dims, g["dims"])


def test_flipped_winding_is_caught():
    """Planted in the oracle's input: every triangle reversed.  The distances
    stay, the signs flip, and the fixture comparison fails."""
    g = fixture("sdf_cube")
    dims, lower, grid = CC.mesh_sdf(g["points"], g["tris"][:, ::-1], float(g["cell"]), float(g["fatten"]))
    with pytest.raises(AssertionError):
        compare_tables("sdf_cube", g, dims, lower, grid)


def test_cube_closed_form():
    """The reference's test cube (half size 1): the box SDF
    |max(|q| - 1, 0)| + min(max_k(|q_k| - 1), 0), and at nodes facing a face
    the gradient is that face's normal."""
    g = fixture("sdf_cube")
    cell = float(g["cell"])
    dims, lower, grid = CC.mesh_sdf(g["points"], g["tris"], cell, 1.05)
    q = node_points(dims, lower, cell).astype(np.float64)
    e = np.abs(q) - 1.0
    exact = np.linalg.norm(np.maximum(e, 0), axis=1) + np.minimum(e.max(1), 0)
    assert np.abs(grid[:, 3] - exact).max() < 2e-6
    # inside: nodes at (0.5, 0.5, 0.5)-like points have winding number 1
    assert (grid[exact < -1e-3, 3] < 0).all() and (grid[exact > 1e-3, 3] > 0).all()
    for ax in range(3):
        o = [k for k in range(3) if k != ax]
        face = (np.abs(q[:, o]).max(1) < 1 - 2 * cell) & (q[:, ax] > 1 + 2 * cell)
        face &= q[:, ax] < q[:, ax].max() - cell / 2  # one-sided rows at the table edge excluded
        assert face.sum() > 10
        n = np.zeros(3)
        n[ax] = 1
        assert np.abs(grid[face, :3] - n).max() < 1e-5, ax


def test_single_triangle_against_sampling():
    """One triangle: unsigned distance (its winding number stays below 0.5
    off its plane) against the minimum over a dense barycentric sampling,
    which lies above the true distance by at most the sampling spacing."""
    rng = np.random.default_rng(5)
    tri = rng.uniform(-0.5, 0.5, (3, 3)).astype(np.float32)
    dims, lower, grid = CC.mesh_sdf(tri, np.array([[0, 1, 2]], np.uint32), 0.1, 0.3)
    q = node_points(dims, lower, 0.1).astype(np.float64)
    m = 400
    u, v = np.meshgrid(np.linspace(0, 1, m + 1), np.linspace(0, 1, m + 1))
    keep = u + v <= 1
    bar = np.stack([1 - u[keep] - v[keep], u[keep], v[keep]], 1)
    pts = bar @ tri.astype(np.float64)
    edge = max(np.linalg.norm(tri[i] - tri[j]) for i, j in ((0, 1), (1, 2), (0, 2)))
    sampled = np.full(len(q), np.inf)
    for a in range(0, len(q), 256):
        d = np.linalg.norm(q[a:a + 256, None, :] - pts[None], axis=2)
        sampled[a:a + 256] = d.min(1)
    assert (grid[:, 3] > 0).all()
    diff = sampled - grid[:, 3]
    assert diff.min() > -1e-5 and diff.max() < 1.5 * edge / m, (diff.min(), diff.max())


# ---------------- mesh-slicer tables (computeModSdf) ----------------

SLICER_FIXTURES = ["slicer_plane_obj", "slicer_hemisphere_obj", "slicer_curved_plane_obj", "slicer_sphere_cap_offorigin"]
# Rows the reference drops although their closest point is on the surface
# (its tree walk marks a node off the surface when any candidate accepted on
# the way lies on the boundary, src/geometry_utils/aabb_tree.cpp:724-760) and
# the exact rule keeps.  Measured on the fixtures: 0, 812, 44 and 63 nodes.
SLICER_EXTRA_ROWS = {"slicer_plane_obj": 0, "slicer_hemisphere_obj": 812, "slicer_curved_plane_obj": 44,
                     "slicer_sphere_cap_offorigin": 63}


def compare_slicer(name, g, dims, lower, rows):
    """Against the reference fixture: same dims, lower bound and spine row;
    every row the reference keeps is kept, with |distance| equal up to tie
    rounding and the same gradient where the neighbourhood is kept; rows kept
    here but dropped by the reference bounded by SLICER_EXTRA_ROWS."""
    assert np.array_equal(dims, g["dims"]) and np.array_equal(lower, g["lower"])
    ref = g["rows"]
    assert np.array_equal(rows[0], ref[0], equal_nan=True)
    rn, on = np.isnan(ref[1:, 3]), np.isnan(rows[1:, 3])
    assert not (on & ~rn).any(), (name, (on & ~rn).sum())
    assert (rn & ~on).sum() <= SLICER_EXTRA_ROWS[name], (name, (rn & ~on).sum())
    both = ~rn & ~on
    assert np.abs(rows[1:, 3][both] - ref[1:, 3][both]).max() <= 1e-6
    assert np.abs(rows[1:, :3][both] - ref[1:, :3][both]).max() <= 2e-4
    assert np.isnan(rows[1:][on]).all() and not np.isnan(rows[1:][~on]).any()


@pytest.mark.parametrize("name", SLICER_FIXTURES)
def test_slicer_oracle_matches_reference_fixture(name):
    g = fixture(name)
    dims, lower, rows = CC.mesh_slicer_sdf(g["points"], g["tris"], float(g["cell"]), float(g["fatten"]), g["spine"])
    assert np.array_equal(dims, g["dims"]) and np.array_equal(lower, g["lower"])
    assert bit_equal(rows, g["rows"]), name


@pytest.mark.parametrize("name", SLICER_FIXTURES)
def test_exact_rule_slicer_oracle_against_reference_fixture(name):
    g = fixture(name)
    dims, lower, rows = CC.mesh_slicer_sdf(g["points"], g["tris"], float(g["cell"]), float(g["fatten"]), g["spine"],
                                           rule="exact")
    compare_slicer(name, g, dims, lower, rows)


def test_slicer_plane_closed_form():
    """The reference's 2 m square sheet (assets/plane.obj, y = 0) with the
    spine x < 0 of its slicer test: distance |y| over the square and to its
    rim beyond, rows kept over the square and on the spine side only."""
    g = fixture("slicer_plane_obj")
    cell = float(g["cell"])
    dims, lower, rows = CC.mesh_slicer_sdf(g["points"], g["tris"], cell, float(g["fatten"]), g["spine"])
    q = node_points(dims, lower, cell).astype(np.float64)
    over = (np.abs(q[:, 0]) <= 1) & (np.abs(q[:, 2]) <= 1)
    keep = over | (q[:, 0] < 0)
    assert np.array_equal(~np.isnan(rows[1:, 3]), keep)
    rim = np.maximum(np.abs(q[:, [0, 2]]) - 1, 0)
    exact = np.sqrt(q[:, 1] ** 2 + (rim ** 2).sum(1))
    assert np.abs(rows[1:, 3][keep] - exact[keep]).max() < 1e-6
    # over the square, away from its rim and the table's faces, the gradient is the sheet's normal
    inner = over & (np.abs(q[:, 0]) < 1 - cell) & (np.abs(q[:, 2]) < 1 - cell) & (np.abs(q[:, 1]) > cell / 2)
    inner &= (q[:, 1] > q[:, 1].min() + cell / 2) & (q[:, 1] < q[:, 1].max() - cell / 2)
    assert inner.sum() > 100
    n = np.zeros((inner.sum(), 3))
    n[:, 1] = np.sign(q[inner, 1])
    assert np.abs(rows[1:][inner, :3] - n).max() < 1e-6


# ---------------- the library's host-side entry points (no GPU) ----------------

def _desc(lib, pts, tris, cell, fatten):
    import crmpm_hip as H

    dims = (ctypes.c_int * 3)()
    lower = (ctypes.c_float * 3)()
    rc = lib.crHipMeshSdfDesc(H._vp(pts.ctypes.data), len(pts), H._vp(tris.ctypes.data), len(tris), cell, fatten,
                              dims, lower)
    return rc, np.array(dims[:]), np.array(lower[:], np.float32)


@pytest.mark.parametrize("name", FIXTURES)
def test_capi_mesh_sdf_desc_matches_reference(name):
    import crmpm_hip as H

    g = fixture(name)
    rc, dims, lower = _desc(H.load_library(), np.ascontiguousarray(g["points"]), np.ascontiguousarray(g["tris"]),
                            float(g["cell"]), float(g["fatten"]))
    assert rc == 0 and np.array_equal(dims, g["dims"]) and np.array_equal(lower, g["lower"])


def test_capi_mesh_sdf_argument_errors():
    import crmpm_hip as H

    lib = H.load_library()
    g = fixture("sdf_cube")
    pts, tris = np.ascontiguousarray(g["points"]), np.ascontiguousarray(g["tris"])
    assert _desc(lib, pts, tris, 0.0, 0.1)[0] == 1
    assert _desc(lib, pts, tris, 0.1, -1.0)[0] == 1
    bad = tris.copy()
    bad[3, 1] = 8
    assert _desc(lib, pts, bad, 0.1, 0.1)[0] == 1
    nanp = pts.copy()
    nanp[2, 0] = np.nan
    assert _desc(lib, nanp, tris, 0.1, 0.1)[0] == 1
    assert _desc(lib, pts, tris, 1e-9, 0.1)[0] == 5  # 2^31 rows or more
    dims = (ctypes.c_int * 3)()
    lower = (ctypes.c_float * 3)()
    args = (H._vp(pts.ctypes.data), len(pts), H._vp(tris.ctypes.data), len(tris), 0.2, 0.3, dims, lower)
    # a table row buffer too small: refused before any device work
    assert lib.crHipMeshSdf(*args, H._vp(16), 10, None) == 5
    # a flat mesh without fattening has one node across its plane
    flat = np.ascontiguousarray(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32))
    one = np.ascontiguousarray(np.array([[0, 1, 2]], np.uint32))
    rc, d, _ = _desc(lib, flat, one, 0.1, 0.0)
    assert rc == 0 and d[2] == 1
    args = (H._vp(flat.ctypes.data), 3, H._vp(one.ctypes.data), 1, 0.1, 0.0, dims, lower)
    assert lib.crHipMeshSdf(*args, H._vp(16), 1 << 20, None) == 1
    # the slicer table's capacity counts its spine row; no spine is an error
    sp = H._f4(float("nan"), 0, 0, 0)
    args = (H._vp(pts.ctypes.data), len(pts), H._vp(tris.ctypes.data), len(tris), 0.2, 0.3)
    assert lib.crHipMeshSlicerSdf(*args, sp, dims, lower, None, 0, None) == 0
    count = dims[0] * dims[1] * dims[2]
    assert lib.crHipMeshSlicerSdf(*args, sp, dims, lower, H._vp(16), count, None) == 5
    assert lib.crHipMeshSlicerSdf(*args, None, dims, lower, H._vp(16), count + 1, None) == 1


def test_load_obj_reads_faces_like_the_reference_loader(tmp_path):
    """crmpm_hip.load_obj: v lines, the first three references of f lines
    (1-based, texture/normal indices dropped), everything else skipped."""
    import crmpm_hip as H

    p = tmp_path / "tri.obj"
    p.write_text("# a comment\no quad\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\n"
                 "f 1/1/1 2/1/1 3/1/1\nf 1//1 3//1 4//1\nf 4 3 2\n")
    pts, tris = H.load_obj(str(p))
    assert pts.dtype == np.float32 and tris.dtype == np.uint32
    np.testing.assert_array_equal(pts, [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]])
    np.testing.assert_array_equal(tris, [[0, 1, 2], [0, 2, 3], [3, 2, 1]])
