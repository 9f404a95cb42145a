"""Mesh -> SDF tables on the device (crHipMeshSdf, crHipMeshSlicerSdf and
their WithRule variants) against the oracle's restatement of SdfBuilder
(src/geometry_utils/sdf_builder.cpp:84-217, 242-357) and the reference-built
fixtures (tests/golden/sdf_*.npz, slicer_*.npz), through the C-ABI, under
both rules: the reference's (the default: its AABB tree's winding
approximation for the sign, its warm-started walk for the slicer's rows) and
the exact one; a large subdivided icosahedron against the closed-form
sphere."""
import os
import sys
import time

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "cressim-mpm_amd", "python"))

import cpu_checkers as CC  # noqa: E402
import crmpm_scenes as S  # noqa: E402
from test_mesh_sdf_oracle import (FIXTURES, SLICER_FIXTURES, bit_equal, compare_slicer, compare_tables,  # noqa: E402
                                  fixture, node_points, sign_stable)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", FIXTURES)
def test_mesh_sdf_reference_rule_matches_fixture(name):
    """The default rule against the reference's own table: distances equal
    up to the rounding of ties between triangles (the closest triangle is
    found by a brute-force minimum here, by the tree walk there), signs from
    the same tree walk in the same order, so equal everywhere but where the
    winding sum sits at 0.5 to the rounding of the device's atan2f."""
    import crmpm_hip as H

    g = fixture(name)
    cell, fatten = float(g["cell"]), float(g["fatten"])
    print(f"{name}: {len(g['tris'])} triangles, dims {tuple(g['dims'])}", flush=True)
    dims, lower, out = H.mesh_sdf(g["points"], g["tris"], cell, fatten)
    grid = out.cpu().numpy()
    assert np.array_equal(dims, g["dims"]) and np.array_equal(lower, g["lower"])
    ref = g["grid"]
    scale = max(1.0, float(np.abs(ref[:, 3]).max()))
    assert np.abs(np.abs(grid[:, 3]) - np.abs(ref[:, 3])).max() <= 1e-6 * scale
    mism = np.sign(grid[:, 3]) != np.sign(ref[:, 3])
    assert mism.sum() <= (2 if name in ("sdf_hemisphere_obj", "sdf_curved_plane_obj") else 0), mism.sum()
    stable = sign_stable(grid[:, 3], dims) & sign_stable(ref[:, 3], dims) & ~mism
    assert np.abs(grid[stable, :3] - ref[stable, :3]).max() <= 2e-4
    # the oracle's reference rule is the fixture bit for bit
    od, ol, ogrid = CC.mesh_sdf(g["points"], g["tris"], cell, fatten)
    assert np.array_equal(dims, od) and np.array_equal(lower, ol) and bit_equal(ogrid, ref)


@pytest.mark.parametrize("name", FIXTURES)
def test_mesh_sdf_exact_rule_matches_oracle_and_reference(name):
    import crmpm_hip as H

    g = fixture(name)
    cell, fatten = float(g["cell"]), float(g["fatten"])
    print(f"{name} (exact rule): {len(g['tris'])} triangles, dims {tuple(g['dims'])}", flush=True)
    dims, lower, out = H.mesh_sdf(g["points"], g["tris"], cell, fatten, rule="exact")
    grid = out.cpu().numpy()
    compare_tables(name, g, dims, lower, grid)
    # against the oracle's exact rule: the same triangle order and arithmetic,
    # so the distances are bit-identical; the winding number is summed in
    # another order, which may move a sign only where it is 0.5 to rounding
    od, ol, ogrid = CC.mesh_sdf(g["points"], g["tris"], cell, fatten, rule="exact")
    assert np.array_equal(dims, od) and np.array_equal(lower, ol)
    np.testing.assert_array_equal(np.abs(grid[:, 3]), np.abs(ogrid[:, 3]))
    mism = np.sign(grid[:, 3]) != np.sign(ogrid[:, 3])
    assert mism.sum() <= (2 if name in ("sdf_hemisphere_obj", "sdf_curved_plane_obj") else 0), mism.sum()
    stable = sign_stable(grid[:, 3], dims) & sign_stable(ogrid[:, 3], dims) & ~mism
    np.testing.assert_array_equal(grid[stable, :3], ogrid[stable, :3])


def test_mesh_sdf_large_icosphere_closed_form():
    """20,480 triangles on a 96^3 table: distance within the tessellation's
    gap of the sphere's, signs equal away from the surface, gradients radial."""
    import crmpm_hip as H
    import torch

    r = 0.1
    pts, tris = S.icosphere(r, 5)
    cell, fatten = 0.0024, 0.0145
    H.mesh_sdf(pts, tris[:64], cell, fatten)  # load the code object
    torch.cuda.synchronize()
    t = time.perf_counter()
    dims, lower, out = H.mesh_sdf(pts, tris, cell, fatten)
    ms = (time.perf_counter() - t) * 1e3
    print(f"icosphere {len(tris)} triangles, dims {tuple(dims)}: {ms:.1f} ms", flush=True)
    assert tuple(dims) == (96, 96, 96)
    grid = out.cpu().numpy().astype(np.float64)
    q = node_points(dims, lower, cell).astype(np.float64)
    rq = np.linalg.norm(q, axis=1)
    exact = rq - r
    # the polyhedron lies between the sphere and the sphere shrunk by its
    # largest face sag, so its distance is the sphere's within that sag
    a, b, c = (pts[tris[:, k]].astype(np.float64) for k in range(3))
    n = np.cross(b - a, c - a)
    sag = r - np.abs((n * a).sum(1) / np.linalg.norm(n, axis=1)).min()
    gap = sag + 2e-6
    assert np.abs(grid[:, 3] - exact).max() < gap
    away = np.abs(exact) > gap
    assert (np.sign(grid[away, 3]) == np.sign(exact[away])).all()
    far = (np.abs(exact) > 3 * cell) & (rq > 3 * cell)
    radial = q[far] / rq[far, None]
    cos = (grid[far, :3] * radial).sum(1)
    assert cos.min() > 0.999, cos.min()


def test_mesh_sdf_table_drives_the_lattice():
    """The device-built table as a triangle-mesh geometry: the particle
    lattice over it equals the oracle's lattice over the oracle's table."""
    import crmpm_hip as H

    g = fixture("sdf_icosphere")
    cell, fatten = float(g["cell"]), float(g["fatten"])
    dims, lower, out = H.mesh_sdf(g["points"], g["tris"], cell, fatten)
    od, ol, ogrid = CC.mesh_sdf(g["points"], g["tris"], cell, fatten)
    grid = out.cpu().numpy()

    def obj(table):
        sdf = (tuple(int(d) for d in dims), tuple(float(x) for x in lower), cell, table)
        return S.lattice_object(S.TRIMESH, (0, 0, 0, 0), 0.005, 1e-4, position=(0.5, 0.5, 0.5), sdf=sdf)

    n_ref, ref = CC.particle_lattice(obj(ogrid))
    n, pts = H.particle_lattice(obj(grid))
    assert n == n_ref and n > 0
    np.testing.assert_array_equal(pts.cpu().numpy(), ref)


@pytest.mark.parametrize("name", SLICER_FIXTURES)
def test_mesh_slicer_sdf_reference_rule_is_the_reference(name):
    """Mesh-slicer table under the default rule: the reference's walk, node
    after node, each warm-started from its predecessor's closest triangle,
    with the same boundary test and the same NaN rows -- bit for bit the
    reference's own table and the oracle's."""
    import crmpm_hip as H

    g = fixture(name)
    cell, fatten = float(g["cell"]), float(g["fatten"])
    print(f"{name}: {len(g['tris'])} triangles, dims {tuple(g['dims'])}", flush=True)
    dims, lower, out = H.mesh_slicer_sdf(g["points"], g["tris"], cell, fatten, g["spine"])
    rows = out.cpu().numpy()
    assert np.array_equal(dims, g["dims"]) and np.array_equal(lower, g["lower"])
    assert bit_equal(rows, g["rows"]), (name, int((np.isnan(rows[:, 3]) != np.isnan(g["rows"][:, 3])).sum()))
    od, ol, orows = CC.mesh_slicer_sdf(g["points"], g["tris"], cell, fatten, g["spine"])
    assert bit_equal(rows, orows)


@pytest.mark.parametrize("name", SLICER_FIXTURES)
def test_mesh_slicer_sdf_exact_rule_matches_oracle_and_reference(name):
    """Mesh-slicer table under the exact rule: bit-identical to the oracle's
    exact rule (same closest triangle, same boundary test, same NaN rows), and
    against the reference fixture up to its documented extra rows."""
    import crmpm_hip as H

    g = fixture(name)
    cell, fatten = float(g["cell"]), float(g["fatten"])
    print(f"{name} (exact rule): {len(g['tris'])} triangles, dims {tuple(g['dims'])}", flush=True)
    dims, lower, out = H.mesh_slicer_sdf(g["points"], g["tris"], cell, fatten, g["spine"], rule="exact")
    rows = out.cpu().numpy()
    compare_slicer(name, g, dims, lower, rows)
    od, ol, orows = CC.mesh_slicer_sdf(g["points"], g["tris"], cell, fatten, g["spine"], rule="exact")
    assert np.array_equal(dims, od) and np.array_equal(lower, ol)
    np.testing.assert_array_equal(rows, orows)
