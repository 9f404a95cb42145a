"""Write the golden fixtures of the MPM substep from the reference's own CPU
solvers (oracle/_ref/libcrmpm_ref.so, compiled from the reference sources by
oracle/Makefile target `ref`).

Each substep fixture holds only outputs; the inputs are regenerated from the
seeded scene builders in crmpm_scenes and pinned by a SHA-256 of the input
arrays.  The mesh-SDF fixtures (sdf_*.npz) hold the mesh with the table, from
the reference's host SDF builder; two of the meshes are the reference's own
test assets (assets/*.obj), read here as data.
Run from the repository root:  python tests/golden/make_golden.py [name ...]
(with names, only those fixtures are written).
"""
import copy
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "cressim-mpm_amd", "python"))

import cpu_checkers as CC  # noqa: E402
import crmpm_scenes as S  # noqa: E402

SHAPES = ["plane", "plane+sphere+box", "plane+capsule", "plane+slicer", "plane+mesh", "plane+meshslicer"]
STEPS = 10


def cases():
    for kind, name in ((S.MLS, "mls"), (S.PB, "pb"), (S.STD, "std")):
        for shapes in SHAPES:
            yield f"{name}_{shapes.replace('+', '_')}", S.small_scene(kind, 7 + kind, shapes=shapes), STEPS
    yield "mls_cube16_100", S.config1(n_side=16, steps=100), 100
    yield "pb_slab24x6x24_30", S.config2(nx=24, ny=6, nz=24, steps=30), 30


REF_ASSETS = "/root/reference/assets"

# The cube of the reference's winding-number test (tests/test_trimesh/main.cpp:162-205).
CUBE_POINTS = np.array([[1, 1, -1], [1, -1, -1], [1, 1, 1], [1, -1, 1], [-1, 1, -1], [-1, -1, -1], [-1, 1, 1],
                        [-1, -1, 1]], np.float32)
CUBE_TRIS = np.array([[4, 2, 0], [2, 7, 3], [6, 5, 7], [1, 7, 5], [0, 3, 1], [4, 1, 5], [4, 6, 2], [2, 6, 7],
                      [6, 4, 5], [1, 3, 7], [0, 2, 3], [4, 0, 1]], np.uint32)


from crmpm_hip import load_obj  # noqa: E402


def sdf_cases():
    """(name, points, triangles, cellSize, fatten) of the mesh-SDF fixtures:
    the reference's own test meshes (assets/sphere.obj at the cell size and
    fattening of tests/test_trimesh/main.cpp:258-260 and a finer one, the open
    assets/hemisphere.obj, the winding-number test cube) and two generated
    meshes, a subdivided icosahedron and a torus."""
    sphere = load_obj(os.path.join(REF_ASSETS, "sphere.obj"))
    yield "sdf_sphere_obj", *sphere, 0.2, 0.05
    yield "sdf_sphere_obj_fine", *sphere, 0.1, 0.05
    yield "sdf_hemisphere_obj", *load_obj(os.path.join(REF_ASSETS, "hemisphere.obj")), 0.1, 0.1
    yield "sdf_cube", CUBE_POINTS, CUBE_TRIS, 0.2, 0.3
    yield "sdf_icosphere", *S.icosphere(0.1, 3), 0.01, 0.02
    yield "sdf_torus", *S.torus(0.1, 0.03, 48, 24), 0.01, 0.01
    # an open, curved sheet: its winding number crosses 0.5 away from the surface
    yield "sdf_curved_plane_obj", *load_obj(os.path.join(REF_ASSETS, "curved_plane.obj")), 0.1, 0.1
    # duplicated triangles (exact ties) and degenerate ones (a segment, a point)
    pts, tris = S.icosphere(0.1, 2)
    tris = np.vstack([tris, tris[:5], [[0, 0, 1]], [[3, 3, 3]]]).astype(np.uint32)
    yield "sdf_degenerate", pts, tris, 0.02, 0.03
    # a mesh whose box excludes the origin: the reference's triangle boxes
    # start from a zero box, so its table still spans the origin
    yield "sdf_icosphere_offorigin", *_off_origin_sphere(), 0.02, 0.01


def _off_origin_sphere():
    pts, tris = S.icosphere(0.05, 2)
    return (pts + np.array([0.25, 0.1, -0.2], np.float32)).astype(np.float32), tris


def slicer_cases():
    """(name, points, triangles, cellSize, fatten, spine area) of the
    mesh-slicer table fixtures: assets/plane.obj with the cell size, fattening
    and spine of the reference's slicer-geometry test
    (tests/test_sdf_geometry/main.cpp:59-63), the open hemisphere without a
    spine and the curved sheet with one."""
    plane = load_obj(os.path.join(REF_ASSETS, "plane.obj"))
    yield "slicer_plane_obj", *plane, 0.1, 0.2, (-1.0, 0.0, 0.0, 0.0)
    yield "slicer_hemisphere_obj", *load_obj(os.path.join(REF_ASSETS, "hemisphere.obj")), 0.1, 0.1, \
        (float("nan"), 0.0, 0.0, 0.0)
    yield "slicer_curved_plane_obj", *load_obj(os.path.join(REF_ASSETS, "curved_plane.obj")), 0.1, 0.1, \
        (0.0, 0.0, 1.0, -0.5)
    # an open cap of the off-origin sphere (its first 150 triangles) with a spine
    pts, tris = _off_origin_sphere()
    yield "slicer_sphere_cap_offorigin", pts, tris[:150], 0.02, 0.01, (0.0, 0.0, 1.0, -0.2)


def input_digest(sc):
    h = hashlib.sha256()
    for k in ("xm", "v", "params", "type", "mask"):
        h.update(np.ascontiguousarray(sc[k]).tobytes())
    h.update(repr(sc["shape_list"]).encode())
    h.update(repr(sc["geoms"]).encode())
    h.update(repr(sc["frames"]).encode())
    if len(sc["sdf_data"]) > 1:  # SDF tables beyond the empty placeholder row
        for k in ("sdf_dims", "sdf_lbcs", "sdf_data"):
            h.update(np.ascontiguousarray(sc[k]).tobytes())
    return h.hexdigest()


def run_ref(sc, steps):
    r = CC.CpuSolver(copy.deepcopy(sc), "ref")
    r.compute_initial_data()
    v0 = r.state()["V0"]
    r.run(steps)
    st = r.state()
    f, t = r.shape_force()
    r.close()
    return st, v0, f, t


def main():
    if not CC.ref_available():
        sys.exit("oracle/_ref/libcrmpm_ref.so is missing: build it with `make -C oracle ref`")
    only = set(sys.argv[1:])
    for name, sc, steps in cases():
        if only and name not in only:
            continue
        st, v0, f, t = run_ref(sc, steps)
        out = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(out, xm=st["xm"].astype(np.float32), v=st["v"].astype(np.float32),
                            F=st["F"].astype(np.float32), C=st["C"].astype(np.float32),
                            V0=st["V0"].astype(np.float32), V0_init=v0.astype(np.float32),
                            force=f.astype(np.float32), torque=t.astype(np.float32),
                            steps=np.int32(steps), digest=np.array(input_digest(sc)))
        print("wrote", out)
    # Mesh -> SDF tables from the reference's host builder (SdfBuilder,
    # src/geometry_utils/sdf_builder.cpp), mesh included.
    for name, pts, tris, cell, fatten in sdf_cases():
        if only and name not in only:
            continue
        dims, lower, grid = CC.mesh_sdf(pts, tris, cell, fatten, which="reference")
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), points=pts, tris=tris, cell=np.float32(cell),
                            fatten=np.float32(fatten), dims=dims, lower=lower, grid=grid)
        print(f"wrote {name}.npz", tuple(dims))
    for name, pts, tris, cell, fatten, spine in slicer_cases():
        if only and name not in only:
            continue
        dims, lower, rows = CC.mesh_slicer_sdf(pts, tris, cell, fatten, spine, which="reference")
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), points=pts, tris=tris, cell=np.float32(cell),
                            fatten=np.float32(fatten), spine=np.asarray(spine, np.float32), dims=dims, lower=lower,
                            rows=rows)
        print(f"wrote {name}.npz", tuple(dims))
    if only and "svd3" not in only:
        return
    # SVD: the matrix of the reference's own SVD test (tests/test_svd/main.cpp:36,
    # column-major constructor) plus random, repeated-singular-value, negative
    # determinant and near-singular matrices.
    rng = np.random.default_rng(1234)
    mats = [np.array([[1, 3, 2], [2, 1, 9], [5, 4, 3]], np.float32)]
    mats += list(rng.standard_normal((60, 3, 3)).astype(np.float32))
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    mats.append((q @ np.diag([2.0, 2.0, 0.5]) @ q.T).astype(np.float32))
    mats.append(np.diag([1.0, 1.0, 1.0]).astype(np.float32))
    mats.append(np.diag([-1.0, 2.0, 3.0]).astype(np.float32))
    mats.append(np.array([[1, 2, 3], [2, 4, 6], [1, 0, 1]], np.float32))
    mats.append(np.array([[1e-4, 0, 0], [0, 1, 0], [0, 0, 1e4]], np.float32))
    A = np.stack(mats)
    U, Vt, Sg = CC.svd3("ref", A)
    np.savez_compressed(os.path.join(HERE, "svd3.npz"), A=A, U=U.astype(np.float32), Vt=Vt.astype(np.float32),
                        S=Sg.astype(np.float32))
    print("wrote svd3.npz")


if __name__ == "__main__":
    main()
