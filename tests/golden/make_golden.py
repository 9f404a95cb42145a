"""Write the golden fixtures of the MPM substep from the reference's own CPU
solvers (oracle/_ref/libcrmpm_ref.so, compiled from the reference sources by
oracle/Makefile target `ref`).

Each fixture holds only outputs; the inputs are regenerated from the seeded
scene builders in crmpm_scenes and pinned by a SHA-256 of the input arrays.
Run from the repository root:  python tests/golden/make_golden.py
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
    for name, sc, steps in cases():
        st, v0, f, t = run_ref(sc, steps)
        out = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(out, xm=st["xm"].astype(np.float32), v=st["v"].astype(np.float32),
                            F=st["F"].astype(np.float32), C=st["C"].astype(np.float32),
                            V0=st["V0"].astype(np.float32), V0_init=v0.astype(np.float32),
                            force=f.astype(np.float32), torque=t.astype(np.float32),
                            steps=np.int32(steps), digest=np.array(input_digest(sc)))
        print("wrote", out)
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
