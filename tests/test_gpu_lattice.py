"""Particle lattice on the device (crHipParticleLattice) against the oracle's
restatement of ParticleObjectImpl::initialize
(src/engine/particle_object_impl.cpp:36-88): bit-exact counts, order and
positions, through the C-ABI."""
import ctypes
import math
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "cressim-mpm_amd", "python"))

import cpu_checkers as CC  # noqa: E402
import crmpm_scenes as S  # noqa: E402

pytestmark = pytest.mark.gpu

_a = math.radians(35.0)
_ax = np.array([0.0, 0.6, 0.8])
QUAT = tuple(_ax * math.sin(_a / 2)) + (math.cos(_a / 2),)

OBJECTS = {
    "box": lambda: S.lattice_object(S.BOX, (0.1, 0.05, 0.08, 0), 0.01, 2e-3, position=(0.3, 0.2, 0.1)),
    "box_posed": lambda: S.lattice_object(S.BOX, (0.07, 0.03, 0.05, 0), 0.004, 1e-4, position=(0.5, 0.3, 0.5),
                                          rotation=QUAT, inv_scale=(0.5, 1.0, 2.0)),
    "sphere": lambda: S.lattice_object(S.SPHERE, (0.1, 0, 0, 0), 0.005, 1e-4, position=(0.5, 0.5, 0.5)),
    "capsule": lambda: S.lattice_object(S.CAPSULE, (0.03, 0.08, 0, 0), 0.004, 1e-4, position=(0.2, 0.3, 0.4),
                                        rotation=QUAT, inv_scale=(1.0, 0.8, 1.25)),
    "mesh_table": lambda: S.lattice_object(S.TRIMESH, (0, 0, 0, 0), 0.004, 1e-4, position=(0.5, 0.5, 0.5),
                                           sdf=S.sphere_sdf_grid(0.05, 0.01, 17)),
}


@pytest.mark.parametrize("name", list(OBJECTS))
def test_lattice_bit_exact(name):
    import crmpm_hip as H

    o = OBJECTS[name]()
    n_ref, ref = CC.particle_lattice(o)
    n, pts = H.particle_lattice(o)
    print(f"{name}: {n} particles", flush=True)
    assert n == n_ref and n > 0
    np.testing.assert_array_equal(pts.cpu().numpy(), ref)


def test_lattice_capacity_and_count_only():
    import crmpm_hip as H

    o = OBJECTS["sphere"]()
    n_ref, ref = CC.particle_lattice(o)
    n0, _ = H.particle_lattice(o, capacity=0)
    assert n0 == n_ref
    n, pts = H.particle_lattice(o, capacity=n_ref // 3)
    assert n == n_ref and len(pts) == n_ref // 3
    np.testing.assert_array_equal(pts.cpu().numpy(), ref[: n_ref // 3])


def test_lattice_argument_errors():
    import crmpm_hip as H

    lib = H.load_library()
    f3, f4 = ctypes.c_float * 3, ctypes.c_float * 4
    n = ctypes.c_uint(7)

    def call(gtype, p0, lo, hi, spacing, inv=(1, 1, 1)):
        return lib.crHipParticleLattice(gtype, f4(*p0), None, f3(*lo), f3(*hi), spacing, f3(0, 0, 0), f4(0, 0, 0, 1),
                                        f3(*inv), 1.0, None, 0, ctypes.byref(n), None)

    assert call(S.BOX, (0.1, 0.1, 0.1, 0), (-0.1,) * 3, (0.1,) * 3, 0.0) == 1          # spacing <= 0
    assert call(S.BOX, (0.1, 0.1, 0.1, 0), (-0.1,) * 3, (0.1,) * 3, 0.01, (1, 0, 1)) == 1  # invScale <= 0
    assert call(S.TRIMESH, (0, 0, 0, 0), (-0.1,) * 3, (0.1,) * 3, 0.01) == 1          # table geometry without tables
    big = 3.4e38
    assert call(S.PLANE, (0, 0, 0, 0), (-big, -big, -big), (big, 0, big), 0.01) == 5   # unbounded: capacity
