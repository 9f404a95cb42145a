"""crHipParticleLatticeNoAlloc (the particle lattice with its per-block
counts in a fixed device array instead of an allocation) against
crHipParticleLattice, through the C-ABI: bit-identical counts and particles
on two calls in a row, the second smaller than the first (the array keeps
the first call's contents), and the capacity error past 2^28 lattice
points."""
import ctypes
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "cressim-mpm_amd", "python"))

import crmpm_scenes as S  # noqa: E402

pytestmark = pytest.mark.gpu


def test_no_alloc_lattice_matches_allocating_entry_point():
    import crmpm_hip as H

    # 2.5M lattice points (1,221 blocks of counts), then 18k points: the second
    # call reads none of the first call's leftover counts
    big = S.lattice_object(S.SPHERE, (0.2, 0, 0, 0), 0.003, 1e-4, position=(0.5, 0.5, 0.5))
    small = S.lattice_object(S.BOX, (0.05, 0.03, 0.04, 0), 0.004, 2e-4, position=(0.3, 0.2, 0.1))
    for obj in (big, small, big):
        n_ref, ref = H.particle_lattice(obj)
        n, out = H.particle_lattice(obj, no_alloc=True)
        print(f"lattice: {n} particles (allocating entry point {n_ref})", flush=True)
        assert n == n_ref and n > 0
        assert np.array_equal(out.cpu().numpy(), ref.cpu().numpy())
    # a partial capacity writes the first particles only
    n, part = H.particle_lattice(small, capacity=100, no_alloc=True)
    _, ref = H.particle_lattice(small, capacity=100)
    assert np.array_equal(part.cpu().numpy(), ref.cpu().numpy())


def test_no_alloc_lattice_rejects_lattices_beyond_its_array():
    """A lattice of 2^28 points or more is refused with CRMPM_HIP_ERR_CAPACITY
    before any launch (the host computes the lattice size)."""
    import torch

    import crmpm_hip as H

    lib = H.load_library()
    # a 0.35 m box at 1 mm spacing: (0.704 / 0.001)^3 = 3.5e8 lattice points
    obj = S.lattice_object(S.BOX, (0.35, 0.35, 0.35, 0), 0.001, 1e-6, position=(0.5, 0.5, 0.5))
    sdf = H.GeometrySdfData(None, None, None)
    n = ctypes.c_uint(7)
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.crHipParticleLatticeNoAlloc(int(obj["type"]), H._f4(*obj["params0"]), ctypes.byref(sdf),
                                         H._f3(*obj["bounds_min"]), H._f3(*obj["bounds_max"]), float(obj["spacing"]),
                                         H._f3(*obj["position"]), H._f4(*obj["rotation"]), H._f3(*obj["inv_scale"]),
                                         float(obj["mass"]), None, 0, ctypes.byref(n), H._vp(stream))
    assert rc == 5, rc
