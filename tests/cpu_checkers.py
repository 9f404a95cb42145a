"""ctypes drivers for the CPU checkers (test infrastructure).

`OracleSolver` runs the CPU restatement (oracle/liboracle.so, FP32 or FP64);
`RefSolver` runs the reference's own CPU solvers compiled from its sources
(oracle/_ref/libcrmpm_ref.so, FP32 only, present only where it was built).
Both take a scene dict from crmpm_scenes and expose the same calls.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "cressim-mpm_amd", "python"))
import crmpm_scenes as S  # noqa: E402

ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libcrmpm_ref.so")

_f = ctypes.POINTER(ctypes.c_float)
_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)
_u = ctypes.POINTER(ctypes.c_uint)
_b = ctypes.POINTER(ctypes.c_ubyte)
_vp = ctypes.c_void_p


def _declare(lib, pre):
    sig = {
        "create": (_vp, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, _f, _f]),
        "destroy": (None, [_vp]),
        "grid_dims": (None, [_vp, _i]),
        "set_gravity": (None, [_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
        "set_dt": (None, [_vp, ctypes.c_float]),
        "set_iterations": (None, [_vp, ctypes.c_int]),
        "set_contact_model": (None, [_vp, ctypes.c_int]),
        "set_particles": (None, [_vp, ctypes.c_int, _f, _f, _f, _i, _b]),
        "set_num_active": (None, [_vp, ctypes.c_uint]),
        "set_shapes": (None, [_vp, ctypes.c_int, _i, ctypes.c_int, _i, _i] + [_f] * 9),
        "set_geometries": (None, [_vp, ctypes.c_int, _i, _f]),
        "set_sdf": (None, [_vp, ctypes.c_int, _i, _f, ctypes.c_int, _f]),
        "reset_shape_force": (None, [_vp]),
        "get_shape_force": (None, [_vp, ctypes.c_int, _d, _d]),
        "compute_initial_data": (None, [_vp, ctypes.c_uint, _u]),
        "step": (None, [_vp, ctypes.c_int]),
        "get_particles": (None, [_vp, ctypes.c_int, _d, _d, _d, _d, _d]),
        "set_deformation": (None, [_vp, ctypes.c_int, _f, _f]),
        "get_grid": (None, [_vp, _d]),
        "base_cells": (None, [_vp, ctypes.c_int, _i]),
        "svd3": (None, [ctypes.c_int, ctypes.c_int, _f, _d, _d, _d]),
        "particle_lattice": (ctypes.c_uint, [ctypes.c_int, _f, ctypes.c_int, _i, _f, ctypes.c_int, _f, _f, _f,
                                             ctypes.c_float, _f, _f, _f, ctypes.c_float, _f, ctypes.c_uint]),
        "mesh_sdf": (ctypes.c_int, [_f, ctypes.c_int, _u, ctypes.c_int, ctypes.c_float, ctypes.c_float, _i, _f, _f,
                                    ctypes.c_int]),
        "mesh_slicer_sdf": (ctypes.c_int, [_f, ctypes.c_int, _u, ctypes.c_int, ctypes.c_float, ctypes.c_float, _f,
                                           _i, _f, _f, ctypes.c_int]),
        "mesh_sdf_exact": (ctypes.c_int, [_f, ctypes.c_int, _u, ctypes.c_int, ctypes.c_float, ctypes.c_float, _i, _f,
                                          _f, ctypes.c_int]),
        "mesh_slicer_sdf_exact": (ctypes.c_int, [_f, ctypes.c_int, _u, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                                 _f, _i, _f, _f, ctypes.c_int]),
        "stress": (None, [ctypes.c_int, _f, _i, _f, _d]),
        "pb_constraint": (None, [ctypes.c_int, ctypes.c_float, _f, _i, _f, _f, _d]),
    }
    fns = {}
    for name, (res, args) in sig.items():
        if not hasattr(lib, pre + name):
            continue
        fn = getattr(lib, pre + name)
        fn.restype = res
        fn.argtypes = args
        fns[name] = fn
    return fns


def build_oracle():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"])


_LIBS = {}


def load(which):
    if which in _LIBS:
        return _LIBS[which]
    if which == "oracle":
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        lib = ctypes.CDLL(ORACLE_SO)
        _LIBS[which] = _declare(lib, "orc_")
    else:
        if not os.path.exists(REF_SO):
            return None
        lib = ctypes.CDLL(REF_SO)
        _LIBS[which] = _declare(lib, "ref_")
    return _LIBS[which]


def ref_available():
    return os.path.exists(REF_SO)


def P(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class CpuSolver:
    """One CPU solver instance (oracle or reference) configured from a scene."""

    def __init__(self, scene, which="oracle", fp64=False):
        self.fn = load(which)
        if self.fn is None:
            raise RuntimeError("reference CPU build not present")
        self.scene = scene
        n = len(scene["xm"])
        self.n = n
        lo = np.ascontiguousarray(scene["lo"], np.float32)
        hi = np.ascontiguousarray(scene["hi"], np.float32)
        self.h = self.fn["create"](scene["kind"], int(fp64), n, scene["dx"], P(lo, _f), P(hi, _f))
        if not self.h:
            raise RuntimeError("create failed")
        dims = np.zeros(3, np.int32)
        self.fn["grid_dims"](self.h, P(dims, _i))
        self.dims = tuple(int(x) for x in dims)
        self.fn["set_gravity"](self.h, *scene["gravity"])
        self.fn["set_dt"](self.h, scene["dt"])
        self.fn["set_iterations"](self.h, scene["iterations"])
        self.fn["set_contact_model"](self.h, scene.get("contact_model", 0))
        self._keep = []
        self.upload_particles()
        self.upload_shapes()
        self.fn["set_num_active"](self.h, n)

    def upload_particles(self):
        sc = self.scene
        arrs = [np.ascontiguousarray(sc[k]) for k in ("xm", "v", "params")]
        t = np.ascontiguousarray(sc["type"], np.int32)
        m = np.ascontiguousarray(sc["mask"], np.uint8)
        self.fn["set_particles"](self.h, self.n, P(arrs[0], _f), P(arrs[1], _f), P(arrs[2], _f), P(t, _i), P(m, _b))

    def upload_shapes(self):
        sc = self.scene
        t = S.shape_tables(sc)
        keep = [t]
        cap = len(t["type"])
        self.fn["set_geometries"](self.h, len(t["gtype"]), P(t["gtype"], _i), P(t["gparams"], _f))
        dims = np.ascontiguousarray(sc["sdf_dims"], np.int32)
        lb = np.ascontiguousarray(sc["sdf_lbcs"], np.float32)
        data = np.ascontiguousarray(sc["sdf_data"], np.float32)
        keep += [dims, lb, data]
        self.fn["set_sdf"](self.h, len(dims), P(dims, _i), P(lb, _f), len(data), P(data, _f))
        self.fn["set_shapes"](self.h, len(t["ids"]), P(t["ids"], _i), cap, P(t["type"], _i), P(t["geom"], _i),
                              *[P(np.ascontiguousarray(t[k]), _f) for k in S.SHAPE_F4])
        self._keep = keep
        self.cap = cap

    def compute_initial_data(self, idx=None):
        if idx is None:
            idx = np.arange(self.n, dtype=np.uint32)
        idx = np.ascontiguousarray(idx, np.uint32)
        self.fn["compute_initial_data"](self.h, len(idx), P(idx, _u))

    def step(self, n=1):
        self.fn["step"](self.h, n)

    def run(self, steps=None):
        """Run the scene's substeps, applying kinematic frames between them."""
        sc = self.scene
        steps = sc.get("steps", 1) if steps is None else steps
        if not sc["frames"]:
            self.step(steps)
            return
        done = 0
        for fr in sc["frames"]:
            if done >= steps:
                break
            S.apply_frame(sc, fr)
            self.upload_shapes()
            k = min(fr["substeps"], steps - done)
            self.step(k)
            done += k

    def state(self):
        n = self.n
        xm = np.zeros((n, 4))
        v = np.zeros((n, 3))
        F = np.zeros((n, 9))
        C = np.zeros((n, 9))
        V0 = np.zeros(n)
        self.fn["get_particles"](self.h, n, P(xm, _d), P(v, _d), P(F, _d), P(C, _d), P(V0, _d))
        return dict(xm=xm, v=v, F=F, C=C, V0=V0)

    def grid(self):
        g = np.zeros((self.dims[0] * self.dims[1] * self.dims[2], 4))
        self.fn["get_grid"](self.h, P(g, _d))
        return g

    def base_cells(self):
        b = np.zeros((self.n, 3), np.int32)
        self.fn["base_cells"](self.h, self.n, P(b, _i))
        return b

    def shape_force(self):
        f = np.zeros((self.cap, 4))
        t = np.zeros((self.cap, 4))
        self.fn["get_shape_force"](self.h, self.cap, P(f, _d), P(t, _d))
        return f, t

    def close(self):
        if self.h:
            self.fn["destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def particle_lattice(obj, cap=None):
    """ParticleObjectImpl::initialize through the oracle: obj is a dict from
    crmpm_scenes.lattice_object; returns (count, float4 particles)."""
    fn = load("oracle")
    f32 = lambda a: np.ascontiguousarray(a, np.float32)  # noqa: E731
    dims = np.ascontiguousarray(obj["sdf_dims"], np.int32)
    lbcs, data = f32(obj["sdf_lbcs"]), f32(obj["sdf_data"])
    p0, lo, hi = f32(obj["params0"]), f32(obj["bounds_min"]), f32(obj["bounds_max"])
    pos, rot, inv = f32(obj["position"]), f32(obj["rotation"]), f32(obj["inv_scale"])
    n = fn["particle_lattice"](int(obj["type"]), P(p0, _f), len(dims), P(dims, _i), P(lbcs, _f), len(data), P(data, _f),
                               P(lo, _f), P(hi, _f), float(obj["spacing"]), P(pos, _f), P(rot, _f), P(inv, _f),
                               float(obj["mass"]), None, 0)
    m = n if cap is None else min(n, cap)
    if m == 0:
        return n, np.zeros((0, 4), np.float32)
    out = np.zeros((m, 4), np.float32)
    fn["particle_lattice"](int(obj["type"]), P(p0, _f), len(dims), P(dims, _i), P(lbcs, _f), len(data), P(data, _f),
                           P(lo, _f), P(hi, _f), float(obj["spacing"]), P(pos, _f), P(rot, _f), P(inv, _f),
                           float(obj["mass"]), P(out, _f), m)
    return n, out[:m]


def mesh_sdf(points, tris, cell, fatten, which="oracle", rule="reference"):
    """Triangle mesh -> SDF table through the oracle or the reference's own
    host builder (oracle/_ref): returns dims, lower bound and the (n, 4)
    float32 rows (gradient, signed distance).  The oracle's rule "exact" signs
    by the exact winding number instead of the reference's approximation."""
    fn = load(which)
    if fn is None:
        raise RuntimeError("reference CPU build not present")
    name = "mesh_sdf" if rule == "reference" else "mesh_sdf_exact"
    pts = np.ascontiguousarray(points, np.float32)
    tr = np.ascontiguousarray(tris, np.uint32)
    dims = np.zeros(3, np.int32)
    lower = np.zeros(3, np.float32)
    args = (P(pts, _f), len(pts), P(tr, _u), len(tr), float(cell), float(fatten), P(dims, _i), P(lower, _f))
    n = fn[name](*args, None, 0)
    if n < 0:
        raise ValueError("triangle vertex index out of range")
    grid = np.zeros((n, 4), np.float32)
    fn[name](*args, P(grid, _f), n)
    return dims, lower, grid


def mesh_slicer_sdf(points, tris, cell, fatten, spine, which="oracle", rule="reference"):
    """Triangle mesh -> mesh-slicer table through the oracle or the
    reference's host builder: dims, lower bound and (n + 1, 4) float32 rows
    (spine area, then gradient|unsigned distance, NaN off the surface).  The
    oracle's rule "exact" drops a node only by its final closest point."""
    fn = load(which)
    if fn is None:
        raise RuntimeError("reference CPU build not present")
    name = "mesh_slicer_sdf" if rule == "reference" else "mesh_slicer_sdf_exact"
    pts = np.ascontiguousarray(points, np.float32)
    tr = np.ascontiguousarray(tris, np.uint32)
    sp = np.ascontiguousarray(spine, np.float32)
    dims = np.zeros(3, np.int32)
    lower = np.zeros(3, np.float32)
    args = (P(pts, _f), len(pts), P(tr, _u), len(tr), float(cell), float(fatten), P(sp, _f), P(dims, _i),
            P(lower, _f))
    n = fn[name](*args, None, 0)
    if n < 0:
        raise ValueError("triangle vertex index out of range")
    rows = np.zeros((n + 1, 4), np.float32)
    fn[name](*args, P(rows, _f), n + 1)
    return dims, lower, rows


def svd3(which, A, fp64=False):
    """A: (n, 3, 3) row-major matrices -> U, Vt (n,3,3), S (n,3)."""
    fn = load(which)
    A = np.asarray(A, np.float32)
    n = len(A)
    cm = np.ascontiguousarray(A.transpose(0, 2, 1).reshape(n, 9))
    U = np.zeros((n, 9))
    Vt = np.zeros((n, 9))
    Sg = np.zeros((n, 3))
    fn["svd3"](int(fp64), n, P(cm, _f), P(U, _d), P(Vt, _d), P(Sg, _d))
    return U.reshape(n, 3, 3).transpose(0, 2, 1), Vt.reshape(n, 3, 3).transpose(0, 2, 1), Sg
